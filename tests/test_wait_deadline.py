"""The host loop's polling waits have a deadline (fsibe_2d_amd/csrc/wait.hpp,
used by fsibe_run's waits for a step's dt1 and results and by the batch
loop): a wait whose event never completes returns FSIBE_EHIP naming the
stage instead of spinning forever.  The wait is compiled with g++ into a
CPU program with an injected event query and clock, so the timeout path is
exercised with no GPU."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FSIBE_OK, FSIBE_EHIP = 0, 3


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("wait") / "wait_driver")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-o", exe, os.path.join(HERE, "cpp", "wait_driver.cpp")],
                   check=True)
    return exe


def _run(exe, env=None):
    e = dict(os.environ)
    e.pop("FSIBE_WAIT_TIMEOUT_MS", None)
    e.update(env or {})
    out = subprocess.run([exe], check=True, capture_output=True, text=True, env=e).stdout
    return {line.split()[0]: line.split(None, 4)[1:] for line in out.splitlines()}


def test_wait_that_never_completes_times_out(driver):
    r = _run(driver)
    rc, t, polls, msg = r["never"]
    assert int(rc) == FSIBE_EHIP
    # gives up within one clock check (64 polls) of the deadline
    assert 250.0 < float(t) <= 250.0 + 64 * 0.5 + 1.0
    assert int(polls) > 1
    assert "the step's results" in msg and "not complete after" in msg


def test_wait_that_completes_returns_ok(driver):
    rc, t, polls, msg = _run(driver)["done"]
    assert int(rc) == FSIBE_OK and int(polls) == 1000 and msg == "-"


def test_runtime_error_is_reported_with_the_stage(driver):
    rc, t, polls, msg = _run(driver)["failed"]
    assert int(rc) == FSIBE_EHIP and int(polls) == 10
    assert msg.startswith("waiting for dt1:") and "illegal memory access" in msg


def test_deadline_from_the_environment(driver):
    assert float(_run(driver)["timeout"][0]) == 120000.0
    assert float(_run(driver, {"FSIBE_WAIT_TIMEOUT_MS": "1500"})["timeout"][0]) == 1500.0
    assert float(_run(driver, {"FSIBE_WAIT_TIMEOUT_MS": "-3"})["timeout"][0]) == 120000.0
