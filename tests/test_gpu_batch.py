"""Batched simulations (fsibe_batch_*): members of uneven grid sizes and
parameters advanced together must reproduce, bit for bit, the same
simulations run one at a time through fsibe_run (the batched kernels run the
same device code with the same launch shapes), and the oracle to the usual
bars; with prescribed beams and with beams solved by the solid solver."""
import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _members():
    from fsibe_2d_amd import Params

    return [
        Params(nx=64, ny=64, ns=32, dt_fixed=2e-4),
        Params(nx=96, ny=64, ns=40, nu=2e-3, heave_freq=1.7, dt_fixed=1.5e-4),
        Params(nx=128, ny=128, ns=64, nu=1e-3, heave_freq=0.6, dt_fixed=1e-4),
        Params(nx=64, ny=64, ns=48, nu=4e-3, heave_amp=0.05, dt_fixed=2e-4),
        Params(nx=128, ny=128, ns=70, iViscous=0, dt_fixed=1e-4),
        Params(nx=96, ny=64, ns=33, iMotion=3, AngleBeam=8.0, dt_fixed=1.5e-4),
    ]


def _noise(oracle_mod, p, seed):
    o = oracle_mod.Oracle(nx=p.nx, ny=p.ny, iBeam=0, iSponge=0, xl=p.xl, yl=p.yl)
    return o.init_noise(seed=seed, amp=1e-3, kmax2=64)


def test_batch_equals_single_runs(oracle_mod):
    from fsibe_2d_amd.solver import Batch, FluidSim

    ps = _members()
    vks = [_noise(oracle_mod, p, i) for i, p in enumerate(ps)]
    b = Batch(ps, vks)
    steps = 9
    b.run(steps)
    for i, (p, vk) in enumerate(zip(ps, vks)):
        g = FluidSim(p)
        g.set_vortk(vk)
        g.run(steps)
        m = b.members[i]
        print(f"[batch member {i}] {p.nx}x{p.ny} time {m.time_state()[0]:.6f}", flush=True)
        assert m.time_state() == g.time_state()
        for k in ("vortk", "u1", "u2", "vor_mean", "mask", "maskvx", "maskvy"):
            assert np.array_equal(m.field(k), g.field(k)), (i, k)
        so, go = m.step_out(), g.step_out()
        for k in ("dt1", "Force", "Force_unst", "Force_press", "energies", "bsi"):
            assert so[k] == go[k], (i, k, so[k], go[k])
        mb, gb = m.beam(), g.beam()
        for k in ("x", "y", "pressure", "tau"):
            assert np.array_equal(mb[k], gb[k]), (i, k)
        g.close()
    b.close()


def test_batch_member_against_oracle(oracle_mod):
    """one member of a mixed batch against the oracle's plain run"""
    from fsibe_2d_amd.solver import Batch

    ps = _members()
    vks = [_noise(oracle_mod, p, i) for i, p in enumerate(ps)]
    b = Batch(ps, vks)
    b.run(12)
    p = ps[2]
    o = oracle_mod.Oracle(oracle_mod.make_params(**{k: getattr(p, k) for k in (
        "nx", "ny", "ns", "nu", "dt_fixed", "iMotion", "heave_amp", "heave_freq", "iViscous")}))
    o.set_vortk(vks[2])
    o.run(12)
    err = rel_l2(b.members[2].field("vortk"), o.field("vortk"))
    print(f"[batch vs oracle] vortk rel L2 {err:.2e}", flush=True)
    assert err < 1e-10
    f_g, f_o = np.array(b.members[2].step_out()["Force"]), np.array(o.beam_arrays()["Force"])
    assert np.all(np.abs(f_g - f_o) <= 1e-9 * np.maximum(np.abs(f_o), 1e-6 * np.abs(f_o).max()))
    b.close()


@pytest.mark.parametrize("coupled", [False, True])
def test_batch_of_benchmark_members(coupled, oracle_mod):
    """8 members of the batched workload (grids 256..768) against single runs"""
    from fsibe_2d_amd import batch_configs
    from fsibe_2d_amd.solver import Batch, FluidSim

    ps = batch_configs(8, coupled=coupled)
    vks = [_noise(oracle_mod, p, 100 + i) for i, p in enumerate(ps)]
    b = Batch(ps, vks)
    b.run(5)
    for i in (0, 3, 7):
        g = FluidSim(ps[i])
        g.set_vortk(vks[i])
        g.run(5)
        assert np.array_equal(b.members[i].field("vortk"), g.field("vortk")), i
        assert b.members[i].step_out()["Force"] == g.step_out()["Force"], i
        assert np.array_equal(b.members[i].beam()["theta"], g.beam()["theta"]), i
        g.close()
    b.close()


def test_batch_with_solved_beams(oracle_mod):
    """members whose beams the solid solver moves (iFLUSI = 1) beside
    prescribed ones: each bit-identical to its own coupled run"""
    from fsibe_2d_amd import Params
    from fsibe_2d_amd.solver import Batch, FluidSim

    ps = _members()
    ps[0] = ps[0].replace(iFLUSI=1, mue=0.6, eta=0.05)
    ps[2] = ps[2].replace(iFLUSI=1, mue=1.5, eta=0.2, TimeMethodSolid=2, sigma=0.3)
    ps[5] = ps[5].replace(iFLUSI=1, iMotion=600, AngleBeam=0.0)
    vks = [_noise(oracle_mod, p, 20 + i) for i, p in enumerate(ps)]
    b = Batch(ps, vks)
    steps = 11
    b.run(steps)
    for i in (0, 1, 2, 5):
        g = FluidSim(ps[i])
        g.set_vortk(vks[i])
        g.run(steps)
        m = b.members[i]
        print(f"[batch fsi member {i}] iFLUSI {ps[i].iFLUSI} theta max {np.abs(g.beam()['theta']).max():.3e}",
              flush=True)
        assert m.time_state() == g.time_state()
        assert np.array_equal(m.field("vortk"), g.field("vortk")), i
        for k in ("x", "y", "theta", "pressure", "tau"):
            assert np.array_equal(m.beam()[k], g.beam()[k]), (i, k)
        so, go = m.step_out(), g.step_out()
        for k in ("dt1", "Force", "Force_unst", "Force_press", "ibes_iter"):
            assert so[k] == go[k], (i, k)
        if ps[i].iFLUSI == 1:
            assert np.array_equal(m.solid.get("theta_dot"), g.solid.get("theta_dot"))
        g.close()
    b.close()


@pytest.mark.parametrize("coupled", [False, True])
def test_graph_replays_match_eager_launches(coupled, oracle_mod):
    """the batch step replayed from captured HIP graphs against the same
    batch launching every stage, step by step: after every replay (with the
    host's reads and synchronisations between replays) each member is
    bit-identical, and a later call of several steps too"""
    from fsibe_2d_amd import batch_configs
    from fsibe_2d_amd.solver import Batch

    ps = batch_configs(6, coupled=coupled)
    vks = [_noise(oracle_mod, p, 40 + i) for i, p in enumerate(ps)]
    a = Batch(ps, vks)
    e = Batch(ps, vks, graph=False)
    assert a.describe()["graph"] == 1 and e.describe()["graph"] == 0
    for call in range(7):
        n = 1 if call < 5 else 3
        a.run(n)
        e.run(n)
        for i in range(len(ps)):
            ma, me = a.members[i], e.members[i]
            assert ma.time_state() == me.time_state(), (call, i)
            assert np.array_equal(ma.field("vortk"), me.field("vortk")), (call, i)
            so, eo = ma.step_out(), me.step_out()
            for k in ("dt1", "Force", "Force_unst", "Force_press"):
                assert so[k] == eo[k], (call, i, k)
            assert np.array_equal(ma.beam()["theta"], me.beam()["theta"]), (call, i)
    d = a.describe()
    print(f"[batch graph coupled={coupled}] captures {d['captures']} graph steps {d['graph_steps']}", flush=True)
    assert d["graph_steps"] == 4 + 2 * 3
    # two graphs per size group, captured again only when a stage outgrows them
    assert 2 * len(d["groups"]) <= d["captures"] <= 8 * len(d["groups"])
    a.close()
    e.close()


def test_every_column_tile_choice_gives_identical_results(oracle_mod):
    """every launch choice the autotuner may pick (compile-time or runtime
    FFT plan, 1 .. 16 columns per tile) for every column stage of the four
    grid sizes of the batched workload: the members stay bit-identical to
    the single path's choices"""
    from fsibe_2d_amd import FsibeError, Params
    from fsibe_2d_amd.solver import Batch

    ps = [Params(nx=n, ny=n, ns=48, dt_fixed=2e-4 * 256 / n, nu=2e-3) for n in (256, 384, 512, 768)]
    vks = [_noise(oracle_mod, p, 60 + i) for i, p in enumerate(ps)]
    steps = 3
    ref = Batch(ps, vks, autotune=False)
    ref.run(steps)
    want = [m.field("vortk") for m in ref.members]
    f_want = [m.step_out()["Force"] for m in ref.members]
    groups = ref.describe()["groups"]
    ref.close()
    assert [g["nx"] for g in groups] == [256, 384, 512, 768]
    tried = 0
    for ct in (True, False):
        for tx in (1, 2, 4, 8, 16):
            b = Batch(ps, vks)
            pinned = 0
            for gi in range(len(groups)):
                for stage in (1, 4, 5, 7):
                    try:
                        b.set_tiles(gi, stage, ct, tx)
                        pinned += 1
                    except FsibeError:
                        pass
            if pinned:
                b.run(steps)
                for i, m in enumerate(b.members):
                    assert np.array_equal(m.field("vortk"), want[i]), (ct, tx, i)
                    assert m.step_out()["Force"] == f_want[i], (ct, tx, i)
                tried += pinned
            print(f"[tiles ct={ct} tx={tx}] {pinned} stages pinned", flush=True)
            b.close()
    assert tried >= 40


def test_autotuner_report_and_refusals(oracle_mod):
    from fsibe_2d_amd import FsibeError, batch_configs
    from fsibe_2d_amd.solver import Batch

    ps = batch_configs(8)
    b = Batch(ps, [None] * len(ps))
    d = b.describe()
    assert d["tuned"] == 1
    for g in d["groups"]:
        for k in ("K1", "K4", "K5", "K7"):
            st = g["column_stages"][k]
            assert st["candidates"] >= 3 and st["us"] > 0 and st["us"] <= st["static_us"], (g["nx"], k, st)
    with pytest.raises(FsibeError):
        b.set_tiles(0, 3, True, 1)  # not a column stage
    with pytest.raises(FsibeError):
        b.set_tiles(len(d["groups"]), 1, True, 1)
    with pytest.raises(FsibeError):
        b.set_tiles(0, 1, True, 8)  # no compile-time plan of that width
    b.configure(autotune=False)
    assert b.describe()["tuned"] == 0
    b.close()
