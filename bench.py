#!/usr/bin/env python3
"""Benchmark: FSI steps/s of the FP64 fluid-structure step at 1024^2 on one MI355X.

A step is one pass of the reference's time loop body (time_step.f90:178-341)
over the 1024^2 workload: create_mask -> EvolveFluidAB2 (cal_nlk, AB2 update,
pressure, cal_drag, cal_energy, time_avg_vort) -> GetForces (stresses, Delta
interpolation, surface integral) -> GravityImpulse + SolidSolverWrapper (the
flexible beam, mass ratio 0.5, stiffness 0.05, clamped at a leading edge
heaving 0.1 sin(2 pi t), solved by Newton iterations on the host) -> swap,
with all fields resident in HBM.  Inputs: band-limited Gaussian vorticity
noise (seed 0), synthetic.  Also reported: the same step with the beam
following the heave rigidly (prescribed kinematics, no solid solve), and the
host solid solver's own time per step.

Prints one JSON line.  roofline: the kernel with the largest total time,
its algorithmic bytes per launch (compulsory reads + writes of its operands)
divided by its mean duration measured with HIP events on the simulation's
stream.  cpu_baseline: the CPU oracle (C/OpenMP restatement of the fluid step
and numpy restatement of the solid solver) on a bounded sample of the same
workload, on this host's cores.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# kernel classes of fsibe_profile_read; 4, 8 and 12 are host phases (the
# dt/AB2 scalars run inside rows_nonlinear, and the host adds the row sums
# and per-node force terms the kernels leave in the pinned result record)
KERNELS = ["create_mask", "startup_velocity", "cols_vortk_n0", "rows_nonlinear", "host_beam_solve",
           "cols_ab2_update", "cols_vortk_n1", "rows_velocity_pressure", "host_mask_geometry",
           "cols_pressure", "rows_pressure_box", "delta_interp_forces", "host_wait_results"]


def alg_bytes(name: str, nx: int, ny: int, visc: bool, box_rows: int, box_cols: int,
              mask_px: int = 0) -> float:
    """Compulsory HBM bytes of one launch: every operand read once and every
    result written once (S = one nx*ny FP64 field = 8 nx ny bytes; a
    half-spectrum field of ny x nx/2 complex values is also S)."""
    S = 8.0 * nx * ny
    rowS = 8.0 * nx
    return {
        "create_mask": 3 * 8.0 * mask_px,           # mask, mask_vx, mask_vy on the mask box
        "cols_vortk_n0": 3 * S,                     # W0 -> A, B
        "rows_nonlinear": 7 * S,                    # A, B, u1, u2, chi_s -> Rc, Rpx
        "cols_ab2_update": 6 * S,                   # Rc, Rpx, W0, nlk(n1) -> nlk(n0), W1
        "cols_vortk_n1": (5 if visc else 4) * S,    # W1 -> omega, psi, i ky psi (, -ky^2 psi)
        "rows_velocity_pressure": 9 * S,            # 3 inputs + vor_mean r/w -> u1, u2, Rq, Rqy
        "cols_pressure": 2 * S + box_rows * rowS,   # Rq, Rqy -> q on the beam rows
        "rows_pressure_box": box_rows * rowS * (6 if visc else 3) + 3 * 8.0 * box_rows * box_cols,
    }.get(name, 0.0)


def hip_sync():
    """hipDeviceSynchronize of the HIP runtime the library links
    (libamdhip64.so.7); torch bundles a second runtime, which is not used
    here"""
    from fsibe_2d_amd._lib import lib

    lib()
    C.CDLL("libamdhip64.so.7").hipDeviceSynchronize()


# kernel names in rocprofv3 output for each class of fsibe_profile_read
ROCPROF_NAME = {"cols_vortk_n0": "k_cols_inv_fields", "rows_nonlinear": "k_rows_nl",
                "cols_ab2_update": "k_cols_ab2", "cols_vortk_n1": "k_cols_inv_fields",
                "rows_velocity_pressure": "k_rows_press", "cols_pressure": "k_cols_press",
                "rows_pressure_box": "k_rows_box", "create_mask": "k_mask", "delta_interp_forces": "k_forces"}


def read_rocprof_avg(kernel: str, n: int):
    """mean duration (us) of `kernel` in the committed rocprofv3 kernel-trace
    summary profiles/kernel_trace_fsi_<n>.txt, or None"""
    path = os.path.join(ROOT, "profiles", f"kernel_trace_fsi_{n}.txt")
    name = ROCPROF_NAME.get(kernel)
    if not name or not os.path.exists(path) or kernel in ("cols_vortk_n0", "cols_vortk_n1"):
        return None
    for line in open(path):
        f = line.split()
        if len(f) >= 5 and (f[4] == name or f[4].startswith(name + "<")):
            return float(f[2])
    return None


def read_traffic(kernel: str, n: int):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (profiles/pmc_<kernel>_<n>.txt), or None."""
    path = os.path.join(ROOT, "profiles", f"pmc_{kernel}_{n}.txt")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def _csv_rows(path):
    import csv

    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _find(d: str, suffix: str) -> str:
    for root, _, files in os.walk(d):
        for fn in sorted(files):
            if fn.endswith(suffix):
                return os.path.join(root, fn)
    raise FileNotFoundError(f"no *{suffix} under {d}")


def _short(name: str) -> str:
    """fsibe::k_rows_press<fsibe::RowCT<512, 256> >(...) -> k_rows_press<RowCT<512, 256>>"""
    name = name.split("(")[0].replace("fsibe::", "").replace(" >", ">")
    return name.replace("void ", "").strip()


def summarize_trace(d: str) -> str:
    """Per-kernel dispatch count / total / mean duration from a
    `rocprofv3 --kernel-trace --stats --output-format csv` directory; kernels
    are keyed by name and grid size (one kernel serves several launches)."""
    rows = _csv_rows(_find(d, "kernel_trace.csv"))
    acc = {}
    for r in rows:
        g = r.get("Grid_Size") or "x".join(r.get(f"Grid_Size_{c}", "1") for c in "XYZ")
        k = f"{_short(r['Kernel_Name'])} grid={g}"
        v = acc.setdefault(k, [0, 0.0])
        v[0] += 1
        v[1] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
    tot = sum(v[1] for v in acc.values())
    out = [" calls   total_us    avg_us    pct  kernel"]
    for k, v in sorted(acc.items(), key=lambda kv: -kv[1][1]):
        out.append(f"{v[0]:6d} {v[1]:10.1f} {v[1] / v[0]:9.3f} {100 * v[1] / tot:6.2f}  {k}")
    return "\n".join(out)


def _pmc_means(d: str, counter: str):
    """mean counter value (KiB) per dispatch for each kernel name"""
    acc = {}
    for r in _csv_rows(_find(d, "counter_collection.csv")):
        if r["Counter_Name"] != counter:
            continue
        k = f"{_short(r['Kernel_Name'])} grid={r['Grid_Size']}"
        v = acc.setdefault(k, [0.0, 0])
        v[0] += float(r["Counter_Value"])
        v[1] += 1
    return {k: (v[0] / v[1], v[1]) for k, v in acc.items()}


def summarize_pmc(fetch_dir: str, write_dir: str, cfetch_dir: str, cwrite_dir: str, copy_mib: int) -> str:
    """HBM bytes per dispatch from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
    runs, each scaled by the factor that makes a device-to-device copy of
    copy_mib MiB read and write exactly copy_mib MiB."""
    ref_kib = copy_mib * 1024.0

    def factor(d, counter):
        m = _pmc_means(d, counter)
        big = max(v[0] for v in m.values())  # the copy dispatches dominate
        return ref_kib / big, big

    ff, fk = factor(cfetch_dir, "FETCH_SIZE")
    wf, wk = factor(cwrite_dir, "WRITE_SIZE")
    fm, wm = _pmc_means(fetch_dir, "FETCH_SIZE"), _pmc_means(write_dir, "WRITE_SIZE")
    out = [f"calibration: copy of {copy_mib} MiB reads FETCH_SIZE {fk:.1f} KiB -> factor {ff:.4f}; "
           f"WRITE_SIZE {wk:.1f} KiB -> factor {wf:.4f}",
           " dispatches  fetch_KiB  write_KiB  traffic_bytes_per_dispatch  kernel"]
    for k in sorted(fm, key=lambda k: -fm[k][0]):
        f = fm[k][0] * ff
        w = wm.get(k, (0.0, 0))[0] * wf
        out.append(f"{fm[k][1]:11d} {f:10.1f} {w:10.1f} {1024.0 * (f + w):27.0f}  {k}")
    return "\n".join(out)


SQ_COUNTERS = ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_LDS", "SQ_LDS_BANK_CONFLICT", "SQ_WAIT_INST_LDS",
               "SQ_ACTIVE_INST_VALU", "SQ_BUSY_CYCLES")


def summarize_sq(d: str) -> str:
    """per kernel (name and grid) of a rocprofv3 --pmc run of the SQ
    counters: dispatches, waves per dispatch, and per wave the VALU and LDS
    instructions, LDS bank-conflict cycles, cycles waiting to issue an LDS
    instruction and cycles issuing VALU instructions; SQ_BUSY_CYCLES per
    dispatch"""
    acc, disp = {}, {}
    for r in _csv_rows(_find(d, "counter_collection.csv")):
        k = f"{_short(r['Kernel_Name'])} grid={r['Grid_Size']}"
        acc.setdefault(k, {}).setdefault(r["Counter_Name"], 0.0)
        acc[k][r["Counter_Name"]] += float(r["Counter_Value"])
        disp.setdefault(k, set()).add(r["Dispatch_Id"])
    out = [" disp  waves/disp  valu/wave  lds/wave  bankconf/wave  waitlds/wave  activevalu/wave  busy/disp  kernel"]
    for k in sorted(acc, key=lambda k: -acc[k].get("SQ_BUSY_CYCLES", 0.0)):
        c, n = acc[k], len(disp[k])
        w = c.get("SQ_WAVES", 0.0)
        if w <= 0:
            continue
        out.append(f"{n:5d} {w / n:11.0f} {c.get('SQ_INSTS_VALU', 0) / w:10.0f} {c.get('SQ_INSTS_LDS', 0) / w:9.0f} "
                   f"{c.get('SQ_LDS_BANK_CONFLICT', 0) / w:14.0f} {c.get('SQ_WAIT_INST_LDS', 0) / w:13.0f} "
                   f"{c.get('SQ_ACTIVE_INST_VALU', 0) / w:16.0f} {c.get('SQ_BUSY_CYCLES', 0) / n:10.0f}  {k}")
    return "\n".join(out)


def fft_anchor(n: int, threads: int, reps: int = 10):
    """ms per 2-D r2c transform of an n x n FP64 field: the oracle's FFT and
    scipy.fft.rfft2 with the same number of workers, on this host."""
    import numpy as np
    from oracle import oracle

    oracle.set_threads(threads)
    o = oracle.Oracle(nx=n, ny=n, iBeam=0, iSponge=0)
    f = np.random.default_rng(0).normal(size=(n, n))
    out = np.zeros((n, 2 * (n // 2 + 1)))
    o.L.o_rfft2(o.h, f, out)
    t0 = time.perf_counter()
    for _ in range(reps):
        o.L.o_rfft2(o.h, f, out)
    t_or = (time.perf_counter() - t0) / reps
    o.close()
    t_sp = None
    try:
        import scipy.fft

        scipy.fft.rfft2(f, workers=threads)
        t0 = time.perf_counter()
        for _ in range(reps):
            scipy.fft.rfft2(f, workers=threads)
        t_sp = (time.perf_counter() - t0) / reps
    except ImportError:
        pass
    return {"n": n, "threads": threads, "oracle_rfft2_ms": round(1e3 * t_or, 3),
            "scipy_rfft2_ms": round(1e3 * t_sp, 3) if t_sp else None,
            "oracle_over_scipy": round(t_or / t_sp, 2) if t_sp else None}


def cpu_baseline(p, seconds: float, threads: int):
    """coupled steps/s of the oracle: its fluid step (C/OpenMP) and its solid
    solver (numpy) in the loop of time_step.f90:178-230"""
    from oracle import oracle
    from oracle.solid import SolidOracle

    oracle.set_threads(threads)
    o = oracle.Oracle(oracle.make_params(**{k: getattr(p, k) for k in (
        "nx", "ny", "ns", "nu", "dt_fixed", "iMotion", "heave_amp", "heave_freq", "iFLUSI")}))
    so = SolidOracle(o, mue=p.mue, eta=p.eta, sigma=p.sigma, grav=p.grav, TimeMethodSolid=p.TimeMethodSolid,
                     iImpulse=p.iImpulse)
    o.set_vortk(o.init_noise(seed=0, amp=1e-3, kmax2=64))
    state = {"time": 0.0, "dt0": 0.0, "k": 0}

    def step():
        t = state["time"]
        o.set_beam_arrays(x=so.x, y=so.y, vx=so.vx, vy=so.vy, theta=so.theta, theta_dot=so.theta_dot)
        o.host_create_mask(t)
        dt1 = o.host_evolve(t, explicit=True) if state["k"] == 0 else o.host_evolve(t, state["dt0"])
        o.host_get_forces(t)
        b = o.beam_arrays()
        so.pressure_new, so.tau_new = b["pressure"].copy(), b["tau"].copy()
        so.step(t, dt1)
        o.host_swap()
        state.update(time=t + dt1, dt0=dt1, k=state["k"] + 1)

    step()
    step()  # startup + one AB2 step (warm-up, not timed)
    steps, t0 = 0, time.perf_counter()
    while True:
        step()
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 50:
            break
    o.close()
    return steps / el, steps, el


def host_cpus():
    """CPU model, os.cpu_count() and the CPUs this process may run on"""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "os_cpu_count": os.cpu_count(), "usable_cpus": usable}


def turbulence_bench(n: int, steps: int, warmup: int = 5):
    """The solver-only workload: n^2 decaying 2-D turbulence (no mask, no
    sponge, fluid at rest, nu = 1e-5, dt = 1e-4, domain 2 pi) from the
    random-phase start (k0 = 12, unit energy, seed 1); steps/s with the
    initial field resident in HBM, and the fraction of 8 TB/s on the
    workload's algorithmic bytes 272 n^2 per step (6 transforms of 4 passes
    of 8 n^2 bytes + 10 auxiliary field passes)."""
    from fsibe_2d_amd import turbulence_4096
    from fsibe_2d_amd.inputs import decaying_turbulence
    from fsibe_2d_amd.solver import FluidSim

    p = turbulence_4096(n)
    sim = FluidSim(p)
    sim.set_vortk(decaying_turbulence(n, seed=1))
    sim.run(warmup)
    hip_sync()
    t0 = time.perf_counter()
    sim.run(steps)
    hip_sync()
    el = time.perf_counter() - t0
    out = sim.step_out()
    sim.close()
    bytes_step = 272.0 * n * n
    rate = steps / el
    return {"workload": f"turbulence_{n}: {n}^2 solver-only decaying 2-D turbulence, domain 2 pi, nu=1e-5, "
                        f"dt=1e-4, E(k) ~ k^4 exp(-2 (k/12)^2) unit energy, seed 1",
            "value": round(rate, 3), "unit": "steps/s", "steps": steps, "warmup": warmup,
            "ms_per_step": round(1e3 * el / steps, 5),
            "alg_bytes_per_step": bytes_step,
            "step_alg_GBps": round(bytes_step * rate / 1e9, 1),
            "step_alg_frac": round(bytes_step * rate / 1e9 / HBM_PEAK_GBS, 4),
            "kinetic_energy_last": out["energies"][1]}


def solid_step_us(p, steps: int = 2000):
    """host time of one solid-solver step (GravityImpulse + SolidSolverWrapper)
    of the workload's beam under a steady load, on this host"""
    import numpy as np

    from fsibe_2d_amd.solid import Solid

    b = Solid(p)
    s = np.linspace(0.0, 1.0, p.ns)
    b.set("pressure_new", 0.3 * s)
    t = 0.0
    for _ in range(20):
        b.step(t, p.dt_fixed)
        t += p.dt_fixed
    t0 = time.perf_counter()
    for _ in range(steps):
        b.step(t, p.dt_fixed)
        t += p.dt_fixed
    el = time.perf_counter() - t0
    it = b.info()["iter"]
    b.close()
    return round(1e6 * el / steps, 2), it


def batch_bench(nsims: int, steps: int, warmup: int = 3):
    """sims*steps/s of the batched workload (batch_configs: uneven grids
    256..768, uneven parameters) advanced as one batch, and of the same
    simulations run one after another on the GPU."""
    from fsibe_2d_amd import batch_configs
    from fsibe_2d_amd.inputs import band_limited_noise
    from fsibe_2d_amd.solver import Batch, FluidSim
    from fsibe_2d_amd.spectral import Spectral

    ps = batch_configs(nsims)
    vks = []
    for i, p in enumerate(ps):
        sp = Spectral(p.nx, p.ny, p.xl, p.yl, p.nu)
        vks.append(band_limited_noise(sp, sp.dealiase_mask(), seed=2 + i, amp=1e-3, kmax2=64))
        sp.close()
    t0 = time.perf_counter()
    b = Batch(ps, vks)  # measures the column stages' launch choices
    t_create = time.perf_counter() - t0

    def timed(nsteps):
        b.run(warmup)
        hip_sync()
        t0 = time.perf_counter()
        b.run(nsteps)
        hip_sync()
        return time.perf_counter() - t0

    # captured graphs with tuned tiles (the default) and eager launches with
    # the same tiles, alternated twice (best of each), then eager launches
    # with the single path's tiles
    t_b = t_eager = float("inf")
    for _ in range(2):
        b.configure(graph=True)
        t_b = min(t_b, timed(steps))
        b.configure(graph=False)
        t_eager = min(t_eager, timed(steps))
    desc = b.describe()
    b.configure(autotune=False)
    t_static = timed(steps)
    b.close()
    t_s = 0.0
    for p, vk in zip(ps, vks):
        g = FluidSim(p)
        g.set_vortk(vk)
        g.run(warmup)
        hip_sync()
        t0 = time.perf_counter()
        g.run(steps)
        hip_sync()
        t_s += time.perf_counter() - t0
        g.close()
    grids = {}
    for p in ps:
        grids[p.nx] = grids.get(p.nx, 0) + 1
    v_b, v_s = nsims * steps / t_b, nsims * steps / t_s
    return {"metric": f"sims*steps/s, {nsims} FSI simulations with uneven grids and parameters",
            "value": round(v_b, 1), "unit": "sims*steps/s", "sequential_gpu_value": round(v_s, 1),
            "speedup_vs_sequential": round(v_b / v_s, 3), "sims": nsims, "steps": steps, "warmup": warmup,
            "grids": {str(k): v for k, v in sorted(grids.items())}, "dtype": "f64", "data": "synthetic",
            "step_mode": "two captured HIP graphs per step, column tiles from the autotuner",
            "timing": "best of two alternated runs per step mode",
            "eager_tuned_value": round(nsims * steps / t_eager, 1),
            "eager_static_tiles_value": round(nsims * steps / t_static, 1),
            "create_and_tune_s": round(t_create, 3), "graph_captures": desc["captures"],
            "tiles": {str(g["nx"]): {k: f"{'ct' if v['plan'] == 'compile-time' else 'rt'}x{v['tx']}"
                                     f" {v['us']:.1f}us (static {v['static_us']:.1f}us)"
                                     for k, v in g["column_stages"].items()} for g in desc["groups"]}}


def adjoint_bench(steps: int = 1000, chk: int = 50, smooth: bool = True, bchk: int = 10, bslots: int = 10):
    """the differentiation workload (adjoint_512: 512^2, Re 500, dt 1e-4,
    motion 601 with coefficients 0.05 N(0,1) of seed 3): wall time of J
    alone and of J + dJ/dc (forward with checkpoints every chk steps,
    recompute of each segment, reverse sweep), each gradient component
    against the central difference of J with h = 1e-6 and 1e-7, and the
    gradient along a random unit direction against the same (h = 1e-6).
    With the continuous mask also the bounded-memory form: units of bchk
    steps, bslots snapshot slots, the binomial schedule."""
    import numpy as np

    from fsibe_2d_amd.adjoint import Adjoint
    from fsibe_2d_amd.inputs import band_limited_noise, normals
    from fsibe_2d_amd.params import adjoint_512, adjoint_coefficients
    from fsibe_2d_amd.spectral import Spectral

    p = adjoint_512(smooth_cutoff=smooth)
    sp = Spectral(p.nx, p.ny, p.xl, p.yl, p.nu)
    vk = band_limited_noise(sp, sp.dealiase_mask(), seed=0, amp=1e-3, kmax2=64)
    sp.close()
    c = np.array(adjoint_coefficients(3))
    a = Adjoint(p, vk, steps, chk)
    a.gradient(c)  # warm-up: first launches and LDS attributes
    t_f = []
    for _ in range(3):
        t0 = time.perf_counter()
        J = a.objective(c)
        t_f.append(time.perf_counter() - t0)
    t_g = []
    for _ in range(3):
        t0 = time.perf_counter()
        Jg, g = a.gradient(c)
        t_g.append(time.perf_counter() - t0)
    tm = a.timings()
    d = np.array(normals(4, 8))
    d /= np.linalg.norm(d)
    h = 1e-6
    fd = (a.objective(c + h * d) - a.objective(c - h * d)) / (2 * h)
    comp = {}
    for hh in (1e-6, 1e-7):
        fdk = np.zeros(8)
        for k in range(8):
            e = np.zeros(8)
            e[k] = hh
            fdk[k] = (a.objective(c + e) - a.objective(c - e)) / (2 * hh)
        comp[f"{hh:g}"] = float(np.max(np.abs(fdk - g) / np.abs(fdk)))
    a.close()
    f, gr = min(t_f), min(t_g)
    bino = None
    if smooth and bchk > 0:
        # the same gradient with bounded memory: units of bchk steps and
        # bslots snapshot slots, reversed by the binomial (revolve) schedule
        from fsibe_2d_amd.adjoint import plan_cost

        b = Adjoint(p, vk, steps, bchk, snapshots=bslots)
        b.gradient(c)
        t_b = []
        for _ in range(2):
            t0 = time.perf_counter()
            Jb, gb = b.gradient(c)
            t_b.append(time.perf_counter() - t0)
        st = b.stats()
        b.close()
        nu, nu_d = -(-steps // bchk), -(-steps // chk)
        mib = lambda n_f: round(n_f * p.nx * p.ny * 8 / 2**20, 1)
        bino = {"chk": bchk, "snapshot_slots": st["slots"], "units": nu,
                "units_advanced": plan_cost(nu, bslots), "steps_advanced": st["steps_advanced"],
                "steps_recorded": st["steps_recorded"], "gradient_ms": round(1e3 * min(t_b), 3),
                "cost_ratio": round(min(t_b) / f, 3),
                "identical_to_default": bool(Jb == Jg and np.array_equal(gb, g)),
                "memory_mib": {"snapshots": mib(7 * st["slots"]), "tape": mib(7.5 * bchk)},
                "default_memory_mib": {"snapshots": mib(7 * (nu_d - 1)), "tape": mib(7.5 * chk)}}
    return {"workload": f"{p.nx}^2, Re=500, dt=1e-4, {steps} steps, motion 601 (8 coefficients, "
                        f"0.05 N(0,1) seed 3), J = (1/T) sum(-Force(1) dt), checkpoints every {chk} steps, "
                        + ("mask cut-offs and cal_drag window moved out (iSmoothCutoff=1)" if smooth
                           else "the reference's mask cut-offs and cal_drag window"),
            "iSmoothCutoff": int(smooth), "J": J, "J_of_gradient_run": Jg, "grad": [float(v) for v in g],
            "forward_ms": round(1e3 * f, 3), "gradient_ms": round(1e3 * gr, 3),
            "cost_ratio": round(gr / f, 3), "reverse_ms": round(tm["reverse_ms"], 3),
            "fd_max_rel_err_per_component": comp,
            "directional_fd": {"h": h, "adjoint": float(g @ d), "fd": fd,
                               "rel_err": abs(float(g @ d) - fd) / abs(fd)},
            "binomial": bino}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1024, help="grid size (1024 or 2048)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--batch-sims", type=int, default=64, help="members of the batched workload (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=1000, help="steps per simulation of the batched workload")
    ap.add_argument("--adjoint-steps", type=int, default=1000, help="steps of the adjoint leg (0: skip)")
    ap.add_argument("--turb-n", type=int, default=4096, help="grid of the solver-only turbulence leg")
    ap.add_argument("--turb-steps", type=int, default=200, help="steps of the turbulence leg (0: skip)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 20 device-to-device copies of MIB MiB (PMC traffic calibration)")
    ap.add_argument("--summarize-trace", metavar="DIR",
                    help="print the per-kernel table of a rocprofv3 --kernel-trace --stats csv directory")
    ap.add_argument("--summarize-pmc", nargs=4, metavar=("FETCH", "WRITE", "COPY_FETCH", "COPY_WRITE"),
                    help="print HBM bytes per dispatch from four rocprofv3 --pmc csv directories")
    ap.add_argument("--summarize-sq", metavar="DIR",
                    help="print per-kernel SQ counter figures of a rocprofv3 --pmc csv directory")
    ap.add_argument("--copy-mib", type=int, default=64)
    ap.add_argument("--batch-only", action="store_true",
                    help="only the batched workload (--batch-sims, --batch-steps); prints its JSON")
    ap.add_argument("--turb-only", action="store_true",
                    help="only the solver-only turbulence workload (--turb-n, --turb-steps); prints its JSON")
    args = ap.parse_args(argv)
    if args.summarize_trace:
        print(summarize_trace(args.summarize_trace))
        return 0
    if args.summarize_sq:
        print(summarize_sq(args.summarize_sq))
        return 0
    if args.summarize_pmc:
        print(summarize_pmc(*args.summarize_pmc, args.copy_mib))
        return 0
    if args.calibrate_copy:
        import torch

        n = args.calibrate_copy * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(20):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": n * 8, "copies": 20}))
        return 0
    if args.batch_only:
        print(json.dumps(batch_bench(args.batch_sims, args.batch_steps)))
        return 0
    if args.turb_only:
        print(json.dumps(turbulence_bench(args.turb_n, args.turb_steps)))
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "a large simulation would shard by row slabs and column slabs with one all-to-all between "
              "passes, and batches by member, neither of which is built here", file=sys.stderr)
        return 2

    from fsibe_2d_amd import fsi_large
    from fsibe_2d_amd._lib import check, lib
    from fsibe_2d_amd.inputs import band_limited_noise
    from fsibe_2d_amd.solver import FluidSim
    from fsibe_2d_amd.spectral import Spectral

    n = args.n
    p = fsi_large(n)
    sp = Spectral(n, n, p.xl, p.yl, p.nu)
    vk = band_limited_noise(sp, sp.dealiase_mask(), seed=0, amp=1e-3, kmax2=64)
    sp.close()
    sim = FluidSim(p)
    sim.set_vortk(vk)

    sim.run(args.warmup)
    hip_sync()
    t0 = time.perf_counter()
    out = sim.run(args.steps)
    hip_sync()
    el = time.perf_counter() - t0
    steps_per_s = args.steps / el

    # the same step with the beam following the heave rigidly (iFLUSI = 0)
    pr = fsi_large(n, coupled=False)
    simr = FluidSim(pr)
    simr.set_vortk(vk)
    simr.run(args.warmup)
    hip_sync()
    t0 = time.perf_counter()
    simr.run(args.steps)
    hip_sync()
    el_r = time.perf_counter() - t0
    simr.close()
    solid_us, solid_iter = solid_step_us(p)

    roofline = None
    prof = {}
    if not args.no_profile:
        L = lib()
        nk = len(KERNELS)

        def timed_pass(classes):
            """HIP-event times of the kernel classes in `classes` (bit mask)
            on the simulation's stream, over args.steps steps"""
            check(L.fsibe_profile_enable(sim.h, classes))
            hip_sync()
            t1 = time.perf_counter()
            sim.run(args.steps)
            hip_sync()
            el = time.perf_counter() - t1
            ms = (C.c_double * nk)()
            cnt = (C.c_longlong * nk)()
            L.fsibe_profile_read(sim.h, ms, cnt, nk)
            check(L.fsibe_profile_enable(sim.h, 0))
            return el, list(ms), list(cnt)

        el_prof, ms, cnt = timed_pass(-1)
        fb = sim.box("press")[1]
        mb = sim.box("mask")[1]
        box_rows, box_cols = fb[3] - fb[2] + 1, fb[1] - fb[0] + 1
        mask_px = max(0, mb[1] - mb[0] + 1) * max(0, mb[3] - mb[2] + 1)
        for i, name in enumerate(KERNELS):
            if cnt[i]:
                avg_us = 1e3 * ms[i] / cnt[i]
                b = alg_bytes(name, n, n, True, box_rows, box_cols, mask_px)
                prof[name] = {"avg_us": round(avg_us, 3), "launches": int(cnt[i]),
                              "alg_bytes": b, "GBps": round(b / (avg_us * 1e-6) / 1e9, 1) if b else None}
        # dominant kernel: largest total time among the HBM-streaming kernels;
        # timed again alone so that no other event pair perturbs it
        hbm = [k for k in prof if prof[k]["alg_bytes"] and k != "create_mask"]
        dom = max(hbm, key=lambda k: prof[k]["avg_us"] * prof[k]["launches"])
        di = KERNELS.index(dom)
        _, ms1, cnt1 = timed_pass(1 << di)
        avg1 = 1e3 * ms1[di] / cnt1[di]
        b = prof[dom]["alg_bytes"]
        ach = b / (avg1 * 1e-6) / 1e9
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": read_traffic(dom, n),
                    "avg_us": round(avg1, 3), "launches": int(cnt1[di]), "alg_bytes_per_launch": b,
                    "rocprof_avg_us": read_rocprof_avg(dom, n)}
        prof["_profiled_ms_per_step"] = round(1e3 * el_prof / args.steps, 4)

    cpu = None
    anchor = None
    if not args.no_cpu:
        hc = host_cpus()
        omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        # OMP_NUM_THREADS when the environment sets it (the host's CPU share
        # for this process), else every CPU the process may run on
        threads = omp if omp > 0 else hc["usable_cpus"]
        rate, k, sec = cpu_baseline(p, args.cpu_seconds, threads)
        anchor = fft_anchor(n, threads)
        ratio = anchor["oracle_over_scipy"]
        cpu = {"value": round(rate, 4), "unit": "steps/s", "cores": threads, "kind": "port",
               "threads_from": "OMP_NUM_THREADS" if omp > 0 else "sched_getaffinity", **hc,
               "sample": f"{k} coupled AB2 steps of the same {n}^2 FSI workload after startup + 1 step, "
                         f"{sec:.1f} s, C/OpenMP fluid oracle with {threads} threads, numpy solid oracle"
                         + (f"; one {n}^2 r2c FFT: oracle {anchor['oracle_rfft2_ms']} ms, "
                            f"scipy.fft.rfft2 ({threads} workers) {anchor['scipy_rfft2_ms']} ms, "
                            f"time ratio {ratio}" if ratio else "")}

    batched = batch_bench(args.batch_sims, args.batch_steps) if args.batch_sims > 0 else None
    adjoint = adjoint_ref = None
    if args.adjoint_steps > 0:
        adjoint = adjoint_bench(args.adjoint_steps, smooth=True)
        adjoint_ref = adjoint_bench(args.adjoint_steps, smooth=False)

    turb = turbulence_bench(args.turb_n, args.turb_steps) if args.turb_steps > 0 else None

    step_bytes = 848.0 * n * n  # reference's per-step algorithmic bytes (19 transforms + 30 passes)
    rec = {
        "metric": f"FSI steps/s at {n}^2 FP64",
        "value": round(steps_per_s, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * el / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"fsi_{n}: {n}^2 periodic grid, domain 4x4, Re=1000, flexible beam of 128 "
                               f"nodes (mass ratio 0.5, stiffness 0.05, BDF2) clamped at a leading edge "
                               f"heaving 0.1 sin(2 pi t), dt=5e-5, iSponge=3, iViscous=1, forces and beam "
                               f"solve every step",
                   "grid": n, "ns": 128, "dt": 5e-5, "iFLUSI": 1},
        "prescribed_motion": {"value": round(args.steps / el_r, 3), "unit": "steps/s",
                              "ms_per_step": round(1e3 * el_r / args.steps, 5),
                              "note": "same step, beam following the heave rigidly (iFLUSI=0): the host "
                                      "prepares step n+1 while step n runs"},
        "host_solid_step_us": solid_us,
        "host_solid_newton_iters": solid_iter,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_fft_anchor": anchor,
        "batched": batched,
        "turbulence": turb,
        "adjoint": adjoint,
        "adjoint_reference_cutoffs": adjoint_ref,
        "step_alg_GBps": round(step_bytes * steps_per_s / 1e9, 1),
        "step_alg_frac": round(step_bytes * steps_per_s / 1e9 / HBM_PEAK_GBS, 4),
        "last_step": {"dt1": out["dt1"], "Force": out["Force"], "Force_press": out["Force_press"]},
        "kernels": prof,
    }
    print(json.dumps(rec))
    sim.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
