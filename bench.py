"""Benchmark: FP64 gridpoint-updates/s of the BSSN RK4 method-of-lines step.

Workload (plain BSSN vacuum forward): BSSN in the Cartesian reference metric
on a 192^3 interior grid (NGHOSTS=3, [-10,10]^3), 24 evolved FP64
gridfunctions, 4th-order FD with upwinded shift advection, 1+log lapse,
covariant Gamma-driver shift (eta=2), RK4 with dt = 0.5*dx; the state is the
flat metric plus 8 seeded Gaussians (amplitude 1e-3, width 1-3) in every
field. One step = one RK4 step = per substep Ricci_eval, the fused RHS + RK
update kernel, the ghost-zone fill and det(gammabar) enforcement, replayed
from a hipGraph. A gridpoint-update is one RHS evaluation at one interior
point (4 per RK4 step), the reference's gp/s.

Also reported: the standalone RHS throughput on the same state, the
Brill-Lindquist head-on evolution on Spherical 72x12x2 (50 RK4 steps, the
latency-bound small case), per-kernel HIP-event timings and the roofline of
the dominant kernel, and the CPU baseline: the C oracle built with the
reference's optimized flags, OpenMP over the host cores, on a bounded sample.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 192] [--no-cpu]
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
METRIC = "FP64 gridpoint-updates/s: BSSN RK4 forward, adjoint, batched vs plain GPU"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def physical_cores(limit=16):
    """One logical CPU per physical core among those this process may use (sibling
    hyperthreads dropped), at most ``limit``: the reference pins its OpenMP runs with
    taskset to cpu_count/2 cores when SMT is on (cmdline_helper.py:127-145)."""
    seen, picked = set(), []
    for c in sorted(os.sched_getaffinity(0)):
        try:
            with open("/sys/devices/system/cpu/cpu%d/topology/core_id" % c) as f:
                core = f.read().strip()
            with open("/sys/devices/system/cpu/cpu%d/topology/physical_package_id" % c) as f:
                pkg = f.read().strip()
        except OSError:
            core, pkg = str(c), "0"
        if (pkg, core) not in seen:
            seen.add((pkg, core))
            picked.append(c)
    return picked[:limit]


def cpu_model():
    """'model name' from /proc/cpuinfo plus the host's socket/core/thread counts."""
    model, sockets, logical = "", set(), 0
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and not model:
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("physical id"):
                    sockets.add(line.split(":", 1)[1].strip())
                elif line.startswith("processor"):
                    logical += 1
    except OSError:
        pass
    return "%s (%d socket(s), %d logical CPUs on the host)" % (model or "unknown", max(1, len(sockets)), logical)


def cpu_baseline_worker(n, min_seconds, max_steps):
    """Runs in a child process (fresh OpenMP runtime, pinned by the parent): oracle RK4 steps of
    the benchmark itself, the seeded flat + Gaussian BSSN state on the n^3 Cartesian grid."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import torch
    from oracle_lib import Oracle, OracleGrid
    from nrpytutorial_amd.workloads import bssn_gaussian_state_torch
    og = OracleGrid(Oracle("bssn_cartesian", fast=True), (n, n, n), {})
    y0 = np.zeros((len(og.o.evol),) + og.shape)
    bssn_gaussian_state_torch(torch.from_numpy(y0), [torch.from_numpy(x) for x in og.xx], og.o.evol, 1)
    og.apply_bcs(y0)
    og.enforce_detg(y0)
    dt = 0.5 * og.p.dxx0
    y = y0.copy()
    og.mol_step(y, dt, 1)  # warm-up
    steps, el = 0, 0.0
    while steps < max_steps and el < min_seconds:
        y[...] = y0
        t0 = time.perf_counter()
        og.mol_step(y, dt, 2)
        el += time.perf_counter() - t0
        steps += 2
        if not np.isfinite(y).all():
            raise RuntimeError("oracle state became non-finite")
    return {"gp_per_s": 4.0 * n ** 3 * steps / el, "steps": steps, "seconds": el}


def cpu_baseline(n=192, min_seconds=10.0, max_steps=8):
    """The C oracle built with the reference's optimized flags, timed on the host on the
    benchmark's own workload (n^3, seeded state, RK4 steps of dt = 0.5 dx), one OpenMP
    thread per physical core, pinned."""
    cpus = physical_cores(int(os.environ.get("NRPY_CPU_BASELINE_THREADS", 16)))
    cores = len(cpus)
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close", OMP_PLACES="cores")
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-worker", str(n),
                        str(min_seconds), str(max_steps)], env=env, capture_output=True, text=True, timeout=900,
                       preexec_fn=lambda: os.sched_setaffinity(0, cpus))
    if r.returncode != 0:
        raise RuntimeError("cpu baseline failed: " + r.stderr[-2000:])
    res = json.loads(r.stdout.strip().splitlines()[-1])
    return {"value": res["gp_per_s"], "unit": "gridpoint-updates/s", "cores": cores, "kind": "port",
            "sample": "C oracle (gcc -Ofast -fopenmp -march=native -funroll-loops) on the benchmark workload: BSSN "
                      "Cartesian %d^3 seeded flat + Gaussians, %d RK4 steps (dt = 0.5 dx) in %.1f s, OMP_NUM_THREADS=%d "
                      "pinned to CPUs %s (one per physical core), CPU: %s"
                      % (n, res["steps"], res["seconds"], cores, ",".join(map(str, cpus)), cpu_model())}


def measured_traffic(kernel, n, outputs):
    """HBM bytes per launch of ``kernel`` at n^3 from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the plain-copy calibration), or None when
    there is none for this kernel and output group."""
    path = os.path.join(ROOT, "profiles", "%s_bssn%d_traffic.json" % (kernel, n))
    try:
        with open(path) as f:
            d = json.load(f)
        if d.get("outputs") != list(outputs):
            return None
        return d["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def time_kernel(torch, fn, reps):
    """Average duration (s) of fn() launched on torch's current stream, HIP events."""
    fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    e.synchronize()
    return s.elapsed_time(e) / 1000.0 / reps


def bench_bl(torch):
    """Brill-Lindquist head-on, Spherical 72x12x2, RMAX=7.5, CFL 0.5, 50 RK4 steps."""
    from nrpytutorial_amd.runtime import Bundle, Grid, Evolution
    b = Bundle("bssn_spherical_sym2")
    pars = dict(BH1_mass=0.5, BH2_mass=0.5, BH1_posn_z=0.5, BH2_posn_z=-0.5, RMAX=7.5, eta=2.0)
    g = Grid(b, (72, 12, 2), **pars)
    ev = Evolution(g)
    ev.initial_data()
    dt = g.find_timestep(0.5)
    ev.step(dt, 5)
    torch.cuda.synchronize()
    ev.initial_data()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev.step(dt, 50)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    H = g.interior(ev.hamiltonian())
    l2 = float(torch.sqrt(torch.mean(H * H)))
    return {"gp_per_s": 4.0 * 72 * 12 * 2 * 50 / el, "seconds": el, "H_L2": l2, "dt": dt}


def bench_adjoint(torch, N=128, nsteps=200):
    """Adjoint scalar wave: 128^3 on [-10,10]^3, 200 RK4 steps, J = 0.5 sum uu(T)^2 dx^3,
    gradient w.r.t. 16 bump amplitudes through the checkpointed autograd evolution.
    Reports forward-step and adjoint-step times (the adjoint step includes recomputing
    its three stage states, and the checkpointed backward re-runs each forward step)."""
    from nrpytutorial_amd.runtime import Bundle, Grid, Evolution
    from nrpytutorial_amd.adjoint import RK4Adjoint, RK4Evolve
    from nrpytutorial_amd.workloads import scalarwave_bump_parameters, scalarwave_bump_basis_torch
    b = Bundle("scalarwave_cartesian")
    g = Grid(b, (N, N, N), wavespeed=1.0)
    ev = Evolution(g)
    adj = RK4Adjoint(ev)
    amps, cents = scalarwave_bump_parameters(3)
    basis = scalarwave_bump_basis_torch(g.xx_dev, cents)
    dt = 0.5 * g.params.dxx0
    dV = g.params.dxx0 * g.params.dxx1 * g.params.dxx2

    def loss(a):
        uu0 = 2.0 + (a[:, None, None, None] * basis).sum(0)
        yT = RK4Evolve().apply(torch.stack([uu0, torch.zeros_like(uu0)]), ev, adj, dt, nsteps)
        return 0.5 * (g.interior(yT[0]) ** 2).sum() * dV

    a = torch.tensor(amps, dtype=torch.float64, device="cuda", requires_grad=True)
    loss(a).backward()   # warm-up (graph capture, allocations)
    torch.cuda.synchronize()
    a = torch.tensor(amps, dtype=torch.float64, device="cuda", requires_grad=True)
    t0 = time.perf_counter()
    J = loss(a)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    J.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    # one forward step and one adjoint step in isolation
    reps = 20
    ev.step(dt, 1)
    y = ev.y_n.clone()
    lam = torch.zeros_like(y)
    g.interior(lam).normal_()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    ev.step(dt, reps)
    e.record()
    e.synchronize()
    t_fwd = s.elapsed_time(e) / 1000.0 / reps
    s.record()
    for _ in range(reps):
        adj.step_back(y, lam, dt)
    e.record()
    e.synchronize()
    t_adj = s.elapsed_time(e) / 1000.0 / reps
    return {"grid": [N, N, N], "steps": nsteps, "forward_s": t1 - t0, "backward_s": t2 - t1,
            "forward_step_ms": 1e3 * t_fwd, "adjoint_step_ms": 1e3 * t_adj, "adjoint_over_forward": t_adj / t_fwd,
            "backward_over_forward": (t2 - t1) / (t1 - t0),
            "gp_per_s_adjoint_steps": 4.0 * N ** 3 / t_adj, "J": float(J.detach())}


def bench_batched(torch, ninst=256, nsteps=100, seed=0):
    """Brill-Lindquist sweep: ninst head-on collisions (masses in [0.3,0.7], separation in
    [0.6,2.0], Nr in [48,96], Ntheta in [8,16], Nphi=2, RMAX 7.5, CFL 0.5), nsteps RK4 steps
    each; batched (one graph, autotuned tile) vs the same instances run one after another
    with their own graphs. Throughput counts 4 RHS evaluations per interior point per step."""
    from nrpytutorial_amd.runtime import Bundle
    from nrpytutorial_amd.batch import BatchedEvolution, brill_lindquist_sweep, autotune
    b = Bundle("bssn_spherical_sym2")
    inst = brill_lindquist_sweep(ninst, seed=seed)
    npts = sum(n[0] * n[1] * n[2] for n, _ in inst)
    be = BatchedEvolution(b, inst)
    best, times = autotune(be)
    be.initial_data()
    be.step(2)
    be.initial_data()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    be.step(nsteps)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    H_batch = be.hamiltonian_l2()
    # sequential plain-GPU runs of the same instances (graphs built outside the timer)
    for ev, dt in zip(be.evs, be.dts):
        ev.initial_data()
        ev.step(dt, 1, use_graph=True)
        ev.initial_data()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for ev, dt in zip(be.evs, be.dts):
        ev.step(dt, nsteps, use_graph=True)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    H_seq = [float(torch.sqrt(torch.mean(ev.g.interior(ev.hamiltonian()) ** 2))) for ev in be.evs]
    maxrel = max(abs(a - c) / abs(c) for a, c in zip(H_batch, H_seq))
    for ev in be.evs:
        ev.free_graph()
    return {"instances": ninst, "steps": nsteps, "points": npts, "tile": list(best),
            "batched_s": t_batch, "sequential_s": t_seq, "speedup": t_seq / t_batch,
            "batched_gp_per_s": 4.0 * npts * nsteps / t_batch, "sequential_gp_per_s": 4.0 * npts * nsteps / t_seq,
            "batched_rk4_steps_per_s": ninst * nsteps / t_batch, "H_L2_max_rel_diff": maxrel,
            "autotune_ms": {"x".join(map(str, k)): 1e3 * v for k, v in times.items()}}


def bench_membound(torch, reps=10):
    """The memory-bound fused RHS + RK4 kernels (substep 1): scalar wave at 256^3 and vacuum
    Maxwell at 192^3, HIP-event averages. Algorithmic bytes per interior point: every
    evolved gridfunction read once, and per output gridfunction y_n and the running total
    read, the stage output and the running total written: 8*(NGF + 4*NGF) B."""
    from nrpytutorial_amd.runtime import Bundle, Grid, Evolution, ptr, check
    out = {}
    for name, N, pars in (("scalarwave_cartesian", 256, {}),
                          ("maxwell_cartesian", 192, dict(xmin=-4.0, xmax=4.0, ymin=-4.0, ymax=4.0, zmin=-4.0,
                                                          zmax=4.0, time=0.4))):
        b = Bundle(name)
        g = Grid(b, (N, N, N), **pars)
        ev = Evolution(g)
        ev.initial_data()
        ev.k_odd.copy_(ev.y_n)
        dt = 0.5 * g.params.dxx0
        s = ev.s()
        f = b.lib.rhs_eval_rk4_substep
        # substep 1: stage input k_odd, y_n and the running total read, stage output k_even
        t = time_kernel(torch, lambda: check(f(1, g.rfm, g.pp(), ptr(ev.auxevol), ptr(ev.k_odd), ptr(ev.y_n),
                                               ptr(ev.rt), ptr(ev.k_even), dt, s), "rhs_eval_rk4_substep"), reps)
        ngf = len(b.evol)
        bpp = 8 * (ngf + 4 * ngf)
        gbs = bpp * N ** 3 / t / 1e9
        out[name] = {"grid": [N, N, N], "avg_launch_ms": 1e3 * t, "bytes_per_point": bpp, "achieved_GBs": gbs,
                     "hbm_frac": gbs / HBM_PEAK_GBS}
        del ev, g
    # reference: a 1 GiB device-to-device copy (read + write) on the same device
    src = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    t = time_kernel(torch, lambda: dst.copy_(src), reps)
    out["device_copy_1GiB"] = {"avg_ms": 1e3 * t, "GBs": 2 * src.numel() * 8 / t / 1e9}
    return out


def bench_next_rows(torch, reps=10):
    """The rows beyond the BSSN hot path, each on one GPU:
    * Sommerfeld outer boundary (the Maxwell Start-to-Finish choice): vacuum Maxwell 192^3 on
      [-4,4]^3, apply_bcs_sommerfeld on the RHS (HIP-event average) and a whole RK4 step of
      MoL_step_forward_in_time_sommerfeld (fused: interior RHS + update, then the condition
      with the update at the outer points), of the unfused step in the reference's order,
      and of the extrapolation-condition step (graph replay) for comparison;
    * curvilinear scalar wave, RK4 steps replayed from a hipGraph: SinhSphericalv2
      128 x 64 x 128 and SymTP 128 x 64 x 128 (2-D precomputed reference-metric tables).
    Rates are RK4 gridpoint-updates/s (4 RHS evaluations per interior point per step)."""
    from nrpytutorial_amd.runtime import Bundle, Grid, Evolution
    out = {}
    b = Bundle("maxwell_cartesian")
    N = 192
    g = Grid(b, (N, N, N), xmin=-4.0, xmax=4.0, ymin=-4.0, ymax=4.0, zmin=-4.0, zmax=4.0, time=0.4)
    ev = Evolution(g)
    ev.initial_data()
    dt = g.find_timestep(0.5)
    t_bc = time_kernel(torch, lambda: ev.apply_bcs_sommerfeld(ev.y_n, ev.k_even), reps)
    t_step = time_kernel(torch, lambda: ev.step_sommerfeld(dt, 1), reps)
    t_unfused = time_kernel(torch, lambda: ev.step_sommerfeld(dt, 1, unfused=True), reps)
    ev.initial_data()
    t_extrap = time_kernel(torch, lambda: ev.step(dt, 1), reps)
    out["sommerfeld_maxwell_192"] = {"apply_bcs_sommerfeld_ms": 1e3 * t_bc, "rk4_step_ms": 1e3 * t_step,
                                     "unfused_rk4_step_ms": 1e3 * t_unfused,
                                     "extrapolation_rk4_step_ms": 1e3 * t_extrap,
                                     "over_extrapolation": t_step / t_extrap, "gp_per_s": 4.0 * N ** 3 / t_step}
    ev.free_graph()
    del ev, g
    for name, Nxx, pars in (("scalarwave_sinhsphericalv2", (128, 64, 128), dict(AMPL=10.0)),
                            ("scalarwave_symtp", (128, 64, 128), dict(AMAX=10.0))):
        b = Bundle(name)
        g = Grid(b, Nxx, **pars)
        ev = Evolution(g)
        ev.initial_data()
        dt = g.find_timestep(0.5)
        ev.step(dt, 2)
        t = time_kernel(torch, lambda: ev.step(dt, 1), reps)
        out[name] = {"grid": list(Nxx), "rk4_step_ms": 1e3 * t, "gp_per_s": 4.0 * Nxx[0] * Nxx[1] * Nxx[2] / t}
        ev.free_graph()
        del ev, g
    return out


def bench_adjoint_bssn(torch, reps=5):
    """Adjoint of one BSSN RK4 step (BSSNStepAdjoint.step_back: recomputation of the four
    stages, then the transposed det(gammabar), ghost fill, RHS and Ricci) against the
    forward step (hipGraph replay) on the same state: Brill-Lindquist data on Spherical
    64x16x2 and the seeded flat + Gaussian state on a Cartesian 96^3 grid."""
    from nrpytutorial_amd.runtime import Bundle, Grid, Evolution
    from nrpytutorial_amd.adjoint import BSSNStepAdjoint
    from nrpytutorial_amd.workloads import bssn_gaussian_state_torch
    out = {}
    cases = (("spherical_bl_64x16x2", "bssn_spherical_sym2", (64, 16, 2),
              dict(BH1_mass=0.5, BH2_mass=0.5, BH1_posn_z=0.5, BH2_posn_z=-0.5, RMAX=7.5, eta=2.0)),
             ("cartesian_96", "bssn_cartesian", (96, 96, 96), {}))
    for label, name, Nxx, pars in cases:
        b = Bundle(name)
        g = Grid(b, Nxx, **pars)
        ev = Evolution(g)
        if name == "bssn_cartesian":
            bssn_gaussian_state_torch(ev.y_n, g.xx_dev, b.evol, seed=1)
            dt = 0.5 * g.params.dxx0
        else:
            ev.initial_data()
            dt = g.find_timestep(0.5)
        y0 = ev.y_n.clone()
        adj = BSSNStepAdjoint(ev)
        lam = torch.randn(ev.y_n.shape, dtype=torch.float64, device="cuda",
                          generator=torch.Generator(device="cuda").manual_seed(3))
        t_fwd = time_kernel(torch, lambda: ev.step(dt, 1), reps)
        t_rec = time_kernel(torch, lambda: adj.recompute(y0, dt), reps)
        t_adj = time_kernel(torch, lambda: adj.step_back(y0, lam, dt), reps)
        out[label] = {"grid": list(Nxx), "forward_step_ms": 1e3 * t_fwd, "adjoint_step_ms": 1e3 * t_adj,
                      "recompute_ms": 1e3 * t_rec, "adjoint_over_forward": t_adj / t_fwd}
        ev.free_graph()
        del adj, ev, g
        torch.cuda.empty_cache()
    return out


def bench_maxwell_curvilinear(torch, nsteps=20):
    """Vacuum Maxwell in rescaled SinhCylindrical coordinates at the sizes the reference
    times (Tutorial-Start_to_Finish-Solving_Maxwells_Equations_in_Vacuum-Curvilinear.ipynb:
    50x4x100 and 80x4x160, sinh_width 0.4, CFL 0.5, toroidal dipole): RK4 steps with the
    Sommerfeld condition (the driver's choice), and with quadratic extrapolation replayed
    from a hipGraph. Rates are the reference's gp/s (4 RHS evaluations per point per step)."""
    from nrpytutorial_amd.runtime import Bundle, Grid, Evolution
    out = {}
    b = Bundle("maxwell_sinhcylindrical")
    for Nxx in ((50, 4, 100), (80, 4, 160)):
        g = Grid(b, Nxx, SINHWRHO=0.4, SINHWZ=0.4)
        ev = Evolution(g)
        ev.initial_data(apply_bcs=False)
        dt = g.find_timestep(0.5)
        n = Nxx[0] * Nxx[1] * Nxx[2]
        t_som = time_kernel(torch, lambda: ev.step_sommerfeld(dt, nsteps), 1) / nsteps
        ev.initial_data()
        ev.step(dt, 1)
        t_ext = time_kernel(torch, lambda: ev.step(dt, nsteps), 1) / nsteps
        out["x".join(map(str, Nxx))] = {"sommerfeld_step_ms": 1e3 * t_som, "sommerfeld_gp_per_s": 4.0 * n / t_som,
                                        "extrapolation_graph_step_ms": 1e3 * t_ext,
                                        "extrapolation_gp_per_s": 4.0 * n / t_ext}
        ev.free_graph()
        del ev, g
    return out


def bench_etk(torch, ev, reps, t_ricci, t_rhs):
    """The Einstein Toolkit plugin form (bundle bssn_etk) on the BSSN state of the main run:
    each gridfunction its own array (the slices of that state: cctk_ash = cctk_lsh, three
    ghost zones), one RHS evaluation = Baikal_driver_pt1_BSSN_Ricci + Baikal_driver_pt2_BSSN_RHSs,
    against Ricci_eval + rhs_eval of the SENR layout on the same state. Also checks that
    both forms give the same values (to rounding for the lane-move group, see below)."""
    from nrpytutorial_amd.etk import Baikal, cctk_grid
    g = ev.g
    G = g.bundle.NGHOSTS
    k = Baikal()
    s = ev.s()
    gh = cctk_grid((g.shape[2], g.shape[1], g.shape[0]), (G, G, G), (g.params.dxx0, g.params.dxx1, g.params.dxx2))
    params = k.new_params(eta=g.params.eta)
    rbar = torch.zeros((len(k.auxevol),) + g.shape, dtype=torch.float64, device="cuda")
    rhs = torch.zeros((len(k.evol),) + g.shape, dtype=torch.float64, device="cuda")
    evol = [ev.y_n[q] for q in range(len(k.evol))]
    rbar_l = [rbar[q] for q in range(len(k.auxevol))]
    rhs_l = [rhs[q] for q in range(len(k.evol))]
    t_r = time_kernel(torch, lambda: k.ricci(gh, params, evol, rbar_l, s), reps)
    t_h = time_kernel(torch, lambda: k.rhss(gh, params, evol, rbar_l, rhs_l, s), reps)
    torch.cuda.synchronize()
    # bit for bit, except the outputs of SENR kernels that build their x-neighbours by lane
    # moves (bundle meta "xshift_kernels"): the compiler contracts their reordered statements
    # into FMAs differently, so those agree to rounding (1e-13 of each gridfunction's max norm)
    b = g.bundle
    moved = {n for q, grp in enumerate(b.meta["rhs_groups"])
             if "rhs_eval_g%d" % q in b.meta.get("xshift_kernels", []) for n in grp}
    ri, rs = g.interior(rhs), g.interior(ev.k_even)
    same = bool(torch.equal(g.interior(rbar), g.interior(ev.auxevol)))
    for q, gf in enumerate(b.evol):
        if gf in moved:
            same = same and float((ri[q] - rs[q]).abs().max()) <= 1e-13 * float(rs[q].abs().max())
        else:
            same = same and bool(torch.equal(ri[q], rs[q]))
    npts = g.Nxx[0] * g.Nxx[1] * g.Nxx[2]
    return {"grid": list(g.Nxx), "ricci_ms": 1e3 * t_r, "rhss_ms": 1e3 * t_h,
            "senr_ricci_ms": 1e3 * t_ricci, "senr_rhs_eval_ms": 1e3 * t_rhs,
            "rhs_evals_gp_per_s": npts / (t_r + t_h), "senr_rhs_evals_gp_per_s": npts / (t_ricci + t_rhs),
            "matches_senr_kernels": same}


def summarize_trace(csv_path, last=0, match=None, top=20):
    """Text summary of a rocprofv3 --kernel-trace CSV (no GPU needed): over the kernels whose
    name matches the regular expression `match` (all if None), and of those the last `last`
    (0 = all), per-kernel calls, average and total duration, the window, the time the GPU
    ran a kernel (overlaps merged) and the gaps between consecutive kernels."""
    import csv
    import re
    rows = []
    with open(csv_path) as f:
        for r in csv.DictReader(f):
            if match is None or re.search(match, r["Kernel_Name"]):
                rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    rows.sort()
    if last:
        rows = rows[-last:]
    agg = {}
    for t0, t1, name in rows:
        a = agg.setdefault(name, [0, 0])
        a[0] += 1
        a[1] += t1 - t0
    busy, gaps, end = 0, [], rows[0][0]
    for t0, t1, _ in rows:
        if t0 > end:
            gaps.append(t0 - end)
        busy += max(0, t1 - max(t0, end))
        end = max(end, t1)
    win = end - rows[0][0]
    gaps += [0] * (len(rows) - 1 - len(gaps))
    gaps.sort()
    out = ["%d kernels: window %.3f ms, GPU busy %.3f ms (%.1f%%), gap between kernels mean %.2f us, "
           "median %.2f us" % (len(rows), win / 1e6, busy / 1e6, 100.0 * busy / win,
                               sum(gaps) / max(len(gaps), 1) / 1e3, gaps[len(gaps) // 2] / 1e3 if gaps else 0.0)]
    for name, (n, tot) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
        out.append("  %-70s n=%5d avg %9.1f us  total %9.2f ms" % (name[:70], n, tot / n / 1e3, tot / 1e6))
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=192)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-bl", action="store_true", help="skip the Brill-Lindquist 72x12x2 line")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint scalar-wave line")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched Brill-Lindquist sweep line")
    ap.add_argument("--no-etk", action="store_true", help="skip the Einstein Toolkit plugin-form line")
    ap.add_argument("--no-next", action="store_true",
                    help="skip the Sommerfeld Maxwell 192^3 and curvilinear scalar-wave line")
    ap.add_argument("--no-adjoint-bssn", action="store_true", help="skip the BSSN adjoint-step line")
    ap.add_argument("--no-maxwell-curvi", action="store_true",
                    help="skip the curvilinear Maxwell 50x4x100 / 80x4x160 line")
    ap.add_argument("--no-membound", action="store_true",
                    help="skip the memory-bound scalar-wave 256^3 / Maxwell 192^3 kernel line")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 1 GiB device copy (scale for the HBM byte counters)")
    ap.add_argument("--cpu-baseline-worker", nargs=3)
    ap.add_argument("--summarize-trace", metavar="CSV",
                    help="print a summary of a rocprofv3 kernel-trace CSV and exit (no GPU used)")
    ap.add_argument("--trace-last", type=int, default=0,
                    help="with --summarize-trace: only the last N (matching) kernels of the trace")
    ap.add_argument("--trace-match", metavar="REGEX",
                    help="with --summarize-trace: only kernels whose name matches REGEX")
    a = ap.parse_args()
    if a.summarize_trace:
        print(summarize_trace(a.summarize_trace, a.trace_last, a.trace_match))
        return 0
    if a.cpu_baseline_worker:
        n, sec, mx = a.cpu_baseline_worker
        print(json.dumps(cpu_baseline_worker(int(n), float(sec), int(mx))))
        return 0
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each box has one MI355X); "
              "run with --gpus 1", file=sys.stderr)
        return 2
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 1
    from nrpytutorial_amd.runtime import Bundle, Grid, Evolution, ptr, check
    from nrpytutorial_amd.workloads import bssn_gaussian_state_torch

    N = a.n
    b = Bundle("bssn_cartesian")
    g = Grid(b, (N, N, N))
    ev = Evolution(g)
    bssn_gaussian_state_torch(ev.y_n, g.xx_dev, b.evol, seed=1)
    ev.apply_bcs(ev.y_n)
    ev.post_bcs(ev.y_n)
    dt = 0.5 * g.params.dxx0
    torch.cuda.synchronize()
    log("grid %d^3 (%d points incl. ghosts), dt=%g" % (N, g.Ntot, dt))

    ev.step(dt, a.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev.step(dt, a.steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    gp = 4.0 * N ** 3 * a.steps / el
    if not torch.isfinite(ev.y_n).all():
        raise RuntimeError("non-finite state after the timed steps")
    log("RK4: %d steps in %.4f s -> %.4e gp/s, %.3f ms/step" % (a.steps, el, gp, 1e3 * el / a.steps))

    # Per-kernel timings (HIP events on torch's current stream = the launch stream).
    s = ev.s()
    reps = 10
    cost = b.meta["cost"]
    npts = N ** 3
    ngf = len(b.evol)
    groups = b.meta["rhs_groups"]
    t_ricci = time_kernel(torch, lambda: check(b.lib.Ricci_eval(g.rfm, g.pp(), ptr(ev.y_n), ptr(ev.auxevol), s),
                                               "Ricci"), reps)
    t_group = []
    for k in range(len(groups)):
        f = getattr(b.lib, "rhs_eval_rk4_substep_g%d" % k)
        f.restype = ctypes.c_int
        f.argtypes = b.lib.rhs_eval_rk4_substep.argtypes
        t_group.append(time_kernel(torch, lambda f=f: check(f(
            1, g.rfm, g.pp(), ptr(ev.auxevol), ptr(ev.k_odd), ptr(ev.y_n), ptr(ev.rt), ptr(ev.k_even), dt, s),
            "rhs_eval_rk4_substep"), reps))
    rhs_tmp = ev.k_even
    t_rhs = time_kernel(torch, lambda: check(b.lib.rhs_eval(g.rfm, g.pp(), ptr(ev.auxevol), ptr(ev.y_n),
                                                            ptr(rhs_tmp), s), "rhs_eval"), reps)
    t_bc = time_kernel(torch, lambda: ev.apply_bcs(ev.k_odd, composed=True), reps)
    t_detg = time_kernel(torch, lambda: ev.post_bcs(ev.k_odd), reps)
    # Dominant kernel: the slowest output group of the fused RHS + RK4 update (substep 1).
    # Algorithmic traffic per interior point: every gridfunction the kernel reads (evolved,
    # and RbarDD where the group needs it; bundle meta "rhs_group_inputs") read once, and per
    # output gridfunction y_n and the running total read, the stage output and the running
    # total written.
    kdom = max(range(len(groups)), key=lambda k: t_group[k])
    n_out0 = len(groups[kdom])
    dom_flops = cost["rhs_eval"]["groups"][kdom]["flops"] + 6 * n_out0
    dom_bytes = 8 * (b.meta["rhs_group_inputs"][kdom] + 4 * n_out0)
    t_dom = t_group[kdom]
    dom_name = "rhs_eval_rk4_substep_g%d" % kdom
    dom_gbs = dom_bytes * npts / t_dom / 1e9
    dom_tf = dom_flops * npts / t_dom / 1e12
    fused_flops = cost["rhs_eval"]["flops"] + 6 * ngf
    t_fused = sum(t_group)
    rhs_only_gp = npts / (t_ricci + t_rhs)
    log("kernel avg: Ricci %.3f ms (%.2f TF/s), fused RHS+RK groups %s ms (slowest group: %.0f GB/s algorithmic, "
        "%.2f TF/s), rhs_eval %.3f ms, ghost fill %.3f ms, detg %.3f ms"
        % (1e3 * t_ricci, cost["Ricci_eval"]["flops"] * npts / t_ricci / 1e12,
           "+".join("%.3f" % (1e3 * t) for t in t_group), dom_gbs, dom_tf, 1e3 * t_rhs, 1e3 * t_bc, 1e3 * t_detg))
    traffic = measured_traffic(dom_name, N, groups[kdom])
    etk = None
    if not a.no_etk:
        etk = bench_etk(torch, ev, reps, t_ricci, t_rhs)
        log("ETK plugin form %d^3: Ricci %.3f ms + RHSs %.3f ms (SENR layout %.3f + %.3f ms), "
            "%.3e RHS-evaluation gp/s; same values as the SENR kernels: %s"
            % (N, etk["ricci_ms"], etk["rhss_ms"], etk["senr_ricci_ms"], etk["senr_rhs_eval_ms"],
               etk["rhs_evals_gp_per_s"], etk["matches_senr_kernels"]))
    if a.calibrate_copy:
        # A plain device copy of known size, for scaling the HBM byte counters.
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: %d bytes read, %d bytes written per launch" % (src.numel() * 8, src.numel() * 8))
    bl = None
    if not a.no_bl:
        bl = bench_bl(torch)
        log("BL 72x12x2: 50 steps in %.4f s -> %.3e gp/s, H-L2 %.6e" % (bl["seconds"], bl["gp_per_s"], bl["H_L2"]))

    batched = None
    if not a.no_batched:
        batched = bench_batched(torch)
        log("batched BL sweep: %d instances x %d steps: batched %.3f s (%.3e gp/s, tile %s) vs sequential %.3f s "
            "(%.3e gp/s): %.1fx; max rel diff of H-L2 %.1e" % (batched["instances"], batched["steps"],
                                                               batched["batched_s"], batched["batched_gp_per_s"],
                                                               batched["tile"], batched["sequential_s"],
                                                               batched["sequential_gp_per_s"], batched["speedup"],
                                                               batched["H_L2_max_rel_diff"]))
    nxt = None
    if not a.no_next:
        nxt = bench_next_rows(torch)
        som = nxt["sommerfeld_maxwell_192"]
        log("Sommerfeld Maxwell 192^3: apply_bcs_sommerfeld %.3f ms, RK4 step %.3f ms (%.3e gp/s; unfused %.3f ms; "
            "%.2fx the extrapolation step's %.3f ms); curvilinear "
            "scalar wave RK4 steps: %s" % (som["apply_bcs_sommerfeld_ms"], som["rk4_step_ms"], som["gp_per_s"],
                                          som["unfused_rk4_step_ms"], som["over_extrapolation"],
                                          som["extrapolation_rk4_step_ms"],
                                          ", ".join("%s %s %.3f ms (%.3e gp/s)" % (k, "x".join(map(str, v["grid"])),
                                                                                v["rk4_step_ms"], v["gp_per_s"])
                                                    for k, v in nxt.items() if "grid" in v)))
    membound = None
    if not a.no_membound:
        membound = bench_membound(torch)
        log("memory-bound fused RHS+RK4 kernels: " + ", ".join(
            "%s %d^3 %.3f ms, %.0f GB/s algorithmic (%.0f%% of HBM)" % (k, v["grid"][0], v["avg_launch_ms"],
                                                                       v["achieved_GBs"], 100 * v["hbm_frac"])
            for k, v in membound.items() if "grid" in v) +
            "; device copy %.0f GB/s" % membound["device_copy_1GiB"]["GBs"])
    adjoint = None
    if not a.no_adjoint:
        adjoint = bench_adjoint(torch)
        log("adjoint scalar wave 128^3: forward step %.3f ms, adjoint step %.3f ms (%.2fx); 200-step loss %.3f s, "
            "gradient %.3f s" % (adjoint["forward_step_ms"], adjoint["adjoint_step_ms"], adjoint["adjoint_over_forward"],
                                 adjoint["forward_s"], adjoint["backward_s"]))
    adjoint_bssn = None
    if not a.no_adjoint_bssn:
        adjoint_bssn = bench_adjoint_bssn(torch)
        log("BSSN adjoint step: " + ", ".join(
            "%s forward %.3f ms, adjoint %.3f ms (%.2fx; recompute %.3f ms)"
            % (k, v["forward_step_ms"], v["adjoint_step_ms"], v["adjoint_over_forward"], v["recompute_ms"])
            for k, v in adjoint_bssn.items()))
    maxwell_curvi = None
    if not a.no_maxwell_curvi:
        maxwell_curvi = bench_maxwell_curvilinear(torch)
        log("Maxwell SinhCylindrical: " + ", ".join(
            "%s Sommerfeld step %.3f ms (%.3e gp/s), extrapolation graph step %.3f ms (%.3e gp/s)"
            % (k, v["sommerfeld_step_ms"], v["sommerfeld_gp_per_s"], v["extrapolation_graph_step_ms"],
               v["extrapolation_gp_per_s"]) for k, v in maxwell_curvi.items()))
    cpu = None
    if not a.no_cpu:
        try:
            cpu = cpu_baseline()
            log("CPU baseline: %.3e gp/s (%s)" % (cpu["value"], cpu["sample"]))
        except Exception as e:  # reported, never replaces the GPU number
            log("CPU baseline failed: %s" % e)
    out = {
        "metric": METRIC, "value": gp, "unit": "gridpoint-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * el / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded flat + 8 Gaussians per gridfunction)",
        "config": {"workload": "BSSN vacuum RK4 step, Cartesian reference metric, %d^3 interior, FD4 + upwinding, "
                               "1+log lapse, covariant Gamma-driver (eta=2), hipGraph-replayed substeps" % N,
                   "grid": [N, N, N], "NGHOSTS": 3, "gridfunctions": ngf, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": dom_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom_gbs / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "%s (fused BSSN RHS of %s + RK4 update, substep 1)"
                               % (dom_name, ",".join(groups[kdom])),
                     "bytes_per_point": dom_bytes, "points_per_launch": npts, "avg_launch_ms": 1e3 * t_dom,
                     "flops_per_point": dom_flops, "achieved_fp64_TFLOPs": dom_tf,
                     "fp64_frac": dom_tf / FP64_PEAK_TFLOPS,
                     "note": "algorithmic intensity %.1f FLOP/B is below the FP64 ridge (%.1f FLOP/B), so HBM is "
                             "the roof; FP64 VALU work, no matrix cores" % (dom_flops / dom_bytes,
                                                                           FP64_PEAK_TFLOPS * 1e3 / HBM_PEAK_GBS)},
        "cpu_baseline": cpu,
        "extra": {"rhs_only_gp_per_s": rhs_only_gp,
                  "kernel_ms": {"Ricci_eval": 1e3 * t_ricci, "rhs_eval": 1e3 * t_rhs, "ghost_fill": 1e3 * t_bc,
                                "detg": 1e3 * t_detg,
                                **{"rhs_eval_rk4_substep_g%d" % k: 1e3 * t for k, t in enumerate(t_group)}},
                  "kernel_ms_note": "each figure is one kernel launched eagerly %d times back to back on the launch "
                                    "stream, HIP events around all of them, divided by %d: it includes the gaps "
                                    "between eager launches, which the graph-replayed step (ms_per_step) does not "
                                    "have, so 4 x the per-substep sum exceeds ms_per_step by about that overhead"
                                    % (reps, reps),
                  "fused_substep_fp64_TFLOPs": fused_flops * npts / t_fused / 1e12,
                  "fused_flops_per_point": fused_flops, "ricci_flops_per_point": cost["Ricci_eval"]["flops"],
                  "rhs_groups": groups, "bl_72x12x2_50steps": bl, "adjoint_scalarwave_128": adjoint,
                  "batched_bl_sweep": batched, "memory_bound_kernels": membound, "etk_plugin_form": etk,
                  "sommerfeld_and_curvilinear": nxt, "adjoint_bssn": adjoint_bssn,
                  "maxwell_sinhcylindrical": maxwell_curvi},
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
