"""Throughput of the hybrid-PIC timestep and of the potential solve on one MI355X.

One step = one fused pass of push + collide + drift/wrap + deposit over every ion,
followed by the potential solve (RHS + GMRES/multigrid), i.e. src/pic.c:246-258.
The default run measures every workload of the hot path and prints one JSON line:

* headline (fp64_128cube_100ppc): 128^3 periodic grid, 100 ions per cell (209,715,200
  ions), FP64, initial density 1 + 1e-3 sin(2 pi 4 x / L), uniform B along z with
  Omega_i dt = 0.05, E0 along x, nu dt = 0.02, FP64 solve at rtol 1e-10;
* "mixed": the same workload with FP32 particles and the mixed-precision solve, with its
  errors against the FP64 path;
* "solve": the potential solve of that workload (FP64 and mixed), its HBM roofline and
  the oracle's FP64 GMRES(30)+multigrid on the same field on the host;
* legs "c0" (32^3 x 64 ions/cell FP64, cold neutrals, 10 steps), "c3" (256 standalone
  solves of smoothed log-normal fields, grid edges from {32, ..., 192}) and "c4" (16
  independent mixed-precision runs of uneven size, stepped as one lockstep ensemble).

Every leg carries the HBM roofline of its dominant kernel (HIP events on the stream the
kernel runs on) and a CPU baseline: the oracle on a bounded sample of the same workload,
median of 3 on one host core, and one replica per host core for the aggregate.
Progress lines go to stderr.
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
BYTES_PER_PARTICLE = 96  # FP64 SoA: read x,y,z,vx,vy,vz, write them back
BYTES_PER_PARTICLE_MIXED = 56  # int32 cell + 3 FP32 offsets + 3 FP32 velocities, read and written
BYTES_PER_NODE = 40      # read phi (8 B), write n, Gx, Gy, Gz (32 B)
T_START = time.perf_counter()


def log(msg):
    print(f"[bench {time.perf_counter() - T_START:7.1f}s] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="all",
                    choices=["all", "cpu32", "fp64", "mixed", "batch", "ensemble", "2d"],
                    help="all = every workload below in one line; cpu32 = 32^3 x 64 ions/cell FP64; fp64 = "
                         "128^3 x 100 ions/cell FP64 (with a mixed leg); mixed = the 128^3 workload on the "
                         "mixed-precision path only; batch = 256 standalone solves; ensemble = 16 independent runs")
    ap.add_argument("--legs", default="c0,c3,c4", help="all: which extra legs to run")
    ap.add_argument("--problems", type=int, default=256, help="batch: number of fields")
    ap.add_argument("--slots", type=int, default=0,
                    help="batch: 0 = lockstep (one launch per phase for all fields), else concurrent solver slots")
    ap.add_argument("--ppc", type=int, default=0, help="override ions per cell")
    ap.add_argument("--grid", type=int, default=0, help="override grid edge")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--cpu-cores", type=int, default=0, help="host cores for the per-core aggregate (0 = all, <= 16)")
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--tile-z", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--split", type=int, default=0)
    ap.add_argument("--collide", default="philox", choices=["philox", "off"])
    ap.add_argument("--mg-pre", type=int, default=0)
    ap.add_argument("--inner-rtol", type=float, default=0.0, help="mixed solve: inner FP32 tolerance (0 = library's)")
    ap.add_argument("--mg-post", type=int, default=0)
    ap.add_argument("--mg-sweep", type=int, default=0, choices=[0, 1, 2],
                    help="multigrid sweep kernel: 0 by level size, 1 one-shot boxes, 2 marching along z")
    ap.add_argument("--autotune", action="store_true",
                    help="pick the tile shape and block size of each particle path with ecsprt_amd.autotune")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only run a device-to-device copy of this many GiB (HBM counter calibration)")
    ap.add_argument("--sequential", action="store_true",
                    help="ensemble with --profile-steps: step the runs one after another on one host thread")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="run only this many steps (batch: passes) without the extras (for rocprofv3)")
    return ap.parse_args()


# ----------------------------------------------------------------------------- host side
def host_cores(args):
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    n = min(16, n)
    return max(1, min(n, args.cpu_cores) if args.cpu_cores else n)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def median_of(fn, n=3):
    """Median wall seconds of n calls of fn() (each returns its own timed seconds)."""
    return float(np.median([fn() for _ in range(n)]))


def replicas(fn, n):
    """Wall seconds of n concurrent calls fn(0..n-1) on n threads (the oracle's ctypes
    calls release the interpreter lock, so the replicas run on separate cores)."""
    from concurrent.futures import ThreadPoolExecutor
    go = threading.Barrier(n + 1)

    def run(i):
        go.wait()
        fn(i)

    with ThreadPoolExecutor(n) as ex:
        futs = [ex.submit(run, i) for i in range(n)]
        go.wait()
        t0 = time.perf_counter()
        for f in futs:
            f.result()
        return time.perf_counter() - t0


def uniform_ions(p, n, rng):
    pos = np.column_stack([rng.uniform(0, p.Lx, n), rng.uniform(0, p.Ly, n), rng.uniform(0, p.Lz, n)])
    vel = rng.normal(size=(n, 3)) * np.array([p.viTx, p.viTy, p.viTz])
    return pos, vel


def baseline(value, unit, sample, agg_value, cores, extra=None):
    out = {"value": value, "unit": unit, "cores": 1, "kind": "port", "sample": sample,
           "runs": "median of 3", "cpu": cpu_model(),
           "per_core_replicas": {"value": agg_value, "unit": unit, "cores": cores,
                                 "note": "independent replicas run at once, one per host core (sizes in the sample)"}}
    if extra:
        out.update(extra)
    return out


def cpu_particles(wl, cores, n=4_000_000, steps=2):
    """Oracle push + collide (NR stream) + drift/wrap + deposit on a uniform sample of ions
    in the workload's grid and physics with a random potential."""
    import oracle
    p = oracle.setup_context(**wl.cli)
    rng = np.random.default_rng(11)
    phi = 1e-2 * rng.normal(size=(p.Nz, p.Ny, p.Nx))
    pos0, vel0 = uniform_ions(p, n, rng)
    stage = []

    def one():
        pos, vel = pos0.copy(), vel0.copy()
        t0 = time.perf_counter()
        _, _, t = oracle.particle_steps(p, wl.dt, steps, pos, vel, phi, oracle.NR())
        stage.append(t)
        return time.perf_counter() - t0

    el = median_of(one)
    nrep = max(1, n // cores)
    reps = [uniform_ions(p, nrep, np.random.default_rng(100 + i)) for i in range(cores)]
    wall = replicas(lambda i: oracle.particle_steps(p, wl.dt, steps, reps[i][0], reps[i][1], phi, oracle.NR()), cores)
    t = np.median(np.array(stage), axis=0)
    return baseline(n * steps / el, "particle-steps/s",
                    f"{n} ions uniform in the {p.Nx}x{p.Ny}x{p.Nz} grid, {steps} steps of push+collide(NR stream)"
                    f"+drift/wrap+deposit, oracle -O3; stage s (push,collide,drift,deposit): "
                    + ",".join(f"{x:.2f}" for x in t) + f"; per-core replicas of {nrep} ions",
                    cores * nrep * steps / wall, cores)


def cpu_solve(p_cli, moments, rtol, cores):
    """The oracle's FP64 GMRES(30) + geometric multigrid on one field: ComputeFullRHS3D and
    the solve from a zero guess at the workload's rtol."""
    import oracle
    p = oracle.setup_context(**p_cli)
    b = oracle.full_rhs(p, moments)
    info = {}

    def one():
        t0 = time.perf_counter()
        x, it, rn, ok = oracle.solve_gmres_mg(p, moments, b, rtol)
        el = time.perf_counter() - t0
        info.update(iterations=it, converged=ok, rel_residual=rn / max(np.linalg.norm(b), 1e-300))
        return el

    el = median_of(one)
    wall = replicas(lambda i: oracle.solve_gmres_mg(p, moments, b, rtol), cores)
    return baseline(1e3 * el, "ms per solve",
                    f"the {p.Nx}x{p.Ny}x{p.Nz} field of the GPU run, FP64 GMRES(30)+GMG oracle at rtol {rtol:g} "
                    f"(RHS excluded)", 1e3 * wall / cores, cores,
                    dict(info, higher_is_better=False, per_core_solves_per_s=cores / wall))


# ----------------------------------------------------------------------------- GPU side
def copy_probe(gib, reps=5):
    """hipMemcpy device-to-device of `gib` GiB: achievable HBM bandwidth (read + write)."""
    import ctypes
    from ecsprt_amd import _lib
    L = _lib.load()
    n = int(gib * (1 << 30))
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert L.ecsprt_hip_dmalloc(ctypes.byref(a), n) == 0 and L.ecsprt_hip_dmalloc(ctypes.byref(b), n) == 0
    ms = ctypes.c_double()
    best = 1e30
    for _ in range(reps):
        assert L.ecsprt_hip_d2d(b, a, n, ctypes.byref(ms)) == 0
        best = min(best, ms.value)
    L.ecsprt_hip_dfree(a)
    L.ecsprt_hip_dfree(b)
    return {"copy_bytes": n, "best_ms": best, "read_plus_write_GBs": 2 * n / (best * 1e-3) / 1e9}


def launch_of(p, o, mixed=False):
    """Effective tile shape and block size of the particle pass.  With none given the
    library chooses: 8-cell tiles once a grid holds at least 512 of them, as 8x8x4 cells
    with 512 threads for FP64 particles and 8x8x8 with 1024 for mixed; else 4x4x4 cells
    with 256 threads."""
    n8 = ((p.Nx + 7) // 8) * ((p.Nz + 7) // 8) * ((p.Ny + 7) // 8)
    auto = not (o.get("tile") or o.get("tile_z") or o.get("block"))
    t = o.get("tile") or (8 if n8 >= 512 else 4)
    f64_auto = auto and t == 8 and not mixed
    tz = o.get("tile_z") or (4 if f64_auto else t)
    block = o.get("block") or (256 if t == 4 else 512 if f64_auto else 1024)
    return f"{t}x{t}x{tz}/{block}"


def committed_traffic(kernel, workload_name, launch, args):
    """HBM bytes per launch of `kernel` on this workload and launch shape, from the
    calibrated rocprofv3 FETCH_SIZE/WRITE_SIZE runs summarised in profiles/hbm_traffic.md;
    None for a configuration that was not profiled."""
    if args.ppc or args.grid or args.mg_pre or args.mg_post or args.mg_sweep or args.split or args.collide != "philox":
        return None
    path = os.path.join(HERE, "profiles", "hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic:"):
            f = dict(kv.split("=", 1) for kv in line.split()[1:])
            if f.get("kernel") == kernel and f.get("workload") == workload_name and f.get("launch") == launch:
                return float(f["bytes_per_launch"])
    return None


def roofline(alg_bytes, ms, kernel, traffic=None, **extra):
    ach = alg_bytes / (ms * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
           "traffic": traffic, "kernel": kernel, "algorithmic_bytes_per_launch": alg_bytes, "avg_ms": ms}
    out.update(extra)
    return out


def run_timed(p, opts, wl, nsteps, warm, pic, init=None):
    """Initial conditions, t = 0 collect + solve (src/pic.c:211-215), warm-up, then
    nsteps timed steps.  Returns the context and the per-stage means."""
    ctx = pic.Context(p, opts)
    if init is None:
        ctx.init_particles(wl.np, 2024, wl.amp, wl.k)
    else:
        ctx.set_particles(*init)
    ctx.collect()
    ctx.compute_potential()
    for _ in range(warm):
        ctx.step(wl.dt)
    ctx.synchronize()
    st_all, iters, sbytes = [], [], []
    t0 = time.perf_counter()
    for _ in range(nsteps):
        ctx.step(wl.dt)
        st_all.append(ctx.stage_times())
        iters.append(ctx.solver_stats()[0])
        sbytes.append(ctx.solver_traffic())
    ctx.synchronize()
    el = time.perf_counter() - t0
    st = np.array(st_all)
    m = st.mean(axis=0)
    solve_ms = st[:, 4] + st[:, 5]
    return ctx, {"seconds": el, "k_ms": float(m[0]), "pre_ms": float(m[1]), "rebin_ms": float(m[2]),
                 "rhs_ms": float(m[4]), "solve_ms": float(m[5]), "iters": float(np.mean(iters)),
                 "solve_bytes": float(np.mean(sbytes)), "solve_total_ms": float(solve_ms.mean()),
                 "solve_GBs": float(np.sum(sbytes) / (solve_ms.sum() * 1e-3) / 1e9)}


def rel_err(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def norms(a, b):
    """Max-norm and RMS errors of a against b, each relative to b's max-norm / RMS."""
    d = a - b
    return (float(np.abs(d).max() / max(np.abs(b).max(), 1e-300)),
            float(np.sqrt(np.mean(d * d)) / max(np.sqrt(np.mean(b * b)), 1e-300)))


def mixed_errors(p, wl, pic, _lib, steps=3):
    """Errors of the mixed path (FP32 particles and mixed solve) against the FP64 path
    (which matches the oracle to 1e-12), collisions off so that both follow the same
    deterministic dynamics.

    per_step: each step starts the mixed context from the FP64 state (particles and
    potential), so the error is that of one mixed step.  free_running: both contexts
    advance on their own from the same initial conditions; the reference's gather is
    discontinuous across cell faces, so the few ions that FP32 rounding puts on the
    other side of a face diverge and dominate the max-norm, while the RMS error stays
    at the rounding level."""
    base = dict(collide_mode=_lib.COLLIDE_OFF, ksp_rtol=wl.rtol)
    c64 = pic.Context(p, pic.default_options(**base))
    cmx = pic.Context(p, pic.default_options(**base, particle_precision=_lib.PARTICLES_MIXED,
                                             solve_precision=_lib.SOLVE_MIXED))
    for c in (c64, cmx):
        c.init_particles(wl.np, 2024, wl.amp, wl.k)
        c.collect()
        c.compute_potential()

    def compare(step):
        m64, mmx = c64.get_moments(), cmx.get_moments()
        n = norms(mmx[..., 0], m64[..., 0])
        g = [norms(mmx[..., d], m64[..., d]) for d in (1, 2, 3)]
        f = norms(cmx.get_potential(), c64.get_potential())
        return {"step": step, "n_max": n[0], "n_rms": n[1], "G_max": max(x[0] for x in g),
                "G_rms": max(x[1] for x in g), "phi_max": f[0], "phi_rms": f[1]}

    free = []
    for s in range(steps):
        c64.step(wl.dt)
        cmx.step(wl.dt)
        free.append(compare(s + 1))
    per = []
    for s in range(steps):
        pos, vel = c64.get_particles()
        cmx.set_particles(pos, vel)
        cmx.set_potential(c64.get_potential())
        del pos, vel
        c64.step(wl.dt)
        cmx.step(wl.dt)
        per.append(compare(steps + s + 1))
    c64.close()
    cmx.close()
    return {"per_step": per, "free_running": free,
            "phi_within_1e-6": {"per_step": max(r["phi_max"] for r in per) <= 1e-6,
                                "free_running": max(r["phi_max"] for r in free) <= 1e-6,
                                "free_running_rms": max(r["phi_rms"] for r in free) <= 1e-6}}


# ----------------------------------------------------------------------------------- c0
def leg_c0(args, pic, _lib, cores):
    """32^3 x 64 ions/cell, FP64, Philox collisions with cold neutrals (Tn = Ti/100,
    nu dt = 0.01), solve at rtol 1e-8, 10 steps (median of 3 fresh runs).  CPU baseline:
    the oracle on the same initial ions, 10 full steps (NR-stream collisions and its own
    GMRES+multigrid solve)."""
    import oracle
    from ecsprt_amd import configs
    wl = configs.CPU32
    p = pic.setup_context(**wl.cli)
    opts = dict(collide_mode=_lib.COLLIDE_PHILOX, ksp_rtol=wl.rtol, timing=1)
    runs = []
    for rep in range(3):
        ctx, r = run_timed(p, pic.default_options(**opts), wl, wl.steps, args.warmup, pic)
        r["coll"] = ctx.collision_stats()
        ctx.close()
        runs.append(r)
    runs.sort(key=lambda r: r["seconds"])
    r = runs[1]
    Nc, Ns, Nf, natt = r["coll"]
    alg = BYTES_PER_PARTICLE * wl.np + BYTES_PER_NODE * wl.ncell
    launch = launch_of(p, {})
    out = {
        "metric": "particle-steps/s (push+deposit+collide+solve, 10 steps)",
        "value": wl.np * wl.steps / r["seconds"], "unit": "particle-steps/s", "steps": wl.steps,
        "ms_per_step": 1e3 * r["seconds"] / wl.steps, "dtype": "f64",
        "config": {"workload": wl.name, "grid": [p.Nx, p.Ny, p.Nz], "ions_per_cell": wl.ppc, "np": wl.np,
                   "Omega_i_dt": p.Omega_i * wl.dt, "nu_dt": p.nui * wl.dt, "Tn_over_Ti": 0.01, "ksp_rtol": wl.rtol,
                   "collisions": "Philox statistical mode", "runs": "median of 3 fresh runs"},
        "phi_solve_ms": r["solve_total_ms"], "ksp_iterations_mean": r["iters"],
        "collisions_last_step": {"Nc": Nc, "successes": Ns, "refused": Nf, "attempts": natt},
        "stage_ms_mean": {"fused_particle_kernel": r["k_ms"], "pre": r["pre_ms"], "rebin_fold": r["rebin_ms"],
                          "rhs": r["rhs_ms"], "solve": r["solve_ms"]},
        "roofline": roofline(alg, r["k_ms"], "k_tile<push|collide|drift|deposit|rebin>",
                             committed_traffic("k_tile", wl.name, launch, args), launch=launch),
        "solve_roofline": roofline(r["solve_bytes"], r["solve_total_ms"], "potential solve (rhs + GMRES/GMG, FP64)"),
    }
    if args.no_cpu:
        return out
    log("c0: CPU baseline (oracle, 10 full steps)")
    ctx = pic.Context(p, pic.default_options(**opts))
    ctx.init_particles(wl.np, 2024, wl.amp, wl.k)
    ctx.collect()
    ctx.compute_potential()
    pos0, vel0 = ctx.get_particles()
    phi0 = ctx.get_potential()
    ctx.close()
    po = oracle.setup_context(**wl.cli)
    stage = []

    def run(pos, vel, nsteps, rec):
        phi = phi0.copy()
        t = np.zeros(6)
        for _ in range(nsteps):
            _, m, ts = oracle.particle_steps(po, wl.dt, 1, pos, vel, phi, oracle.NR())
            t[:4] += ts
            a = time.perf_counter()
            b = oracle.full_rhs(po, m)
            c = time.perf_counter()
            phi, _, _, _ = oracle.solve_gmres_mg(po, m, b, wl.rtol)
            t[4] += c - a
            t[5] += time.perf_counter() - c
        if rec is not None:
            rec.append(t)

    def one():
        pos, vel = pos0.copy(), vel0.copy()
        t0 = time.perf_counter()
        run(pos, vel, wl.steps, stage)
        return time.perf_counter() - t0

    el = median_of(one)
    steps_rep = 2
    reps = [(pos0.copy(), vel0.copy()) for _ in range(cores)]
    wall = replicas(lambda i: run(reps[i][0], reps[i][1], steps_rep, None), cores)
    t = np.median(np.array(stage), axis=0)
    out["cpu_baseline"] = baseline(
        wl.np * wl.steps / el, "particle-steps/s",
        f"the full workload: the GPU run's {wl.np} initial ions, {wl.steps} steps of push+collide(NR stream)"
        f"+drift/wrap+deposit+RHS+GMRES/GMG solve at rtol {wl.rtol:g}, oracle -O3; stage s per run "
        "(push,collide,drift,deposit,rhs,solve): " + ",".join(f"{x:.2f}" for x in t)
        + f"; per-core replicas of the full workload for {steps_rep} steps",
        cores * wl.np * steps_rep / wall, cores)
    return out


# ----------------------------------------------------------------------------------- c3
def leg_c3(args, pic, _lib, cores):
    """256 standalone solves of smoothed log-normal density fields, grid edges drawn from
    {32, 48, 64, 96, 128, 192} per axis (seed 3), mixed precision to an FP64 residual of
    1e-8 relative, in lockstep (one launch per phase for all fields) unless --slots is
    given; the FP64 batch on the same fields is the reference for the error.  CPU
    baseline: the oracle's FP64 GMRES(30)+multigrid on the first fields of the draw."""
    import oracle
    from ecsprt_amd import configs
    nprob = args.problems
    dims = configs.batch_dims(nprob)
    cli0 = configs.base_cli(32, E0=0.05)
    p = pic.setup_context(**cli0)
    t0 = time.perf_counter()
    mom = [pic.lognormal_moments(d, seed=1000 + q, sigma=0.3, smooth=3.0) for q, d in enumerate(dims)]
    gen_s = time.perf_counter() - t0
    npts = np.array([d[0] * d[1] * d[2] for d in dims], dtype=np.int64)
    passes = args.profile_steps or 3
    res = {}
    for name, prec in (("mixed", _lib.SOLVE_MIXED), ("fp64", _lib.SOLVE_FP64)):
        phi = [pic.DeviceVector(int(n)) for n in npts]
        b = pic.Batch(p, pic.default_options(ksp_rtol=1e-8, solve_precision=prec), dims, slots=args.slots)
        if not args.profile_steps:
            b.solve(mom, phi)  # warm-up pass
        times = []
        for _ in range(passes):
            t0 = time.perf_counter()
            b.solve(mom, phi)
            times.append(time.perf_counter() - t0)
        it, rn = b.stats()
        res[name] = {"seconds": float(np.median(times)), "iters": it, "phi": phi, "traffic": b.traffic()}
        b.close()
        log(f"c3 {name}: {np.median(times) * 1e3:.1f} ms per batch")
        if args.profile_steps:
            break
    if args.profile_steps:
        for f in res["mixed"]["phi"]:
            f.free()
        for m in mom:
            m.free()
        return {"profile_passes": passes, "seconds": res["mixed"]["seconds"]}
    errs = []
    for q in range(nprob):
        ref = res["fp64"]["phi"][q].download()
        got = res["mixed"]["phi"][q].download()
        errs.append(float(np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300)))
    for name in res:
        for f in res[name]["phi"]:
            f.free()
    ms, fs = res["mixed"]["seconds"], res["fp64"]["seconds"]
    mode = "lockstep: one launch per phase for all fields" if args.slots == 0 else f"{args.slots} concurrent slots"
    out = {
        "metric": "batched potential solves/s (256 fields, mixed precision, 1e-8 FP64 residual)",
        "value": nprob / ms, "unit": "solves/s", "steps": passes, "ms_per_step": 1e3 * ms, "dtype": "f32 inner / f64 refinement",
        "config": {"workload": "batched_solves_256_fields", "problems": nprob, "points_total": int(npts.sum()),
                   "points_max": int(npts.max()), "scheduling": mode, "ksp_rtol": 1e-8,
                   "density": "exp(0.3 g - 0.045), Gaussian smoothing sigma 3 cells, mean 1; flux = n",
                   "runs": "median of 3 passes after a warm-up pass"},
        "points_per_s": float(npts.sum() / ms),
        "fp64": {"seconds": fs, "solves_per_s": nprob / fs, "mean_iterations": float(res["fp64"]["iters"].mean()),
                 "roofline": roofline(res["fp64"]["traffic"], 1e3 * fs, "batched solve, FP64 (all launches)")},
        "mixed": {"seconds": ms, "mean_iterations": float(res["mixed"]["iters"].mean()),
                  "speedup_vs_fp64": fs / ms, "phi_rel_err_vs_fp64_max": max(errs),
                  "phi_rel_err_vs_fp64_median": float(np.median(errs))},
        "roofline": roofline(res["mixed"]["traffic"], 1e3 * ms, "batched solve, mixed (all launches)",
                             committed_traffic("batch_mixed", "batched_solves_256_fields", "all", args)
                             if nprob == 256 and args.slots == 0 else None,
                             note="algorithmic bytes of every launch of one batch pass / the pass's wall time; "
                                  "traffic: calibrated HBM bytes of the pass's launches"),
        "field_generation_s": gen_s,
    }
    if args.no_cpu:
        for m in mom:
            m.free()
        return out
    log("c3: CPU baseline (oracle GMRES+GMG on a sample of the fields)")
    sample, tot = [], 0
    for q in range(nprob):
        if tot >= 3_000_000:
            break
        sample.append(q)
        tot += int(npts[q])
    fields = []
    for q in sample:
        d = dims[q]
        po = oracle.setup_context(**dict(cli0, Nx=d[0], Ny=d[1], Nz=d[2]))
        m = mom[q].download().reshape(d[2], d[1], d[0], 4)
        fields.append((po, m, oracle.full_rhs(po, m)))
    for m in mom:
        m.free()
    iters = []

    def one():
        t0 = time.perf_counter()
        for po, m, b in fields:
            iters.append(oracle.solve_gmres_mg(po, m, b, 1e-8)[1])
        return time.perf_counter() - t0

    el = median_of(one)
    # per-core: every host core solves its own copy of the sample's fields, all at once
    wall = replicas(lambda i: [oracle.solve_gmres_mg(po, m, b, 1e-8) for po, m, b in fields], cores)
    rate1 = tot / el
    ratec = cores * tot / wall
    mean_pts = float(npts.mean())
    out["cpu_baseline"] = baseline(
        rate1 / mean_pts, "solves/s",
        f"fields 0-{sample[-1]} of the draw ({tot} points, {len(sample)} fields), FP64 GMRES(30)+GMG oracle at rtol "
        f"1e-8, mean {np.mean(iters):.1f} iterations; solves/s projected to the batch's mean field of "
        f"{mean_pts:.0f} points from the sample's points/s",
        ratec / mean_pts, cores, {"points_per_s": rate1, "per_core_points_per_s": ratec})
    return out


# ----------------------------------------------------------------------------------- c4
def leg_c4(args, pic, _lib, cores):
    """16 independent PIC runs of uneven size, mixed precision (particles and solve, rtol
    1e-10 as in the 128^3 workload), stepped as one lockstep ensemble (one host thread: the
    tile passes and the solve phases one launch for all runs, each run's small launches on
    its own stream).  Also timed: the runs stepped concurrently
    on one stream and host thread each, and one by one.  CPU baseline: per run, one oracle
    step on 1/16 of its ions plus the full RHS + solve on the run's own final moments."""
    import oracle
    from ecsprt_amd import configs
    runs = configs.ensemble_runs()
    mixed = dict(particle_precision=_lib.PARTICLES_MIXED, solve_precision=_lib.SOLVE_MIXED)

    def make(r, extra, collide=_lib.COLLIDE_PHILOX, timing=0):
        cli = configs.ensemble_cli(r)
        c = pic.Context(pic.setup_context(**cli),
                        pic.default_options(ksp_rtol=1e-10, collide_mode=collide, timing=timing, **extra))
        c.init_particles(r["ppc"] * r["N"] ** 3, r["seed"], 1e-3, 2)
        c.collect()
        c.compute_potential()
        return c

    ctxs = [make(r, mixed, timing=1) for r in runs]
    nps = [r["ppc"] * r["N"] ** 3 for r in runs]
    total = sum(nps)
    nsteps = args.profile_steps or args.steps
    ens = pic.Ensemble(ctxs)
    for _ in range(0 if args.profile_steps else args.warmup):
        ens.advance(1.0)
    stages = np.zeros(4)
    sbytes_ens = 0.0
    t0 = time.perf_counter()
    for _ in range(nsteps):
        if args.sequential and args.profile_steps:
            for c in ctxs:
                c.step(1.0)
        else:
            ens.advance(1.0)
            if not args.profile_steps:
                stages += ens.stage_times() / nsteps
                sbytes_ens += ens.traffic() / nsteps
    el = time.perf_counter() - t0
    if args.profile_steps:
        ens.close()
        for c in ctxs:
            c.close()
        return {"profile_steps": nsteps, "seconds": el, "sequential": bool(args.sequential)}
    ens.close()
    # The runs stepped concurrently, one stream and host thread per run.
    nt = max(2, nsteps // 4)
    t0 = time.perf_counter()
    for _ in range(nt):
        pic.ensemble_step(ctxs, 1.0)
    for c in ctxs:
        c.synchronize()
    el_thr = (time.perf_counter() - t0) / nt
    # The same runs stepped one after another on this thread, for each kernel's own
    # duration (HIP events on the run's stream, nothing else running).
    kms = np.zeros(len(runs))
    sms = np.zeros(len(runs))
    sbytes = np.zeros(len(runs))
    nk = max(1, nsteps // 4)
    t0 = time.perf_counter()
    for _ in range(nk):
        for i, c in enumerate(ctxs):
            c.step(1.0)
            st = c.stage_times()
            kms[i] += st[0] / nk
            sms[i] += (st[4] + st[5]) / nk
            sbytes[i] += c.solver_traffic() / nk
    for c in ctxs:
        c.synchronize()
    el_seq = (time.perf_counter() - t0) / nk
    final_m = [c.get_moments() for c in ctxs]
    for c in ctxs:
        c.close()
    log("c4: per-run errors of one mixed step against FP64")
    errs = []
    for r in runs:
        a = make(r, {}, _lib.COLLIDE_OFF)
        b = make(r, mixed, _lib.COLLIDE_OFF)
        a.step(1.0)
        b.step(1.0)
        ma, mb = a.get_moments(), b.get_moments()
        fa, fb = a.get_potential(), b.get_potential()
        errs.append({"N": r["N"], "ppc": r["ppc"], "n_rel_max": rel_err(mb[..., 0], ma[..., 0]),
                     "G_rel_max": max(rel_err(mb[..., d], ma[..., d]) for d in (1, 2, 3)),
                     "phi_rel_max": rel_err(fb, fa)})
        a.close()
        b.close()
    alg = sum(BYTES_PER_PARTICLE_MIXED * n + BYTES_PER_NODE * r["N"] ** 3 for n, r in zip(nps, runs))
    out = {
        "metric": "aggregate particle-steps/s (16-run mixed-precision ensemble)",
        "value": total * nsteps / el, "unit": "particle-steps/s", "steps": nsteps,
        "ms_per_step": 1e3 * el / nsteps, "dtype": "f32 particles, f64 fold, mixed solve",
        "config": {"workload": "ensemble_16_runs", "runs": runs, "particles_total": total,
                   "stepping": "lockstep ensemble: one host thread; tile passes and solve phases one launch "
                               "for all runs, each run's small launches on its own stream"},
        "stage_ms_mean": {"tile_passes": stages[0], "pre": stages[1], "rebin_fold": stages[2],
                          "rhs_and_solve": stages[3]},
        "threads_ms_per_step": 1e3 * el_thr, "sequential_ms_per_step": 1e3 * el_seq,
        "speedup_vs_threads": el_thr / (el / nsteps), "speedup_vs_sequential": el_seq / (el / nsteps),
        "roofline": roofline(alg, float(stages[0]), "k_tile_mx_runs (one launch per tile shape: 4x4x4 and 8x8x8)",
                             committed_traffic("k_tile_mx_runs", "ensemble_16_runs", "all", args),
                             note="algorithmic bytes of all runs' ions and nodes / the span of the step's tile-pass "
                                  "launches (HIP events); traffic: calibrated HBM bytes of both launches"),
        "sequential_roofline": roofline(alg, float(kms.sum()), "k_tile_mx (16 launches per step, one per run)",
                                        note="the runs stepped one by one: sum of the runs' k_tile_mx durations"),
        "solve_roofline": roofline(sbytes_ens, float(stages[3]),
                                   "lockstep solve of the 16 runs (forcing, GMRES/GMG, all launches)",
                                   note="algorithmic bytes of the lockstep solve / its span (forcing included in "
                                        "the span, not in the bytes)"),
        "sequential_solve_roofline": roofline(float(sbytes.sum()), float(sms.sum()),
                                              "potential solves, mixed (16 per step, sequential pass, summed spans)"),
        "errors_vs_fp64_one_step": errs,
        "phi_rel_err_vs_fp64_max": max(e["phi_rel_max"] for e in errs),
    }
    if args.no_cpu:
        return out
    log("c4: CPU baseline (oracle on a sample of every run)")
    frac = 16
    work = []
    for i, r in enumerate(runs):
        po = oracle.setup_context(**configs.ensemble_cli(r))
        rng = np.random.default_rng(50 + i)
        pos, vel = uniform_ions(po, nps[i] // frac, rng)
        phi = 1e-2 * rng.normal(size=(po.Nz, po.Ny, po.Nx))
        m = final_m[i]
        work.append((po, pos, vel, phi, m, oracle.full_rhs(po, m)))

    def run_one(w):
        po, pos, vel, phi, m, b = w
        pos, vel = pos.copy(), vel.copy()
        a = time.perf_counter()
        oracle.particle_steps(po, 1.0, 1, pos, vel, phi, oracle.NR())
        c = time.perf_counter()
        oracle.solve_gmres_mg(po, m, b, 1e-10)
        return c - a, time.perf_counter() - c

    parts = []

    def one():
        t = np.array([run_one(w) for w in work])
        parts.append(t.sum(axis=0))
        return float(frac * t[:, 0].sum() + t[:, 1].sum())

    el_cpu = median_of(one)
    # per-core: the runs spread over the host cores (one run per core with 16 cores), all at
    # once; each core's projected step time is its runs' particle time x frac plus solves
    per_core = [0.0] * cores

    def core(i):
        for w in work[i::cores]:
            tp, ts = run_one(w)
            per_core[i] += frac * tp + ts

    replicas(core, cores)
    pt = np.median(np.array(parts), axis=0)
    out["cpu_baseline"] = baseline(
        total / el_cpu, "particle-steps/s",
        f"per run: 1 oracle step of push+collide(NR)+drift/wrap+deposit on 1/{frac} of its ions (time x{frac}) plus "
        f"FP64 GMRES+GMG at rtol 1e-10 on the run's final moments (RHS excluded); one-core sample s (particles, "
        f"solves): {pt[0]:.2f},{pt[1]:.2f}",
        total / max(per_core), cores)
    out["cpu_baseline"]["per_core_replicas"]["note"] = (
        "the 16 runs spread over the host cores and run at once; value from the slowest core's projected step")
    return out


# ----------------------------------------------------------------------------------- 2-D
# Algorithmic bytes of the two-dimensional particle stages, as for the three-dimensional
# step: per ion its position and velocity read and written once (two 16-byte vectors,
# 64 B); per node the potential read (8 B), the fold's three 8-byte sums and 32-byte
# overflow moments read and 32 bytes of moments written.
BYTES_PER_PARTICLE_2D = 64
BYTES_PER_NODE_2D = 8 + 24 + 32 + 32


def cpu_baseline_2d(p2, target_s=15.0):
    """Oracle (1 core) on a bounded sample of the 2-D workload: push, drift/wrap and
    deposit on a subset of its ions (collisions excluded: their NR-stream cost depends
    on Nc, not on the ions)."""
    import oracle
    rng = np.random.default_rng(12)
    phi = 1e-2 * rng.normal(size=(p2.Ny, p2.Nx))

    def run(n, steps):
        pos = np.column_stack([rng.uniform(0, p2.Lx, n), rng.uniform(0, p2.Ly, n)])
        vel = rng.normal(size=(n, 2)) * np.array([p2.viTx, p2.viTy])
        t0 = time.perf_counter()
        for _ in range(steps):
            oracle.boris2d(p2, 1.0, pos, vel, phi)
            oracle.update_positions2d(1.0, pos, vel)
            oracle.apply_bc2d(p2, pos)
            oracle.collect_moments2d(p2, pos, vel)
        return time.perf_counter() - t0

    n = 500_000
    el = run(n, 1)
    n2 = int(min(16_000_000, max(n, n * target_s / max(el, 1e-3) / 2)))
    el = run(n2, 2)
    return {"value": n2 * 2 / el, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": (f"{n2} ions uniform in the {p2.Nx}x{p2.Ny} grid, 2 steps of push+drift/wrap+deposit, "
                       "oracle -O3 on 1 host core")}


def bench_2d(args, pic, _lib):
    """The two-dimensional run: 1024^2 cells, 100 ions per cell, FP64, Philox
    collisions: BorisMoverBz2D, ScatterElastic and UpdatePositions in one pass over
    the tile-ordered ions, Collect2DFluidMoments (binned by tile, which re-orders the
    ions), ComputeFullRHS2D + GMRES/GMG on ComputeFullLHS2D."""
    import oracle
    from ecsprt_amd import configs
    N = args.grid or 1024
    ppc = args.ppc or 100
    cli = configs.base_cli(N, nui=0.02, Tn_over_Ti=1.0, E0=0.05)
    cli.update(Nz=1, dz=1.0)
    p = pic.setup_context(**dict(cli, ndim=2))
    cmode = _lib.COLLIDE_PHILOX if args.collide == "philox" else _lib.COLLIDE_OFF
    opts = pic.default_options(collide_mode=cmode, ksp_rtol=1e-10, timing=1)
    n = ppc * N * N
    ctx = pic.Context(p, opts)
    ctx.init_particles(n, 2024, 1e-3, 4)
    ctx.collect()
    ctx.compute_potential()
    for _ in range(args.warmup):
        ctx.step(1.0)
    ctx.synchronize()
    st_all, iters = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.step(1.0)
        st_all.append(ctx.stage_times())
        iters.append(ctx.solver_stats()[0])
    ctx.synchronize()
    el = time.perf_counter() - t0
    st = np.mean(np.array(st_all), axis=0)
    Nc, Ns, Nf, natt = ctx.collision_stats()
    ctx.close()
    alg = BYTES_PER_PARTICLE_2D * n + BYTES_PER_NODE_2D * N * N
    out = {
        "metric": "particle-steps/s (2-D: push+collide+drift+deposit+solve)",
        "value": n * args.steps / el, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (device Philox initial conditions)",
        "config": {"workload": f"fp64_2d_{N}sq_{ppc}ppc", "grid": [N, N], "ndim": 2, "ions_per_cell": ppc,
                   "np": n, "dt": 1.0, "Omega_i_dt": p.Omega_i, "nu_dt": p.nui, "E0": p.E0, "kappa_e": p.kappa_e,
                   "ksp_rtol": 1e-10, "collisions": "Philox statistical mode" if cmode else "off"},
        "phi_solve_ms": float(st[4] + st[5]), "ksp_iterations_mean": float(np.mean(iters)),
        "collisions_last_step": {"Nc": Nc, "successes": Ns, "refused": Nf, "attempts": natt},
        "stage_ms_mean": {"particle_stages": float(st[0]), "rhs": float(st[4]), "solve": float(st[5])},
        "roofline": roofline(alg, float(st[0]), "k_push2d_tiles+k_bin_hist2d+k_bin_scatter2d+k_deposit2d_tiles+k_fold2d"),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_2d(oracle.setup_context(**cli))
    print(json.dumps(out))


# ----------------------------------------------------------------------------------- main
def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is a design "
                                   "note, not an implemented path)"}))
        sys.exit(2)
    from ecsprt_amd import pic, _lib, configs
    cores = host_cores(args)
    if args.copy_probe:
        print(json.dumps(copy_probe(args.copy_probe)))
        return
    if args.workload == "batch":
        out = leg_c3(args, pic, _lib, cores)
        print(json.dumps(out))
        return
    if args.workload == "ensemble":
        print(json.dumps(leg_c4(args, pic, _lib, cores)))
        return
    if args.workload == "2d":
        bench_2d(args, pic, _lib)
        return
    if args.workload == "cpu32":
        print(json.dumps(leg_c0(args, pic, _lib, cores)))
        return
    wl = configs.PROD128
    if args.grid:
        wl.cli.update(Nx=args.grid, Ny=args.grid, Nz=args.grid)
    if args.ppc:
        wl.ppc = args.ppc
    p = pic.setup_context(**wl.cli)
    cmode = _lib.COLLIDE_PHILOX if args.collide == "philox" else _lib.COLLIDE_OFF
    common = dict(collide_mode=cmode, ksp_rtol=wl.rtol, timing=1, tile=args.tile, tile_z=args.tile_z,
                  block=args.block, split=args.split)
    for key in ("mg_pre", "inner_rtol", "mg_post", "mg_sweep"):
        if getattr(args, key):
            common[key] = getattr(args, key)
    mixed_only = args.workload == "mixed"
    if mixed_only:
        common.update(particle_precision=_lib.PARTICLES_MIXED, solve_precision=_lib.SOLVE_MIXED)
    tuned = {}
    if args.autotune:
        from ecsprt_amd import autotune
        for name, extra in (("fp64", {}), ("mixed", dict(particle_precision=_lib.PARTICLES_MIXED,
                                                          solve_precision=_lib.SOLVE_MIXED))):
            if (name == "fp64" and mixed_only) or (name == "mixed" and args.no_mixed and not mixed_only):
                continue
            o, table = autotune.tune_particles(p, pic.default_options(**dict(common, **extra)), wl.np, amp=wl.amp,
                                               k=wl.k, dt=wl.dt)
            tuned[name] = {"choice": [o.tile, o.tile_z, o.block], "kernel_ms": autotune.describe(table)}
        key = "mixed" if mixed_only else "fp64"
        common.update(tile=tuned[key]["choice"][0], tile_z=tuned[key]["choice"][1], block=tuned[key]["choice"][2])
    nsteps = args.profile_steps or args.steps
    warm = 0 if args.profile_steps else args.warmup
    log(f"{wl.name}{'_mixed' if mixed_only else ''}: {nsteps} steps")
    ctx, r = run_timed(p, pic.default_options(**common), wl, nsteps, warm, pic)
    if args.profile_steps:
        print(json.dumps({"profile_steps": nsteps, "seconds": r["seconds"]}))
        ctx.close()
        return
    Nc, Ns, Nf, natt = ctx.collision_stats()
    final_m = ctx.get_moments()
    ctx.close()
    npart, ncell = wl.np, wl.ncell
    bpp = BYTES_PER_PARTICLE_MIXED if mixed_only else BYTES_PER_PARTICLE
    alg_bytes = bpp * npart + BYTES_PER_NODE * ncell
    particle_ms = r["k_ms"] + r["pre_ms"] + r["rebin_ms"]
    kernel = "k_tile_mx" if mixed_only else "k_tile"
    launch = launch_of(p, common, mixed_only)
    solve_name = "mixed" if mixed_only else "fp64"
    out = {
        "metric": "particle-steps/s (push+deposit+collide); phi-solve ms; mixed vs FP64 @1e-6 err",
        "value": npart * nsteps / r["seconds"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": nsteps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["seconds"] / nsteps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 particles, f64 fold, mixed solve" if mixed_only else "f64",
        "data": "synthetic (device Philox initial conditions)",
        "config": {"workload": wl.name + ("_mixed" if mixed_only else ""), "grid": [p.Nx, p.Ny, p.Nz],
                   "ions_per_cell": wl.ppc, "np": npart, "dt": wl.dt, "Omega_i_dt": p.Omega_i * wl.dt,
                   "nu_dt": p.nui * wl.dt, "E0": p.E0, "kappa_e": p.kappa_e, "ksp_rtol": wl.rtol,
                   "collisions": "Philox statistical mode" if args.collide == "philox" else "off",
                   "operator": "frozen after the first solve", "initial_guess": "zero each solve"},
        "particle_steps_per_s_particle_stages": npart / (particle_ms * 1e-3),
        "phi_solve_ms": r["solve_total_ms"],
        "ksp_iterations_mean": r["iters"],
        "collisions_last_step": {"Nc": Nc, "successes": Ns, "refused": Nf, "attempts": natt},
        "stage_ms_mean": {"fused_particle_kernel": r["k_ms"], "pre": r["pre_ms"], "rebin_fold": r["rebin_ms"],
                          "rhs": r["rhs_ms"], "solve": r["solve_ms"]},
        "roofline": roofline(alg_bytes, r["k_ms"], kernel + "<push|collide|drift|deposit|rebin>",
                             committed_traffic(kernel, wl.name, launch, args), launch=launch),
        "solve": {solve_name: {"ms": r["solve_total_ms"], "iterations": r["iters"],
                               "roofline": roofline(r["solve_bytes"], r["solve_total_ms"],
                                                    f"potential solve (rhs + GMRES/GMG, {solve_name}, all launches)",
                                                    committed_traffic("solve_" + solve_name, wl.name, "all", args),
                                                    note="algorithmic bytes of every launch of one solve / the "
                                                         "solve's span between HIP events; traffic: calibrated HBM "
                                                         "bytes of all the solve's launches")}},
    }
    if not args.no_mixed and not mixed_only:
        # The same workload on the mixed-precision path (FP32 particles, mixed solve).
        log("mixed leg")
        mcommon = dict(common, particle_precision=_lib.PARTICLES_MIXED, solve_precision=_lib.SOLVE_MIXED)
        if "mixed" in tuned:
            mcommon.update(tile=tuned["mixed"]["choice"][0], tile_z=tuned["mixed"]["choice"][1],
                           block=tuned["mixed"]["choice"][2])
        cm, rm = run_timed(p, pic.default_options(**mcommon), wl, nsteps, warm, pic)
        cm.close()
        malg = BYTES_PER_PARTICLE_MIXED * npart + BYTES_PER_NODE * ncell
        mvalue = npart * nsteps / rm["seconds"]
        mlaunch = launch_of(p, mcommon, True)
        log("mixed errors against FP64")
        out["mixed"] = {
            "value": mvalue, "unit": "particle-steps/s", "ms_per_step": 1e3 * rm["seconds"] / nsteps,
            "speedup_vs_fp64": mvalue / out["value"],
            "fused_particle_kernel_ms": rm["k_ms"],
            "particle_kernel_speedup_vs_fp64": r["k_ms"] / rm["k_ms"],
            "phi_solve_ms": rm["solve_total_ms"],
            "solve_speedup_vs_fp64": r["solve_total_ms"] / max(rm["solve_total_ms"], 1e-9),
            "ksp_iterations_mean": rm["iters"],
            "roofline": roofline(malg, rm["k_ms"], "k_tile_mx",
                                 committed_traffic("k_tile_mx", wl.name, mlaunch, args), launch=mlaunch),
            "errors_vs_fp64": mixed_errors(p, wl, pic, _lib),
        }
        out["solve"]["mixed"] = {"ms": rm["solve_total_ms"], "iterations": rm["iters"],
                                 "speedup_vs_fp64": r["solve_total_ms"] / max(rm["solve_total_ms"], 1e-9),
                                 "roofline": roofline(rm["solve_bytes"], rm["solve_total_ms"],
                                                      "potential solve (rhs + GMRES/GMG, mixed, all launches)",
                                                      committed_traffic("solve_mixed", wl.name, "all", args))}
    if tuned:
        out["autotune"] = tuned
    if not args.no_cpu:
        log("CPU baselines: particles and solve of the 128^3 workload")
        out["cpu_baseline"] = cpu_particles(wl, cores)
        out["solve"]["cpu_baseline"] = cpu_solve(wl.cli, final_m, wl.rtol, cores)
    if args.workload == "all":
        legs = {}
        for name in [s for s in args.legs.split(",") if s]:
            log(f"leg {name}")
            fn = {"c0": leg_c0, "c3": leg_c3, "c4": leg_c4}[name]
            legs[name] = fn(args, pic, _lib, cores)
        out["legs"] = legs
    log("done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
