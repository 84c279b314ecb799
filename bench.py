"""Throughput of the hybrid-PIC timestep on one MI355X.

One step = one fused pass of push + collide + drift/wrap + deposit over every ion,
followed by the potential solve (RHS + GMRES/multigrid), i.e. src/pic.c:246-258.
Default workload: 128^3 periodic grid, 100 ions per cell (209,715,200 ions), FP64,
initial density 1 + 1e-3 sin(2 pi 4 x / L), uniform B along z with Omega_i dt = 0.05,
E0 along x, nu dt = 0.02, FP64 solve at rtol 1e-10.  Synthetic initial conditions are
generated on the device.

Prints one JSON line.  Extra fields: the potential-solve time, the particle-kernel
rate, the mixed-precision solve error against FP64, the HBM roofline of the fused
particle kernel (HIP events on the context's stream) and a CPU baseline from the
oracle timed on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
BYTES_PER_PARTICLE = 96  # FP64 SoA: read x,y,z,vx,vy,vz, write them back
BYTES_PER_PARTICLE_MIXED = 56  # int32 cell + 3 FP32 offsets + 3 FP32 velocities, read and written
BYTES_PER_NODE = 40      # read phi (8 B), write n, Gx, Gy, Gz (32 B)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="fp64", choices=["cpu32", "fp64", "mixed", "batch", "ensemble", "2d"],
                    help="cpu32 = 32^3 x 64 ions/cell FP64; fp64 = 128^3 x 100 ions/cell FP64 (with a mixed "
                         "leg); mixed = the 128^3 workload on the mixed-precision path only; batch = 256 "
                         "standalone solves; ensemble = 16 independent runs")
    ap.add_argument("--problems", type=int, default=256, help="batch: number of fields")
    ap.add_argument("--slots", type=int, default=8, help="batch: concurrent solver slots")
    ap.add_argument("--ppc", type=int, default=0, help="override ions per cell")
    ap.add_argument("--grid", type=int, default=0, help="override grid edge")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision solve check")
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
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="run only this many steps without the extras (for rocprofv3)")
    return ap.parse_args()


def workload(args):
    from ecsprt_amd import configs
    wl = configs.CPU32 if args.workload == "cpu32" else configs.PROD128
    if args.grid:
        wl.cli.update(Nx=args.grid, Ny=args.grid, Nz=args.grid)
    if args.ppc:
        wl.ppc = args.ppc
    return wl


def cpu_baseline(wl, target_s=15.0):
    """Oracle (1 core) on a bounded sample: the workload's grid and physics with a
    subset of its ions, push + collide + drift/wrap + deposit for a few steps."""
    import oracle
    p = oracle.setup_context(**wl.cli)
    rng = np.random.default_rng(11)
    n = 250_000
    steps = 1
    pos = np.column_stack([rng.uniform(0, p.Lx, n), rng.uniform(0, p.Ly, n), rng.uniform(0, p.Lz, n)])
    vel = rng.normal(size=(n, 3)) * np.array([p.viTx, p.viTy, p.viTz])
    phi = 1e-2 * rng.normal(size=(p.Nz, p.Ny, p.Nx))
    nr = oracle.NR()
    t0 = time.perf_counter()
    rc, _, t = oracle.particle_steps(p, wl.dt, steps, pos, vel, phi, nr)
    el = time.perf_counter() - t0
    # scale the sample so that the measured part takes about target_s
    n2 = int(min(8_000_000, max(n, n * target_s / max(el, 1e-3) / 2)))
    steps2 = 2
    pos = np.column_stack([rng.uniform(0, p.Lx, n2), rng.uniform(0, p.Ly, n2), rng.uniform(0, p.Lz, n2)])
    vel = rng.normal(size=(n2, 3)) * np.array([p.viTx, p.viTy, p.viTz])
    t0 = time.perf_counter()
    rc, _, t = oracle.particle_steps(p, wl.dt, steps2, pos, vel, phi, nr)
    el = time.perf_counter() - t0
    return {
        "value": n2 * steps2 / el,
        "unit": "particle-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"{n2} ions uniform in the {p.Nx}x{p.Ny}x{p.Nz} grid, {steps2} steps of push+collide"
                   f"(NR stream)+drift/wrap+deposit, oracle -O3 on 1 host core; stage s: "
                   + ",".join(f"{x:.2f}" for x in t)),
        "rc": rc,
    }


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


def run_timed(p, opts, wl, nsteps, warm, pic):
    """Initial conditions, t = 0 collect + solve (src/pic.c:211-215), warm-up, then
    nsteps timed steps.  Returns the per-stage means and the context."""
    ctx = pic.Context(p, opts)
    ctx.init_particles(wl.np, 2024, wl.amp, wl.k)
    ctx.collect()
    ctx.compute_potential()
    for _ in range(warm):
        ctx.step(wl.dt)
    ctx.synchronize()
    st_all, iters = [], []
    t0 = time.perf_counter()
    for _ in range(nsteps):
        ctx.step(wl.dt)
        st_all.append(ctx.stage_times())
        iters.append(ctx.solver_stats()[0])
    ctx.synchronize()
    el = time.perf_counter() - t0
    st = np.mean(np.array(st_all), axis=0)
    return ctx, {"seconds": el, "k_ms": float(st[0]), "pre_ms": float(st[1]), "rebin_ms": float(st[2]),
                 "rhs_ms": float(st[4]), "solve_ms": float(st[5]), "iters": float(np.mean(iters))}


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
    return {"per_step": per, "free_running": free}


def bench_batch(args, pic, _lib):
    """256 standalone solves of smoothed log-normal density fields, grid edges drawn from
    {32, 48, 64, 96, 128, 192} per axis (seed 3), mixed precision to an FP64 residual of
    1e-8 relative; the FP64 batch on the same fields is the reference."""
    from ecsprt_amd import configs
    nprob = args.problems
    dims = configs.batch_dims(nprob)
    p = pic.setup_context(**configs.base_cli(32, E0=0.05))
    t0 = time.perf_counter()
    mom = [pic.lognormal_moments(d, seed=1000 + q, sigma=0.3, smooth=3.0) for q, d in enumerate(dims)]
    gen_s = time.perf_counter() - t0
    npts = np.array([d[0] * d[1] * d[2] for d in dims], dtype=np.int64)
    res = {}
    for name, prec in (("mixed", _lib.SOLVE_MIXED), ("fp64", _lib.SOLVE_FP64)):
        phi = [pic.DeviceVector(int(n)) for n in npts]
        b = pic.Batch(p, pic.default_options(ksp_rtol=1e-8, solve_precision=prec), dims, slots=args.slots)
        b.solve(mom, phi)  # warm-up pass (first-use costs of the slots)
        t0 = time.perf_counter()
        b.solve(mom, phi)
        el = time.perf_counter() - t0
        it, rn = b.stats()
        res[name] = {"seconds": el, "iters": it, "phi": phi}
        b.close()
    errs = []
    for q in range(nprob):
        ref = res["fp64"]["phi"][q].download()
        got = res["mixed"]["phi"][q].download()
        errs.append(float(np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300)))
    for name in res:
        for f in res[name]["phi"]:
            f.free()
    for m in mom:
        m.free()
    ms, fs = res["mixed"]["seconds"], res["fp64"]["seconds"]
    out = {
        "metric": "batched potential solves/s (256 fields, mixed precision, 1e-8 FP64 residual)",
        "value": nprob / ms, "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": 1e3 * ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 inner / f64 refinement", "data": "synthetic (device log-normal fields)",
        "config": {"workload": "batched_solves_256_fields", "problems": nprob, "points_total": int(npts.sum()),
                   "points_max": int(npts.max()), "slots": args.slots, "ksp_rtol": 1e-8,
                   "density": "exp(0.3 g - 0.045), Gaussian smoothing sigma 3 cells, mean 1; flux = n"},
        "points_per_s": float(npts.sum() / ms),
        "fp64": {"seconds": fs, "solves_per_s": nprob / fs, "mean_iterations": float(res["fp64"]["iters"].mean())},
        "mixed": {"seconds": ms, "mean_iterations": float(res["mixed"]["iters"].mean()),
                  "speedup_vs_fp64": fs / ms, "phi_rel_err_vs_fp64_max": max(errs),
                  "phi_rel_err_vs_fp64_median": float(np.median(errs))},
        "field_generation_s": gen_s,
    }
    print(json.dumps(out))


def bench_ensemble(args, pic, _lib):
    """16 independent PIC runs of uneven size, mixed precision (particles and solve, rtol
    1e-10 as in the 128^3 workload), stepped concurrently (one stream and host thread per run)."""
    from ecsprt_amd import configs
    runs = configs.ensemble_runs()
    mixed = dict(particle_precision=_lib.PARTICLES_MIXED, solve_precision=_lib.SOLVE_MIXED)

    def make(r, extra, collide=_lib.COLLIDE_PHILOX):
        cli = configs.ensemble_cli(r)
        c = pic.Context(pic.setup_context(**cli), pic.default_options(ksp_rtol=1e-10, collide_mode=collide, **extra))
        c.init_particles(r["ppc"] * r["N"] ** 3, r["seed"], 1e-3, 2)
        c.collect()
        c.compute_potential()
        return c

    ctxs = [make(r, mixed) for r in runs]
    nps = [r["ppc"] * r["N"] ** 3 for r in runs]
    nsteps = args.profile_steps or args.steps
    for _ in range(0 if args.profile_steps else args.warmup):
        pic.ensemble_step(ctxs, 1.0)
    for c in ctxs:
        c.synchronize()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        pic.ensemble_step(ctxs, 1.0)
    for c in ctxs:
        c.synchronize()
    el = time.perf_counter() - t0
    if args.profile_steps:
        print(json.dumps({"profile_steps": nsteps, "seconds": el}))
        return
    # the same runs stepped one after another, for the concurrency gain
    t0 = time.perf_counter()
    for _ in range(max(1, nsteps // 4)):
        for c in ctxs:
            c.step(1.0)
    for c in ctxs:
        c.synchronize()
    el_seq = (time.perf_counter() - t0) / max(1, nsteps // 4)
    for c in ctxs:
        c.close()
    # per-run error of one mixed step against the FP64 path from the same state
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
    total = sum(nps)
    out = {
        "metric": "aggregate particle-steps/s (16-run mixed-precision ensemble)",
        "value": total * nsteps / el, "unit": "particle-steps/s", "n_gpus": 1, "steps": nsteps, "warmup": args.warmup,
        "ms_per_step": 1e3 * el / nsteps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 particles, f64 fold, mixed solve", "data": "synthetic (device Philox initial conditions)",
        "config": {"workload": "ensemble_16_runs", "runs": runs, "particles_total": total,
                   "concurrency": "one stream and host thread per run"},
        "sequential_ms_per_step": 1e3 * el_seq, "concurrency_speedup": el_seq / (el / nsteps),
        "errors_vs_fp64_one_step": errs,
    }
    print(json.dumps(out))


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
    achieved = alg / (st[0] * 1e-3) / 1e9
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
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_push2d_tiles+k_bin_hist2d+k_bin_scatter2d+k_deposit2d_tiles+k_fold2d",
                     "algorithmic_bytes_per_launch": alg},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_2d(oracle.setup_context(**cli))
    print(json.dumps(out))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is a design "
                                   "note, not an implemented path)"}))
        sys.exit(2)
    from ecsprt_amd import pic, _lib
    if args.copy_probe:
        print(json.dumps(copy_probe(args.copy_probe)))
        return
    if args.workload == "batch":
        bench_batch(args, pic, _lib)
        return
    if args.workload == "ensemble":
        bench_ensemble(args, pic, _lib)
        return
    if args.workload == "2d":
        bench_2d(args, pic, _lib)
        return
    wl = workload(args)
    p = pic.setup_context(**wl.cli)
    cmode = _lib.COLLIDE_PHILOX if args.collide == "philox" else _lib.COLLIDE_OFF
    common = dict(collide_mode=cmode, ksp_rtol=wl.rtol, timing=1, tile=args.tile, tile_z=args.tile_z,
                  block=args.block, split=args.split)
    if args.mg_pre:
        common["mg_pre"] = args.mg_pre
    if args.inner_rtol:
        common["inner_rtol"] = args.inner_rtol
    if args.mg_post:
        common["mg_post"] = args.mg_post
    if args.mg_sweep:
        common["mg_sweep"] = args.mg_sweep
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
    ctx, r = run_timed(p, pic.default_options(**common), wl, nsteps, warm, pic)
    if args.profile_steps:
        print(json.dumps({"profile_steps": nsteps, "seconds": r["seconds"]}))
        ctx.close()
        return
    Nc, Ns, Nf, natt = ctx.collision_stats()
    ctx.close()
    npart, ncell = wl.np, wl.ncell
    bpp = BYTES_PER_PARTICLE_MIXED if mixed_only else BYTES_PER_PARTICLE
    alg_bytes = bpp * npart + BYTES_PER_NODE * ncell
    achieved = alg_bytes / (r["k_ms"] * 1e-3) / 1e9
    particle_ms = r["k_ms"] + r["pre_ms"] + r["rebin_ms"]
    kernel = "k_tile_mx" if mixed_only else "k_tile"
    traffic = committed_traffic(kernel, wl.name, launch_of(p, common, mixed_only), args)
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
        "phi_solve_ms": r["solve_ms"] + r["rhs_ms"],
        "ksp_iterations_mean": r["iters"],
        "collisions_last_step": {"Nc": Nc, "successes": Ns, "refused": Nf, "attempts": natt},
        "stage_ms_mean": {"fused_particle_kernel": r["k_ms"], "pre": r["pre_ms"], "rebin_fold": r["rebin_ms"],
                          "rhs": r["rhs_ms"], "solve": r["solve_ms"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kernel + "<push|collide|drift|deposit|rebin>", "launch": launch_of(p, common, mixed_only),
                     "algorithmic_bytes_per_launch": alg_bytes},
    }
    if not args.no_mixed and not mixed_only:
        # The same workload on the mixed-precision path (FP32 particles, mixed solve).
        mcommon = dict(common, particle_precision=_lib.PARTICLES_MIXED, solve_precision=_lib.SOLVE_MIXED)
        if "mixed" in tuned:
            mcommon.update(tile=tuned["mixed"]["choice"][0], tile_z=tuned["mixed"]["choice"][1],
                           block=tuned["mixed"]["choice"][2])
        cm, rm = run_timed(p, pic.default_options(**mcommon), wl, nsteps, warm, pic)
        cm.close()
        malg = BYTES_PER_PARTICLE_MIXED * npart + BYTES_PER_NODE * ncell
        mach = malg / (rm["k_ms"] * 1e-3) / 1e9
        mvalue = npart * nsteps / rm["seconds"]
        out["mixed"] = {
            "value": mvalue, "unit": "particle-steps/s", "ms_per_step": 1e3 * rm["seconds"] / nsteps,
            "speedup_vs_fp64": mvalue / out["value"],
            "fused_particle_kernel_ms": rm["k_ms"],
            "particle_kernel_speedup_vs_fp64": r["k_ms"] / rm["k_ms"],
            "phi_solve_ms": rm["solve_ms"] + rm["rhs_ms"],
            "solve_speedup_vs_fp64": (r["solve_ms"] + r["rhs_ms"]) / max(rm["solve_ms"] + rm["rhs_ms"], 1e-9),
            "ksp_iterations_mean": rm["iters"],
            "roofline": {"bound": "hbm", "achieved": mach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": mach / HBM_PEAK_GBS,
                         "traffic": committed_traffic("k_tile_mx", wl.name, launch_of(p, mcommon, True), args),
                         "kernel": "k_tile_mx", "launch": launch_of(p, mcommon, True),
                         "algorithmic_bytes_per_launch": malg},
            "errors_vs_fp64": mixed_errors(p, wl, pic, _lib),
        }
    if tuned:
        out["autotune"] = tuned
    if not args.no_cpu:
        cb = cpu_baseline(wl)
        cb.pop("rc", None)
        out["cpu_baseline"] = cb
    print(json.dumps(out))


if __name__ == "__main__":
    main()
