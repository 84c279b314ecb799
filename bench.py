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
BYTES_PER_NODE = 40      # read phi (8 B), write n, Gx, Gy, Gz (32 B)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c1", choices=["c0", "c1"])
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
    ap.add_argument("--mg-post", type=int, default=0)
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only run a device-to-device copy of this many GiB (HBM counter calibration)")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="run only this many steps without the extras (for rocprofv3)")
    return ap.parse_args()


def workload(args):
    from ecsprt_amd import configs
    wl = configs.C1 if args.workload == "c1" else configs.C0
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
    wl = workload(args)
    p = pic.setup_context(**wl.cli)
    cmode = _lib.COLLIDE_PHILOX if args.collide == "philox" else _lib.COLLIDE_OFF
    opts = pic.default_options(collide_mode=cmode, ksp_rtol=wl.rtol, timing=1,
                               tile=args.tile, tile_z=args.tile_z, block=args.block, split=args.split)
    if args.mg_pre:
        opts.mg_pre = args.mg_pre
    if args.mg_post:
        opts.mg_post = args.mg_post
    ctx = pic.Context(p, opts)
    npart = wl.np
    ncell = wl.ncell
    ctx.init_particles(npart, 2024, wl.amp, wl.k)
    # t = 0: collect and the first solve (src/pic.c:211-215), which builds the operator.
    ctx.collect()
    ctx.compute_potential()
    nsteps = args.profile_steps or args.steps
    warm = 0 if args.profile_steps else args.warmup
    for _ in range(warm):
        ctx.step(wl.dt)
    ctx.synchronize()
    k_ms, pre_ms, rebin_ms, rhs_ms, solve_ms, iters = [], [], [], [], [], []
    t0 = time.perf_counter()
    for _ in range(nsteps):
        ctx.step(wl.dt)
        st = ctx.stage_times()
        k_ms.append(st[0])
        pre_ms.append(st[1])
        rebin_ms.append(st[2])
        rhs_ms.append(st[4])
        solve_ms.append(st[5])
        iters.append(ctx.solver_stats()[0])
    ctx.synchronize()
    el = time.perf_counter() - t0
    if args.profile_steps:
        print(json.dumps({"profile_steps": nsteps, "seconds": el}))
        ctx.close()
        return
    Nc, Ns, Nf, natt = ctx.collision_stats()
    ms_step = 1e3 * el / nsteps
    value = npart * nsteps / el
    kmean = float(np.mean(k_ms))
    alg_bytes = BYTES_PER_PARTICLE * npart + BYTES_PER_NODE * ncell
    achieved = alg_bytes / (kmean * 1e-3) / 1e9
    particle_ms = kmean + float(np.mean(pre_ms)) + float(np.mean(rebin_ms))
    out = {
        "metric": "particle-steps/s (push+deposit+collide); phi-solve ms; mixed vs FP64 @1e-6 err",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": nsteps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (device Philox initial conditions)",
        "config": {"workload": wl.name, "grid": [p.Nx, p.Ny, p.Nz], "ions_per_cell": wl.ppc, "np": npart,
                   "dt": wl.dt, "Omega_i_dt": p.Omega_i * wl.dt, "nu_dt": p.nui * wl.dt, "E0": p.E0,
                   "kappa_e": p.kappa_e, "ksp_rtol": wl.rtol, "collisions": "Philox statistical mode" if args.collide == "philox" else "off",
                   "operator": "frozen after the first solve", "initial_guess": "zero each solve"},
        "particle_steps_per_s_particle_stages": npart / (particle_ms * 1e-3),
        "phi_solve_ms": float(np.mean(solve_ms)) + float(np.mean(rhs_ms)),
        "ksp_iterations_mean": float(np.mean(iters)),
        "collisions_last_step": {"Nc": Nc, "successes": Ns, "refused": Nf, "attempts": natt},
        "stage_ms_mean": {"fused_particle_kernel": kmean, "pre": float(np.mean(pre_ms)),
                          "rebin_fold": float(np.mean(rebin_ms)), "rhs": float(np.mean(rhs_ms)),
                          "solve": float(np.mean(solve_ms))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_tile<push|collide|drift|deposit|rebin>",
                     "algorithmic_bytes_per_launch": alg_bytes},
    }
    if not args.no_mixed:
        m = ctx.get_moments()
        phi64 = ctx.get_potential()
        optm = pic.default_options(collide_mode=_lib.COLLIDE_OFF, ksp_rtol=wl.rtol, timing=1,
                                   solve_precision=_lib.SOLVE_MIXED)
        cm = pic.Context(p, optm)
        cm.set_moments(m)
        cm.lhs()
        cm.compute_potential()  # warm
        cm.compute_potential()
        st = cm.stage_times()
        phim = cm.get_potential()
        # FP64 reference solve on the same moments (the timed context froze its operator at t=0)
        c64 = pic.Context(p, pic.default_options(collide_mode=_lib.COLLIDE_OFF, ksp_rtol=wl.rtol, timing=1))
        c64.set_moments(m)
        c64.compute_potential()
        c64.compute_potential()
        st64 = c64.stage_times()
        phi64b = c64.get_potential()
        out["mixed_solve"] = {"phi_rel_err_vs_fp64": float(np.abs(phim - phi64b).max() / np.abs(phi64b).max()),
                              "mixed_solve_ms": float(st[4] + st[5]), "fp64_solve_ms": float(st64[4] + st64[5]),
                              "speedup": float((st64[4] + st64[5]) / max(st[4] + st[5], 1e-9)),
                              "iterations": cm.solver_stats()[0], "fp64_iterations": c64.solver_stats()[0]}
        del phi64
        cm.close()
        c64.close()
    ctx.close()
    if not args.no_cpu:
        cb = cpu_baseline(wl)
        cb.pop("rc", None)
        out["cpu_baseline"] = cb
    print(json.dumps(out))


if __name__ == "__main__":
    main()
