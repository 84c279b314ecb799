"""Autotuner for the particle pass: tile shape and workgroup size per workload.

The fused tile kernel is compiled for three tile shapes (8x8x8, 8x8x4 and 4x4x4
cells) and runs with 128-1024 threads per workgroup.  Which one is fastest depends on
the ions per tile (occupancy against LDS and re-binning traffic), the particle
precision and whether collisions are on.  ``tune_particles`` times every valid
candidate on the workload itself, a few fused steps each, and returns the fastest
options; results are cached per workload key for the life of the process.
"""
import time

from . import _lib, pic

# (tile, tile_z, block): tile edge along x and y, along z, threads per workgroup.
CANDIDATES = [(8, 8, 1024), (8, 8, 512), (8, 4, 1024), (8, 4, 512), (4, 4, 512), (4, 4, 256)]

_cache = {}


def valid_candidates(params, candidates=CANDIDATES):
    """Candidates whose tiles fit the grid (at least one tile per axis of the tile edge)."""
    out = []
    for tile, tz, block in candidates:
        if params.Nx >= tile and params.Ny >= tile and params.Nz >= tz:
            out.append((tile, tz, block))
    return out


def workload_key(params, options, np_):
    return (params.Nx, params.Ny, params.Nz, int(np_), options.particle_precision, options.collide_mode)


def tune_particles(params, options, np_, seed=2024, amp=0.0, k=0, steps=3, candidates=CANDIDATES, dt=1.0):
    """Time `steps` fused steps (the particle kernel's HIP-event time) of each candidate
    on a fresh context with the workload's initial conditions.  Returns (best options,
    table of {candidate: mean kernel ms})."""
    key = workload_key(params, options, np_)
    if key in _cache:
        return _cache[key]
    table = {}
    for tile, tz, block in valid_candidates(params, candidates):
        o = pic.default_options()
        for name, _ in _lib.Options._fields_:
            setattr(o, name, getattr(options, name))
        o.tile, o.tile_z, o.block, o.timing = tile, tz, block, 1
        ctx = pic.Context(params, o)
        try:
            ctx.init_particles(np_, seed, amp, k)
            ctx.collect()
            ctx.compute_potential()
            ctx.step(dt)  # first-use costs
            ms = 0.0
            for _ in range(steps):
                ctx.step(dt)
                ms += ctx.stage_times()[0]
            table[(tile, tz, block)] = ms / steps
        finally:
            ctx.close()
    best = min(table, key=table.get)
    o = pic.default_options()
    for name, _ in _lib.Options._fields_:
        setattr(o, name, getattr(options, name))
    o.tile, o.tile_z, o.block = best
    _cache[key] = (o, table)
    return o, table


def describe(table):
    return {f"{t}x{t}x{tz}/{b}": round(ms, 3) for (t, tz, b), ms in sorted(table.items(), key=lambda kv: kv[1])}


if __name__ == "__main__":
    import argparse
    import json

    from . import configs

    ap = argparse.ArgumentParser(description="time the particle-pass candidates on a workload")
    ap.add_argument("--workload", default="fp64", choices=["cpu32", "fp64"])
    ap.add_argument("--mixed", action="store_true")
    ap.add_argument("--steps", type=int, default=3)
    a = ap.parse_args()
    wl = configs.CPU32 if a.workload == "cpu32" else configs.PROD128
    p = pic.setup_context(**wl.cli)
    opts = pic.default_options(ksp_rtol=wl.rtol)
    if a.mixed:
        opts.particle_precision = _lib.PARTICLES_MIXED
        opts.solve_precision = _lib.SOLVE_MIXED
    best, table = tune_particles(p, opts, wl.np, amp=wl.amp, k=wl.k, steps=a.steps, dt=wl.dt)
    print(json.dumps({"best": f"{best.tile}x{best.tile}x{best.tile_z}/{best.block}", "ms": describe(table)}))
