"""Mixed-precision particle path (int32 cell + FP32 offset/velocity, FP32 push, drift
and deposit, fixed-point LDS accumulation) against the FP64 oracle.

Bounds, each stated against the oracle's FP64 result:
- storage round trip: offsets and velocities carry one FP32 rounding (2^-24 relative);
  per-cell counts are exact;
- push: velocity error <= 2e-6 of the largest speed (FP32 potential, FP32 Boris);
- drift: position error <= 2e-6 cells per step, cell counts equal up to ions within
  that distance of a face;
- deposit: moments within 2e-6 of their largest magnitude (each weight rounded once in
  FP32, sums exact in 64-bit fixed point).
"""
import numpy as np
import pytest

import oracle
from ecsprt_amd import configs, pic, _lib

pytestmark = pytest.mark.gpu

MIXED = dict(particle_precision=_lib.PARTICLES_MIXED)


def say(*a):
    print(*a, flush=True)


def make(wl, seed=1, **opt):
    p = pic.setup_context(**wl.cli)
    po = oracle.setup_context(**wl.cli)
    o = pic.default_options(**{**MIXED, **opt})
    ctx = pic.Context(p, o)
    rng = np.random.default_rng(seed)
    n = wl.np
    pos = np.column_stack([rng.uniform(0, po.Lx, n), rng.uniform(0, po.Ly, n), rng.uniform(0, po.Lz, n)])
    vel = rng.normal(size=(n, 3)) * np.array([po.viTx, po.viTy, po.viTz])
    return ctx, po, pos, vel


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def periodic_err(a, b, L):
    d = np.abs(a - b)
    return np.minimum(d, L - d).max()


def f32(a):
    return a.astype(np.float32).astype(np.float64)


def test_mixed_storage_round_trip():
    say("mixed round trip")
    wl = configs.small(12, 8)
    ctx, po, pos, vel = make(wl, tile=4)
    ctx.set_particles(pos, vel)
    gpos, gvel = ctx.get_particles()
    np.testing.assert_array_equal(gvel, f32(vel))
    cell = np.floor(pos)
    assert np.abs((gpos - cell) - (pos - cell)).max() <= 2.0 ** -24
    np.testing.assert_array_equal(ctx.cell_counts(), oracle.cell_counts(po, pos))
    ctx.close()


@pytest.mark.parametrize("N,ppc,tile", [(12, 8, 4), (16, 6, 8)])
def test_mixed_push(N, ppc, tile):
    say(f"mixed push N={N} tile={tile}")
    wl = configs.small(N, ppc)
    ctx, po, pos, vel = make(wl, tile=tile)
    rng = np.random.default_rng(2)
    phi = rng.normal(size=(N, N, N))
    ctx.set_particles(pos, vel)
    ctx.set_potential(phi)
    ctx.push(0.9)
    _, gvel = ctx.get_particles()
    ovel = vel.copy()
    oracle.boris(po, 0.9, pos, ovel, phi)
    err = rel(gvel, ovel)
    say("  velocity error / max speed =", err)
    assert err <= 2e-6
    # the push changed velocities by far more than the bound
    assert rel(vel, ovel) > 1e-2
    ctx.close()


@pytest.mark.parametrize("N,ppc,tile", [(12, 8, 4), (10, 5, 8)])
def test_mixed_drift_wrap_rebin(N, ppc, tile):
    say(f"mixed drift N={N} tile={tile}")
    wl = configs.small(N, ppc)
    ctx, po, pos, vel = make(wl, tile=tile)
    vel *= 2.5
    ctx.set_particles(pos, vel)
    vel = f32(vel)  # the oracle drifts with the stored FP32 velocities
    for step in range(3):
        ctx.update_positions(1.0)
        oracle.update_positions(1.0, pos, vel)
        oracle.apply_bc(po, pos)
        gpos, gvel = ctx.get_particles()
        np.testing.assert_array_equal(gvel, vel)
        err = periodic_err(gpos, pos, po.Lx)
        assert err <= 2e-6 * (step + 1), (step, err)
        gc, oc = ctx.cell_counts(), oracle.cell_counts(po, pos)
        near = np.sum(np.abs(pos - np.round(pos)) < 1e-5)
        assert np.abs(gc - oc).sum() <= 2 * near, (step, np.abs(gc - oc).sum(), near)
    assert ctx.num_particles == wl.np
    ctx.close()


@pytest.mark.parametrize("N,ppc,tile", [(12, 8, 4), (16, 6, 8), (9, 7, 4)])
def test_mixed_deposit(N, ppc, tile):
    say(f"mixed deposit N={N} tile={tile}")
    wl = configs.small(N, ppc)
    ctx, po, pos, vel = make(wl, tile=tile)
    ctx.set_particles(pos, vel)
    ctx.collect()
    gm = ctx.get_moments()
    om = oracle.collect_moments(po, pos, vel)
    for d in range(4):
        e = rel(gm[..., d], om[..., d])
        say(f"  moment {d} error {e:.2e}")
        assert e <= 2e-6, d
    assert gm[..., 0].sum() == pytest.approx(wl.np, rel=1e-8)
    ctx.close()


def test_mixed_deposit_is_order_independent():
    """Fixed-point accumulation: the same ions in another order give identical moments."""
    say("mixed deposit order independence")
    wl = configs.small(12, 8)
    ctx, po, pos, vel = make(wl, tile=4)
    ctx.set_particles(pos, vel)
    ctx.collect()
    m1 = ctx.get_moments()
    perm = np.random.default_rng(9).permutation(len(pos))
    ctx.set_particles(np.ascontiguousarray(pos[perm]), np.ascontiguousarray(vel[perm]))
    ctx.collect()
    np.testing.assert_array_equal(ctx.get_moments(), m1)
    ctx.close()


def test_mixed_spill_and_fast_ions():
    """Tiny tile slack (spill and rebuild) and a few ions faster than the fixed-point
    flux range (they deposit through the FP64 overflow path) over several fused steps."""
    say("mixed spills and fast ions")
    wl = configs.small(12, 6)
    ctx, po, pos, vel = make(wl, tile=4, slack=1e-9, collide_mode=_lib.COLLIDE_OFF,
                             operator_mode=_lib.OPERATOR_REBUILD, ksp_rtol=1e-8)
    dt = 0.05
    rng = np.random.default_rng(3)
    u = rng.normal(size=(5, 3))
    vel[:5] = 150.0 * u / np.linalg.norm(u, axis=1, keepdims=True)  # above vmax = 128, 7.5 cells per step
    phi = 0.05 * rng.normal(size=(12, 12, 12))
    ctx.set_particles(pos, vel)
    vel = f32(vel)
    for _ in range(3):
        ctx.set_potential(phi)
        ctx.step(dt)
        oracle.boris(po, dt, pos, vel, phi)
        oracle.update_positions(dt, pos, vel)
        oracle.apply_bc(po, pos)
    gpos, gvel = ctx.get_particles()
    assert rel(gvel, vel) <= 1e-5
    assert periodic_err(gpos, pos, po.Lx) <= 1e-5
    assert rel(ctx.get_moments(), oracle.collect_moments(po, pos, vel)) <= 1e-5
    assert ctx.num_particles == wl.np
    ctx.close()


@pytest.mark.parametrize("tile,tile_z,block", [(8, 8, 1024), (8, 4, 512), (4, 4, 256)])
def test_mixed_fused_step_vs_fp64(tile, tile_z, block):
    """One fused step (push, drift, deposit, solve) from the same state: mixed
    particles against the FP64 path, which matches the oracle to 1e-12."""
    say(f"mixed fused vs FP64 tile {tile}x{tile}x{tile_z}")
    wl = configs.small(16, 20)
    p = pic.setup_context(**wl.cli)
    common = dict(tile=tile, tile_z=tile_z, block=block, collide_mode=_lib.COLLIDE_OFF,
                  operator_mode=_lib.OPERATOR_REBUILD, ksp_rtol=1e-10)
    c64 = pic.Context(p, pic.default_options(**common))
    cmx = pic.Context(p, pic.default_options(**common, **MIXED))
    c64.init_particles(wl.np, 5, 0.1, 2)
    pos, vel = c64.get_particles()
    cmx.set_particles(pos, vel)
    for c in (c64, cmx):
        c.collect()
        c.compute_potential()
    errs = []
    for _ in range(2):
        c64.step(1.0)
        cmx.step(1.0)
        m64, mmx = c64.get_moments(), cmx.get_moments()
        e = [rel(mmx[..., d], m64[..., d]) for d in range(4)]
        ephi = rel(cmx.get_potential(), c64.get_potential())
        errs.append((e, ephi))
        say("  moments", ["%.1e" % x for x in e], "phi %.2e" % ephi)
    # From one state the first step's errors are rounding-level.  Afterwards ions that
    # FP32 rounding moved across a cell face see the reference's discontinuous gather;
    # the second step's moments stay close in the mean.
    assert max(errs[0][0]) <= 1e-5
    assert errs[0][1] <= 1e-5
    m64, mmx = c64.get_moments(), cmx.get_moments()
    assert np.sqrt(np.mean((mmx - m64) ** 2)) <= 1e-5 * np.sqrt(np.mean(m64 ** 2))
    c64.close()
    cmx.close()


def test_mixed_rejects_predrawn_collisions():
    wl = configs.small(8, 2)
    p = pic.setup_context(**wl.cli)
    with pytest.raises(pic.HipError) as e:
        pic.Context(p, pic.default_options(collide_mode=_lib.COLLIDE_PREDRAWN, **MIXED))
    assert e.value.code == _lib.ERR_ARG


def test_mixed_philox_collisions_match_fp64():
    """Philox collisions on mixed particles draw the same per-ion numbers as the FP64
    path, so from one state the two agree ion by ion except where an FP32 rounding
    flips an acceptance test."""
    say("mixed philox collisions")
    N, ppc = 16, 64
    cli = configs.base_cli(N, nui=0.1, Tn_over_Ti=1.0, B0=0.0)
    cli["vn0x"] = 0.7
    p = pic.setup_context(**cli)
    base = dict(collide_mode=_lib.COLLIDE_PHILOX, ksp_rtol=1e-6, operator_mode=_lib.OPERATOR_REBUILD)
    cmx = pic.Context(p, pic.default_options(**base, **MIXED))
    c64 = pic.Context(p, pic.default_options(**base))
    rng = np.random.default_rng(8)
    n = ppc * N ** 3
    pos = rng.uniform(0, N, (n, 3))
    vel = rng.normal(size=(n, 3))
    for c in (cmx, c64):
        c.set_particles(pos, vel)
        c.set_potential(np.zeros((N, N, N)))
        c.step(1.0)
    Nc, Ns, Nf, _ = cmx.collision_stats()
    Nc6, Ns6, _, _ = c64.collision_stats()
    say(f"  Nc={Nc} mixed successes={Ns} fp64 successes={Ns6}")
    assert Nc == Nc6 and Nf == 0
    assert abs(Ns - Ns6) <= 0.01 * Ns6 + 5
    _, vm = cmx.get_particles()
    _, v6 = c64.get_particles()
    moved = np.abs(v6 - vel).max(axis=1) > 1e-3
    assert moved.sum() > 0.5 * Ns6
    differ = np.abs(vm - v6).max(axis=1) > 1e-4
    say(f"  collided {moved.sum()}, differing ions {differ.sum()}")
    assert differ.sum() <= 0.01 * moved.sum() + 5
    np.testing.assert_allclose(vm.mean(axis=0), v6.mean(axis=0), atol=1e-3)
    cmx.close()
    c64.close()


@pytest.mark.parametrize("dx,dy,dz", [(0.5, 0.25, 2.0), (0.7, 1.3, 0.9)])
def test_mixed_stages_with_cell_widths(dx, dy, dz):
    """Mixed ions keep offsets in cell units, so non-unit and non-power-of-two widths
    enter only through the per-axis drift and field scalings: one push, drift and
    deposit against the oracle, to the bounds stated at the top of this file."""
    say(f"mixed cell widths {dx},{dy},{dz}")
    cli = configs.base_cli(12, E0=0.05)
    cli.update(dx=dx, dy=dy, dz=dz)
    p = pic.setup_context(**cli)
    po = oracle.setup_context(**cli)
    ctx = pic.Context(p, pic.default_options(tile=4, collide_mode=_lib.COLLIDE_OFF, **MIXED))
    rng = np.random.default_rng(22)
    n = 6 * 12 ** 3
    pos = np.column_stack([rng.uniform(0, po.Lx, n), rng.uniform(0, po.Ly, n), rng.uniform(0, po.Lz, n)])
    vel = rng.normal(size=(n, 3)) * np.array([dx, dy, dz])
    phi = 0.05 * rng.normal(size=(12, 12, 12))
    ctx.set_particles(pos, vel)
    gpos0, gvel0 = ctx.get_particles()  # the FP32-rounded state the GPU starts from
    ctx.set_potential(phi)
    ctx.push(1.0)
    ctx.update_positions(1.0)
    ctx.collect()
    opos, ovel = gpos0.copy(), gvel0.copy()
    oracle.boris(po, 1.0, opos, ovel, phi)
    oracle.update_positions(1.0, opos, ovel)
    oracle.apply_bc(po, opos)
    gpos, gvel = ctx.get_particles()
    assert rel(gvel, ovel) <= 2e-6
    d = np.abs(gpos - opos)
    L = np.array([po.Lx, po.Ly, po.Lz])
    d = np.minimum(d, L - d) / np.array([dx, dy, dz])
    assert d.max() <= 4e-6  # cells
    assert rel(ctx.get_moments(), oracle.collect_moments(po, gpos, gvel)) <= 2e-6
    assert ctx.num_particles == n
    ctx.close()
