"""GPU parity of the particle stages against the oracle (through the C ABI)."""
import numpy as np
import pytest

import oracle
from ecsprt_amd import configs, pic, _lib

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def make(wl, seed=1, **opt):
    p = pic.setup_context(**wl.cli)
    po = oracle.setup_context(**wl.cli)
    o = pic.default_options(**opt)
    ctx = pic.Context(p, o)
    rng = np.random.default_rng(seed)
    n = wl.np
    pos = np.column_stack([rng.uniform(0, po.Lx, n), rng.uniform(0, po.Ly, n), rng.uniform(0, po.Lz, n)])
    vel = rng.normal(size=(n, 3)) * np.array([po.viTx, po.viTy, po.viTz])
    return ctx, po, pos, vel


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("N,ppc,tile", [(12, 8, 4), (16, 6, 8), (20, 4, 8), (9, 10, 4)])
def test_push_parity(N, ppc, tile):
    say(f"push parity N={N} ppc={ppc} tile={tile}")
    wl = configs.small(N, ppc)
    ctx, po, pos, vel = make(wl, tile=tile)
    rng = np.random.default_rng(2)
    phi = rng.normal(size=(N, N, N))
    ctx.set_particles(pos, vel)
    ctx.set_potential(phi)
    ctx.push(0.9)
    gpos, gvel = ctx.get_particles()
    ovel = vel.copy()
    oracle.boris(po, 0.9, pos, ovel, phi)
    np.testing.assert_array_equal(gpos, pos)
    assert rel(gvel, ovel) <= 1e-12
    say("  max |dv| =", np.abs(gvel - ovel).max())
    ctx.close()


@pytest.mark.parametrize("N,ppc,tile", [(12, 8, 4), (16, 6, 8), (10, 5, 8)])
def test_drift_wrap_rebin_parity(N, ppc, tile):
    say(f"drift parity N={N} ppc={ppc} tile={tile}")
    wl = configs.small(N, ppc)
    ctx, po, pos, vel = make(wl, tile=tile)
    vel *= 2.5  # many particles leave their tile and wrap
    ctx.set_particles(pos, vel)
    for step in range(3):
        ctx.update_positions(1.0)
        oracle.update_positions(1.0, pos, vel)
        oracle.apply_bc(po, pos)
        gpos, gvel = ctx.get_particles()
        np.testing.assert_array_equal(gpos, pos)
        np.testing.assert_array_equal(gvel, vel)
        np.testing.assert_array_equal(ctx.cell_counts(), oracle.cell_counts(po, pos))
    assert ctx.num_particles == wl.np
    ctx.close()


@pytest.mark.parametrize("N,ppc,tile", [(12, 8, 4), (16, 6, 8), (9, 7, 4), (24, 2, 8)])
def test_deposit_parity(N, ppc, tile):
    say(f"deposit parity N={N} ppc={ppc} tile={tile}")
    wl = configs.small(N, ppc)
    ctx, po, pos, vel = make(wl, tile=tile)
    ctx.set_particles(pos, vel)
    ctx.collect()
    gm = ctx.get_moments()
    om = oracle.collect_moments(po, pos, vel)
    for d in range(4):
        assert rel(gm[..., d], om[..., d]) <= 1e-12, d
    assert gm[..., 0].sum() == pytest.approx(wl.np, rel=1e-12)
    ctx.close()


def test_deposit_edge_positions():
    say("deposit edge positions")
    wl = configs.small(8, 1)
    ctx, po, pos, vel = make(wl, tile=4)
    L = po.Lx
    edge = np.array([[0.0, 0.0, 0.0], [np.nextafter(L, 0), 3.0, 7.5], [4.0, np.nextafter(L, 0), 0.0],
                     [7.999, 7.999, 7.999], [4.0, 4.0, 4.0]])
    pos[:len(edge)] = edge
    ctx.set_particles(pos, vel)
    ctx.collect()
    om = oracle.collect_moments(po, pos, vel)
    gm = ctx.get_moments()
    assert rel(gm, om) <= 1e-12
    np.testing.assert_array_equal(ctx.cell_counts(), oracle.cell_counts(po, pos))
    ctx.close()


def test_stage_sequence_many_steps_with_spills():
    """Push, drift+rebin, deposit for several steps with tiny tile capacity slack so
    that tiles overflow and the spill/rebuild path runs."""
    say("stage sequence with spills")
    wl = configs.small(12, 6)
    ctx, po, pos, vel = make(wl, tile=4, slack=1e-9)
    rng = np.random.default_rng(3)
    phi = 0.05 * rng.normal(size=(12, 12, 12))
    ctx.set_particles(pos, vel)
    ctx.set_potential(phi)
    for step in range(4):
        ctx.push(1.0)
        ctx.update_positions(1.0)
        ctx.collect()
        oracle.boris(po, 1.0, pos, vel, phi)
        oracle.update_positions(1.0, pos, vel)
        oracle.apply_bc(po, pos)
    gpos, gvel = ctx.get_particles()
    assert rel(gpos, pos) <= 1e-12
    assert rel(gvel, vel) <= 1e-12
    om = oracle.collect_moments(po, pos, vel)
    assert rel(ctx.get_moments(), om) <= 1e-12
    ctx.close()


def test_fused_particle_pass_matches_stages():
    """ecsprt_hip_step with collisions off equals push, drift, wrap, deposit."""
    say("fused step vs stages")
    wl = configs.small(16, 6)
    ctx, po, pos, vel = make(wl, tile=8, collide_mode=_lib.COLLIDE_OFF, ksp_rtol=1e-10,
                             operator_mode=_lib.OPERATOR_REBUILD)
    rng = np.random.default_rng(4)
    phi = 0.05 * rng.normal(size=(16, 16, 16))
    phi -= phi.mean()
    ctx.set_particles(pos, vel)
    ctx.set_potential(phi)
    ctx.step(1.0)
    oracle.boris(po, 1.0, pos, vel, phi)
    oracle.update_positions(1.0, pos, vel)
    oracle.apply_bc(po, pos)
    om = oracle.collect_moments(po, pos, vel)
    gpos, gvel = ctx.get_particles()
    assert rel(gvel, vel) <= 1e-12
    assert rel(gpos, pos) <= 1e-12
    gm = ctx.get_moments()
    assert rel(gm, om) <= 1e-12
    # the solve that followed used those moments
    b = oracle.full_rhs(po, gm)
    x = oracle.solve_sparse(po, gm, b)
    gphi = ctx.get_potential()
    assert rel(gphi, x) <= 1e-7
    ctx.close()


def test_count_changed_is_an_error():
    say("count changed error")
    wl = configs.small(8, 2)
    ctx, po, pos, vel = make(wl, tile=4)
    ctx.set_particles(pos, vel)
    vel_fast = np.zeros_like(vel)
    vel_fast[0] = [3 * po.Lx, 0, 0]  # leaves the domain after one wrap
    ctx.set_particles(pos, vel_fast)
    with pytest.raises(pic.HipError) as e:
        ctx.update_positions(1.0)
    assert e.value.code == _lib.ERR_COUNT_CHANGED
    ctx.close()


def test_apply_bc_device_aos():
    say("apply_bc on device AoS")
    wl = configs.small(8, 1)
    ctx, po, pos, vel = make(wl, tile=4)
    L = po.Lx
    x = np.array([[L + 0.5, -0.5, 3.0], [-1e-300, L, 2 * L + 1], [np.nextafter(L, 0), 0.0, -0.25]])
    dv = pic.DeviceVector(x.size)
    dv.upload(x)
    ctx.apply_bc_device(dv.ptr, x.shape[0])
    got = dv.download().reshape(-1, 3)
    oracle.apply_bc(po, x)
    np.testing.assert_array_equal(got, x)
    dv.free()
    ctx.close()


def test_device_init_statistics():
    say("device initial conditions")
    wl = configs.small(16, 20)
    p = pic.setup_context(**wl.cli)
    ctx = pic.Context(p, pic.default_options())
    ctx.init_particles(wl.np, 7, 0.2, 2)
    pos, vel = ctx.get_particles()
    assert ctx.num_particles == wl.np
    assert pos.min() >= 0 and pos.max() < 16
    np.testing.assert_allclose(vel.mean(axis=0), 0, atol=0.02)
    np.testing.assert_allclose(vel.std(axis=0), 1, atol=0.02)
    # x density follows 1 + 0.2 sin(2 pi 2 x / L)
    h, _ = np.histogram(pos[:, 0], bins=32, range=(0, 16))
    xc = (np.arange(32) + 0.5) * 0.5
    expect = wl.np / 32 * (1 + 0.2 * np.sin(2 * np.pi * 2 * xc / 16))
    assert np.abs(h - expect).max() < 5 * np.sqrt(expect.max())
    ctx.close()


@pytest.mark.parametrize("dx,dy,dz", [(0.5, 0.25, 2.0), (0.7, 1.3, 0.9)])
def test_stages_with_cell_widths(dx, dy, dz):
    """Power-of-two widths take the multiply-by-reciprocal path, others divide; both
    must equal the reference's pos/dx arithmetic."""
    say(f"cell widths {dx},{dy},{dz}")
    cli = configs.base_cli(12, E0=0.05)
    cli.update(dx=dx, dy=dy, dz=dz)
    p = pic.setup_context(**cli)
    po = oracle.setup_context(**cli)
    ctx = pic.Context(p, pic.default_options(tile=4, collide_mode=_lib.COLLIDE_OFF))
    rng = np.random.default_rng(21)
    n = 6 * 12 ** 3
    pos = np.column_stack([rng.uniform(0, po.Lx, n), rng.uniform(0, po.Ly, n), rng.uniform(0, po.Lz, n)])
    vel = rng.normal(size=(n, 3)) * np.array([dx, dy, dz])  # about a cell per step
    phi = 0.05 * rng.normal(size=(12, 12, 12))
    ctx.set_particles(pos, vel)
    ctx.set_potential(phi)
    for _ in range(2):
        ctx.push(1.0)
        ctx.update_positions(1.0)
        oracle.boris(po, 1.0, pos, vel, phi)
        oracle.update_positions(1.0, pos, vel)
        oracle.apply_bc(po, pos)
    ctx.collect()
    gpos, gvel = ctx.get_particles()
    np.testing.assert_array_equal(gpos, pos)
    np.testing.assert_array_equal(gvel, vel)
    np.testing.assert_array_equal(ctx.cell_counts(), oracle.cell_counts(po, pos))
    assert rel(ctx.get_moments(), oracle.collect_moments(po, pos, vel)) <= 1e-12
    ctx.close()


@pytest.mark.parametrize("tile,tile_z,block", [(8, 4, 512), (4, 4, 128), (8, 8, 256)])
def test_fused_step_tile_shapes(tile, tile_z, block):
    """The fused pass for every tile shape and block size equals the stage sequence."""
    say(f"fused step tile {tile}x{tile}x{tile_z} block {block}")
    wl = configs.small(20, 5)
    ctx, po, pos, vel = make(wl, tile=tile, tile_z=tile_z, block=block, collide_mode=_lib.COLLIDE_OFF,
                             ksp_rtol=1e-8, operator_mode=_lib.OPERATOR_REBUILD)
    rng = np.random.default_rng(5)
    phi = 0.05 * rng.normal(size=(20, 20, 20))
    ctx.set_particles(pos, vel)
    for _ in range(3):
        ctx.set_potential(phi)
        ctx.step(1.0)
        oracle.boris(po, 1.0, pos, vel, phi)
        oracle.update_positions(1.0, pos, vel)
        oracle.apply_bc(po, pos)
    gpos, gvel = ctx.get_particles()
    np.testing.assert_array_equal(gpos, pos)
    np.testing.assert_array_equal(gvel, vel)
    assert rel(ctx.get_moments(), oracle.collect_moments(po, pos, vel)) <= 1e-12
    ctx.close()


@pytest.mark.parametrize("prec,tile,coll", [(_lib.PARTICLES_FP64, 8, False), (_lib.PARTICLES_FP64, 4, True),
                                            (_lib.PARTICLES_MIXED, 8, True), (_lib.PARTICLES_MIXED, 4, False)])
def test_fused_deposit_of_far_movers(prec, tile, coll):
    """Ions that move 3-6 cells in a step leave their source tile's deposit region
    (tile plus two cells); their deposit is deferred to their destination tile.  After
    each fused step the moments equal the oracle's deposit of the returned ions (FP64:
    1e-12; mixed: 2e-6 of the largest moment, FP32 weights), with tight tile slack so
    that spills mix with the deferred deposits."""
    say(f"far movers prec={prec} tile={tile} collisions={coll}")
    N, ppc = 16, 8
    cli = configs.base_cli(N, nui=0.1 if coll else 0.0, B0=0.02)
    po = oracle.setup_context(**cli)
    o = pic.default_options(particle_precision=prec, tile=tile, slack=0.02, ksp_rtol=1e-8,
                            operator_mode=_lib.OPERATOR_REBUILD,
                            collide_mode=_lib.COLLIDE_PHILOX if coll else _lib.COLLIDE_OFF)
    ctx = pic.Context(pic.setup_context(**cli), o)
    rng = np.random.default_rng(12)
    n = ppc * N ** 3
    pos = rng.uniform(0, N, (n, 3))
    vel = rng.normal(size=(n, 3))
    far = rng.random(n) < 0.3
    u = rng.normal(size=(far.sum(), 3))
    vel[far] = rng.uniform(3.0, 6.0, (far.sum(), 1)) * u / np.linalg.norm(u, axis=1, keepdims=True)
    ctx.set_particles(pos, vel)
    ctx.set_potential(np.zeros((N, N, N)))
    tol = 1e-12 if prec == _lib.PARTICLES_FP64 else 2e-6
    for step in range(3):
        ctx.step(1.0)
        gpos, gvel = ctx.get_particles()
        om = oracle.collect_moments(po, gpos, gvel)
        gm = ctx.get_moments()
        e = rel(gm, om)
        say(f"  step {step} moment error {e:.2e}")
        assert e <= tol
        # FP32 weights of one ion sum to 1 within a few FP32 roundings
        assert gm[..., 0].sum() == pytest.approx(n, rel=1e-10 if prec == _lib.PARTICLES_FP64 else 1e-6)
    assert ctx.num_particles == n
    if coll:
        assert ctx.collision_stats()[1] > 0
    ctx.close()


def test_fp64_deposit_is_order_independent():
    """FP64 deposit regions accumulate in exact fixed point: the same ions in another
    order give identical moments, bit for bit."""
    say("fp64 deposit order independence")
    wl = configs.small(12, 8)
    ctx, po, pos, vel = make(wl, tile=4)
    ctx.set_particles(pos, vel)
    ctx.collect()
    m1 = ctx.get_moments()
    perm = np.random.default_rng(9).permutation(len(pos))
    ctx.set_particles(np.ascontiguousarray(pos[perm]), np.ascontiguousarray(vel[perm]))
    ctx.collect()
    np.testing.assert_array_equal(ctx.get_moments(), m1)
    assert rel(m1, oracle.collect_moments(po, pos, vel)) <= 1e-12
    ctx.close()


def test_fp64_fast_ions_deposit():
    """Ions with a velocity component above the fixed-point flux range (vmax = 128 here)
    take the FP64 global-atomic path; fused steps still match the oracle to 1e-12."""
    say("fp64 fast ions")
    wl = configs.small(12, 6)
    ctx, po, pos, vel = make(wl, tile=4, collide_mode=_lib.COLLIDE_OFF, operator_mode=_lib.OPERATOR_REBUILD,
                             ksp_rtol=1e-8)
    dt = 0.05
    rng = np.random.default_rng(4)
    u = rng.normal(size=(7, 3))
    vel[:7] = 150.0 * u / np.linalg.norm(u, axis=1, keepdims=True)
    phi = 0.05 * rng.normal(size=(12, 12, 12))
    ctx.set_particles(pos, vel)
    for _ in range(3):
        ctx.set_potential(phi)
        ctx.step(dt)
        oracle.boris(po, dt, pos, vel, phi)
        oracle.update_positions(dt, pos, vel)
        oracle.apply_bc(po, pos)
    gpos, gvel = ctx.get_particles()
    assert rel(gvel, vel) <= 1e-12
    assert rel(gpos, pos) <= 1e-12
    assert rel(ctx.get_moments(), oracle.collect_moments(po, pos, vel)) <= 1e-12
    ctx.close()


def test_reference_initial_conditions():
    """ecsprt_hip_init_reference reproduces the reference's own initial particles: ran3
    positions times L and vT*gasdev + v0 velocities, both streams from seed -12345, as the
    oracle's restatement of the Numerical Recipes generators draws them."""
    say("reference initial conditions")
    cli = configs.base_cli(10, viT=1.3)
    cli.update(vi0x=0.2, vi0z=-0.1)
    p = pic.setup_context(**cli)
    po = oracle.setup_context(**cli)
    n = 4000
    ctx = pic.Context(p, pic.default_options(tile=4))
    ctx.init_reference(n)
    gpos, gvel = ctx.get_particles()
    s3, _ = oracle.NR().stream("ran3", -12345, 3 * n)
    sg, _ = oracle.NR().stream("gasdev", -12345, 3 * n)
    pos = s3.astype(np.float64).reshape(n, 3) * np.array([po.Lx, po.Ly, po.Lz])
    vel = np.array([po.viTx, po.viTy, po.viTz]) * sg.astype(np.float64).reshape(n, 3) + \
        np.array([po.vi0x, po.vi0y, po.vi0z])
    np.testing.assert_array_equal(gpos, pos)
    np.testing.assert_array_equal(gvel, vel)
    np.testing.assert_array_equal(ctx.cell_counts(), oracle.cell_counts(po, pos))
    ctx.close()


@pytest.mark.parametrize("dist", ["sobol", "sinusoidal"])
def test_reference_position_distributions(dist):
    """ecsprt_hip_init_distribution places ions as the reference's SobolDistribution and
    sinusoidal Rejection do (positions bit-equal to the oracle's), with the reference's
    gasdev velocities; the unimplemented PDistTypes are refused."""
    say(f"reference position distribution {dist}")
    cli = configs.base_cli(12, viT=0.7)
    cli.update(dx=0.5, dz=0.25)
    p = pic.setup_context(**cli)
    po = oracle.setup_context(**cli)
    n = 6000
    ctx = pic.Context(p, pic.default_options(tile=4))
    ctx.init_distribution(n, dist)
    gpos, gvel = ctx.get_particles()
    pos = oracle.sobol_positions(po, n) if dist == "sobol" else oracle.sinusoidal_positions(po, n)[0]
    sg, _ = oracle.NR().stream("gasdev", -12345, 3 * n)
    vel = np.array([po.viTx, po.viTy, po.viTz]) * sg.astype(np.float64).reshape(n, 3) + \
        np.array([po.vi0x, po.vi0y, po.vi0z])
    np.testing.assert_array_equal(gpos, pos)
    np.testing.assert_array_equal(gvel, vel)
    np.testing.assert_array_equal(ctx.cell_counts(), oracle.cell_counts(po, pos))
    for bad in (_lib.PDIST_REVERSE, _lib.PDIST_UNIFORM, _lib.PDIST_GAUSSIAN, 17):
        with pytest.raises(pic.HipError) as e:
            ctx.init_distribution(n, bad)
        assert e.value.code == _lib.ERR_ARG
    ctx.close()


@pytest.mark.parametrize("prec", [_lib.PARTICLES_FP64, _lib.PARTICLES_MIXED])
def test_many_fused_steps_conserve(prec):
    """Forty fused steps (collisions, heating, late deposits, re-binning) on a 64^3 grid:
    the ion count and the deposited charge stay exact, every id is present once, and the
    collision counts track the reference's Nc."""
    say(f"40 fused steps, precision {prec}")
    N, ppc = 64, 16
    cli = configs.base_cli(N, nui=0.02, Tn_over_Ti=1.0, E0=0.05)
    p = pic.setup_context(**cli)
    ctx = pic.Context(p, pic.default_options(particle_precision=prec, solve_precision=_lib.SOLVE_MIXED
                                             if prec == _lib.PARTICLES_MIXED else _lib.SOLVE_FP64,
                                             ksp_rtol=1e-8))
    n = ppc * N ** 3
    ctx.init_particles(n, 77, 1e-3, 4)
    ctx.collect()
    ctx.compute_potential()
    tol = 1e-10 if prec == _lib.PARTICLES_FP64 else 1e-6
    for step in range(40):
        ctx.step(1.0)
        if step % 10 == 9:
            m = ctx.get_moments()
            assert m[..., 0].sum() == pytest.approx(n, rel=tol)
            assert ctx.num_particles == n
            Nc, Ns, Nf, natt = ctx.collision_stats()
            assert Nf == 0 and abs(Ns - Nc) <= 0.1 * Nc + 50, (Nc, Ns)
            say(f"  step {step}: Nc {Nc} successes {Ns} attempts {natt}, solve iterations {ctx.solver_stats()[0]}")
    pos, vel = ctx.get_particles()
    assert np.isfinite(pos).all() and np.isfinite(vel).all()
    assert (pos >= 0).all() and (pos <= N).all()
    np.testing.assert_array_equal(ctx.cell_counts().sum(), n)
    ctx.close()


@pytest.mark.parametrize("prec", [_lib.PARTICLES_FP64, _lib.PARTICLES_MIXED])
def test_steps_independent_of_workgroups_per_tile(prec):
    """A tile's ions may be split over several workgroups (the default below 512 tiles).
    Each workgroup's fixed-point region is summed exactly by the fold and every draw is
    keyed by ion and step, so three fused steps with Philox collisions give the same ions,
    moments, potential and collision counts bit for bit whatever the split."""
    say(f"split independence precision={prec}")
    wl = configs.small(16, 12, nui=0.05)
    runs = []
    for split in (1, 3, 0):
        ctx, po, pos, vel = make(wl, seed=5, split=split, particle_precision=prec,
                                 collide_mode=_lib.COLLIDE_PHILOX, ksp_rtol=1e-10)
        ctx.set_particles(pos, vel)
        ctx.collect()
        ctx.compute_potential()
        for _ in range(3):
            ctx.step(wl.dt)
        gpos, gvel = ctx.get_particles()
        runs.append((gpos, gvel, ctx.get_moments(), ctx.get_potential(), ctx.collision_stats()))
        ctx.close()
    ref = runs[0]
    assert ref[4][1] > 0  # collisions happened
    for r in runs[1:]:
        for a, b in zip(r[:4], ref[:4]):
            np.testing.assert_array_equal(a, b)
        assert r[4] == ref[4]
