"""GPU parity of the two-dimensional run (cli.ndim = 2) against the oracle, through the
C ABI: BorisMoverBz2D, the drift and wrap, Collect2DFluidMoments, ComputeFullRHS2D and
ComputeFullLHS2D (MatMult and the GMRES/GMG solve), the two-dimensional ScatterElastic
and the reference's initial conditions in two dimensions."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle
from ecsprt_amd import configs, pic, _lib

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def cli2d(N=12, Ny=None, **kw):
    cli = configs.base_cli(N, **kw)
    cli.update(Ny=Ny or N, Nz=1, dz=1.0)
    return cli


def make(cli, n, seed=1, **opt):
    p = pic.setup_context(**dict(cli, ndim=2))
    po = oracle.setup_context(**cli)
    ctx = pic.Context(p, pic.default_options(**opt))
    rng = np.random.default_rng(seed)
    pos = np.column_stack([rng.uniform(0, po.Lx, n), rng.uniform(0, po.Ly, n)])
    vel = rng.normal(size=(n, 2)) * np.array([po.viTx, po.viTy])
    return ctx, po, pos, vel


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("N,Ny,dx,sort", [(12, 12, 1.0, False), (16, 10, 0.5, True), (9, 13, 0.3, True)])
def test_push2d_and_drift_bit_exact(N, Ny, dx, sort):
    say(f"2-D push/drift N={N} Ny={Ny} dx={dx} sort={sort}")
    cli = cli2d(N, Ny, B0=0.2)
    cli.update(dx=dx, dy=0.75)
    ctx, po, pos, vel = make(cli, 5000)
    phi = np.random.default_rng(2).normal(size=(po.Ny, po.Nx))
    # ions on the upper faces and at the last cell exercise the one-sided differences
    pos[:4] = [[0.0, 0.0], [po.Lx * (1 - 1e-16), po.Ly * 0.5], [po.Lx - 0.5 * po.dx, po.Ly - 0.25 * po.dy],
               [0.5 * po.dx, po.Ly * (1 - 1e-16)]]
    ctx.set_particles(pos, vel)
    if sort:
        ctx.collect()  # the deposit stores the ions in tile order; results are still by set index
    ctx.set_potential(phi)
    ctx.push(0.9)
    gpos, gvel = ctx.get_particles()
    v = vel.copy()
    oracle.boris2d(po, 0.9, pos, v, phi)
    np.testing.assert_array_equal(gvel, v)
    ctx.update_positions(0.25)  # |v| dt below the smallest L, as a wrap-once boundary needs
    gpos, gvel = ctx.get_particles()
    p2 = pos.copy()
    oracle.update_positions2d(0.25, p2, v)
    oracle.apply_bc2d(po, p2)
    np.testing.assert_array_equal(gpos, p2)
    ctx.close()


def test_collect2d_parity_and_order_independence():
    say("2-D collect")
    cli = cli2d(14, 11)
    cli.update(dx=0.5)
    ctx, po, pos, vel = make(cli, 20000, seed=3)
    vel[:3] *= 500.0  # fast ions take the FP64 overflow path
    ctx.set_particles(pos, vel)
    ctx.collect()
    m = ctx.get_moments()
    ref = oracle.collect_moments2d(po, pos, vel)
    assert m.shape == (po.Ny, po.Nx, 4)
    assert rel(m, ref) <= 1e-12
    assert (m[..., 3] == 0).all()
    assert m[..., 0].sum() == pytest.approx(len(pos), rel=1e-12)
    perm = np.random.default_rng(9).permutation(len(pos))
    ctx.set_particles(pos[perm], vel[perm])
    ctx.collect()
    m2 = ctx.get_moments()
    slow = np.ones(len(pos), bool)
    slow[:3] = False
    # fixed-point sums: bit-equal whatever the ion order, the three fast ions aside
    ctx.set_particles(pos[slow], vel[slow])
    ctx.collect()
    a = ctx.get_moments()
    ctx.set_particles(pos[slow][::-1].copy(), vel[slow][::-1].copy())
    ctx.collect()
    np.testing.assert_array_equal(ctx.get_moments(), a)
    assert rel(m2, ref) <= 1e-12
    ctx.close()


def test_rhs2d_matmult_and_solve():
    say("2-D rhs, matmult, solve")
    cli = cli2d(24, 20, kappa_e=4.0, E0=0.05)
    cli.update(dy=0.75)
    p = pic.setup_context(**dict(cli, ndim=2))
    po = oracle.setup_context(**cli)
    rng = np.random.default_rng(5)
    m = np.zeros((po.Ny, po.Nx, 4))
    m[..., 0] = 64 + 8 * rng.normal(size=(po.Ny, po.Nx))
    m[..., 1:3] = 8 * rng.normal(size=(po.Ny, po.Nx, 2))
    rtol = 1e-10
    ctx = pic.Context(p, pic.default_options(operator_mode=_lib.OPERATOR_REBUILD, ksp_rtol=rtol))
    ctx.set_moments(m)
    ctx.compute_potential()
    b = oracle.full_rhs2d(po, m)
    assert rel(ctx.get_forcing(), b) <= 1e-12
    x = rng.normal(size=(po.Ny, po.Nx))
    xd, yd = pic.DeviceVector(x.size), pic.DeviceVector(x.size)
    xd.upload(x)
    ctx.matmult(xd, yd)
    assert rel(yd.download().reshape(x.shape), oracle.lhs2d_apply(po, m, x)) <= 1e-12
    xd.free()
    yd.free()
    gphi = ctx.get_potential()
    r = b - oracle.lhs2d_apply(po, m, gphi)
    it, _, _ = ctx.solver_stats()
    say(f"  iterations {it}, ||r||/||b|| = {np.linalg.norm(r) / np.linalg.norm(b):.2e}")
    assert np.linalg.norm(r) <= 1.01 * rtol * np.linalg.norm(b)
    A = oracle.lhs2d_csr(po, m).tolil()
    A[0, :] = 0.0
    A[0, 0] = 1.0
    rhs = b.reshape(-1).copy()
    rhs[0] = 0.0
    xs = spla.spsolve(A.tocsc(), rhs)
    xs -= xs.mean()
    assert rel(gphi.reshape(-1) - gphi.mean(), xs) < 1e4 * rtol
    ctx.close()


@pytest.mark.parametrize("sort", [False, True])
def test_collide2d_predrawn_matches_reference_stream(sort):
    say(f"2-D predrawn collisions sort={sort}")
    cli = cli2d(10, nui=0.05, Tn_over_Ti=2.0)
    ctx, po, pos, vel = make(cli, 4000, seed=7, collide_mode=_lib.COLLIDE_PREDRAWN)
    v_nr = vel.copy()
    rc, nl, Ns, Nf, log = oracle.scatter_elastic_nr_nd(po, 2, 1.0, v_nr, oracle.NR(), log_cap=200000)
    assert rc == 0 and nl == len(log)
    ctx.set_particles(pos, vel)
    if sort:
        ctx.collect()  # records name ions by set index, whatever order they are stored in
    ctx.set_attempts(log)
    ctx.collide(1.0)
    _, gvel = ctx.get_particles()
    assert rel(gvel, v_nr) <= 1e-12
    Nc, ns, nf, att = ctx.collision_stats()
    assert (Nc, ns, nf, att) == (oracle.collision_count(po, 4000, 1.0), Ns, Nf, nl)
    assert ctx.L.ecsprt_hip_collide(3, 1.0, ctx.ptr) == _lib.ERR_ARG  # a 3-D call on a 2-D context
    ctx.close()


def test_step2d_philox_conserves_and_collides():
    """Staged two-dimensional steps with Philox collisions: the ion count and the
    deposited charge are exact, the collisions track Nc and the potential solves."""
    say("2-D steps")
    N = 32
    cli = cli2d(N, nui=0.02, E0=0.05)
    p = pic.setup_context(**dict(cli, ndim=2))
    ctx = pic.Context(p, pic.default_options(collide_mode=_lib.COLLIDE_PHILOX, ksp_rtol=1e-8))
    n = 64 * N * N
    ctx.init_particles(n, 11, 1e-2, 2)
    ctx.collect()
    ctx.compute_potential()
    for step in range(10):
        ctx.step(1.0)
        m = ctx.get_moments()
        assert m[..., 0].sum() == pytest.approx(n, rel=1e-12)
        Nc, Ns, Nf, att = ctx.collision_stats()
        assert Nf == 0 and abs(Ns - Nc) <= 0.1 * Nc + 50, (Nc, Ns)
        it, rn, bn = ctx.solver_stats()
        assert rn <= 1.01e-8 * bn
    assert np.isfinite(ctx.get_potential()).all()
    assert ctx.num_particles == n
    ctx.close()


@pytest.mark.parametrize("N,Ny,dx,viT,collide", [(40, 40, 1.0, 1.0, True), (48, 20, 0.5, 1.0, True),
                                                  (70, 36, 0.3, 3.0, False)])
def test_step2d_independent_of_storage_order(N, Ny, dx, viT, collide):
    """Philox draws are keyed by the index an ion was set with, so a step gives the same
    ions whether they start in set order or in the tile order a deposit leaves them in,
    and the fused step (push, collisions, drift and the count of the ions' new tiles over
    tile-ordered ions) gives the same ions and moments as the stages called one by one.
    Cases: one tile row (Ny < 32), partial tiles and cells of 0.3, hot ions of which some
    move further than a neighbouring tile."""
    say(f"2-D steps, storage order and fused vs staged N={N} Ny={Ny} dx={dx} viT={viT} collide={collide}")
    cli = cli2d(N, Ny, nui=0.05 if collide else 0.0, E0=0.05, viT=viT)
    cli.update(dx=dx)
    runs, phi0 = [], None
    for how in ("step", "unsorted", "staged"):
        p = pic.setup_context(**dict(cli, ndim=2))
        mode = _lib.COLLIDE_PHILOX if collide else _lib.COLLIDE_OFF
        ctx = pic.Context(p, pic.default_options(collide_mode=mode, ksp_rtol=1e-10,
                                                 operator_mode=_lib.OPERATOR_REBUILD))
        ctx.init_particles(32 * N * Ny, 5, 1e-2, 2)
        if how == "unsorted":
            ctx.set_potential(phi0)
        else:
            ctx.collect()
            ctx.compute_potential()
            phi0 = ctx.get_potential()
        out = []
        for step in range(3):
            if how == "staged":
                ctx.push(1.0)
                ctx.collide(1.0)
                ctx.update_positions(1.0)
                ctx.collect()
                ctx.compute_potential()
            else:
                ctx.step(1.0)
            out.append((*ctx.get_particles(), ctx.get_moments(), ctx.collision_stats()))
        runs.append(out)
        ctx.close()
    for other in runs[1:]:
        for (pa, va, ma, ca), (pb, vb, mb, cb) in zip(runs[0], other):
            np.testing.assert_array_equal(pa, pb)
            np.testing.assert_array_equal(va, vb)
            np.testing.assert_array_equal(ma, mb)
            assert ca == cb
            assert (ca[0] > 0 and ca[1] > 0) == collide


@pytest.mark.parametrize("dist", ["normal", "sobol", "sinusoidal"])
def test_reference_initial_conditions_2d(dist):
    say(f"2-D initial conditions {dist}")
    cli = cli2d(12, viT=0.8)
    cli.update(dx=0.5, vi0x=0.1)
    p = pic.setup_context(**dict(cli, ndim=2))
    po = oracle.setup_context(**cli)
    n = 3000
    ctx = pic.Context(p, pic.default_options())
    ctx.init_distribution(n, dist)
    gpos, gvel = ctx.get_particles()
    L = np.array([po.Lx, po.Ly])
    if dist == "normal":
        s3, _ = oracle.NR().stream("ran3", -12345, 2 * n)
        pos = s3.astype(np.float64).reshape(n, 2) * L
    elif dist == "sobol":
        pos = oracle.sobol_points(n, 2) * L
    else:
        pos, _ = oracle.sinusoidal_positions(po, n, ndim=2)
    sg, _ = oracle.NR().stream("gasdev", -12345, 2 * n)
    vel = np.array([po.viTx, po.viTy]) * sg.astype(np.float64).reshape(n, 2) + np.array([po.vi0x, po.vi0y])
    np.testing.assert_array_equal(gpos, pos)
    np.testing.assert_array_equal(gvel, vel)
    counts = ctx.cell_counts()
    assert counts.shape == (po.Ny, po.Nx) and counts.sum() == n
    ctx.close()
