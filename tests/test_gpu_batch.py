"""Batched standalone potential solves and the synthetic density generator."""
import numpy as np
import pytest

import oracle
from ecsprt_amd import configs, pic, _lib

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def numpy_lognormal(dims, seed, sigma, smooth, fs):
    """Restatement of the generator: Philox normals per node, exp, periodic Gaussian
    smoothing per axis, mean 1."""
    Nx, Ny, Nz = dims
    n = Nx * Ny * Nz
    g = np.empty(n)
    key = [seed & 0xFFFFFFFF, seed >> 32]
    for q in range(n):
        r = pic.philox4x32_10([q & 0xFFFFFFFF, q >> 32, 0x106E, 0], key)
        u1 = ((r[0] >> 5) * 67108864.0 + (r[1] >> 6) + 0.5) / 9007199254740992.0
        u2 = ((r[2] >> 5) * 67108864.0 + (r[3] >> 6)) / 9007199254740992.0
        g[q] = np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)
    a = np.exp(sigma * g - 0.5 * sigma * sigma).reshape(Nz, Ny, Nx)
    R = int(np.ceil(4 * smooth))
    w = np.exp(-0.5 * np.arange(-R, R + 1) ** 2 / smooth ** 2)
    w /= w.sum()
    for axis in (2, 1, 0):  # x, y, z
        a = sum(w[d + R] * np.roll(a, -d, axis=axis) for d in range(-R, R + 1))
    a *= a.size / a.sum()
    m = np.empty(a.shape + (4,))
    m[..., 0] = a
    m[..., 1:] = fs * a[..., None]
    return m


def test_lognormal_moments_match_restatement():
    dims = (6, 5, 4)
    dv = pic.lognormal_moments(dims, seed=(7 << 32) | 3, sigma=0.3, smooth=1.0, flux_scale=0.5)
    got = dv.download().reshape(dims[2], dims[1], dims[0], 4)
    dv.free()
    ref = numpy_lognormal(dims, (7 << 32) | 3, 0.3, 1.0, 0.5)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0)
    assert got[..., 0].mean() == pytest.approx(1.0, rel=1e-14)


def test_lognormal_statistics():
    """sigma = 0.3 without smoothing: log n is Gaussian with standard deviation 0.3."""
    dims = (40, 40, 40)
    dv = pic.lognormal_moments(dims, seed=11, sigma=0.3, smooth=0.0)
    m = dv.download().reshape(-1, 4)
    dv.free()
    ln = np.log(m[:, 0])
    assert ln.std() == pytest.approx(0.3, rel=0.02)
    np.testing.assert_allclose(m[:, 1:], m[:, :1] * np.ones((1, 3)))


DIMS = [(12, 10, 8), (16, 16, 8), (8, 12, 16), (10, 10, 10), (24, 8, 8), (6, 9, 12)]


def fields(dims, base_seed=100):
    return [pic.lognormal_moments(d, seed=base_seed + q, sigma=0.3, smooth=2.0) for q, d in enumerate(dims)]


def c3_params():
    return pic.setup_context(**configs.base_cli(8, E0=0.05))


@pytest.mark.parametrize("slots", [1, 3])
def test_batch_matches_single_context_solves(slots):
    """Each problem of an FP64 batch equals the same solve on its own context."""
    say(f"batch vs single, slots={slots}")
    p = c3_params()
    opts = pic.default_options(ksp_rtol=1e-10)
    mom = fields(DIMS)
    phi = [pic.DeviceVector(d[0] * d[1] * d[2]) for d in DIMS]
    b = pic.Batch(p, opts, DIMS, slots=slots)
    b.solve(mom, phi)
    it, rn = b.stats()
    say("  iterations", it.tolist())
    assert (it > 0).all()
    for q, d in enumerate(DIMS):
        cli = configs.base_cli(8, E0=0.05)
        cli.update(Nx=d[0], Ny=d[1], Nz=d[2])
        ctx = pic.Context(pic.setup_context(**cli), pic.default_options(ksp_rtol=1e-10))
        m = mom[q].download().reshape(d[2], d[1], d[0], 4)
        ctx.set_moments(m)
        ctx.compute_potential()
        ref = ctx.get_potential().reshape(-1)
        ctx.close()
        np.testing.assert_array_equal(phi[q].download(), ref)
    # and against the oracle's sparse direct solve for two of them
    for q in (0, 3):
        d = DIMS[q]
        cli = configs.base_cli(8, E0=0.05)
        cli.update(Nx=d[0], Ny=d[1], Nz=d[2])
        po = oracle.setup_context(**cli)
        m = mom[q].download().reshape(d[2], d[1], d[0], 4)
        x = oracle.solve_sparse(po, m, oracle.full_rhs(po, m)).reshape(-1)
        got = phi[q].download()
        assert np.abs(got - x).max() <= 1e-7 * np.abs(x).max()
    b.close()
    for v in mom + phi:
        v.free()


def test_batch_mixed_precision_error():
    """Mixed-precision batch (FP32 GMRES + multigrid inside FP64 refinement) against
    the FP64 batch on the same fields, both to rtol 1e-10."""
    say("batch mixed vs FP64")
    p = c3_params()
    mom = fields(DIMS, base_seed=200)
    out = {}
    for prec in (_lib.SOLVE_FP64, _lib.SOLVE_MIXED):
        phi = [pic.DeviceVector(d[0] * d[1] * d[2]) for d in DIMS]
        b = pic.Batch(p, pic.default_options(ksp_rtol=1e-10, solve_precision=prec), DIMS, slots=4)
        b.solve(mom, phi)
        out[prec] = [f.download() for f in phi]
        b.close()
        for f in phi:
            f.free()
    for q in range(len(DIMS)):
        ref = out[_lib.SOLVE_FP64][q]
        err = np.abs(out[_lib.SOLVE_MIXED][q] - ref).max() / np.abs(ref).max()
        say(f"  problem {q} {DIMS[q]} error {err:.2e}")
        assert err <= 1e-6
    for v in mom:
        v.free()


def test_batch_rejects_bad_dims():
    p = c3_params()
    with pytest.raises(pic.HipError):
        pic.Batch(p, pic.default_options(), [(8, 8, 2)], slots=1)


def test_ensemble_step_matches_sequential_steps():
    """Concurrent ensemble steps of uneven runs equal the same runs stepped one by one."""
    say("ensemble vs sequential")
    runs = [(12, 6, _lib.PARTICLES_MIXED, 0.05), (16, 4, _lib.PARTICLES_FP64, 0.08), (8, 10, _lib.PARTICLES_MIXED, 0.02)]

    def make(N, ppc, prec, B0, seed):
        cli = configs.base_cli(N, B0=B0, nui=0.02)
        o = pic.default_options(particle_precision=prec, ksp_rtol=1e-10, collide_mode=_lib.COLLIDE_PHILOX)
        c = pic.Context(pic.setup_context(**cli), o)
        c.init_particles(ppc * N ** 3, seed, 0.05, 1)
        c.collect()
        c.compute_potential()
        return c

    ens = [make(*r, seed=q + 1) for q, r in enumerate(runs)]
    seq = [make(*r, seed=q + 1) for q, r in enumerate(runs)]
    for _ in range(2):
        pic.ensemble_step(ens, 1.0)
        for c in seq:
            c.step(1.0)
    for a, b in zip(ens, seq):
        pa, va = a.get_particles()
        pb, vb = b.get_particles()
        np.testing.assert_allclose(pa, pb, rtol=0, atol=1e-11)
        np.testing.assert_allclose(va, vb, rtol=0, atol=1e-11)
        np.testing.assert_allclose(a.get_moments(), b.get_moments(), rtol=1e-11, atol=1e-11)
        assert a.collision_stats()[1] > 0
    for c in ens + seq:
        c.close()


LOCKSTEP_DIMS = DIMS + [(32, 32, 32), (48, 32, 64), (64, 24, 40), (20, 28, 36)]


@pytest.mark.parametrize("prec", [_lib.SOLVE_FP64, _lib.SOLVE_MIXED])
def test_lockstep_batch_matches_single_context_solves(prec):
    """The lockstep batch (slots = 0: every kernel launched once for all problems, tails at
    level 0 and beyond, marching and box sweeps mixed in one launch) solves each problem as
    its own context does: the same method, summed in other partial orders, so within the
    accuracy rtol 1e-10 allows; every true residual meets rtol."""
    say(f"lockstep batch vs single contexts, precision {prec}")
    p = c3_params()
    opts = pic.default_options(ksp_rtol=1e-10, solve_precision=prec)
    mom = fields(LOCKSTEP_DIMS, base_seed=300)
    phi = [pic.DeviceVector(d[0] * d[1] * d[2]) for d in LOCKSTEP_DIMS]
    b = pic.Batch(p, opts, LOCKSTEP_DIMS, slots=0)
    b.solve(mom, phi)
    it, rn = b.stats()
    say("  iterations", it.tolist(), "traffic GB", b.traffic() / 1e9)
    assert (it > 0).all() and b.traffic() > 0
    for q, d in enumerate(LOCKSTEP_DIMS):
        cli = configs.base_cli(8, E0=0.05)
        cli.update(Nx=d[0], Ny=d[1], Nz=d[2])
        m = mom[q].download().reshape(d[2], d[1], d[0], 4)
        ctx = pic.Context(pic.setup_context(**cli), pic.default_options(ksp_rtol=1e-10, solve_precision=prec))
        ctx.set_moments(m)
        ctx.compute_potential()
        ref = ctx.get_potential().reshape(-1)
        ctx.close()
        got = phi[q].download()
        err = np.abs(got - ref).max() / np.abs(ref).max()
        po = oracle.setup_context(**cli)
        bq = oracle.full_rhs(po, m)
        r = bq - oracle.lhs_apply(po, m, got.reshape(d[2], d[1], d[0]))
        rr = np.linalg.norm(r) / np.linalg.norm(bq)
        say(f"  problem {q} {d}: vs single context {err:.1e}, ||b - A phi||/||b|| {rr:.1e}")
        assert err <= 1e-7
        assert rr <= 1.01e-10
        assert abs(got.mean()) <= 1e-12 * np.abs(got).max()
    b.close()
    for v in mom + phi:
        v.free()


def test_lockstep_batch_repeats_bitwise():
    """Two solves of the same batch give the same potentials bit for bit (fixed reduction
    orders, no atomics)."""
    p = c3_params()
    dims = LOCKSTEP_DIMS[-4:]
    mom = fields(dims, base_seed=400)
    out = []
    b = pic.Batch(p, pic.default_options(ksp_rtol=1e-8, solve_precision=_lib.SOLVE_MIXED), dims, slots=0)
    for _ in range(2):
        phi = [pic.DeviceVector(d[0] * d[1] * d[2]) for d in dims]
        b.solve(mom, phi)
        out.append([f.download() for f in phi])
        for f in phi:
            f.free()
    for a, c in zip(*out):
        np.testing.assert_array_equal(a, c)
    b.close()
    for v in mom:
        v.free()
