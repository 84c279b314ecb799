"""GPU parity of the potential solve: RHS (src/rhs.c:259-373), the 11-point operator
(src/lhs.c:337-505) as a MATSHELL, the multigrid PCShell and the KSP solve."""
import numpy as np
import pytest

import oracle
from ecsprt_amd import configs, pic, _lib

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def moments_case(N, seed, shape=None, kappa_e=4.0, E0=0.05):
    cli = configs.base_cli(N, kappa_e=kappa_e, E0=E0)
    if shape:
        cli.update(Nx=shape[2], Ny=shape[1], Nz=shape[0])
    p = pic.setup_context(**cli)
    po = oracle.setup_context(**cli)
    rng = np.random.default_rng(seed)
    sh = (po.Nz, po.Ny, po.Nx)
    m = np.empty(sh + (4,))
    m[..., 0] = 64 + 8 * rng.normal(size=sh)
    m[..., 1:] = 8 * rng.normal(size=sh + (3,))
    return p, po, m


@pytest.mark.parametrize("shape", [(8, 8, 8), (12, 10, 6), (16, 16, 16)])
def test_rhs_and_matmult_parity(shape):
    say(f"rhs/matmult {shape}")
    p, po, m = moments_case(0, 1, shape=shape)
    ctx = pic.Context(p, pic.default_options(operator_mode=_lib.OPERATOR_REBUILD))
    ctx.set_moments(m)
    n = m[..., 0].size
    b = pic.DeviceVector(n)
    ctx.rhs(b)
    gb = b.download().reshape(m.shape[:3])
    ob = oracle.full_rhs(po, m)
    assert rel(gb, ob) <= 1e-12
    ctx.lhs()
    x = np.random.default_rng(2).normal(size=m.shape[:3])
    xd, yd = pic.DeviceVector(n), pic.DeviceVector(n)
    xd.upload(x)
    ctx.matmult(xd, yd)
    gy = yd.download().reshape(m.shape[:3])
    oy = oracle.lhs_apply(po, m, x)
    assert rel(gy, oy) <= 1e-12
    ctx.remove_mean(xd)
    np.testing.assert_allclose(xd.download().reshape(x.shape), x - x.mean(), atol=1e-14)
    for v in (b, xd, yd):
        v.free()
    ctx.close()


@pytest.mark.parametrize("N,pc", [(8, _lib.PC_GMG), (16, _lib.PC_GMG), (16, _lib.PC_JACOBI), (12, _lib.PC_NONE),
                                  (24, _lib.PC_GMG)])
def test_ksp_solve_matches_sparse_direct(N, pc):
    say(f"ksp N={N} pc={pc}")
    p, po, m = moments_case(N, 3)
    rtol = 1e-10
    ctx = pic.Context(p, pic.default_options(operator_mode=_lib.OPERATOR_REBUILD, pc_type=pc, ksp_rtol=rtol))
    ctx.set_moments(m)
    ctx.compute_potential()
    gphi = ctx.get_potential()
    b = oracle.full_rhs(po, m)
    x = oracle.solve_sparse(po, m, b)
    r = b - oracle.lhs_apply(po, m, gphi)
    it, rn, bn = ctx.solver_stats()
    say(f"  iters={it} ||r||/||b||={np.linalg.norm(r) / np.linalg.norm(b):.2e}")
    assert np.linalg.norm(r) <= 1.01 * rtol * np.linalg.norm(b)
    assert abs(gphi.mean()) < 1e-12 * np.abs(gphi).max()
    assert rel(gphi, x) < 1e4 * rtol
    ctx.close()


def test_uniform_density_fft_solution():
    say("uniform density FFT")
    N = 16
    p, po, m = moments_case(N, 4)
    m[..., 0] = 64.0
    ctx = pic.Context(p, pic.default_options(operator_mode=_lib.OPERATOR_REBUILD, ksp_rtol=1e-12))
    ctx.set_moments(m)
    ctx.compute_potential()
    gphi = ctx.get_potential()
    b = oracle.full_rhs(po, m)
    k = 2 * np.pi * np.fft.fftfreq(N)
    KZ, KY, KX = np.meshgrid(k, k, k, indexing="ij")
    lam = po.scale * 64.0 * ((2 * np.cos(KX) - 2) + (2 * np.cos(KY) - 2) + (1 + po.kappa_e ** 2) * (2 * np.cos(KZ) - 2))
    lam[0, 0, 0] = 1.0
    xh = np.fft.fftn(b) / lam
    xh[0, 0, 0] = 0
    assert rel(gphi, np.real(np.fft.ifftn(xh))) < 1e-9
    ctx.close()


@pytest.mark.parametrize("prec,rtol,tol", [(_lib.SOLVE_FP64, 1e-10, 1e-9), (_lib.SOLVE_MIXED, 1e-6, 1e-6)])
def test_ksp_first_step_solution_drops_the_queued_step(prec, rtol, tol):
    """A Fourier mode is an eigenvector of the uniform-density operator, so
    unpreconditioned GMRES meets its tolerance in its first Arnoldi step; the solver has
    by then queued the second step on the residual estimates, and that step must not
    reach the solution, which is b / lambda.  In the mixed solve this is the first FP32
    cycle (its tolerance 1e-3; the FP32 operator's rounding on this smooth mode leaves
    about 3e-5, which a second refinement cycle removes)."""
    say(f"first-step solve, precision {prec}")
    N = 16
    p, po, m = moments_case(N, 7)
    m[..., 0] = 64.0
    m[..., 1:] = 0.0
    ctx = pic.Context(p, pic.default_options(operator_mode=_lib.OPERATOR_REBUILD, pc_type=_lib.PC_NONE,
                                             ksp_rtol=rtol, solve_precision=prec))
    ctx.set_moments(m)
    ctx.lhs()
    i = np.arange(N)
    b = np.broadcast_to(np.cos(2 * np.pi * i / N)[None, None, :], (N, N, N)).copy()
    lam = po.scale * 64.0 * (2 * np.cos(2 * np.pi / N) - 2)
    # the mode is an eigenvector of the oracle's operator
    assert rel(oracle.lhs_apply(po, m, b), lam * b) < 1e-12
    n = b.size
    bd, xd = pic.DeviceVector(n), pic.DeviceVector(n)
    bd.upload(b.reshape(-1))
    ctx.ksp_solve(bd, xd)
    x = xd.download().reshape(b.shape)
    it, rn, bn = ctx.solver_stats()
    say(f"  iters={it} ||r||/||b||={rn / bn:.2e}")
    if prec == _lib.SOLVE_FP64:
        assert it == 1
    assert rn <= rtol * bn
    assert rel(x, b / lam) < tol
    bd.free()
    xd.free()
    ctx.close()


def test_frozen_operator_semantics():
    """KSPSetComputeOperators: the operator is built at the first solve only."""
    say("frozen operator")
    p, po, m = moments_case(8, 5)
    ctx = pic.Context(p, pic.default_options(operator_mode=_lib.OPERATOR_FROZEN, ksp_rtol=1e-11))
    ctx.set_moments(m)
    ctx.compute_potential()
    m2 = m.copy()
    m2[..., 0] *= 1.5
    m2[..., 1:] *= 0.5
    ctx.set_moments(m2)
    ctx.compute_potential()
    gphi = ctx.get_potential()
    b2 = oracle.full_rhs(po, m2)
    x_frozen = oracle.solve_sparse(po, m, b2)   # operator from the first moments
    assert rel(gphi, x_frozen) < 1e-7
    ctx.close()


def test_pcapply_is_a_contraction():
    say("pcapply")
    p, po, m = moments_case(16, 6)
    ctx = pic.Context(p, pic.default_options(operator_mode=_lib.OPERATOR_REBUILD))
    ctx.set_moments(m)
    ctx.lhs()
    n = m[..., 0].size
    b = oracle.full_rhs(po, m).reshape(-1)
    x_true = oracle.solve_sparse(po, m, b.reshape(m.shape[:3])).reshape(-1)
    rd, zd = pic.DeviceVector(n), pic.DeviceVector(n)
    rd.upload(b)
    ctx.pcapply(rd, zd)
    z = zd.download()
    z -= z.mean()
    # one V-cycle from zero removes most of the error
    assert np.linalg.norm(z - x_true) < 0.5 * np.linalg.norm(x_true)
    rd.free()
    zd.free()
    ctx.close()


def test_mixed_precision_solve_error():
    say("mixed precision solve")
    N = 16
    p, po, m = moments_case(N, 7)
    o64 = pic.default_options(operator_mode=_lib.OPERATOR_REBUILD, ksp_rtol=1e-10)
    o32 = pic.default_options(operator_mode=_lib.OPERATOR_REBUILD, ksp_rtol=1e-10,
                              solve_precision=_lib.SOLVE_MIXED)
    res = []
    for o in (o64, o32):
        ctx = pic.Context(p, o)
        ctx.set_moments(m)
        ctx.compute_potential()
        res.append(ctx.get_potential())
        say("  iters", ctx.solver_stats())
        ctx.close()
    assert rel(res[1], res[0]) < 1e-6


@pytest.mark.parametrize("prec", [_lib.SOLVE_FP64, _lib.SOLVE_MIXED])
def test_marching_and_box_sweeps_agree(prec):
    """The z-marching smoother and the one-shot box smoother do the same arithmetic: a
    solve with either on every level gives the same potential bit for bit, on a grid
    with partial xy tiles and a z extent that is not a multiple of the chunk."""
    say(f"marching vs box sweeps, precision {prec}")
    cli = configs.base_cli(8, E0=0.05)
    cli.update(Nx=40, Ny=20, Nz=14)
    p = pic.setup_context(**cli)
    rng = np.random.default_rng(21)
    m = np.empty((14, 20, 40, 4))
    m[..., 0] = 1.0 + 0.2 * rng.random((14, 20, 40))
    m[..., 1:] = 0.1 * rng.normal(size=(14, 20, 40, 3))
    out = []
    for mode in (1, 2):
        ctx = pic.Context(p, pic.default_options(ksp_rtol=1e-10, solve_precision=prec, mg_sweep=mode))
        ctx.set_moments(m)
        ctx.compute_potential()
        out.append((ctx.get_potential(), ctx.solver_stats()[0]))
        ctx.close()
    assert out[0][1] == out[1][1]
    np.testing.assert_array_equal(out[0][0], out[1][0])


@pytest.mark.parametrize("prec", [_lib.SOLVE_FP64, _lib.SOLVE_MIXED])
def test_gram_schmidt_once_and_twice(prec):
    """Classical Gram-Schmidt once per Arnoldi step (PETSc's default) and with one
    reorthogonalisation both reach the true-residual tolerance and the same potential to
    within the tolerance's reach; the preconditioned Krylov basis stays orthogonal enough
    at these restart lengths that the iteration counts differ by at most one."""
    say(f"Gram-Schmidt once vs twice, precision {prec}")
    cli = configs.base_cli(8, E0=0.05, kappa_e=4.0)
    cli.update(Nx=48, Ny=40, Nz=24)
    p = pic.setup_context(**cli)
    po = oracle.setup_context(**cli)
    rng = np.random.default_rng(8)
    m = np.empty((24, 40, 48, 4))
    m[..., 0] = 1.0 + 0.3 * rng.random((24, 40, 48))
    m[..., 1:] = 0.1 * rng.normal(size=(24, 40, 48, 3))
    rtol = 1e-10
    out = []
    for refine in (0, 1):
        ctx = pic.Context(p, pic.default_options(ksp_rtol=rtol, solve_precision=prec, gmres_refine=refine,
                                                 operator_mode=_lib.OPERATOR_REBUILD))
        ctx.set_moments(m)
        ctx.compute_potential()
        phi = ctx.get_potential()
        b = oracle.full_rhs(po, m)
        r = b - oracle.lhs_apply(po, m, phi)
        assert np.linalg.norm(r) <= 1.01 * rtol * np.linalg.norm(b)
        out.append((phi - phi.mean(), ctx.solver_stats()[0]))
        ctx.close()
    say(f"  iterations once {out[0][1]}, twice {out[1][1]}")
    assert abs(out[0][1] - out[1][1]) <= 1
    assert np.abs(out[0][0] - out[1][0]).max() <= 1e4 * rtol * np.abs(out[1][0]).max()
