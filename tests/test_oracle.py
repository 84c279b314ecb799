"""Pins of the CPU oracle against mathematics the reference's algorithm fixes.

The reference cannot be built here (PETSc is absent) and its own tests only check
exit codes, so each oracle function is checked against closed forms, invariants,
published known answers or an independent library, chosen so that a dropped term,
a wrong sign or a swapped index fails.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle
from ecsprt_amd import configs


def params(N=8, **kw):
    cli = configs.base_cli(N, **kw)
    return oracle.setup_context(**cli)


# ---------------------------------------------------------------------------- setup
def test_setup_context_derived_coefficients():
    p = oracle.setup_context(Nx=16, Ny=8, Nz=4, dx=0.5, dy=2.0, dz=1.0, B0=0.05, qi=1.0, mi=1.0,
                             nue=1e9, Te=11604.0, viTx=1.0, viTy=2.0, viTz=2.0, mn=2.0, Tn=-1.0)
    assert (p.Lx, p.Ly, p.Lz) == (8.0, 16.0, 4.0)
    assert p.x1 == 8.0 and p.x0 == 0.0
    omega_e = oracle.Q * 0.05 / oracle.ME
    assert p.kappa_e == pytest.approx(omega_e / 1e9, rel=1e-15)
    assert p.scale == pytest.approx(1.0 / (1.0 + p.kappa_e ** 2), rel=1e-15)
    assert p.viT == pytest.approx(3.0, rel=1e-15)  # sqrt(1 + 4 + 4)
    assert p.Ti == pytest.approx(0.5 * 1.0 / oracle.KB * 9.0, rel=1e-15)
    assert p.Tn == p.Ti  # default neutral temperature = ion temperature
    assert p.vnT == pytest.approx(math.sqrt(2 * oracle.KB * p.Tn / 2.0), rel=1e-15)
    # dx <= 0: L from the bounds
    q = oracle.setup_context(Nx=10, Ny=10, Nz=10, x0=-1.0, x1=3.0)
    assert q.dx == pytest.approx(0.4) and q.Lx == pytest.approx(4.0)


# ----------------------------------------------------------------------- field gather
def test_gather_weight_swap_x():
    """phi = i^2: E_x numerator = 4*(il + w_z), not 4*(il + w_x) (vector-math.c:645-651)."""
    N = 10
    i = np.arange(N, dtype=float)
    phi = np.broadcast_to(i ** 2, (N, N, N)).copy()
    rng = np.random.default_rng(1)
    for _ in range(50):
        x, y, z = rng.uniform(2, N - 3), rng.uniform(0, N - 1), rng.uniform(0, N - 1)
        f = oracle.difference_vector3d(phi, x, y, z)
        il, wz = int(x), z - int(z)
        assert f[0] == pytest.approx(4 * (il + wz), rel=1e-13)
        assert f[1] == 0.0 and f[2] == 0.0


def test_gather_weight_swap_z():
    N = 10
    k = np.arange(N, dtype=float)
    phi = np.broadcast_to((k ** 2)[:, None, None], (N, N, N)).copy()
    rng = np.random.default_rng(2)
    for _ in range(50):
        x, y, z = rng.uniform(0, N - 1), rng.uniform(0, N - 1), rng.uniform(2, N - 3)
        f = oracle.difference_vector3d(phi, x, y, z)
        assert f[2] == pytest.approx(4 * (int(z) + (x - int(x))), rel=1e-13)
        assert f[0] == 0.0 and f[1] == 0.0


def test_gather_y_component_plain_weights():
    N = 10
    j = np.arange(N, dtype=float)
    phi = np.broadcast_to((j ** 2)[None, :, None], (N, N, N)).copy()
    x, y, z = 3.3, 4.6, 5.8
    f = oracle.difference_vector3d(phi, x, y, z)
    # y-difference interpolated with wyh along y: 4*(jl + wy)
    assert f[1] == pytest.approx(4 * (4 + 0.6), rel=1e-13)


def test_gather_linear_is_twice_slope():
    N = 12
    i = np.arange(N, dtype=float)
    phi = np.broadcast_to(0.75 * i, (N, N, N)).copy()
    f = oracle.difference_vector3d(phi, 5.3, 5.5, 5.7)
    assert f[0] == pytest.approx(2 * 0.75, rel=1e-14)


def test_gather_last_cell_one_sided():
    """In the last x-cell the upper node uses 3F[N]-4F[N-1]+F[N-2], exact for quadratics."""
    N = 8
    c = N - 1
    i = np.arange(N, dtype=float)
    q = (i - c) ** 2
    q[0] = (N - c) ** 2  # periodic image of node N carries the quadratic's value at N
    phi = np.broadcast_to(q, (N, N, N)).copy()
    for wz in (0.0, 0.25, 0.8):
        f = oracle.difference_vector3d(phi, N - 1 + 0.4, 3.2, 2.0 + wz)
        # node N-1: central F[N]-F[N-2] = 2*q'(N-1) = 0; node N: one-sided = 2*q'(N) = 4
        assert f[0] == pytest.approx(4 * wz, abs=1e-13)


# ---------------------------------------------------------------------------- Boris
def test_boris_rotation_constant_phi():
    p = params(8, B0=0.3)
    rng = np.random.default_rng(3)
    n = 200
    pos = rng.uniform(0, 8, (n, 3))
    vel = rng.normal(size=(n, 3))
    v0 = vel.copy()
    phi = np.full((8, 8, 8), 3.0)
    dt = 0.7
    oracle.boris(p, dt, pos, vel, phi)
    t = 0.5 * p.qi / p.mi * dt * p.B0
    th = 2 * math.atan(t)
    exp_x = math.cos(th) * v0[:, 0] + math.sin(th) * v0[:, 1]
    exp_y = -math.sin(th) * v0[:, 0] + math.cos(th) * v0[:, 1]
    np.testing.assert_allclose(vel[:, 0], exp_x, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(vel[:, 1], exp_y, rtol=1e-13, atol=1e-14)
    np.testing.assert_array_equal(vel[:, 2], v0[:, 2])


def test_boris_half_kick_only():
    p = params(12, B0=0.0)
    N = 12
    i = np.arange(N, dtype=float)
    a = 0.3
    phi = np.broadcast_to(a * i, (N, N, N)).copy()
    rng = np.random.default_rng(4)
    pos = np.column_stack([rng.uniform(2, N - 3, 50), rng.uniform(0, N, 50), rng.uniform(0, N, 50)])
    vel = rng.normal(size=(50, 3))
    v0 = vel.copy()
    dt = 0.8
    oracle.boris(p, dt, pos, vel, phi)
    Ex = -a / p.dx
    np.testing.assert_allclose(vel[:, 0] - v0[:, 0], p.qi / p.mi * dt / 2 * Ex, rtol=1e-12)
    np.testing.assert_allclose(vel[:, 1:], v0[:, 1:], rtol=0, atol=1e-15)


# ------------------------------------------------------------------- drift and wrap
def test_drift_exact():
    rng = np.random.default_rng(5)
    pos = rng.uniform(0, 8, (100, 3))
    vel = rng.normal(size=(100, 3))
    exp = pos + vel * 0.25
    oracle.update_positions(0.25, pos, vel)
    np.testing.assert_array_equal(pos, exp)


def test_wrap_edge_cases():
    p = params(8)
    L = p.Lx
    ulp = np.nextafter(L, 0) - L
    x = np.array([L + ulp, -1e-300, L, L + 0.5, -0.5, 2.5 * L, -1.5 * L, np.nan])
    pos = np.column_stack([x, np.full_like(x, 1.0), np.full_like(x, 1.0)])
    oracle.apply_bc(p, pos)
    got = pos[:, 0]
    assert got[0] == L + ulp            # L - ulp stays
    assert got[1] == 0.0                # -tiny: +L rounds to L, then >= L wraps to 0
    assert got[2] == 0.0                # L -> 0
    assert got[3] == 0.5
    assert got[4] == L - 0.5
    assert got[5] == 1.5 * L            # one wrap only: still outside
    assert got[6] == -0.5 * L
    assert np.isnan(got[7])
    assert oracle.count_outside(p, pos) == 3


# ---------------------------------------------------------------------------- deposit
def _cic_numpy(p, pos, vel):
    N = (p.Nz, p.Ny, p.Nx)
    m = np.zeros(N + (4,))
    x = pos[:, 0] / p.dx
    y = pos[:, 1] / p.dy
    z = pos[:, 2] / p.dz
    i0, j0, k0 = np.floor(x).astype(int), np.floor(y).astype(int), np.floor(z).astype(int)
    fx, fy, fz = x - i0, y - j0, z - k0
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                w = (fx if a else 1 - fx) * (fy if b else 1 - fy) * (fz if c else 1 - fz)
                ii, jj, kk = (i0 + a) % p.Nx, (j0 + b) % p.Ny, (k0 + c) % p.Nz
                np.add.at(m[..., 0], (kk, jj, ii), w)
                for d in range(3):
                    np.add.at(m[..., 1 + d], (kk, jj, ii), w * vel[:, d])
    return m


def test_deposit_against_numpy_cic():
    p = oracle.setup_context(Nx=6, Ny=5, Nz=7, dx=0.5, dy=1.0, dz=2.0)
    rng = np.random.default_rng(6)
    n = 3000
    pos = np.column_stack([rng.uniform(0, p.Lx, n), rng.uniform(0, p.Ly, n), rng.uniform(0, p.Lz, n)])
    vel = rng.normal(size=(n, 3))
    m = oracle.collect_moments(p, pos, vel)
    np.testing.assert_allclose(m, _cic_numpy(p, pos, vel), rtol=1e-12, atol=1e-12)
    assert m[..., 0].sum() == pytest.approx(n, rel=1e-13)
    np.testing.assert_allclose(m[..., 1:].sum(axis=(0, 1, 2)), vel.sum(axis=0), rtol=1e-11, atol=1e-10)


def test_deposit_node_centred_particle():
    p = params(8)
    pos = np.array([[3.0, 4.0, 5.0]])
    vel = np.array([[0.5, -1.5, 2.0]])
    m = oracle.collect_moments(p, pos, vel)
    assert m[5, 4, 3, 0] == 1.0
    np.testing.assert_array_equal(m[5, 4, 3, 1:], vel[0])
    assert m[..., 0].sum() == 1.0


def test_cell_counts_histogram():
    p = oracle.setup_context(Nx=6, Ny=5, Nz=4, dx=0.5, dy=1.0, dz=2.0)
    rng = np.random.default_rng(7)
    n = 5000
    pos = np.column_stack([rng.uniform(0, p.Lx, n), rng.uniform(0, p.Ly, n), rng.uniform(0, p.Lz, n)])
    c = oracle.cell_counts(p, pos)
    idx = (np.floor(pos[:, 2] / p.dz).astype(int), np.floor(pos[:, 1] / p.dy).astype(int),
           np.floor(pos[:, 0] / p.dx).astype(int))
    h = np.zeros((4, 5, 6), dtype=np.int64)
    np.add.at(h, idx, 1)
    np.testing.assert_array_equal(c, h)


# --------------------------------------------------------------------------------- RNG
def test_ran1_park_miller_minimal_standard():
    """Park & Miller (1988): from seed 1 the 10000th state is 1043618065."""
    nr = oracle.NR()
    seed = ctypes.c_long(-1)
    L = oracle.lib()
    L.oracle_ran1(ctypes.byref(nr.state), ctypes.byref(seed))  # init: 40 LCG steps, then 1
    for _ in range(10000 - 41):
        L.oracle_ran1(ctypes.byref(nr.state), ctypes.byref(seed))
    assert seed.value == 1043618065


def test_ran1_outputs_are_shuffled_lcg_states():
    nr = oracle.NR()
    out, _ = nr.stream("ran1", -7, 2000)
    # independent LCG in Python
    s = 7
    states = []
    for _ in range(2100):
        s = (16807 * s) % 2147483647
        states.append(s)
    pool = np.array(states, dtype=np.float64) / 2147483647.0
    pool32 = pool.astype(np.float32)
    assert np.isin(out, pool32).all()
    assert abs(out.mean() - 0.5) < 0.02


def test_ran3_lagged_subtraction_identity():
    nr = oracle.NR()
    L = oracle.lib()
    seed = ctypes.c_long(-12345)
    xs, outs = [], []
    for _ in range(400):
        outs.append(L.oracle_ran3(ctypes.byref(nr.state), ctypes.byref(seed)))
        xs.append(nr.state.r3_ma[nr.state.r3_inext])
    xs = np.array(xs, dtype=np.int64)
    for n in range(55, 400):
        assert xs[n] == (xs[n - 55] - xs[n - 24]) % 1000000000
    np.testing.assert_array_equal(np.array(outs, dtype=np.float32), (xs * 1e-9).astype(np.float32))
    assert seed.value == 1


def test_gasdev_polar_identity_and_moments():
    nr_g = oracle.NR()
    g, _ = nr_g.stream("gasdev", -99, 20000)
    nr_u = oracle.NR()
    u, _ = nr_u.stream("ran1", -99, 60000)
    # rebuild the accepted pairs from the independent ran1 stream
    pairs = []
    k = 0
    while len(pairs) < 10000:
        v1 = np.float32(2.0 * float(u[k]) - 1.0)
        v2 = np.float32(2.0 * float(u[k + 1]) - 1.0)
        k += 2
        rsq = np.float32(v1 * v1 + v2 * v2)
        if rsq >= 1.0 or rsq == 0.0:
            continue
        pairs.append((v1, v2, rsq))
    v1, v2, rsq = (np.array(a, dtype=np.float64) for a in zip(*pairs))
    a, b = g[0::2].astype(np.float64), g[1::2].astype(np.float64)  # v2*fac, then v1*fac
    np.testing.assert_allclose(a * a + b * b, -2.0 * np.log(rsq), rtol=3e-6)
    np.testing.assert_allclose(b / a, v1 / v2, rtol=3e-6)
    assert abs(g.mean()) < 0.03 and abs(g.var() - 1) < 0.04


# -------------------------------------------------------------------------- collisions
def _col_params(**kw):
    cli = configs.base_cli(8, nui=0.05, Tn_over_Ti=1.0)
    cli.update(kw)
    return oracle.setup_context(**cli)


def test_collision_count_cases():
    p = _col_params()
    vrm = 4 * p.viT
    # reference formula by hand
    np_ = 100000
    expect = int((np_ * p.nui * 1.0 * vrm) / ((p.vnT + p.vn0) * math.sqrt(p.mn / p.mi)))
    assert oracle.collision_count(p, np_, 1.0) == expect
    # saturation at np
    hot = _col_params(nui=100.0)
    assert oracle.collision_count(hot, 1000, 1.0) == 1000
    # cold neutrals: the quotient exceeds INT_MAX -> INT_MIN -> clamped to 0
    cold = _col_params(Tn=1e-40)
    assert oracle.collision_count(cold, 10 ** 6, 1.0) == 0
    # vnT = vn0 = 0: division by zero -> inf -> 0
    zero = _col_params(Tn=0.0)
    assert oracle.collision_count(zero, 1000, 1.0) == 0


def _random_attempts(n, rng):
    att = np.zeros(n, dtype=oracle.ATTEMPT_DTYPE)
    att["u_ip"] = rng.random(n, dtype=np.float32)
    att["g"] = rng.normal(size=(n, 3)).astype(np.float32)
    att["u_acc1"] = rng.random(n, dtype=np.float32)
    att["u_acc2"] = rng.random(n, dtype=np.float32)
    att["u_z"] = rng.random(n, dtype=np.float32)
    att["u_phi"] = rng.random(n, dtype=np.float32)
    return att


def test_collide_one_centre_of_mass_invariants():
    p = _col_params(mn=2.5)
    rng = np.random.default_rng(8)
    att = _random_attempts(400, rng)
    att["u_acc1"] = 0.0  # always accept
    vn0 = np.array([p.vn0x, p.vn0y, p.vn0z])
    n_ok = 0
    for a in att:
        vi = rng.normal(size=3)
        vf, acc, fail = oracle.collide_one(p, a, vi)
        assert acc
        if fail:
            continue
        n_ok += 1
        vn = a["g"].astype(np.float64) * p.vnT + vn0
        vc = (vi * p.mi + vn * p.mn) / (p.mi + p.mn)
        vcr = np.linalg.norm(vi - vc)
        assert np.linalg.norm(vf - vc) == pytest.approx(vcr, rel=1e-12)
        vn_after = vc - (p.mi / p.mn) * (vf - vc)
        np.testing.assert_allclose(p.mi * vf + p.mn * vn_after, p.mi * vi + p.mn * vn, rtol=1e-12, atol=1e-12)
        e0 = p.mi * vi @ vi + p.mn * vn @ vn
        e1 = p.mi * vf @ vf + p.mn * vn_after @ vn_after
        assert e1 == pytest.approx(e0, rel=1e-12)
        vr = vi - vn
        cos_scatter = (vf - vc) @ vr / (vcr * np.linalg.norm(vr))
        assert cos_scatter == pytest.approx(2.0 * float(a["u_z"]) - 1.0, abs=1e-10)
    assert n_ok > 300


def test_collide_acceptance_uses_two_draws():
    """Acceptance is u1*u2 < vrr^2/vrm^2 (PetscSqr expands to two ran3 calls)."""
    p = _col_params()
    vi = np.array([0.4, 0.0, 0.0])
    a = _random_attempts(1, np.random.default_rng(9))[0]
    a["g"] = 0.0
    vrr = np.linalg.norm(vi - np.array([p.vn0x, p.vn0y, p.vn0z]))
    thr = vrr ** 2 / (4 * p.viT) ** 2
    a["u_acc1"], a["u_acc2"] = np.float32(math.sqrt(thr) * 1.2), np.float32(math.sqrt(thr) * 0.5)
    _, acc, _ = oracle.collide_one(p, a, vi)
    assert acc  # the product is below thr although u1^2 would not be
    a["u_acc2"] = np.float32(math.sqrt(thr) * 1.2)
    _, acc, _ = oracle.collide_one(p, a, vi)
    assert not acc


@pytest.mark.parametrize("vi", [(0.4, -0.3, 0.9), (0.05, 0.02, -1.1), (1.2, 0.0, 0.0)])
def test_collide_scattering_direction_is_isotropic(vi):
    """For fixed ion and neutral velocities, the post-collision direction e = (vf - vc)/vcr
    over uniform (u_z, u_phi) is uniform on the sphere (collisions.c:100-132: cos(theta) =
    2 u_z - 1, azimuth 2 pi u_phi, rotated into the frame of vr): mean 0 and second moments
    I/3.  A midpoint grid in (u_z, u_phi) makes the quadrature accurate to O(1/n^2); a wrong
    azimuth scale (pi u_phi), polar map or rotation leaves a mean or a covariance off."""
    p = _col_params(mn=1.7)
    vi = np.array(vi)
    a = _random_attempts(1, np.random.default_rng(12))[0]
    a["g"] = np.array([0.3, -0.8, 0.5], dtype=np.float32)
    a["u_acc1"], a["u_acc2"] = 0.0, 0.0  # always accepted
    vn = a["g"].astype(np.float64) * p.vnT + np.array([p.vn0x, p.vn0y, p.vn0z])
    vc = (vi * p.mi + vn * p.mn) / (p.mi + p.mn)
    vcr = np.linalg.norm(vi - vc)
    n = 96
    us = ((np.arange(n) + 0.5) / n).astype(np.float32)
    es = []
    for uz in us:
        for uphi in us:
            a["u_z"], a["u_phi"] = uz, uphi
            vf, acc, fail = oracle.collide_one(p, a, vi)
            assert acc and not fail
            es.append((vf - vc) / vcr)
    e = np.array(es)
    np.testing.assert_allclose(np.linalg.norm(e, axis=1), 1.0, rtol=1e-12)
    np.testing.assert_allclose(e.mean(axis=0), 0.0, atol=2e-3)
    np.testing.assert_allclose(e.T @ e / len(e), np.eye(3) / 3.0, atol=2e-3)


def test_scatter_nr_log_replays_through_predrawn_and_counts_draws():
    p = _col_params(nui=0.05)
    rng = np.random.default_rng(10)
    n = 3000
    vel0 = rng.normal(size=(n, 3))
    nr = oracle.NR()
    vel_a = vel0.copy()
    rc, nl, ns, nf, log = oracle.scatter_elastic_nr(p, 1.0, vel_a, nr, log_cap=100000)
    Nc = oracle.collision_count(p, n, 1.0)
    assert rc == 0 and ns == Nc and nl == len(log)
    vel_b = vel0.copy()
    rc2, used, ns2, nf2 = oracle.scatter_elastic_predrawn(p, 1.0, vel_b, log, Nc)
    assert rc2 == 0 and used == nl and ns2 == ns and nf2 == nf
    np.testing.assert_array_equal(vel_a, vel_b)
    # draw accounting: 3 ran3 per attempt + 2 more per accepted attempt
    accepted = int((log["u_z"] >= 0).sum())
    n_ran3 = 3 * nl + 2 * accepted
    ref = oracle.NR()
    s = ctypes.c_long(-12345)
    L = oracle.lib()
    for _ in range(n_ran3):
        L.oracle_ran3(ctypes.byref(ref.state), ctypes.byref(s))
    assert ref.state.r3_inext == nr.state.r3_inext
    assert list(ref.state.r3_ma) == list(nr.state.r3_ma)


def test_scatter_abort_after_ten_refusals():
    p = _col_params(mn=1.0)
    n = 100
    vel = np.zeros((n, 3))
    att = _random_attempts(50, np.random.default_rng(11))
    att["u_acc1"] = 0.0
    att["g"] = 1e3  # neutrals so fast that every result exceeds 10 thermal speeds
    rc, used, ns, nf = oracle.scatter_elastic_predrawn(p, 1.0, vel, att, 20)
    assert rc == 2 and nf == 10 and ns == 0 and used == 10


# -------------------------------------------------------------------------- operator
def _moments_from_density(n):
    m = np.zeros(n.shape + (4,))
    m[..., 0] = n
    return m


def test_lhs_zero_row_and_column_sums():
    import scipy.sparse as sp
    p = params(8, kappa_e=4.0)
    rng = np.random.default_rng(12)
    n = 64 + rng.normal(scale=5, size=(8, 8, 8))
    rowptr, cols, vals = oracle.lhs_csr(p, _moments_from_density(n))
    A = sp.csr_matrix((vals, cols, rowptr), shape=(512, 512))
    scale = np.abs(vals).max()
    assert np.abs(np.asarray(A.sum(axis=1))).max() < 1e-12 * scale
    assert np.abs(np.asarray(A.sum(axis=0))).max() < 1e-12 * scale


def _fft_symbol(p, n0):
    kx = 2 * np.pi * np.fft.fftfreq(p.Nx)
    ky = 2 * np.pi * np.fft.fftfreq(p.Ny)
    kz = 2 * np.pi * np.fft.fftfreq(p.Nz)
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    lam = (2 * np.cos(KX) - 2) / p.dx ** 2 + (2 * np.cos(KY) - 2) / p.dy ** 2 + \
        (1 + p.kappa_e ** 2) * (2 * np.cos(KZ) - 2) / p.dz ** 2
    return p.scale * n0 * lam


def test_lhs_uniform_density_is_anisotropic_laplacian_fft():
    p = oracle.setup_context(**dict(configs.base_cli(10, kappa_e=3.0), Nx=10, Ny=6, Nz=8, dy=0.5))
    n0 = 64.0
    m = _moments_from_density(np.full((p.Nz, p.Ny, p.Nx), n0))
    x = np.random.default_rng(13).normal(size=(p.Nz, p.Ny, p.Nx))
    y = oracle.lhs_apply(p, m, x)
    y_fft = np.real(np.fft.ifftn(np.fft.fftn(x) * _fft_symbol(p, n0)))
    np.testing.assert_allclose(y, y_fft, rtol=1e-10, atol=1e-10 * np.abs(y).max())


def test_lhs_converges_to_continuum_operator():
    """A phi -> scale*[d_x(n d_x phi) + d_y(n d_y phi) + (1+k^2) d_z(n d_z phi)
    + k (d_y n d_x phi - d_x n d_y phi)] at second order."""
    kappa = 2.0
    errs = []
    for N in (16, 32):
        cli = configs.base_cli(N, kappa_e=kappa)
        cli.update(dx=1.0 / N, dy=1.0 / N, dz=1.0 / N)
        p = oracle.setup_context(**cli)
        g = np.arange(N) / N
        Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
        tp = 2 * np.pi
        n = 2 + np.sin(tp * X) * np.cos(tp * Y) + 0.5 * np.sin(tp * Z)
        phi = np.cos(tp * X + 0.3) * np.sin(tp * Y) + np.cos(tp * Z)
        nx = tp * np.cos(tp * X) * np.cos(tp * Y)
        ny = -tp * np.sin(tp * X) * np.sin(tp * Y)
        nz = 0.5 * tp * np.cos(tp * Z)
        px = -tp * np.sin(tp * X + 0.3) * np.sin(tp * Y)
        py = tp * np.cos(tp * X + 0.3) * np.cos(tp * Y)
        pz = -tp * np.sin(tp * Z)
        pxx = -tp ** 2 * np.cos(tp * X + 0.3) * np.sin(tp * Y)
        pyy = -tp ** 2 * np.cos(tp * X + 0.3) * np.sin(tp * Y)
        pzz = -tp ** 2 * np.cos(tp * Z)
        cont = p.scale * (n * pxx + nx * px + n * pyy + ny * py + (1 + kappa ** 2) * (n * pzz + nz * pz)
                          + kappa * (ny * px - nx * py))
        y = oracle.lhs_apply(p, _moments_from_density(n), phi)
        errs.append(np.abs(y - cont).max() / np.abs(cont).max())
    # second-order convergence (a wrong term would leave an O(1) error)
    assert errs[1] < 0.3 * errs[0] and errs[1] < 1e-2


def test_lhs_sparse_solve_matches_fft_for_uniform_density():
    p = params(8, kappa_e=4.0)
    n0 = 50.0
    m = _moments_from_density(np.full((8, 8, 8), n0))
    b = np.random.default_rng(14).normal(size=(8, 8, 8))
    b -= b.mean()
    x = oracle.solve_sparse(p, m, b)
    sym = _fft_symbol(p, n0)
    bh = np.fft.fftn(b)
    sym[0, 0, 0] = 1.0
    xh = bh / sym
    xh[0, 0, 0] = 0.0
    np.testing.assert_allclose(x, np.real(np.fft.ifftn(xh)), rtol=1e-9, atol=1e-12)


def test_laplacian_operator_fft():
    p = oracle.setup_context(Nx=6, Ny=8, Nz=10, dx=0.5, dy=1.0, dz=2.0)
    x = np.random.default_rng(15).normal(size=(10, 8, 6))
    y = oracle.laplacian_apply(p, x)
    kx = 2 * np.pi * np.fft.fftfreq(6)
    ky = 2 * np.pi * np.fft.fftfreq(8)
    kz = 2 * np.pi * np.fft.fftfreq(10)
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    lam = (2 * np.cos(KX) - 2) / 0.25 + (2 * np.cos(KY) - 2) + (2 * np.cos(KZ) - 2) / 4.0
    np.testing.assert_allclose(y, np.real(np.fft.ifftn(np.fft.fftn(x) * lam)), atol=1e-12)


# ------------------------------------------------------------------------------ RHS
def test_rhs_closed_form_sinusoids():
    N = 16
    cli = configs.base_cli(N, kappa_e=3.0, E0=0.7)
    cli.update(dx=0.5, dy=0.25, dz=1.0)
    p = oracle.setup_context(**cli)
    i = np.arange(N)
    Z, Y, X = np.meshgrid(i * p.dz, i * p.dy, i * p.dx, indexing="ij")
    kx, ky, kz = 2 * np.pi / p.Lx, 2 * np.pi * 2 / p.Ly, 2 * np.pi / p.Lz
    a, b, c = 0.3, 0.2, 0.1
    n = 5 + a * np.sin(kx * X) + b * np.sin(ky * Y) + c * np.sin(kz * Z)
    G = np.stack([np.cos(kx * X), np.sin(ky * Y), 2 * np.cos(kz * Z)], axis=-1)
    m = np.concatenate([n[..., None], G], axis=-1)
    rhs = oracle.full_rhs(p, m)
    kap = p.kappa_e
    cth = p.gammaT * oracle.KB * p.Te / oracle.Q
    cG = (1 + kap ** 2) * p.me * p.nue / oracle.Q
    D1 = lambda k, h: np.sin(k * h) / h          # centred first difference of sin/cos
    D2 = lambda k, h: -4 * np.sin(k * h / 2) ** 2 / h ** 2
    dndx = a * np.cos(kx * X) * D1(kx, p.dx)
    dndy = b * np.cos(ky * Y) * D1(ky, p.dy)
    lap = a * np.sin(kx * X) * D2(kx, p.dx) + b * np.sin(ky * Y) * D2(ky, p.dy) + \
        (1 + kap ** 2) * c * np.sin(kz * Z) * D2(kz, p.dz)
    div = -np.sin(kx * X) * D1(kx, p.dx) + np.cos(ky * Y) * D1(ky, p.dy) - 2 * np.sin(kz * Z) * D1(kz, p.dz)
    exp = p.scale * ((dndy - kap * dndx) * p.E0 + cth * lap + cG * div)
    exp -= exp.mean()
    np.testing.assert_allclose(rhs, exp, rtol=1e-9, atol=1e-12 * np.abs(exp).max())
    assert abs(rhs.mean()) < 1e-14 * np.abs(rhs).max()


def test_sinusoidal_rhs_closed_form():
    p = oracle.setup_context(Nx=8, Ny=8, Nz=8, x1=1.0, y1=1.0, z1=1.0, n0=2.0, B0=0.0)
    b = oracle.sinusoidal_rhs(p)
    c = np.cos(2 * np.pi * (np.arange(8) + 0.5) / 8)
    exp = 2.0 * p.scale * c[:, None, None] * c[None, :, None] * c[None, None, :]
    np.testing.assert_allclose(b, exp - exp.mean(), atol=1e-14)


# ---- initial position distributions (src/positions.c, src/random.c:233-289) ----------

# Numerical Recipes (2nd edition, section 7.7) sobseq table: degree s, interior
# coefficients of the primitive polynomial (bit s-1-l holds a_l), initial odd m_1..m_s.
_NR_SOBOL = [(1, 0, [1]), (2, 1, [1, 1]), (3, 1, [1, 3, 7]), (3, 2, [1, 3, 3]), (4, 1, [1, 1, 3, 13]),
             (4, 4, [1, 1, 5, 9])]


def _sobol_direct(n, dim, reference_rule=False):
    """Points 1..n of one Sobol dimension built from the direction numerators m_j and the
    Gray code of the point index (Bratley & Fox, ACM TOMS 14 (1988) 88-100):
    m_j = 2 a_1 m_{j-1} ^ ... ^ 2^(s-1) a_(s-1) m_(j-s+1) ^ 2^s m_(j-s) ^ m_(j-s).
    reference_rule: every set a_l contributes 2 m_(j-1), as src/random.c:263-266 does."""
    s, ip, m0 = _NR_SOBOL[dim - 1]
    m = [0] + list(m0)
    for j in range(s + 1, 31):
        v = (m[j - s] << s) ^ m[j - s]
        for l in range(1, s):
            if (ip >> (s - 1 - l)) & 1:
                v ^= (m[j - 1] << 1) if reference_rule else (m[j - l] << l)
        m.append(v)
    vdir = [0] + [m[j] << (30 - j) for j in range(1, 31)]
    out = np.empty(n)
    for i in range(1, n + 1):
        g, x, b = i ^ (i >> 1), 0, 1
        while g:
            if g & 1:
                x ^= vdir[b]
            g >>= 1
            b += 1
        out[i - 1] = x / 2.0 ** 30
    return out


def test_sobol_matches_published_construction():
    """Dimensions whose polynomial has at most one interior coefficient at l = 1 (1, 2, 4,
    6) follow the published recurrence exactly; dimensions 3 and 5 follow the reference's
    own rule (it XORs m_(j-1) where the published algorithm XORs m_(j-l)) and differ from
    the published sequence."""
    n = 4096
    pts = oracle.sobol_points(n, 6)
    for d in (1, 2, 4, 6):
        np.testing.assert_array_equal(pts[:, d - 1], _sobol_direct(n, d))
    for d in (3, 5):
        np.testing.assert_array_equal(pts[:, d - 1], _sobol_direct(n, d, reference_rule=True))
        assert (pts[:, d - 1] != _sobol_direct(n, d)).any()
    np.testing.assert_array_equal(pts[0], [0.5] * 6)


def test_sobol_dimensions_are_stratified():
    """Every direction numerator is odd, so the first 2^m points of each dimension put
    exactly one point in each interval [i 2^-m, (i+1) 2^-m) (the point x = 0 of index 0
    is not emitted, so 2^m - 1 points leave only [0, 2^-m) empty)."""
    m = 12
    pts = oracle.sobol_points(2 ** m - 1, 6)
    for d in range(6):
        cnt = np.bincount((pts[:, d] * 2 ** m).astype(int), minlength=2 ** m)
        assert cnt[0] == 0 and (cnt[1:] == 1).all()


def test_rand48_matches_glibc_drand48():
    """The restated rand48 stream equals glibc's srand48/drand48, value for value."""
    libc = ctypes.CDLL("libc.so.6")
    libc.drand48.restype = ctypes.c_double
    libc.srand48.argtypes = [ctypes.c_long]
    for seed in (7, 0, 123456789, -5):
        libc.srand48(seed)
        ref = np.array([libc.drand48() for _ in range(2000)])
        np.testing.assert_array_equal(oracle.rand48_stream(seed, 2000), ref)


def test_sinusoidal_rejection_statistics_and_stream_order():
    """Rejection against 1 + 0.25 sin(2 pi x): acceptance 1/1.25 of the trials, x
    distributed as the density, y and z uniform; the first accepted ions are those of a
    glibc drand48 stream read as (x, y, z, test) per trial."""
    N = 16
    p = oracle.setup_context(Nx=N, Ny=N, Nz=N, dx=0.5, dy=0.5, dz=0.5)
    n = 100000
    pos, trials = oracle.sinusoidal_positions(p, n)
    assert abs(trials / n - 1.25) < 0.015
    L = np.array([p.Lx, p.Ly, p.Lz])
    u = pos / L
    assert (u >= 0).all() and (u < 1).all()
    nb = 20
    edges = np.linspace(0, 1, nb + 1)
    exp = n * (np.diff(edges) - 0.25 / (2 * np.pi) * np.diff(np.cos(2 * np.pi * edges)))
    hx = np.histogram(u[:, 0], edges)[0]
    assert ((hx - exp) ** 2 / exp).sum() < 45  # chi-square, 19 degrees of freedom
    for d in (1, 2):
        h = np.histogram(u[:, d], edges)[0]
        assert ((h - n / nb) ** 2 / (n / nb)).sum() < 45
    libc = ctypes.CDLL("libc.so.6")
    libc.drand48.restype = ctypes.c_double
    libc.srand48(7)
    acc = []
    while len(acc) < 50:
        s = [libc.drand48() for _ in range(3)]
        w = 1.0 + 0.25 * math.sin(2 * math.pi * s[0])
        if w * (1.0 / 1.25) > libc.drand48():
            acc.append(np.array(s) * L)
    np.testing.assert_array_equal(pos[:50], np.array(acc))


def test_sobol_positions_scale_the_sequence():
    p = oracle.setup_context(Nx=8, Ny=4, Nz=16, dx=0.5, dy=2.0, dz=0.25)
    pos = oracle.sobol_positions(p, 1000)
    np.testing.assert_array_equal(pos, oracle.sobol_points(1000, 3) * np.array([p.Lx, p.Ly, p.Lz]))


@pytest.mark.parametrize("shape,rtol", [((12, 12, 12), 1e-10), ((16, 8, 12), 1e-10), ((10, 14, 6), 1e-8)])
def test_cpu_gmres_multigrid_solve(shape, rtol):
    """The CPU baseline solver (GMRES(30) + geometric multigrid, one core) meets rtol on the
    true residual of the assembled operator and agrees with scipy's sparse direct solve of
    the same singular system (mean removed) to the accuracy rtol allows."""
    Nx, Ny, Nz = shape
    cli = configs.base_cli(8, E0=0.05)
    cli.update(Nx=Nx, Ny=Ny, Nz=Nz, dz=0.7)
    po = oracle.setup_context(**cli)
    rng = np.random.default_rng(3)
    m = np.empty((Nz, Ny, Nx, 4))
    m[..., 0] = 50.0 * np.exp(0.2 * rng.normal(size=(Nz, Ny, Nx)))
    m[..., 1:] = rng.normal(size=(Nz, Ny, Nx, 3))
    b = oracle.full_rhs(po, m)
    x, it, rn, ok = oracle.solve_gmres_mg(po, m, b, rtol)
    assert ok and it > 0
    r = b - oracle.lhs_apply(po, m, x)
    assert np.linalg.norm(r) <= 1.0001 * rtol * np.linalg.norm(b)
    assert abs(x.mean()) <= 1e-12 * np.abs(x).max()
    xs = oracle.solve_sparse(po, m, b)
    assert np.abs(x - xs).max() <= 1e4 * rtol * np.abs(xs).max()
