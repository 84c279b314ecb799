"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle restates ECSPRT's hot-path arithmetic in C (ecsprt_oracle.c).  Only the
test suite, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg import
this package, and only as the checker or the timed CPU baseline; the product
(``ecsprt_amd``) never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libecsprt_oracle.so")

Q = 1.6022e-19
MP = 1.6726e-27
ME = 9.1094e-31
KB = 1.3807e-23


class OracleCLI(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("Nx", "Ny", "Nz")] + [
        (n, ctypes.c_double) for n in (
            "dx", "dy", "dz", "x0", "y0", "z0", "x1", "y1", "z1", "B0", "E0", "n0",
            "qi", "mi", "nui", "vi0x", "vi0y", "vi0z", "viTx", "viTy", "viTz", "Ti",
            "nue", "Te", "veTx", "veTy", "veTz", "mn", "vn0x", "vn0y", "vn0z", "Tn")]


class OracleParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("Nx", "Ny", "Nz")] + [
        (n, ctypes.c_double) for n in (
            "dx", "dy", "dz", "Lx", "Ly", "Lz", "x0", "y0", "z0", "x1", "y1", "z1",
            "B0", "E0", "n0",
            "qi", "mi", "nui", "vi0x", "vi0y", "vi0z", "vi0", "viTx", "viTy", "viTz", "viT",
            "Ti", "Omega_i", "kappa_i",
            "qe", "me", "nue", "Te", "veT", "Omega_e", "kappa_e",
            "mn", "vn0x", "vn0y", "vn0z", "vn0", "vnT", "Tn", "scale", "gammaT")]


class OracleAttempt(ctypes.Structure):
    _fields_ = [("u_ip", ctypes.c_float), ("g", ctypes.c_float * 3),
                ("u_acc1", ctypes.c_float), ("u_acc2", ctypes.c_float),
                ("u_z", ctypes.c_float), ("u_phi", ctypes.c_float)]


ATTEMPT_DTYPE = np.dtype([("u_ip", "<f4"), ("g", "<f4", (3,)), ("u_acc1", "<f4"),
                          ("u_acc2", "<f4"), ("u_z", "<f4"), ("u_phi", "<f4")])
assert ATTEMPT_DTYPE.itemsize == ctypes.sizeof(OracleAttempt)


class NRState(ctypes.Structure):
    _fields_ = [("r1_iy", ctypes.c_long), ("r1_iv", ctypes.c_long * 32),
                ("r3_inext", ctypes.c_int), ("r3_inextp", ctypes.c_int), ("r3_iff", ctypes.c_int),
                ("r3_ma", ctypes.c_long * 56), ("g_iset", ctypes.c_int), ("g_gset", ctypes.c_float)]


_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_L = ctypes.POINTER(ctypes.c_long)
_lib = None


class SobolState(ctypes.Structure):
    _fields_ = [("in_", ctypes.c_ulong), ("ix", ctypes.c_ulong * 7), ("iv", ctypes.c_ulong * 181),
                ("fac", ctypes.c_double)]


class Rand48(ctypes.Structure):
    _fields_ = [("x", ctypes.c_ulonglong)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER(OracleParams)
        sig = {
            "oracle_setup_context": (None, [ctypes.POINTER(OracleCLI), P]),
            "oracle_difference_vector3d": (None, [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_double, ctypes.c_double, ctypes.c_double, _D]),
            "oracle_boris_bz3d": (None, [P, ctypes.c_double, ctypes.c_long, _D, _D, _D]),
            "oracle_update_positions": (None, [ctypes.c_double, ctypes.c_long, _D, _D]),
            "oracle_apply_bc": (None, [P, ctypes.c_long, _D]),
            "oracle_count_outside": (ctypes.c_long, [P, ctypes.c_long, _D]),
            "oracle_collect_moments": (None, [P, ctypes.c_long, _D, _D, _D]),
            "oracle_cell_counts": (None, [P, ctypes.c_long, _D, _I]),
            "oracle_nr_init": (None, [ctypes.POINTER(NRState)]),
            "oracle_ran1": (ctypes.c_float, [ctypes.POINTER(NRState), _L]),
            "oracle_ran3": (ctypes.c_float, [ctypes.POINTER(NRState), _L]),
            "oracle_gasdev": (ctypes.c_float, [ctypes.POINTER(NRState), _L]),
            "oracle_getseed": (ctypes.c_long, [ctypes.c_int]),
            "oracle_nr_stream": (ctypes.c_int, [ctypes.c_int, ctypes.c_long, ctypes.c_long,
                                                ctypes.POINTER(ctypes.c_float)]),
            "oracle_collision_count": (ctypes.c_int, [P, ctypes.c_long, ctypes.c_double]),
            "oracle_scatter_elastic_nr": (ctypes.c_int, [P, ctypes.c_double, ctypes.c_long, _D,
                                                         ctypes.POINTER(NRState), ctypes.c_long,
                                                         ctypes.c_void_p, ctypes.c_long, _L, _I, _I]),
            "oracle_scatter_elastic_predrawn": (ctypes.c_int, [P, ctypes.c_double, ctypes.c_long, _D,
                                                               ctypes.c_void_p, ctypes.c_long, ctypes.c_int,
                                                               _L, _I, _I]),
            "oracle_collide_one": (None, [P, ctypes.POINTER(OracleAttempt), _D, _D, _I, _I]),
            "oracle_full_lhs3d_coeffs": (None, [P, _D, _D]),
            "oracle_full_lhs3d_csr": (None, [P, _D, _I, _I, _D]),
            "oracle_full_lhs3d_apply": (None, [P, _D, _D, _D]),
            "oracle_laplacian_lhs3d_apply": (None, [P, _D, _D]),
            "oracle_full_rhs3d": (None, [P, _D, _D]),
            "oracle_sinusoidal_rhs3d": (None, [P, _D]),
            "oracle_remove_mean": (None, [ctypes.c_long, _D]),
            "oracle_sobseq_init": (None, [ctypes.POINTER(SobolState)]),
            "oracle_sobseq_next": (ctypes.c_int, [ctypes.POINTER(SobolState), ctypes.c_int, _D]),
            "oracle_rand48_seed": (None, [ctypes.POINTER(Rand48), ctypes.c_long]),
            "oracle_rand48_next": (ctypes.c_double, [ctypes.POINTER(Rand48)]),
            "oracle_sobol_positions": (ctypes.c_int, [P, ctypes.c_long, _D]),
            "oracle_sinusoidal_positions": (ctypes.c_long, [P, ctypes.c_long, _D]),
            "oracle_sinusoidal_positions_nd": (ctypes.c_long, [P, ctypes.c_int, ctypes.c_long, _D]),
            "oracle_collide_one_nd": (None, [P, ctypes.c_int, ctypes.POINTER(OracleAttempt), _D, _D, _I, _I]),
            "oracle_scatter_elastic_nr_nd": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_double, ctypes.c_long, _D,
                                                            ctypes.POINTER(NRState), ctypes.c_long,
                                                            ctypes.c_void_p, ctypes.c_long, _L, _I, _I]),
            "oracle_scatter_elastic_predrawn_nd": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_double, ctypes.c_long,
                                                                  _D, ctypes.c_void_p, ctypes.c_long, ctypes.c_int,
                                                                  _L, _I, _I]),
            "oracle_difference_vector2d": (None, [_D, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                  _D]),
            "oracle_boris_bz2d": (None, [P, ctypes.c_double, ctypes.c_long, _D, _D, _D]),
            "oracle_update_positions2d": (None, [ctypes.c_double, ctypes.c_long, _D, _D]),
            "oracle_apply_bc2d": (None, [P, ctypes.c_long, _D]),
            "oracle_collect_moments2d": (None, [P, ctypes.c_long, _D, _D, _D]),
            "oracle_full_lhs2d_coeffs": (None, [P, _D, _D]),
            "oracle_full_lhs2d_csr": (None, [P, _D, _I, _I, _D]),
            "oracle_full_lhs2d_apply": (None, [P, _D, _D, _D]),
            "oracle_full_rhs2d": (None, [P, _D, _D]),
            "oracle_solve_gmres_mg": (ctypes.c_int, [P, _D, _D, _D, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                                     _I, _D]),
            "oracle_particle_steps": (ctypes.c_int, [P, ctypes.c_double, ctypes.c_int, ctypes.c_long,
                                                     _D, _D, _D, _D, ctypes.POINTER(NRState),
                                                     ctypes.c_int, _D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def setup_context(**cli_kw):
    """SetUpContext + NormalizeGrid on reference defaults (src/parameters.c:25-346)."""
    cli = OracleCLI()
    defaults = dict(Nx=-1, Ny=-1, Nz=-1, dx=-1.0, dy=-1.0, dz=-1.0, x0=0.0, y0=0.0, z0=0.0,
                    x1=1.0, y1=1.0, z1=1.0, B0=0.0, E0=0.0, n0=1.0, qi=Q, mi=MP, nui=1.0,
                    vi0x=0.0, vi0y=0.0, vi0z=0.0, viTx=0.0, viTy=0.0, viTz=0.0, Ti=0.0,
                    nue=1.0, Te=0.0, veTx=0.0, veTy=0.0, veTz=0.0, mn=None,
                    vn0x=0.0, vn0y=0.0, vn0z=0.0, Tn=-1.0)
    defaults.update(cli_kw)
    if defaults["mn"] is None:
        defaults["mn"] = defaults["mi"]
    for k, v in defaults.items():
        setattr(cli, k, v)
    p = OracleParams()
    lib().oracle_setup_context(ctypes.byref(cli), ctypes.byref(p))
    return p


def params_dict(p):
    return {name: getattr(p, name) for name, _ in p._fields_}


def difference_vector3d(phi, x, y, z):
    Nz, Ny, Nx = phi.shape
    f = np.zeros(3)
    lib().oracle_difference_vector3d(_dp(phi), Nx, Ny, Nz, x, y, z, _dp(f))
    return f


def boris(p, dt, pos, vel, phi):
    lib().oracle_boris_bz3d(ctypes.byref(p), dt, pos.shape[0], _dp(pos), _dp(vel), _dp(phi))


def update_positions(dt, pos, vel):
    lib().oracle_update_positions(dt, pos.shape[0], _dp(pos), _dp(vel))


def apply_bc(p, pos):
    lib().oracle_apply_bc(ctypes.byref(p), pos.shape[0], _dp(pos))


def count_outside(p, pos):
    return lib().oracle_count_outside(ctypes.byref(p), pos.shape[0], _dp(pos))


def collect_moments(p, pos, vel):
    m = np.zeros((p.Nz, p.Ny, p.Nx, 4))
    lib().oracle_collect_moments(ctypes.byref(p), pos.shape[0], _dp(pos), _dp(vel), _dp(m))
    return m


def cell_counts(p, pos):
    c = np.zeros((p.Nz, p.Ny, p.Nx), dtype=np.int32)
    lib().oracle_cell_counts(ctypes.byref(p), pos.shape[0], _dp(pos),
                             c.ctypes.data_as(_I))
    return c


class NR:
    """Reference RNG static state (src/random.c) held explicitly."""

    def __init__(self):
        self.state = NRState()
        lib().oracle_nr_init(ctypes.byref(self.state))

    def ran1(self, seed):
        s = ctypes.c_long(seed.value if isinstance(seed, ctypes.c_long) else seed)
        return lib().oracle_ran1(ctypes.byref(self.state), ctypes.byref(s)), s

    def stream(self, fn, seed, n):
        s = ctypes.c_long(seed)
        f = {"ran1": lib().oracle_ran1, "ran3": lib().oracle_ran3, "gasdev": lib().oracle_gasdev}[fn]
        out = np.empty(n, dtype=np.float32)
        for i in range(n):
            out[i] = f(ctypes.byref(self.state), ctypes.byref(s))
        return out, s.value


def nr_stream(generator, seed, n):
    """n values of ran1, ran3 or gasdev from fresh state (the bulk form of NR.stream)."""
    which = {"ran1": 0, "ran3": 1, "gasdev": 2}[generator]
    out = np.empty(n, dtype=np.float32)
    if lib().oracle_nr_stream(which, seed, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))):
        raise ValueError(generator)
    return out


# ---- two-dimensional variants (cli.ndim = 2) ----------------------------------------

def difference_vector2d(phi, x, y):
    Ny, Nx = phi.shape
    f = np.zeros(2)
    lib().oracle_difference_vector2d(_dp(phi), Nx, Ny, x, y, _dp(f))
    return f


def boris2d(p, dt, pos, vel, phi):
    lib().oracle_boris_bz2d(ctypes.byref(p), dt, pos.shape[0], _dp(pos), _dp(vel), _dp(phi))


def update_positions2d(dt, pos, vel):
    lib().oracle_update_positions2d(dt, pos.shape[0], _dp(pos), _dp(vel))


def apply_bc2d(p, pos):
    lib().oracle_apply_bc2d(ctypes.byref(p), pos.shape[0], _dp(pos))


def collect_moments2d(p, pos, vel):
    m = np.zeros((p.Ny, p.Nx, 4))
    lib().oracle_collect_moments2d(ctypes.byref(p), pos.shape[0], _dp(pos), _dp(vel), _dp(m))
    return m


def lhs2d_coeffs(p, moments):
    c = np.zeros((p.Ny, p.Nx, 9))
    lib().oracle_full_lhs2d_coeffs(ctypes.byref(p), _dp(np.ascontiguousarray(moments)), _dp(c))
    return c


def lhs2d_csr(p, moments):
    import scipy.sparse as sp
    n = p.Nx * p.Ny
    rowptr = np.zeros(n + 1, dtype=np.int32)
    cols = np.zeros(9 * n, dtype=np.int32)
    vals = np.zeros(9 * n)
    lib().oracle_full_lhs2d_csr(ctypes.byref(p), _dp(np.ascontiguousarray(moments)), rowptr.ctypes.data_as(_I),
                                cols.ctypes.data_as(_I), _dp(vals))
    return sp.csr_matrix((vals, cols, rowptr), shape=(n, n))


def lhs2d_apply(p, moments, x):
    y = np.zeros((p.Ny, p.Nx))
    lib().oracle_full_lhs2d_apply(ctypes.byref(p), _dp(np.ascontiguousarray(moments)),
                                  _dp(np.ascontiguousarray(x)), _dp(y))
    return y


def full_rhs2d(p, moments):
    b = np.zeros((p.Ny, p.Nx))
    lib().oracle_full_rhs2d(ctypes.byref(p), _dp(np.ascontiguousarray(moments)), _dp(b))
    return b


def _attempt(att_row):
    a = OracleAttempt()
    a.u_ip = float(att_row["u_ip"])
    for d in range(3):
        a.g[d] = float(att_row["g"][d])
    a.u_acc1, a.u_acc2 = float(att_row["u_acc1"]), float(att_row["u_acc2"])
    a.u_z, a.u_phi = float(att_row["u_z"]), float(att_row["u_phi"])
    return a


def collide_one_nd(p, ndim, att_row, vi):
    """One attempt in ndim (2 or 3) dimensions: (vf[3], accepted, refused)."""
    a = _attempt(att_row)
    v = np.zeros(3)
    v[:ndim] = np.asarray(vi, dtype=np.float64)[:ndim]
    vf = np.zeros(3)
    acc, fail = ctypes.c_int(), ctypes.c_int()
    lib().oracle_collide_one_nd(ctypes.byref(p), ndim, ctypes.byref(a), _dp(v), _dp(vf),
                                ctypes.byref(acc), ctypes.byref(fail))
    return vf, bool(acc.value), bool(fail.value)


def scatter_elastic_nr_nd(p, ndim, dt, vel, nr, seed=None, log_cap=0):
    """ScatterElastic on vel[np, ndim] from the reference's NR stream, as
    scatter_elastic_nr."""
    if seed is None:
        seed = lib().oracle_getseed(0)
    log = np.zeros(max(log_cap, 1), dtype=ATTEMPT_DTYPE)
    nl, ns, nf = ctypes.c_long(), ctypes.c_int(), ctypes.c_int()
    rc = lib().oracle_scatter_elastic_nr_nd(ctypes.byref(p), ndim, dt, vel.shape[0], _dp(vel),
                                            ctypes.byref(nr.state), seed,
                                            log.ctypes.data if log_cap else None, log_cap,
                                            ctypes.byref(nl), ctypes.byref(ns), ctypes.byref(nf))
    return rc, nl.value, ns.value, nf.value, log[:min(nl.value, log_cap)]


def scatter_elastic_predrawn_nd(p, ndim, dt, vel, att, Nc):
    att = np.ascontiguousarray(att, dtype=ATTEMPT_DTYPE)
    nu, ns, nf = ctypes.c_long(), ctypes.c_int(), ctypes.c_int()
    rc = lib().oracle_scatter_elastic_predrawn_nd(ctypes.byref(p), ndim, dt, vel.shape[0], _dp(vel),
                                                  att.ctypes.data, att.shape[0], Nc,
                                                  ctypes.byref(nu), ctypes.byref(ns), ctypes.byref(nf))
    return rc, nu.value, ns.value, nf.value


def sobol_points(n, dim=3):
    """The first n points of the reference's Sobseq (src/random.c:233-289), shape (n, dim)."""
    st = SobolState()
    lib().oracle_sobseq_init(ctypes.byref(st))
    out = np.empty((n, dim))
    x = np.empty(dim)
    for i in range(n):
        if lib().oracle_sobseq_next(ctypes.byref(st), dim, _dp(x)):
            raise RuntimeError("Sobol sequence exhausted")
        out[i] = x
    return out


def rand48_stream(seed, n):
    """n values of the rand48 generator seeded as srand48(seed)."""
    r = Rand48()
    lib().oracle_rand48_seed(ctypes.byref(r), seed)
    return np.array([lib().oracle_rand48_next(ctypes.byref(r)) for _ in range(n)])


def sobol_positions(p, np_):
    """SobolDistribution on one rank (src/positions.c:139-234), pos[np, 3]."""
    pos = np.empty((np_, 3))
    if lib().oracle_sobol_positions(ctypes.byref(p), np_, _dp(pos)):
        raise RuntimeError("Sobol sequence exhausted")
    return pos


def sinusoidal_positions(p, np_, ndim=3):
    """Rejection of SinusoidalDistribution on one rank (src/positions.c:236-388):
    (pos[np, ndim], trials)."""
    pos = np.empty((np_, ndim))
    trials = lib().oracle_sinusoidal_positions_nd(ctypes.byref(p), ndim, np_, _dp(pos))
    return pos, trials


def collision_count(p, np_, dt):
    return lib().oracle_collision_count(ctypes.byref(p), np_, dt)


def scatter_elastic_nr(p, dt, vel, nr, seed=None, log_cap=0):
    if seed is None:
        seed = lib().oracle_getseed(0)
    log = np.zeros(max(log_cap, 1), dtype=ATTEMPT_DTYPE)
    nl, ns, nf = ctypes.c_long(), ctypes.c_int(), ctypes.c_int()
    rc = lib().oracle_scatter_elastic_nr(ctypes.byref(p), dt, vel.shape[0], _dp(vel),
                                         ctypes.byref(nr.state), seed,
                                         log.ctypes.data if log_cap else None, log_cap,
                                         ctypes.byref(nl), ctypes.byref(ns), ctypes.byref(nf))
    return rc, nl.value, ns.value, nf.value, log[:min(nl.value, log_cap)]


def scatter_elastic_predrawn(p, dt, vel, att, Nc):
    att = np.ascontiguousarray(att, dtype=ATTEMPT_DTYPE)
    nu, ns, nf = ctypes.c_long(), ctypes.c_int(), ctypes.c_int()
    rc = lib().oracle_scatter_elastic_predrawn(ctypes.byref(p), dt, vel.shape[0], _dp(vel),
                                               att.ctypes.data, att.shape[0], Nc,
                                               ctypes.byref(nu), ctypes.byref(ns), ctypes.byref(nf))
    return rc, nu.value, ns.value, nf.value


def collide_one(p, att_row, vi):
    a = OracleAttempt()
    a.u_ip = float(att_row["u_ip"])
    for d in range(3):
        a.g[d] = float(att_row["g"][d])
    a.u_acc1, a.u_acc2 = float(att_row["u_acc1"]), float(att_row["u_acc2"])
    a.u_z, a.u_phi = float(att_row["u_z"]), float(att_row["u_phi"])
    vi = np.ascontiguousarray(vi, dtype=np.float64)
    vf = np.zeros(3)
    acc, fail = ctypes.c_int(), ctypes.c_int()
    lib().oracle_collide_one(ctypes.byref(p), ctypes.byref(a), _dp(vi), _dp(vf),
                             ctypes.byref(acc), ctypes.byref(fail))
    return vf, bool(acc.value), bool(fail.value)


def lhs_coeffs(p, moments):
    c = np.zeros((p.Nz * p.Ny * p.Nx, 11))
    lib().oracle_full_lhs3d_coeffs(ctypes.byref(p), _dp(np.ascontiguousarray(moments)), _dp(c))
    return c


def lhs_csr(p, moments):
    n = p.Nx * p.Ny * p.Nz
    rowptr = np.zeros(n + 1, dtype=np.int32)
    cols = np.zeros(11 * n, dtype=np.int32)
    vals = np.zeros(11 * n)
    lib().oracle_full_lhs3d_csr(ctypes.byref(p), _dp(np.ascontiguousarray(moments)),
                                rowptr.ctypes.data_as(_I), cols.ctypes.data_as(_I), _dp(vals))
    return rowptr, cols, vals


def lhs_apply(p, moments, x):
    y = np.zeros_like(x)
    lib().oracle_full_lhs3d_apply(ctypes.byref(p), _dp(np.ascontiguousarray(moments)),
                                  _dp(np.ascontiguousarray(x)), _dp(y))
    return y


def laplacian_apply(p, x):
    y = np.zeros_like(x)
    lib().oracle_laplacian_lhs3d_apply(ctypes.byref(p), _dp(np.ascontiguousarray(x)), _dp(y))
    return y


def full_rhs(p, moments):
    b = np.zeros((p.Nz, p.Ny, p.Nx))
    lib().oracle_full_rhs3d(ctypes.byref(p), _dp(np.ascontiguousarray(moments)), _dp(b))
    return b


def sinusoidal_rhs(p):
    b = np.zeros((p.Nz, p.Ny, p.Nx))
    lib().oracle_sinusoidal_rhs3d(ctypes.byref(p), _dp(b))
    return b


def particle_steps(p, dt, nt, pos, vel, phi, nr, collide=True):
    m = np.zeros((p.Nz, p.Ny, p.Nx, 4))
    t = np.zeros(4)
    rc = lib().oracle_particle_steps(ctypes.byref(p), dt, nt, pos.shape[0], _dp(pos), _dp(vel),
                                     _dp(phi), _dp(m), ctypes.byref(nr.state), int(collide), _dp(t))
    return rc, m, t


def solve_gmres_mg(p, moments, b, rtol, restart=30, maxit=10000):
    """FP64 GMRES(restart) + geometric multigrid on one core (ecsprt_solver.c): A from the
    density of `moments`, zero initial guess, mean removed.  Returns (x, iterations,
    ||b - A x||, converged)."""
    x = np.zeros((p.Nz, p.Ny, p.Nx))
    it, rn = ctypes.c_int(), ctypes.c_double()
    rc = lib().oracle_solve_gmres_mg(ctypes.byref(p), _dp(np.ascontiguousarray(moments)),
                                     _dp(np.ascontiguousarray(b, dtype=np.float64)), _dp(x), rtol, restart, maxit,
                                     ctypes.byref(it), ctypes.byref(rn))
    return x, it.value, rn.value, rc == 0


def solve_sparse(p, moments, b):
    """Exact solve of the assembled operator (scipy sparse LU) for the mean-zero
    solution of the consistent singular system: one equation is replaced by
    phi[0] = 0 (the equations sum to zero) and the mean is removed afterwards."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    n = p.Nx * p.Ny * p.Nz
    rowptr, cols, vals = lhs_csr(p, moments)
    A = sp.csr_matrix((vals, cols, rowptr), shape=(n, n)).tolil()
    A[0, :] = 0.0
    A[0, 0] = 1.0
    rhs = np.ascontiguousarray(b, dtype=np.float64).reshape(-1).copy()
    rhs[0] = 0.0
    x = spla.spsolve(A.tocsc(), rhs)
    x -= x.mean()
    return x.reshape(p.Nz, p.Ny, p.Nx)
