"""Host-side mirror of the reference's hot-path interface over libecsprt_hip.so.

``Context`` holds one device context.  Its methods carry the reference names
and argument meaning:

  push(dt)                 PushFunc  BorisMoverBz3D           src/velocities.c:281
  collide(dt)              CollideFunc ScatterElastic         src/collisions.c:9
  update_velocities(dt)    UpdateVelocities                   src/velocities.c:375
  update_positions(dt)     UpdatePositions + ApplyBCAndMigrate src/positions.c:436
  collect()                CollectFunc Collect3DFluidMoments  src/moments.c:108
  compute_potential()      ComputePotential (KSPSolve)         src/potential.c:6
  step(dt)                 one fused timestep in the order of src/pic.c:246-258

Errors come back as ``HipError`` carrying the library's code, the analogue of a
non-zero PetscErrorCode; the reference's Abort conditions (particle count
changed, ten refused collisions) are such codes.
"""
import ctypes
import os

import numpy as np

from . import _lib


class HipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ecsprt_hip error {code}: {msg}")
        self.code = code


def _check(rc, ctx_ptr=None):
    if rc != _lib.OK:
        msg = _lib.load().ecsprt_hip_last_error(ctx_ptr)
        raise HipError(rc, msg.decode() if msg else "")


def setup_context(**cli_kw):
    """SetUpContext + NormalizeGrid on the reference defaults (src/parameters.c:25-346)."""
    L = _lib.load()
    cli = _lib.CLI()
    defaults = dict(Nx=-1, Ny=-1, Nz=-1, dx=-1.0, dy=-1.0, dz=-1.0, x0=0.0, y0=0.0, z0=0.0,
                    x1=1.0, y1=1.0, z1=1.0, B0=0.0, E0=0.0, n0=1.0, qi=1.6022e-19, mi=1.6726e-27,
                    nui=1.0, vi0x=0.0, vi0y=0.0, vi0z=0.0, viTx=0.0, viTy=0.0, viTz=0.0, Ti=0.0,
                    nue=1.0, Te=0.0, veTx=0.0, veTy=0.0, veTz=0.0, mn=None, vn0x=0.0, vn0y=0.0,
                    vn0z=0.0, Tn=-1.0)
    defaults.update(cli_kw)
    if defaults["mn"] is None:
        defaults["mn"] = defaults["mi"]
    for k, v in defaults.items():
        setattr(cli, k, v)
    p = _lib.Params()
    _check(L.ecsprt_hip_setup_context(ctypes.byref(cli), ctypes.byref(p)))
    return p


def default_options(**kw):
    o = _lib.Options()
    _check(_lib.load().ecsprt_hip_default_options(ctypes.byref(o)))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class DeviceVector:
    """A device vector of doubles owned by the caller (for MATSHELL/PCShell calls)."""

    def __init__(self, n):
        self.n = n
        self.ptr = ctypes.c_void_p()
        _check(_lib.load().ecsprt_hip_dmalloc(ctypes.byref(self.ptr), 8 * n))

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        assert a.size == self.n
        _check(_lib.load().ecsprt_hip_h2d(self.ptr, a.ctypes.data, 8 * self.n))

    def download(self):
        out = np.empty(self.n)
        _check(_lib.load().ecsprt_hip_d2h(out.ctypes.data, self.ptr, 8 * self.n))
        return out

    def free(self):
        if self.ptr:
            _lib.load().ecsprt_hip_dfree(self.ptr)
            self.ptr = ctypes.c_void_p()


class Context:
    def __init__(self, params, options=None):
        self.L = _lib.load()
        self.params = params
        self.options = options if options is not None else default_options()
        self.ptr = ctypes.c_void_p()
        _check(self.L.ecsprt_hip_create(ctypes.byref(params), ctypes.byref(self.options), ctypes.byref(self.ptr)))
        self.ndim = params.ndim if params.ndim in (2, 3) else 3
        # grid arrays: [Nz][Ny][Nx] in 3-D, [Ny][Nx] in 2-D (Nz = 1)
        self.shape = (params.Nz, params.Ny, params.Nx) if self.ndim == 3 else (params.Ny, params.Nx)

    def close(self):
        if self.ptr:
            self.L.ecsprt_hip_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _c(self, rc):
        _check(rc, self.ptr)

    # ---- particles ------------------------------------------------------------
    def set_particles(self, pos, vel):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        assert pos.shape[1:] == (self.ndim,) and vel.shape == pos.shape
        self._c(self.L.ecsprt_hip_set_particles(self.ptr, pos.shape[0], _dp(pos), _dp(vel)))

    def get_particles(self):
        n = self.num_particles
        pos = np.empty((n, self.ndim))
        vel = np.empty((n, self.ndim))
        self._c(self.L.ecsprt_hip_get_particles(self.ptr, _dp(pos), _dp(vel)))
        return pos, vel

    def init_particles(self, np_, seed, amp=0.0, k=0):
        self._c(self.L.ecsprt_hip_init_particles(self.ptr, np_, seed, amp, k))

    def init_reference(self, np_):
        """The reference's own initial conditions (ran3 positions, gasdev velocities,
        seed -12345), generated on the host and binned."""
        self._c(self.L.ecsprt_hip_init_reference(self.ptr, np_))

    def init_distribution(self, np_, pdist):
        """InitializePositions for a PDistType (sobol, normal or sinusoidal; see
        _lib.PDIST_*) followed by the reference's velocities, on the host, binned."""
        if isinstance(pdist, str):
            pdist = {"sobol": _lib.PDIST_SOBOL, "reverse": _lib.PDIST_REVERSE, "normal": _lib.PDIST_NORMAL,
                     "uniform": _lib.PDIST_UNIFORM, "sinusoidal": _lib.PDIST_SINUSOIDAL,
                     "gaussian": _lib.PDIST_GAUSSIAN}[pdist]
        self._c(self.L.ecsprt_hip_init_distribution(self.ptr, np_, pdist))

    @property
    def num_particles(self):
        return self.L.ecsprt_hip_num_particles(self.ptr)

    def cell_counts(self):
        c = np.empty(self.shape, dtype=np.int32)
        self._c(self.L.ecsprt_hip_cell_counts(self.ptr, c.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return c

    # ---- plugin slots and direct entry points -------------------------------------
    def push(self, dt):
        self._c(self.L.ecsprt_hip_push(dt, self.ptr))

    def collide(self, dt):
        self._c(self.L.ecsprt_hip_collide(self.ndim, dt, self.ptr))

    def collect(self):
        self._c(self.L.ecsprt_hip_collect(self.ptr))

    def update_velocities(self, dt):
        self._c(self.L.ecsprt_hip_update_velocities(self.ndim, dt, self.ptr))

    def update_positions(self, dt):
        self._c(self.L.ecsprt_hip_update_positions(self.ndim, dt, self.ptr))

    def compute_potential(self):
        self._c(self.L.ecsprt_hip_compute_potential(self.ptr))

    def step(self, dt):
        self._c(self.L.ecsprt_hip_step(self.ptr, dt))

    def apply_bc_device(self, pos_dev, np_):
        self._c(self.L.ecsprt_hip_apply_bc(np_, pos_dev, None, self.ptr))

    def apply_bc_host(self, pos):
        """BCFunc on a host array pos[np, ndim] in place, as the reference's caller passes
        the swarm's coordinates (src/boundaries.c:67-80)."""
        assert pos.dtype == np.float64 and pos.flags.c_contiguous and pos.shape[1:] == (self.ndim,)
        self._c(self.L.ecsprt_hip_apply_bc(pos.shape[0], pos.ctypes.data, None, self.ptr))

    # ---- linear system ---------------------------------------------------------------
    def lhs(self):
        self._c(self.L.ecsprt_hip_lhs(self.ptr))

    def rhs(self, b_dev=None):
        self._c(self.L.ecsprt_hip_rhs(self.ptr, b_dev.ptr if b_dev is not None else None))

    def matmult(self, x_dev, y_dev):
        self._c(self.L.ecsprt_hip_matmult(self.ptr, x_dev.ptr, y_dev.ptr))

    def pcapply(self, r_dev, z_dev):
        self._c(self.L.ecsprt_hip_pcapply(self.ptr, r_dev.ptr, z_dev.ptr))

    def remove_mean(self, v_dev):
        self._c(self.L.ecsprt_hip_remove_mean(self.ptr, v_dev.ptr))

    def ksp_solve(self, b_dev, x_dev):
        self._c(self.L.ecsprt_hip_ksp_solve(self.ptr, b_dev.ptr, x_dev.ptr))

    # ---- grid vectors --------------------------------------------------------------------
    def get_moments(self):
        m = np.empty(self.shape + (4,))
        self._c(self.L.ecsprt_hip_get_moments(self.ptr, _dp(m)))
        return m

    def set_moments(self, m):
        m = np.ascontiguousarray(m, dtype=np.float64)
        assert m.shape == self.shape + (4,)
        self._c(self.L.ecsprt_hip_set_moments(self.ptr, _dp(m)))

    # ---- on-disk formats (src/file-io.c) ------------------------------------------------------
    def output_fluid_hdf5(self, insert=""):
        """OutputFluidHDF5: "fluid<insert>.hdf" in the working directory."""
        self._c(self.L.ecsprt_hip_output_fluid_hdf5(self.ptr, insert.encode()))

    def output_swarm_binary(self, insert=""):
        """OutputSwarmBinary: "position<insert>.bin" and "velocity<insert>.bin"."""
        self._c(self.L.ecsprt_hip_output_swarm_binary(self.ptr, insert.encode()))

    def load_fluid_quantities(self, flux_scale, inpath):
        """LoadFluidQuantities: density from "density-kji" of inpath, flux = flux_scale * n;
        an empty path switches the forcing to ComputeSinusoidalRHS3D."""
        fs = np.ascontiguousarray(flux_scale, dtype=np.float64)
        assert fs.shape == (3,)
        self._c(self.L.ecsprt_hip_load_fluid_quantities(self.ptr, _dp(fs), os.fsencode(inpath) if inpath else b""))

    def get_potential(self):
        phi = np.empty(self.shape)
        self._c(self.L.ecsprt_hip_get_potential(self.ptr, _dp(phi)))
        return phi

    def set_potential(self, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        assert phi.shape == self.shape
        self._c(self.L.ecsprt_hip_set_potential(self.ptr, _dp(phi)))

    def get_forcing(self):
        b = np.empty(self.shape)
        self._c(self.L.ecsprt_hip_get_forcing(self.ptr, _dp(b)))
        return b

    # ---- collisions ------------------------------------------------------------------------
    def set_attempts(self, att):
        att = np.ascontiguousarray(att)
        assert att.dtype.itemsize == ctypes.sizeof(_lib.Attempt)
        self._c(self.L.ecsprt_hip_set_attempts(self.ptr, att.ctypes.data if att.size else None, att.shape[0]))

    def collision_stats(self):
        Nc, Ns, Nf, na = ctypes.c_int(), ctypes.c_long(), ctypes.c_long(), ctypes.c_long()
        self._c(self.L.ecsprt_hip_collision_stats(self.ptr, ctypes.byref(Nc), ctypes.byref(Ns), ctypes.byref(Nf),
                                                  ctypes.byref(na)))
        return Nc.value, Ns.value, Nf.value, na.value

    # ---- diagnostics ---------------------------------------------------------------------------
    def solver_stats(self):
        it, rn, bn = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        self._c(self.L.ecsprt_hip_solver_stats(self.ptr, ctypes.byref(it), ctypes.byref(rn), ctypes.byref(bn)))
        return it.value, rn.value, bn.value

    def solver_traffic(self):
        """Algorithmic HBM bytes of the last potential solve (forcing included)."""
        b = ctypes.c_double()
        self._c(self.L.ecsprt_hip_solver_traffic(self.ptr, ctypes.byref(b)))
        return b.value

    def stage_times(self):
        ms = np.zeros(6)
        self._c(self.L.ecsprt_hip_stage_times(self.ptr, _dp(ms)))
        return ms

    def synchronize(self):
        self._c(self.L.ecsprt_hip_synchronize(self.ptr))


def collision_count(params, np_, dt):
    return _lib.load().ecsprt_hip_collision_count(ctypes.byref(params), np_, dt)


def pointer_is_device(ptr):
    """True when ptr (an int address or c_void_p) is device memory."""
    flag = ctypes.c_int()
    _check(_lib.load().ecsprt_hip_pointer_is_device(ptr, ctypes.byref(flag)))
    return bool(flag.value)


def poisson_attempts(lam, u):
    """The Philox collision mode's attempt counts for uniforms u at rate lam."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    k = np.empty(u.shape, dtype=np.int32)
    _check(_lib.load().ecsprt_hip_poisson_attempts(lam, u.size, _dp(u), k.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
    return k


def nr_stream(generator, seed, n):
    """The library's ran1 / ran3 / gasdev stream (generator "ran1", "ran3" or "gasdev")."""
    g = {"ran1": 0, "ran3": 1, "gasdev": 2}[generator]
    out = np.empty(n, dtype=np.float32)
    _check(_lib.load().ecsprt_hip_nr_stream(g, seed, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    return out


def philox4x32_10(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    _lib.load().ecsprt_hip_philox4x32_10(c, k, o)
    return list(o)


def philox2x32_10(ctr, key):
    c = (ctypes.c_uint32 * 2)(*ctr)
    o = (ctypes.c_uint32 * 2)()
    _lib.load().ecsprt_hip_philox2x32_10(c, key, o)
    return list(o)


def lognormal_moments(dims, seed, sigma=0.3, smooth=3.0, flux_scale=1.0):
    """Device FluidNode moments of a smoothed log-normal density (mean 1) with flux =
    flux_scale * n.  Returns the DeviceVector (4 doubles per node)."""
    Nx, Ny, Nz = dims
    dv = DeviceVector(4 * Nx * Ny * Nz)
    _check(_lib.load().ecsprt_hip_lognormal_moments(Nx, Ny, Nz, seed, sigma, smooth, flux_scale, dv.ptr))
    return dv


def ensemble_step(contexts, dt):
    """One timestep of every context, run concurrently (ecsprt_hip_ensemble_step)."""
    n = len(contexts)
    arr = (ctypes.c_void_p * n)(*[c.ptr.value for c in contexts])
    rc = _lib.load().ecsprt_hip_ensemble_step(arr, n, dt)
    if rc != _lib.OK:
        for c in contexts:
            msg = _lib.load().ecsprt_hip_last_error(c.ptr)
            if msg:
                raise HipError(rc, msg.decode())
        raise HipError(rc, "")


class Ensemble:
    """Lockstep ensemble over contexts (ecsprt_hip_ensemble_*): each advance is one
    timestep of every run, driven from one thread: one tile-pass launch per group of
    runs that share precision and tile shape, each run's small launches on its own
    stream, one lockstep solve of all potentials; every run's ions and moments equal
    those of stepping it alone."""

    def __init__(self, contexts):
        L = _lib.load()
        self.contexts = list(contexts)
        n = len(self.contexts)
        arr = (ctypes.c_void_p * n)(*[c.ptr.value for c in self.contexts])
        self.ptr = ctypes.c_void_p()
        rc = L.ecsprt_hip_ensemble_create(arr, n, ctypes.byref(self.ptr))
        if rc != _lib.OK:
            msg = L.ecsprt_hip_last_error(None)
            raise HipError(rc, msg.decode() if msg else "")

    def advance(self, dt):
        L = _lib.load()
        rc = L.ecsprt_hip_ensemble_advance(self.ptr, dt)
        if rc != _lib.OK:
            for c in self.contexts:
                msg = L.ecsprt_hip_last_error(c.ptr)
                if msg:
                    raise HipError(rc, msg.decode())
            msg = L.ecsprt_hip_ensemble_last_error(self.ptr)
            raise HipError(rc, msg.decode() if msg else "")

    def traffic(self):
        """Algorithmic HBM bytes of the last advance's solve."""
        b = ctypes.c_double()
        _check(_lib.load().ecsprt_hip_ensemble_traffic(self.ptr, ctypes.byref(b)))
        return b.value

    def stage_times(self):
        """ms of the last advance: tile passes, stages before them, re-binning and folds,
        forcing and solve."""
        ms = np.zeros(4)
        _check(_lib.load().ecsprt_hip_ensemble_stage_times(self.ptr, _dp(ms)))
        return ms

    def close(self):
        if self.ptr:
            _lib.load().ecsprt_hip_ensemble_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()


class Batch:
    """Batched standalone potential solves (one per grid in `dims`): on `slots`
    concurrent streams, or with slots = 0 in lockstep, every kernel launched once for
    all problems; the physical parameters and solver options are shared."""

    def __init__(self, params, options, dims, slots=8):
        L = _lib.load()
        self.dims = [tuple(int(v) for v in d) for d in dims]
        flat = (ctypes.c_int * (3 * len(self.dims)))(*[v for d in self.dims for v in d])
        self.ptr = ctypes.c_void_p()
        rc = L.ecsprt_hip_batch_create(ctypes.byref(params), ctypes.byref(options), len(self.dims), flat, slots,
                                       ctypes.byref(self.ptr))
        _check(rc)

    def solve(self, moments, phi):
        """moments, phi: sequences of DeviceVector (or device pointers) per problem."""
        L = _lib.load()
        n = len(self.dims)
        mp = (ctypes.c_void_p * n)(*[getattr(m, "ptr", m) for m in moments])
        pp = (ctypes.c_void_p * n)(*[getattr(f, "ptr", f) for f in phi])
        rc = L.ecsprt_hip_batch_solve(self.ptr, mp, pp)
        if rc != _lib.OK:
            msg = L.ecsprt_hip_batch_last_error(self.ptr)
            raise HipError(rc, msg.decode() if msg else "")

    def stats(self):
        n = len(self.dims)
        it = np.zeros(n, dtype=np.int32)
        rn = np.zeros(n)
        _check(_lib.load().ecsprt_hip_batch_stats(self.ptr, it.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                                  _dp(rn)))
        return it, rn

    def traffic(self):
        """Algorithmic HBM bytes of the last solve (lockstep mode)."""
        b = ctypes.c_double()
        _check(_lib.load().ecsprt_hip_batch_traffic(self.ptr, ctypes.byref(b)))
        return b.value

    def close(self):
        if self.ptr:
            _lib.load().ecsprt_hip_batch_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()
