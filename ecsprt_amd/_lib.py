"""ctypes declarations of libecsprt_hip.so (include/ecsprt_hip.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback for any entry point.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libecsprt_hip.so")

OK = 0
ERR_ARG = 1
ERR_COLLISION = 2
ERR_COUNT_CHANGED = 3
ERR_HIP = 4
ERR_NOT_CONVERGED = 5
ERR_ATTEMPTS = 6
ERR_NOMEM = 7
ERR_STATE = 8

COLLIDE_OFF, COLLIDE_PREDRAWN, COLLIDE_PHILOX = 0, 1, 2
OPERATOR_FROZEN, OPERATOR_REBUILD = 0, 1
PC_NONE, PC_JACOBI, PC_GMG = 0, 1, 2
SOLVE_FP64, SOLVE_MIXED = 0, 1
PARTICLES_FP64, PARTICLES_MIXED = 0, 1
# PDistType order, src/positions.c:8-10.
PDIST_SOBOL, PDIST_REVERSE, PDIST_NORMAL, PDIST_UNIFORM, PDIST_SINUSOIDAL, PDIST_GAUSSIAN = range(6)

_CLI_DOUBLES = ("dx", "dy", "dz", "x0", "y0", "z0", "x1", "y1", "z1", "B0", "E0", "n0",
                "qi", "mi", "nui", "vi0x", "vi0y", "vi0z", "viTx", "viTy", "viTz", "Ti",
                "nue", "Te", "veTx", "veTy", "veTz", "mn", "vn0x", "vn0y", "vn0z", "Tn")
_PARAM_DOUBLES = ("dx", "dy", "dz", "Lx", "Ly", "Lz", "x0", "y0", "z0", "x1", "y1", "z1",
                  "B0", "E0", "n0",
                  "qi", "mi", "nui", "vi0x", "vi0y", "vi0z", "vi0", "viTx", "viTy", "viTz", "viT",
                  "Ti", "Omega_i", "kappa_i",
                  "qe", "me", "nue", "Te", "veT", "Omega_e", "kappa_e",
                  "mn", "vn0x", "vn0y", "vn0z", "vn0", "vnT", "Tn", "scale", "gammaT")


class CLI(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("Nx", "Ny", "Nz")] + [(n, ctypes.c_double) for n in _CLI_DOUBLES] + \
        [("ndim", ctypes.c_int)]


class Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("Nx", "Ny", "Nz")] + [(n, ctypes.c_double) for n in _PARAM_DOUBLES] + \
        [("ndim", ctypes.c_int)]


class Attempt(ctypes.Structure):
    _fields_ = [("u_ip", ctypes.c_float), ("g", ctypes.c_float * 3), ("u_acc1", ctypes.c_float),
                ("u_acc2", ctypes.c_float), ("u_z", ctypes.c_float), ("u_phi", ctypes.c_float)]


class Options(ctypes.Structure):
    _fields_ = [("tile", ctypes.c_int), ("tile_z", ctypes.c_int), ("halo", ctypes.c_int), ("block", ctypes.c_int),
                ("split", ctypes.c_int), ("slack", ctypes.c_double), ("collide_mode", ctypes.c_int),
                ("seed", ctypes.c_uint64), ("operator_mode", ctypes.c_int), ("initial_guess", ctypes.c_int),
                ("pc_type", ctypes.c_int), ("solve_precision", ctypes.c_int), ("ksp_restart", ctypes.c_int),
                ("ksp_max_it", ctypes.c_int), ("ksp_rtol", ctypes.c_double), ("ksp_atol", ctypes.c_double),
                ("mg_pre", ctypes.c_int), ("mg_post", ctypes.c_int), ("mg_coarse_sweeps", ctypes.c_int),
                ("inner_rtol", ctypes.c_double), ("timing", ctypes.c_int),
                ("particle_precision", ctypes.c_int), ("mg_sweep", ctypes.c_int), ("gmres_refine", ctypes.c_int)]


_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_L = ctypes.POINTER(ctypes.c_long)

# name -> (restype, argtypes); every symbol declared in include/ecsprt_hip.h.
SIGNATURES = {
    "ecsprt_hip_setup_context": (ctypes.c_int, [ctypes.POINTER(CLI), ctypes.POINTER(Params)]),
    "ecsprt_hip_default_options": (ctypes.c_int, [ctypes.POINTER(Options)]),
    "ecsprt_hip_create": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.POINTER(Options), ctypes.POINTER(_P)]),
    "ecsprt_hip_destroy": (ctypes.c_int, [_P]),
    "ecsprt_hip_set_particles": (ctypes.c_int, [_P, ctypes.c_long, _D, _D]),
    "ecsprt_hip_get_particles": (ctypes.c_int, [_P, _D, _D]),
    "ecsprt_hip_init_particles": (ctypes.c_int, [_P, ctypes.c_long, ctypes.c_uint64, ctypes.c_double, ctypes.c_int]),
    "ecsprt_hip_init_reference": (ctypes.c_int, [_P, ctypes.c_long]),
    "ecsprt_hip_init_distribution": (ctypes.c_int, [_P, ctypes.c_long, ctypes.c_int]),
    "ecsprt_hip_num_particles": (ctypes.c_long, [_P]),
    "ecsprt_hip_cell_counts": (ctypes.c_int, [_P, _I]),
    "ecsprt_hip_push": (ctypes.c_int, [ctypes.c_double, _P]),
    "ecsprt_hip_collide": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, _P]),
    "ecsprt_hip_apply_bc": (ctypes.c_int, [ctypes.c_int, _P, _P, _P]),
    "ecsprt_hip_collect": (ctypes.c_int, [_P]),
    "ecsprt_hip_update_velocities": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, _P]),
    "ecsprt_hip_update_positions": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, _P]),
    "ecsprt_hip_compute_potential": (ctypes.c_int, [_P]),
    "ecsprt_hip_step": (ctypes.c_int, [_P, ctypes.c_double]),
    "ecsprt_hip_lhs": (ctypes.c_int, [_P]),
    "ecsprt_hip_rhs": (ctypes.c_int, [_P, _P]),
    "ecsprt_hip_matmult": (ctypes.c_int, [_P, _P, _P]),
    "ecsprt_hip_pcapply": (ctypes.c_int, [_P, _P, _P]),
    "ecsprt_hip_remove_mean": (ctypes.c_int, [_P, _P]),
    "ecsprt_hip_ksp_solve": (ctypes.c_int, [_P, _P, _P]),
    "ecsprt_hip_get_moments": (ctypes.c_int, [_P, _D]),
    "ecsprt_hip_set_moments": (ctypes.c_int, [_P, _D]),
    "ecsprt_hip_get_potential": (ctypes.c_int, [_P, _D]),
    "ecsprt_hip_set_potential": (ctypes.c_int, [_P, _D]),
    "ecsprt_hip_get_forcing": (ctypes.c_int, [_P, _D]),
    "ecsprt_hip_set_attempts": (ctypes.c_int, [_P, _P, ctypes.c_long]),
    "ecsprt_hip_collision_count": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.c_long, ctypes.c_double]),
    "ecsprt_hip_collision_stats": (ctypes.c_int, [_P, _I, _L, _L, _L]),
    "ecsprt_hip_solver_stats": (ctypes.c_int, [_P, _I, _D, _D]),
    "ecsprt_hip_solver_traffic": (ctypes.c_int, [_P, _D]),
    "ecsprt_hip_stage_times": (ctypes.c_int, [_P, _D]),
    "ecsprt_hip_synchronize": (ctypes.c_int, [_P]),
    "ecsprt_hip_last_error": (ctypes.c_char_p, [_P]),
    "ecsprt_hip_pointer_is_device": (ctypes.c_int, [_P, _I]),
    "ecsprt_hip_dmalloc": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_long]),
    "ecsprt_hip_dfree": (ctypes.c_int, [_P]),
    "ecsprt_hip_h2d": (ctypes.c_int, [_P, _P, ctypes.c_long]),
    "ecsprt_hip_d2h": (ctypes.c_int, [_P, _P, ctypes.c_long]),
    "ecsprt_hip_d2d": (ctypes.c_int, [_P, _P, ctypes.c_long, _D]),
    "ecsprt_hip_philox4x32_10": (None, [ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                        ctypes.POINTER(ctypes.c_uint32)]),
    "ecsprt_hip_philox2x32_10": (None, [ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32,
                                        ctypes.POINTER(ctypes.c_uint32)]),
    "ecsprt_hip_poisson_attempts": (ctypes.c_int, [ctypes.c_double, ctypes.c_long, _D, _I]),
    "ecsprt_hip_nr_stream": (ctypes.c_int, [ctypes.c_int, ctypes.c_long, ctypes.c_long,
                                            ctypes.POINTER(ctypes.c_float)]),
    "ecsprt_hip_batch_create": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.POINTER(Options), ctypes.c_int, _I,
                                               ctypes.c_int, ctypes.POINTER(_P)]),
    "ecsprt_hip_batch_destroy": (ctypes.c_int, [_P]),
    "ecsprt_hip_batch_solve": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.POINTER(_P)]),
    "ecsprt_hip_batch_stats": (ctypes.c_int, [_P, _I, _D]),
    "ecsprt_hip_batch_last_error": (ctypes.c_char_p, [_P]),
    "ecsprt_hip_batch_traffic": (ctypes.c_int, [_P, _D]),
    "ecsprt_hip_ensemble_step": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int, ctypes.c_double]),
    "ecsprt_hip_ensemble_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int, ctypes.POINTER(_P)]),
    "ecsprt_hip_ensemble_advance": (ctypes.c_int, [_P, ctypes.c_double]),
    "ecsprt_hip_ensemble_destroy": (ctypes.c_int, [_P]),
    "ecsprt_hip_ensemble_last_error": (ctypes.c_char_p, [_P]),
    "ecsprt_hip_ensemble_traffic": (ctypes.c_int, [_P, _D]),
    "ecsprt_hip_ensemble_stage_times": (ctypes.c_int, [_P, _D]),
    "ecsprt_hip_lognormal_moments": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                                    ctypes.c_double, ctypes.c_double, ctypes.c_double, _P]),
    "ecsprt_hip_lognormal_moments_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                                            ctypes.c_double, ctypes.c_double, ctypes.c_double, _P,
                                                            _P]),
    "ecsprt_hip_output_fluid_hdf5": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "ecsprt_hip_output_swarm_binary": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "ecsprt_hip_load_fluid_quantities": (ctypes.c_int, [_P, _D, ctypes.c_char_p]),
    "ecsprt_io_hdf5_available": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    "ecsprt_io_last_error": (ctypes.c_char_p, []),
    "ecsprt_io_write_fluid_hdf5": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_int, _D, _D, _D]),
    "ecsprt_io_write_hdf5_dataset": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, _L, _D]),
    "ecsprt_io_read_hdf5_dataset": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _I, _L, _D, ctypes.c_long]),
    "ecsprt_io_write_vec_binary": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_long, ctypes.c_int, _D]),
    "ecsprt_io_read_vec_binary": (ctypes.c_int, [ctypes.c_char_p, _D, ctypes.c_long, _L, _I]),
}

_lib = None


def load():
    """Load libecsprt_hip.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m ecsprt_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib
