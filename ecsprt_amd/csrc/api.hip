// extern "C" entry points of libecsprt_hip.so (declared in include/ecsprt_hip.h).
// Each converts internal errors to int codes and records the message on the context.
#include <string.h>

#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "philox.hpp"
#include "pic_math.hpp"

using ecsprt::Error;

namespace {

thread_local std::string g_last_global_error;

template <typename F>
int guarded(ecsprt_hip_ctx *c, F &&f) {
  try {
    f();
    return ECSPRT_OK;
  } catch (const Error &e) {
    if (c) c->last_error = e.msg;
    else g_last_global_error = e.msg;
    return e.code;
  } catch (const std::bad_alloc &) {
    if (c) c->last_error = "host allocation failed";
    return ECSPRT_ERR_NOMEM;
  }
}

ecsprt_hip_ctx *as_ctx(void *p) { return static_cast<ecsprt_hip_ctx *>(p); }

// True only for memory a kernel may dereference as a device address: hipMalloc'd or
// managed.  Host memory, pinned or not, is false.  For an address the runtime does not
// know (pageable host memory) hipPointerGetAttributes fails; its error is consumed so a
// later hipGetLastError does not report it.
bool pointer_is_device(const void *p) {
  hipPointerAttribute_t at{};
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged;
}

void ev_record(ecsprt_hip_ctx *c, int i) {
  if (c->o.timing) ECS_HIP(hipEventRecord(c->ev[i], c->stream));
}

float ev_ms(ecsprt_hip_ctx *c, int a, int b) {
  float ms = 0.f;
  ECS_HIP(hipEventElapsedTime(&ms, c->ev[a], c->ev[b]));
  return ms;
}

// Build the operator when the operator mode asks for it (src/pic.c:192: the
// KSPSetComputeOperators callback runs at the first KSPSolve only).
void maybe_build_operator(ecsprt_hip_ctx *c) {
  if (!c->operator_built || c->o.operator_mode == ECSPRT_OPERATOR_REBUILD) ecsprt::solver_build_operator(c);
}

void compute_potential(ecsprt_hip_ctx *c) {
  maybe_build_operator(c);
  c->solve_bytes = 0.0;
  ecsprt::solver_rhs(c, c->d_rhs);
  ev_record(c, 4);
  int rc = ecsprt::solver_solve(c, c->d_rhs, c->d_phi);
  ev_record(c, 5);
  if (rc != ECSPRT_OK)
    throw Error{rc, "KSP did not converge: ||r|| = " + std::to_string(c->last_rnorm) + " after " +
                        std::to_string(c->last_iters) + " iterations"};
}

}  // namespace

void ecsprt::set_global_error(const std::string &msg) { g_last_global_error = msg; }

extern "C" {

int ecsprt_hip_setup_context(const ecsprt_cli *c, ecsprt_params *p) {
  if (!c || !p) return ECSPRT_ERR_ARG;
  // SetUpContext (src/setup.c:116-271) and NormalizeGrid (src/setup.c:296-346).
  const double Q = 1.6022e-19, ME = 9.1094e-31, KB = 1.3807e-23;
  double tmp;
  memset(p, 0, sizeof(*p));
  p->qe = -Q;
  p->me = ME;
  p->n0 = c->n0;
  p->B0 = c->B0;
  p->E0 = c->E0;
  p->qi = c->qi;
  p->mi = c->mi;
  p->vi0x = c->vi0x;
  p->vi0y = c->vi0y;
  p->vi0z = c->vi0z;
  p->viTx = c->viTx;
  p->viTy = c->viTy;
  p->viTz = c->viTz;
  p->Ti = c->Ti;
  p->nui = c->nui;
  p->nue = c->nue;
  p->Te = c->Te;
  p->mn = c->mn;
  p->vn0x = c->vn0x;
  p->vn0y = c->vn0y;
  p->vn0z = c->vn0z;
  p->Tn = c->Tn;
  if (c->ndim != 0 && c->ndim != 2 && c->ndim != 3) return ECSPRT_ERR_ARG;  // src/setup.c:41
  p->ndim = c->ndim == 2 ? 2 : 3;
  p->Nx = c->Nx > 0 ? c->Nx : 7;
  p->Ny = c->Ny > 0 ? c->Ny : 7;
  p->Nz = p->ndim == 2 ? 1 : (c->Nz > 0 ? c->Nz : 7);  // src/setup.c:178-188
  p->dx = c->dx;
  p->dy = c->dy;
  p->dz = p->ndim == 2 ? 0.0 : c->dz;
  p->x0 = c->x0;
  p->y0 = c->y0;
  p->z0 = c->z0;
  p->x1 = c->x1;
  p->y1 = c->y1;
  p->z1 = c->z1;
  if (p->ndim == 2 && !(p->z1 > p->z0)) {  // one plane of unit thickness; no z coupling in 2-D
    p->z0 = 0.0;
    p->z1 = 1.0;
  }
  if (p->dx <= 0.0) { p->Lx = p->x1 - p->x0; p->dx = p->Lx / (double)p->Nx; }
  else { p->Lx = p->dx * (double)p->Nx; p->x0 = 0.0; p->x1 = p->Lx; }
  if (p->dy <= 0.0) { p->Ly = p->y1 - p->y0; p->dy = p->Ly / (double)p->Ny; }
  else { p->Ly = p->dy * (double)p->Ny; p->y0 = 0.0; p->y1 = p->Ly; }
  if (p->dz <= 0.0) { p->Lz = p->z1 - p->z0; p->dz = p->Lz / (double)p->Nz; }
  else { p->Lz = p->dz * (double)p->Nz; p->z0 = 0.0; p->z1 = p->Lz; }
  p->Omega_e = std::fabs(p->qe * p->B0 / p->me);
  p->Omega_i = std::fabs(p->qi * p->B0 / p->mi);
  p->kappa_e = p->Omega_e / p->nue;
  p->kappa_i = p->Omega_i / p->nui;
  p->scale = 1.0 / (1 + p->kappa_e * p->kappa_e);
  p->vi0 = std::sqrt(p->vi0x * p->vi0x + p->vi0y * p->vi0y + p->vi0z * p->vi0z);
  p->vn0 = std::sqrt(p->vn0x * p->vn0x + p->vn0y * p->vn0y + p->vn0z * p->vn0z);
  if (c->veTx != 0.0 || c->veTy != 0.0 || c->veTz != 0.0) {
    p->veT = std::sqrt(c->veTx * c->veTx + c->veTy * c->veTy + c->veTz * c->veTz);
    p->Te = (0.5 * p->me / KB) * (p->veT * p->veT);
  } else {
    tmp = std::sqrt(2.0 * KB * p->Te / p->me) / 3.0;
    p->veT = std::sqrt(tmp * tmp + tmp * tmp + tmp * tmp);
  }
  if (p->viTx != 0.0 || p->viTy != 0.0 || p->viTz != 0.0) {
    p->viT = std::sqrt(p->viTx * p->viTx + p->viTy * p->viTy + p->viTz * p->viTz);
    p->Ti = (0.5 * p->mi / KB) * (p->viT * p->viT);
  } else {
    tmp = std::sqrt(2.0 * KB * p->Ti / p->mi) / 3.0;
    p->viTx = tmp;
    p->viTy = tmp;
    p->viTz = tmp;
    p->viT = std::sqrt(p->viTx * p->viTx + p->viTy * p->viTy + p->viTz * p->viTz);
  }
  if (p->Tn == -1.0) p->Tn = p->Ti;
  p->vnT = std::sqrt(2.0 * KB * p->Tn / p->mn);
  p->gammaT = 1.0;
  return ECSPRT_OK;
}

int ecsprt_hip_default_options(ecsprt_options *o) {
  if (!o) return ECSPRT_ERR_ARG;
  memset(o, 0, sizeof(*o));
  o->tile = 0;
  o->tile_z = 0;
  o->halo = 2;
  o->block = 0;
  o->split = 0;
  o->slack = 0.25;
  o->collide_mode = ECSPRT_COLLIDE_PHILOX;
  o->seed = 2024;
  o->operator_mode = ECSPRT_OPERATOR_FROZEN;
  o->initial_guess = 0;
  o->pc_type = ECSPRT_PC_GMG;
  o->solve_precision = ECSPRT_SOLVE_FP64;
  o->ksp_restart = 30;
  o->ksp_max_it = 10000;
  o->ksp_rtol = 1e-5;
  o->ksp_atol = 1e-50;
  o->mg_pre = 2;
  o->mg_post = 2;
  o->mg_coarse_sweeps = 50;
  o->inner_rtol = 0.0;
  o->timing = 0;
  o->particle_precision = ECSPRT_PARTICLES_FP64;
  o->mg_sweep = 0;
  o->gmres_refine = 0;
  return ECSPRT_OK;
}

int ecsprt_hip_create(const ecsprt_params *p, const ecsprt_options *opt, ecsprt_hip_ctx **out) {
  if (!p || !out) return ECSPRT_ERR_ARG;
  *out = nullptr;
  auto *c = new ecsprt_hip_ctx();
  c->p = *p;
  if (opt) c->o = *opt;
  else ecsprt_hip_default_options(&c->o);
  int rc = guarded(c, [&] {
    const bool two_d = p->ndim == 2;
    ECS_CHECK(p->ndim == 2 || p->ndim == 3, ECSPRT_ERR_ARG, "ndim must be 2 or 3");
    ECS_CHECK(p->Nx >= 3 && p->Ny >= 3 && (two_d ? p->Nz == 1 : p->Nz >= 3), ECSPRT_ERR_ARG,
              "grid edges must be >= 3 (the 11-point stencil needs distinct neighbours; Nz = 1 in 2-D)");
    ECS_CHECK(!two_d || c->o.particle_precision == ECSPRT_PARTICLES_FP64, ECSPRT_ERR_ARG,
              "two-dimensional runs use FP64 particles");
    ECS_CHECK(p->dx > 0 && p->dy > 0 && p->dz > 0, ECSPRT_ERR_ARG, "cell widths must be positive");
    ECS_CHECK(c->o.particle_precision == ECSPRT_PARTICLES_FP64 || c->o.particle_precision == ECSPRT_PARTICLES_MIXED,
              ECSPRT_ERR_ARG, "particle_precision must be ECSPRT_PARTICLES_FP64 or ECSPRT_PARTICLES_MIXED");
    c->mixed = c->o.particle_precision == ECSPRT_PARTICLES_MIXED;
    ECS_CHECK(c->o.mg_sweep >= 0 && c->o.mg_sweep <= 2, ECSPRT_ERR_ARG, "mg_sweep must be 0, 1 or 2");
    if (c->mixed) {
      ECS_CHECK(p->Nx <= ecsprt::MX_MAX_NX && p->Ny <= ecsprt::MX_MAX_NY && p->Nz <= ecsprt::MX_MAX_NZ, ECSPRT_ERR_ARG,
                "mixed-precision particles need Nx, Ny <= 2048 and Nz <= 1024");
      ECS_CHECK(p->x0 == 0.0 && p->y0 == 0.0 && p->z0 == 0.0, ECSPRT_ERR_ARG,
                "mixed-precision particles need the grid origin at 0");
      ECS_CHECK(c->o.collide_mode != ECSPRT_COLLIDE_PREDRAWN, ECSPRT_ERR_ARG,
                "pre-drawn collision records need FP64 particles");
    }
    ECS_HIP(hipGetDevice(&c->device));
    ECS_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto &e : c->ev) ECS_HIP(hipEventCreate(&e));
    c->nnode = (long)p->Nx * p->Ny * p->Nz;
    ECS_HIP(hipMalloc(&c->d_phi, sizeof(double) * c->nnode));
    ECS_HIP(hipMalloc(&c->d_moments, sizeof(double) * 4 * c->nnode));
    ECS_HIP(hipMalloc(&c->d_rhs, sizeof(double) * c->nnode));
    ECS_HIP(hipMemsetAsync(c->d_phi, 0, sizeof(double) * c->nnode, c->stream));
    ECS_HIP(hipMemsetAsync(c->d_moments, 0, sizeof(double) * 4 * c->nnode, c->stream));
    ECS_HIP(hipMemsetAsync(c->d_rhs, 0, sizeof(double) * c->nnode, c->stream));
    if (!two_d) ecsprt::tiles_init(c);
    ecsprt::solver_create(c);
    ECS_HIP(hipStreamSynchronize(c->stream));
  });
  if (rc != ECSPRT_OK) {
    g_last_global_error = c->last_error;
    ecsprt_hip_destroy(c);
    return rc;
  }
  *out = c;
  return ECSPRT_OK;
}

int ecsprt_hip_destroy(ecsprt_hip_ctx *c) {
  if (!c) return ECSPRT_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  try {
    ecsprt::solver_destroy(c);
  } catch (...) {
  }
  ecsprt::tiles_free(c);
  ecsprt::p2d_free(c);
  (void)hipFree(c->d_phi);
  (void)hipFree(c->d_moments);
  (void)hipFree(c->d_rhs);
  (void)hipFree(c->d_att);
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return ECSPRT_OK;
}

int ecsprt_hip_set_particles(ecsprt_hip_ctx *c, long np, const double *pos, const double *vel) {
  if (!c || !pos || !vel || np <= 0) return ECSPRT_ERR_ARG;
  return guarded(c, [&] {
    if (c->p.ndim == 2) return ecsprt::p2d_set(c, np, pos, vel, hipMemcpyHostToDevice);
    ecsprt::ensure_stage(c, np);
    ECS_HIP(hipMemcpyAsync(c->d_stage, pos, sizeof(double) * 3 * np, hipMemcpyHostToDevice, c->stream));
    ECS_HIP(hipMemcpyAsync(c->d_stage + 3 * np, vel, sizeof(double) * 3 * np, hipMemcpyHostToDevice, c->stream));
    ecsprt::particles_bin_from_stage(c, np, true);
  });
}

int ecsprt_hip_get_particles(ecsprt_hip_ctx *c, double *pos, double *vel) {
  if (!c || !pos || !vel) return ECSPRT_ERR_ARG;
  return guarded(c, [&] {
    ECS_CHECK(c->np > 0, ECSPRT_ERR_STATE, "no particles");
    if (c->p.ndim == 2) return ecsprt::p2d_get(c, pos, vel);
    ecsprt::particles_to_stage(c);
    ECS_HIP(hipMemcpyAsync(pos, c->d_stage, sizeof(double) * 3 * c->np, hipMemcpyDeviceToHost, c->stream));
    ECS_HIP(hipMemcpyAsync(vel, c->d_stage + 3 * c->np, sizeof(double) * 3 * c->np, hipMemcpyDeviceToHost, c->stream));
    ECS_HIP(hipStreamSynchronize(c->stream));
  });
}

int ecsprt_hip_init_particles(ecsprt_hip_ctx *c, long np, uint64_t seed, double amp, int k) {
  if (!c || np <= 0) return ECSPRT_ERR_ARG;
  return guarded(c, [&] {
    if (c->p.ndim == 2) ecsprt::p2d_init_device(c, np, seed, amp, k);
    else ecsprt::particles_init_device(c, np, seed, amp, k);
  });
}

int ecsprt_hip_init_reference(ecsprt_hip_ctx *c, long np) {
  if (!c || np <= 0) return ECSPRT_ERR_ARG;
  return guarded(c, [&] { ecsprt::particles_init_distribution(c, np, ECSPRT_PDIST_NORMAL); });
}

int ecsprt_hip_init_distribution(ecsprt_hip_ctx *c, long np, int pdist) {
  if (!c || np <= 0) return ECSPRT_ERR_ARG;
  return guarded(c, [&] { ecsprt::particles_init_distribution(c, np, pdist); });
}

long ecsprt_hip_num_particles(const ecsprt_hip_ctx *c) { return c ? c->np : 0; }

int ecsprt_hip_cell_counts(ecsprt_hip_ctx *c, int *counts) {
  if (!c || !counts) return ECSPRT_ERR_ARG;
  return guarded(c, [&] {
    int *d = nullptr;
    ECS_HIP(hipMallocAsync((void **)&d, sizeof(int) * c->nnode, c->stream));
    if (c->p.ndim == 2) ecsprt::p2d_cell_counts(c, d);
    else ecsprt::run_cell_counts(c, d);
    ECS_HIP(hipMemcpyAsync(counts, d, sizeof(int) * c->nnode, hipMemcpyDeviceToHost, c->stream));
    ECS_HIP(hipFreeAsync(d, c->stream));
    ECS_HIP(hipStreamSynchronize(c->stream));
  });
}

int ecsprt_hip_push(double dt, void *opts) {
  ecsprt_hip_ctx *c = as_ctx(opts);
  if (!c) return ECSPRT_ERR_ARG;
  return guarded(c, [&] {
    ECS_CHECK(c->np > 0, ECSPRT_ERR_STATE, "no particles");
    ev_record(c, 0);
    if (c->p.ndim == 2) ecsprt::p2d_push(c, dt);
    else ecsprt::run_push(c, dt);
    ev_record(c, 1);
    if (c->o.timing) {
      ECS_HIP(hipStreamSynchronize(c->stream));
      c->stage_ms[0] = ev_ms(c, 0, 1);
    }
  });
}

int ecsprt_hip_collide(int ndim, double dt, void *opts) {
  ecsprt_hip_ctx *c = as_ctx(opts);
  if (!c || ndim != c->p.ndim) return ECSPRT_ERR_ARG;
  return guarded(c, [&] {
    ECS_CHECK(c->np > 0, ECSPRT_ERR_STATE, "no particles");
    if (c->o.collide_mode == ECSPRT_COLLIDE_PREDRAWN) {
      ecsprt::run_collide_predrawn(c, dt);
    } else if (c->o.collide_mode == ECSPRT_COLLIDE_PHILOX && ndim == 2) {
      ecsprt::p2d_collide_philox(c, dt);
      c->step_index++;
    } else if (c->o.collide_mode == ECSPRT_COLLIDE_PHILOX) {
      ecsprt::run_collide_philox(c, dt);
    }
  });
}

int ecsprt_hip_apply_bc(int np, double *pos, double *vel, void *opts) {
  ecsprt_hip_ctx *c = as_ctx(opts);
  (void)vel;
  if (!c || !pos || np < 0) return ECSPRT_ERR_ARG;
  return guarded(c, [&] {
    if (np == 0) return;
    const long n = (long)np * c->p.ndim;
    // The reference's caller hands over the swarm's host array (src/boundaries.c:67-80);
    // a kernel must never dereference host memory, so anything that is not device
    // memory goes through the staging buffer and back.
    double *d = pos;
    const bool on_device = pointer_is_device(pos);
    if (!on_device) {
      ecsprt::ensure_stage(c, (n + 5) / 6);
      d = c->d_stage;
      ECS_HIP(hipMemcpyAsync(d, pos, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    }
    if (c->p.ndim == 2) ecsprt::p2d_apply_bc(c, np, d);
    else ecsprt::run_apply_bc_aos(c, np, d);
    if (!on_device) ECS_HIP(hipMemcpyAsync(pos, d, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    ECS_HIP(hipStreamSynchronize(c->stream));
  });
}

int ecsprt_hip_collect(void *opts) {
  ecsprt_hip_ctx *c = as_ctx(opts);
  if (!c) return ECSPRT_ERR_ARG;
  return guarded(c, [&] {
    ECS_CHECK(c->np > 0, ECSPRT_ERR_STATE, "no particles");
    ev_record(c, 2);
    if (c->p.ndim == 2) ecsprt::p2d_deposit(c);
    else ecsprt::run_deposit(c);
    ev_record(c, 3);
    ECS_HIP(hipStreamSynchronize(c->stream));
    if (c->o.timing) c->stage_ms[3] = ev_ms(c, 2, 3);
  });
}

int ecsprt_hip_update_velocities(int ndim, double dt, void *opts) {
  ecsprt_hip_ctx *c = as_ctx(opts);
  if (!c || ndim != c->p.ndim) return ECSPRT_ERR_ARG;
  int rc = ecsprt_hip_push(dt, opts);
  if (rc != ECSPRT_OK) return rc;
  // src/velocities.c:384-386: collide when the neutrals have mass.
  if (c->p.mn > 0.0 && c->o.collide_mode != ECSPRT_COLLIDE_OFF) return ecsprt_hip_collide(ndim, dt, opts);
  return ECSPRT_OK;
}

int ecsprt_hip_update_positions(int ndim, double dt, void *opts) {
  ecsprt_hip_ctx *c = as_ctx(opts);
  if (!c || ndim != c->p.ndim) return ECSPRT_ERR_ARG;
  return guarded(c, [&] {
    ECS_CHECK(c->np > 0, ECSPRT_ERR_STATE, "no particles");
    if (ndim == 2) ecsprt::p2d_drift(c, dt);
    else ecsprt::run_drift_rebin(c, dt);
  });
}

int ecsprt_hip_compute_potential(void *opts) {
  ecsprt_hip_ctx *c = as_ctx(opts);
  if (!c) return ECSPRT_ERR_ARG;
  return guarded(c, [&] {
    ev_record(c, 3);
    compute_potential(c);
    if (c->o.timing) {
      ECS_HIP(hipStreamSynchronize(c->stream));
      c->stage_ms[4] = ev_ms(c, 3, 4);
      c->stage_ms[5] = ev_ms(c, 4, 5);
    }
  });
}

int ecsprt_hip_step(void *opts, double dt) {
  ecsprt_hip_ctx *c = as_ctx(opts);
  if (!c) return ECSPRT_ERR_ARG;
  return guarded(c, [&] {
    ECS_CHECK(c->np > 0, ECSPRT_ERR_STATE, "no particles");
    if (c->p.ndim == 2) {
      // The reference loop (src/pic.c:246-254): UpdateVelocities (push, collisions),
      // UpdatePositions, collect, potential; push, Philox collisions and drift fused
      // when the ions are in tile order.
      ev_record(c, 0);
      if (!ecsprt::p2d_push_collide_drift(c, dt)) {
        ecsprt::p2d_push(c, dt);
        if (c->p.mn > 0.0 && c->o.collide_mode == ECSPRT_COLLIDE_PHILOX) ecsprt::p2d_collide_philox(c, dt);
        else if (c->p.mn > 0.0 && c->o.collide_mode == ECSPRT_COLLIDE_PREDRAWN) ecsprt::run_collide_predrawn(c, dt);
        ecsprt::p2d_drift(c, dt);
      }
      ecsprt::p2d_deposit(c);
      ev_record(c, 3);
      compute_potential(c);
      c->step_index++;
      if (c->o.timing) {
        ECS_HIP(hipStreamSynchronize(c->stream));
        c->stage_ms[0] = ev_ms(c, 0, 3);
        c->stage_ms[1] = c->stage_ms[2] = c->stage_ms[3] = 0.0;
        c->stage_ms[4] = ev_ms(c, 3, 4);
        c->stage_ms[5] = ev_ms(c, 4, 5);
      }
      return;
    }
    ev_record(c, 0);
    ecsprt::run_fused_particles(c, dt);
    ev_record(c, 3);
    compute_potential(c);
    if (c->o.timing) {
      ECS_HIP(hipStreamSynchronize(c->stream));
      c->stage_ms[0] = ev_ms(c, 6, 1);
      c->stage_ms[1] = ev_ms(c, 0, 6);
      c->stage_ms[2] = ev_ms(c, 1, 2);
      c->stage_ms[3] = 0.0;
      c->stage_ms[4] = ev_ms(c, 3, 4);
      c->stage_ms[5] = ev_ms(c, 4, 5);
    }
  });
}

int ecsprt_hip_lhs(void *opts) {
  ecsprt_hip_ctx *c = as_ctx(opts);
  if (!c) return ECSPRT_ERR_ARG;
  return guarded(c, [&] {
    ecsprt::solver_build_operator(c);
    ECS_HIP(hipStreamSynchronize(c->stream));
  });
}

int ecsprt_hip_rhs(void *opts, double *b) {
  ecsprt_hip_ctx *c = as_ctx(opts);
  if (!c) return ECSPRT_ERR_ARG;
  return guarded(c, [&] {
    ecsprt::solver_rhs(c, b ? b : c->d_rhs);
    ECS_HIP(hipStreamSynchronize(c->stream));
  });
}

int ecsprt_hip_matmult(void *shell, const double *x, double *y) {
  ecsprt_hip_ctx *c = as_ctx(shell);
  if (!c || !x || !y) return ECSPRT_ERR_ARG;
  return guarded(c, [&] {
    ecsprt::solver_matmult(c, x, y);
    ECS_HIP(hipStreamSynchronize(c->stream));
  });
}

int ecsprt_hip_pcapply(void *shell, const double *r, double *z) {
  ecsprt_hip_ctx *c = as_ctx(shell);
  if (!c || !r || !z) return ECSPRT_ERR_ARG;
  return guarded(c, [&] {
    ecsprt::solver_pcapply(c, r, z);
    ECS_HIP(hipStreamSynchronize(c->stream));
  });
}

int ecsprt_hip_remove_mean(void *opts, double *v) {
  ecsprt_hip_ctx *c = as_ctx(opts);
  if (!c || !v) return ECSPRT_ERR_ARG;
  return guarded(c, [&] {
    ecsprt::solver_remove_mean(c, v);
    ECS_HIP(hipStreamSynchronize(c->stream));
  });
}

int ecsprt_hip_ksp_solve(void *opts, const double *b, double *phi) {
  ecsprt_hip_ctx *c = as_ctx(opts);
  if (!c || !b || !phi) return ECSPRT_ERR_ARG;
  int code = ECSPRT_OK;
  int rc = guarded(c, [&] { code = ecsprt::solver_solve(c, b, phi); });
  return rc != ECSPRT_OK ? rc : code;
}

int ecsprt_hip_get_moments(ecsprt_hip_ctx *c, double *m) {
  if (!c || !m) return ECSPRT_ERR_ARG;
  return guarded(c, [&] {
    ECS_HIP(hipMemcpyAsync(m, c->d_moments, sizeof(double) * 4 * c->nnode, hipMemcpyDeviceToHost, c->stream));
    ECS_HIP(hipStreamSynchronize(c->stream));
  });
}

int ecsprt_hip_set_moments(ecsprt_hip_ctx *c, const double *m) {
  if (!c || !m) return ECSPRT_ERR_ARG;
  return guarded(c, [&] {
    ECS_HIP(hipMemcpyAsync(c->d_moments, m, sizeof(double) * 4 * c->nnode, hipMemcpyHostToDevice, c->stream));
    ECS_HIP(hipStreamSynchronize(c->stream));
  });
}

int ecsprt_hip_get_potential(ecsprt_hip_ctx *c, double *phi) {
  if (!c || !phi) return ECSPRT_ERR_ARG;
  return guarded(c, [&] {
    ECS_HIP(hipMemcpyAsync(phi, c->d_phi, sizeof(double) * c->nnode, hipMemcpyDeviceToHost, c->stream));
    ECS_HIP(hipStreamSynchronize(c->stream));
  });
}

int ecsprt_hip_set_potential(ecsprt_hip_ctx *c, const double *phi) {
  if (!c || !phi) return ECSPRT_ERR_ARG;
  return guarded(c, [&] {
    ECS_HIP(hipMemcpyAsync(c->d_phi, phi, sizeof(double) * c->nnode, hipMemcpyHostToDevice, c->stream));
    ECS_HIP(hipStreamSynchronize(c->stream));
  });
}

int ecsprt_hip_get_forcing(ecsprt_hip_ctx *c, double *b) {
  if (!c || !b) return ECSPRT_ERR_ARG;
  return guarded(c, [&] {
    ECS_HIP(hipMemcpyAsync(b, c->d_rhs, sizeof(double) * c->nnode, hipMemcpyDeviceToHost, c->stream));
    ECS_HIP(hipStreamSynchronize(c->stream));
  });
}

int ecsprt_hip_set_attempts(ecsprt_hip_ctx *c, const ecsprt_attempt *att, long n) {
  if (!c || (n > 0 && !att) || n < 0) return ECSPRT_ERR_ARG;
  return guarded(c, [&] {
    if (n > c->att_cap) {
      (void)hipFree(c->d_att);
      c->d_att = nullptr;
      if (hipMalloc(&c->d_att, sizeof(ecsprt_attempt) * n) != hipSuccess) throw Error{ECSPRT_ERR_NOMEM, "attempts"};
      c->att_cap = n;
    }
    if (n > 0)
      ECS_HIP(hipMemcpyAsync(c->d_att, att, sizeof(ecsprt_attempt) * n, hipMemcpyHostToDevice, c->stream));
    ECS_HIP(hipStreamSynchronize(c->stream));
    c->n_att = n;
  });
}

int ecsprt_hip_collision_count(const ecsprt_params *p, long np, double dt) {
  if (!p) return 0;
  return ecsprt::collision_count(p, np, dt);
}

int ecsprt_hip_collision_stats(ecsprt_hip_ctx *c, int *Nc, long *Ns, long *Nf, long *attempts) {
  if (!c) return ECSPRT_ERR_ARG;
  if (Nc) *Nc = c->last_Nc;
  if (Ns) *Ns = c->last_Ns;
  if (Nf) *Nf = c->last_Nf;
  if (attempts) *attempts = c->last_attempts;
  return ECSPRT_OK;
}

int ecsprt_hip_solver_stats(ecsprt_hip_ctx *c, int *iters, double *rnorm, double *bnorm) {
  if (!c) return ECSPRT_ERR_ARG;
  if (iters) *iters = c->last_iters;
  if (rnorm) *rnorm = c->last_rnorm;
  if (bnorm) *bnorm = c->last_bnorm;
  return ECSPRT_OK;
}

int ecsprt_hip_solver_traffic(ecsprt_hip_ctx *c, double *bytes) {
  if (!c || !bytes) return ECSPRT_ERR_ARG;
  *bytes = c->solve_bytes;
  return ECSPRT_OK;
}

int ecsprt_hip_stage_times(ecsprt_hip_ctx *c, double *ms) {
  if (!c || !ms) return ECSPRT_ERR_ARG;
  for (int i = 0; i < 6; ++i) ms[i] = c->stage_ms[i];
  return ECSPRT_OK;
}

int ecsprt_hip_synchronize(ecsprt_hip_ctx *c) {
  if (!c) return ECSPRT_ERR_ARG;
  return guarded(c, [&] { ECS_HIP(hipStreamSynchronize(c->stream)); });
}

const char *ecsprt_hip_last_error(const ecsprt_hip_ctx *c) {
  return c ? c->last_error.c_str() : g_last_global_error.c_str();
}

int ecsprt_hip_pointer_is_device(const void *ptr, int *is_device) {
  if (!is_device) return ECSPRT_ERR_ARG;
  *is_device = ptr && pointer_is_device(ptr) ? 1 : 0;
  return ECSPRT_OK;
}

int ecsprt_hip_dmalloc(void **ptr, long bytes) {
  if (!ptr || bytes <= 0) return ECSPRT_ERR_ARG;
  return hipMalloc(ptr, (size_t)bytes) == hipSuccess ? ECSPRT_OK : ECSPRT_ERR_NOMEM;
}
int ecsprt_hip_dfree(void *ptr) { return hipFree(ptr) == hipSuccess ? ECSPRT_OK : ECSPRT_ERR_HIP; }
int ecsprt_hip_h2d(void *dst, const void *src, long bytes) {
  return hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice) == hipSuccess ? ECSPRT_OK : ECSPRT_ERR_HIP;
}
int ecsprt_hip_d2h(void *dst, const void *src, long bytes) {
  return hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost) == hipSuccess ? ECSPRT_OK : ECSPRT_ERR_HIP;
}

int ecsprt_hip_d2d(void *dst, const void *src, long bytes, double *ms) {
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return ECSPRT_ERR_HIP;
  if (hipEventCreate(&b) != hipSuccess) return ECSPRT_ERR_HIP;
  int rc = ECSPRT_OK;
  if (hipEventRecord(a, 0) != hipSuccess || hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, 0) != hipSuccess ||
      hipEventRecord(b, 0) != hipSuccess || hipEventSynchronize(b) != hipSuccess)
    rc = ECSPRT_ERR_HIP;
  float t = 0.f;
  if (rc == ECSPRT_OK && hipEventElapsedTime(&t, a, b) != hipSuccess) rc = ECSPRT_ERR_HIP;
  if (ms) *ms = t;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return rc;
}

void ecsprt_hip_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  ecsprt::u32x4 c = {{ctr[0], ctr[1], ctr[2], ctr[3]}};
  ecsprt::u32x4 r = ecsprt::philox4x32_10(c, key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}

int ecsprt_hip_poisson_attempts(double lam, long n, const double *u, int *k) {
  if (n < 0 || (n > 0 && (!u || !k))) return ECSPRT_ERR_ARG;
  try {
    const ecsprt::PoissonParams pp = ecsprt::poisson_params(lam);
    for (long i = 0; i < n; ++i) k[i] = ecsprt::poisson_invert(pp, u[i]);
  } catch (const Error &e) {
    g_last_global_error = e.msg;
    return e.code;
  }
  return ECSPRT_OK;
}

void ecsprt_hip_philox2x32_10(const uint32_t ctr[2], uint32_t key, uint32_t out[2]) {
  ecsprt::u32x2 c = {{ctr[0], ctr[1]}};
  ecsprt::u32x2 r = ecsprt::philox2x32_10(c, key);
  out[0] = r.v[0];
  out[1] = r.v[1];
}

}  // extern "C"
