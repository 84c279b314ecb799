// Batched standalone potential solves.  Independent periodic grids of different
// shapes are solved concurrently: a fixed pool of slots, each a solver context with
// its own HIP stream and host thread (the Krylov loop decides convergence on the
// host), takes problems round-robin.  Every slot carves its workspace from one
// device arena sized for the largest problem, so the batch allocates nothing while
// it runs and the slots never serialise on hipMalloc/hipFree.
//
// Also the device generator of the synthetic density fields: log-normal node values
// smoothed by a periodic Gaussian and normalised to mean 1, with flux = scale * n.
#include <algorithm>
#include <cmath>
#include <thread>

#include "common.hpp"
#include "philox.hpp"

namespace ecsprt {

// The generator's kernels address a field as base[q * stride]: the allocating entry point
// works on plain arrays (stride 1), the allocation-free one on the components of the
// caller's FluidNode output (stride 4), with the same arithmetic in the same order.

// n = exp(sigma g - sigma^2 / 2), g ~ N(0, 1) per node: Box-Muller (cosine branch)
// on two 53-bit uniforms from Philox4x32-10, counter (node, node >> 32, 0x106E, 0).
__global__ void k_lognormal(double *out, int so, long n, double sigma, uint32_t k0, uint32_t k1) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const u32x4 c = {{(uint32_t)q, (uint32_t)(q >> 32), 0x106Eu, 0u}};
    const u32x4 r = philox4x32_10(c, k0, k1);
    const double u1 = ((double)(r.v[0] >> 5) * 67108864.0 + (double)(r.v[1] >> 6) + 0.5) * (1.0 / 9007199254740992.0);
    const double u2 = ((double)(r.v[2] >> 5) * 67108864.0 + (double)(r.v[3] >> 6)) * (1.0 / 9007199254740992.0);
    const double g = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    out[q * so] = exp(sigma * g - 0.5 * sigma * sigma);
  }
}

// Normalised weights of a periodic Gaussian of `smooth` cells over offsets -R..R:
// w[d + R] = exp(-d^2 / (2 smooth^2)) / sum, summed in offset order by one thread
// (`smooth` = 0 gives the single weight 1).
__global__ void k_gauss_weights(double *w, int R, double smooth) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double ws = 0.0;
  for (int d = -R; d <= R; ++d) ws += w[d + R] = smooth > 0 ? exp(-0.5 * d * d / (smooth * smooth)) : 1.0;
  for (int d = -R; d <= R; ++d) w[d + R] /= ws;
}

// Periodic convolution along one axis with weights w[0..2R] (offsets -R..R).
__global__ void k_smooth_axis(const double *__restrict__ a, int sa, double *__restrict__ b, int sb, int Nx, int Ny,
                              int Nz, int axis, const double *__restrict__ w, int R) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i >= Nx || j >= Ny) return;
  const int N = axis == 0 ? Nx : (axis == 1 ? Ny : Nz);
  const int at = axis == 0 ? i : (axis == 1 ? j : k);
  const long stride = axis == 0 ? 1 : (axis == 1 ? (long)Nx : (long)Nx * Ny);
  const long base = ((long)k * Ny + j) * Nx + i - (long)at * stride;
  double acc = 0.0;
  for (int d = -R; d <= R; ++d) {
    int t = (at + d) % N;
    if (t < 0) t += N;
    acc += w[d + R] * a[(base + (long)t * stride) * sa];
  }
  b[(((long)k * Ny + j) * Nx + i) * sb] = acc;
}

__global__ void k_block_sums(const double *__restrict__ a, int sa, long n, double *__restrict__ partial) {
  __shared__ double red[256];
  double acc = 0.0;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) acc += a[q * sa];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// FluidNode moments (n, G = fs n per component) with n normalised to mean 1; the sum
// of the node values is the fixed-order sum of the block partial sums.  The thread of
// node q reads a[q * sa] before it writes m[4q .. 4q+3], so a may be a component of m.
__global__ void k_moments_from(const double *a, int sa, long n, const double *__restrict__ partial, int nb, double fs,
                               double *m) {
  double sum = 0.0;
  for (int b = 0; b < nb; ++b) sum += partial[b];
  const double inv = (double)n / sum;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const double v = a[q * sa] * inv;
    m[q * 4] = v;
    m[q * 4 + 1] = fs * v;
    m[q * 4 + 2] = fs * v;
    m[q * 4 + 3] = fs * v;
  }
}

// The whole generator on stream st: the field into `gen` (stride sg), smoothed through two
// buffers (x: gen -> b1, y: b1 -> b2, z: b2 -> b1; the result is in b1 when smoothing,
// gen otherwise), block partial sums, moments.  w holds 2R+1 doubles, partial LN_BLOCKS.
hipError_t lognormal_launch(hipStream_t st, int Nx, int Ny, int Nz, uint64_t seed, double sigma, double smooth,
                                   int R, double fs, double *gen, int sg, double *b1, int s1, double *b2, int s2,
                                   double *w, double *partial, double *moments) {
  const long n = (long)Nx * Ny * Nz;
  const int g = (int)std::min<long>((n + 255) / 256, 65536);
  k_gauss_weights<<<1, 64, 0, st>>>(w, R, smooth);
  k_lognormal<<<g, 256, 0, st>>>(gen, sg, n, sigma, (uint32_t)seed, (uint32_t)(seed >> 32));
  const dim3 g3((Nx + 63) / 64, (Ny + 3) / 4, Nz), b3(64, 4, 1);
  const double *src = gen;
  int ss = sg;
  if (R > 0) {
    k_smooth_axis<<<g3, b3, 0, st>>>(gen, sg, b1, s1, Nx, Ny, Nz, 0, w, R);
    k_smooth_axis<<<g3, b3, 0, st>>>(b1, s1, b2, s2, Nx, Ny, Nz, 1, w, R);
    k_smooth_axis<<<g3, b3, 0, st>>>(b2, s2, b1, s1, Nx, Ny, Nz, 2, w, R);
    src = b1;
    ss = s1;
  }
  k_block_sums<<<LN_BLOCKS, 256, 0, st>>>(src, ss, n, partial);
  k_moments_from<<<g, 256, 0, st>>>(src, ss, n, partial, LN_BLOCKS, fs, moments);
  return hipGetLastError();
}

}  // namespace ecsprt

using ecsprt::Error;

struct ecsprt_hip_batch {
  ecsprt_params p;
  ecsprt_options o;
  int device = 0;
  ecsprt::LockstepBatch *lockstep = nullptr;  // slots == 0: every problem in every launch
  std::vector<int> dims;
  std::vector<ecsprt_hip_ctx *> slots;
  std::vector<ecsprt::Arena> arenas;
  std::vector<int> iters;
  std::vector<double> rnorm;
  std::string last_error;
};

namespace {

void free_slot(ecsprt_hip_ctx *c, ecsprt::Arena &a) {
  if (c && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  if (a.dev) (void)hipFree(a.dev);
  if (a.host) (void)hipHostFree(a.host);
  a = ecsprt::Arena{};
}

}  // namespace

extern "C" {

int ecsprt_hip_lognormal_moments(int Nx, int Ny, int Nz, uint64_t seed, double sigma, double smooth,
                                 double flux_scale, double *moments) {
  if (Nx < 1 || Ny < 1 || Nz < 1 || !moments || !(smooth >= 0)) return ECSPRT_ERR_ARG;
  try {
    const long n = (long)Nx * Ny * Nz;
    hipStream_t st;
    ECS_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    double *a = nullptr, *b = nullptr, *w = nullptr, *partial = nullptr;
    const int R = smooth > 0 ? (int)std::ceil(4.0 * smooth) : 0;
    hipError_t e = hipMalloc(&a, sizeof(double) * n);
    if (e == hipSuccess) e = hipMalloc(&b, sizeof(double) * n);
    if (e == hipSuccess) e = hipMalloc(&w, sizeof(double) * (2 * R + 1));
    if (e == hipSuccess) e = hipMalloc(&partial, sizeof(double) * ecsprt::LN_BLOCKS);
    if (e == hipSuccess)
      e = ecsprt::lognormal_launch(st, Nx, Ny, Nz, seed, sigma, smooth, R, flux_scale, a, 1, b, 1, a, 1, w, partial,
                                   moments);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipFree(w);
    (void)hipFree(partial);
    (void)hipStreamDestroy(st);
    return e == hipSuccess ? ECSPRT_OK : ECSPRT_ERR_HIP;
  } catch (const Error &e) {
    return e.code;
  }
}

int ecsprt_hip_batch_create(const ecsprt_params *p, const ecsprt_options *o, int nprob, const int *dims, int nslots,
                            ecsprt_hip_batch **out) {
  if (!p || !o || !dims || !out || nprob <= 0 || nslots < 0) return ECSPRT_ERR_ARG;
  *out = nullptr;
  auto *b = new ecsprt_hip_batch();
  b->p = *p;
  b->o = *o;
  b->dims.assign(dims, dims + 3 * nprob);
  long nmax = 0;
  for (int q = 0; q < nprob; ++q) {
    if (dims[3 * q] < 3 || dims[3 * q + 1] < 3 || dims[3 * q + 2] < 3) {
      delete b;
      return ECSPRT_ERR_ARG;
    }
    nmax = std::max(nmax, (long)dims[3 * q] * dims[3 * q + 1] * dims[3 * q + 2]);
  }
  b->iters.assign(nprob, 0);
  b->rnorm.assign(nprob, 0.0);
  if (nslots == 0) {
    try {
      ECS_HIP(hipGetDevice(&b->device));
      b->lockstep = ecsprt::lockstep_create(*p, *o, nprob, dims);
    } catch (const Error &e) {
      ecsprt::set_global_error(e.msg);
      delete b;
      return e.code;
    }
    *out = b;
    return ECSPRT_OK;
  }
  nslots = std::min(nslots, nprob);
  try {
    ECS_HIP(hipGetDevice(&b->device));
    b->arenas.resize(nslots);
    b->slots.assign(nslots, nullptr);
    const size_t dev_bytes = ecsprt::solver_arena_bytes(*o, nmax), host_bytes = 1 << 16;
    for (int s = 0; s < nslots; ++s) {
      auto *c = new ecsprt_hip_ctx();
      b->slots[s] = c;
      c->o = *o;
      c->o.timing = 0;
      c->device = b->device;
      c->arena = &b->arenas[s];
      ECS_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
      if (hipMalloc(&b->arenas[s].dev, dev_bytes) != hipSuccess) throw Error{ECSPRT_ERR_NOMEM, "batch arena"};
      b->arenas[s].dev_cap = dev_bytes;
      ECS_HIP(hipHostMalloc((void **)&b->arenas[s].host, host_bytes, hipHostMallocDefault));
      b->arenas[s].host_cap = host_bytes;
    }
  } catch (const Error &e) {
    for (size_t s = 0; s < b->slots.size(); ++s) free_slot(b->slots[s], b->arenas[s]);
    delete b;
    return e.code;
  }
  b->iters.assign(nprob, 0);
  b->rnorm.assign(nprob, 0.0);
  *out = b;
  return ECSPRT_OK;
}

int ecsprt_hip_batch_destroy(ecsprt_hip_batch *b) {
  if (!b) return ECSPRT_OK;
  if (b->lockstep) ecsprt::lockstep_destroy(b->lockstep);
  for (size_t s = 0; s < b->slots.size(); ++s) free_slot(b->slots[s], b->arenas[s]);
  delete b;
  return ECSPRT_OK;
}

int ecsprt_hip_batch_solve(ecsprt_hip_batch *b, double *const *moments, double *const *phi) {
  if (!b || !moments || !phi) return ECSPRT_ERR_ARG;
  if (b->lockstep) {
    try {
      ECS_HIP(hipSetDevice(b->device));
      const int rc = ecsprt::lockstep_solve(*b->lockstep, moments, phi);
      ecsprt::lockstep_stats(*b->lockstep, b->iters.data(), b->rnorm.data(), nullptr);
      if (rc != ECSPRT_OK) b->last_error = "KSP did not converge on every problem of the batch";
      return rc;
    } catch (const Error &e) {
      b->last_error = e.msg;
      return e.code;
    }
  }
  const int nprob = (int)b->iters.size(), nslots = (int)b->slots.size();
  std::vector<int> rc(nslots, ECSPRT_OK);
  std::vector<std::string> msg(nslots);
  auto work = [&](int s) {
    ecsprt_hip_ctx *c = b->slots[s];
    try {
      ECS_HIP(hipSetDevice(b->device));
      for (int q = s; q < nprob && rc[s] == ECSPRT_OK; q += nslots) {
        ecsprt_params &p = c->p;
        p = b->p;
        p.Nx = b->dims[3 * q];
        p.Ny = b->dims[3 * q + 1];
        p.Nz = b->dims[3 * q + 2];
        p.Lx = p.Nx * p.dx;
        p.Ly = p.Ny * p.dy;
        p.Lz = p.Nz * p.dz;
        p.x1 = p.x0 + p.Lx;
        p.y1 = p.y0 + p.Ly;
        p.z1 = p.z0 + p.Lz;
        c->nnode = (long)p.Nx * p.Ny * p.Nz;
        c->d_moments = moments[q];
        c->d_phi = phi[q];
        c->operator_built = false;
        c->arena->reset();
        const int r = ecsprt::solver_standalone(c);
        b->iters[q] = c->last_iters;
        b->rnorm[q] = c->last_rnorm;
        if (r != ECSPRT_OK) {
          rc[s] = r;
          msg[s] = "problem " + std::to_string(q) + " (" + std::to_string(p.Nx) + "x" + std::to_string(p.Ny) + "x" +
                   std::to_string(p.Nz) + "): KSP did not converge, ||r|| = " + std::to_string(c->last_rnorm) +
                   " after " + std::to_string(c->last_iters) + " iterations";
        }
      }
    } catch (const Error &e) {
      rc[s] = e.code;
      msg[s] = e.msg;
    } catch (const std::exception &e) {
      rc[s] = ECSPRT_ERR_STATE;
      msg[s] = e.what();
    }
    c->d_moments = nullptr;
    c->d_phi = nullptr;
  };
  std::vector<std::thread> th;
  for (int s = 1; s < nslots; ++s) th.emplace_back(work, s);
  work(0);
  for (auto &t : th) t.join();
  for (int s = 0; s < nslots; ++s)
    if (rc[s] != ECSPRT_OK) {
      b->last_error = msg[s];
      return rc[s];
    }
  return ECSPRT_OK;
}

int ecsprt_hip_batch_stats(const ecsprt_hip_batch *b, int *iters, double *rnorm) {
  if (!b) return ECSPRT_ERR_ARG;
  for (size_t q = 0; q < b->iters.size(); ++q) {
    if (iters) iters[q] = b->iters[q];
    if (rnorm) rnorm[q] = b->rnorm[q];
  }
  return ECSPRT_OK;
}

const char *ecsprt_hip_batch_last_error(const ecsprt_hip_batch *b) { return b ? b->last_error.c_str() : ""; }

int ecsprt_hip_batch_traffic(const ecsprt_hip_batch *b, double *bytes) {
  if (!b || !bytes) return ECSPRT_ERR_ARG;
  *bytes = 0.0;
  if (b->lockstep) ecsprt::lockstep_stats(*b->lockstep, nullptr, nullptr, bytes);
  return ECSPRT_OK;
}

// One timestep of every context of an ensemble of independent runs (each created by
// ecsprt_hip_create on this device, each with its own stream): the steps run
// concurrently, one host thread per context, because each step's particle-count check,
// collision-rate controller and Krylov loop read results back on the host.
int ecsprt_hip_ensemble_step(ecsprt_hip_ctx *const *ctxs, int n, double dt) {
  if (!ctxs || n <= 0) return ECSPRT_ERR_ARG;
  for (int q = 0; q < n; ++q)
    if (!ctxs[q]) return ECSPRT_ERR_ARG;
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) return ECSPRT_ERR_HIP;
  std::vector<int> rc(n, ECSPRT_OK);
  auto work = [&](int q) {
    if (hipSetDevice(device) != hipSuccess) {
      rc[q] = ECSPRT_ERR_HIP;
      return;
    }
    rc[q] = ecsprt_hip_step(ctxs[q], dt);
  };
  std::vector<std::thread> th;
  for (int q = 1; q < n; ++q) th.emplace_back(work, q);
  work(0);
  for (auto &t : th) t.join();
  for (int q = 0; q < n; ++q)
    if (rc[q] != ECSPRT_OK) return rc[q];
  return ECSPRT_OK;
}

// ---- lockstep ensembles --------------------------------------------------------------

struct ecsprt_hip_ensemble {
  std::vector<ecsprt_hip_ctx *> ctx;
  ecsprt::LockstepBatch *solve = nullptr;
  hipStream_t st = nullptr;
  void *h_args = nullptr, *d_args = nullptr;
  std::vector<char> op_set;  // the batch holds the run's current operator
  double solve_bytes = 0.0;
  std::string last_error;
  // start, tile passes begin, tile passes end, particle stages end, solve end
  hipEvent_t ev[5] = {};
  bool timed = false;  // the events of the last advance are complete
  ecsprt::RunStreams rs;  // the runs' own streams during a step (events owned here)
  ~ecsprt_hip_ensemble() {
    for (hipEvent_t v : ev)
      if (v) (void)hipEventDestroy(v);
    for (hipEvent_t v : rs.done)
      if (v) (void)hipEventDestroy(v);
    if (rs.fork) (void)hipEventDestroy(rs.fork);
    if (solve) ecsprt::lockstep_destroy(solve);
    if (h_args) (void)hipHostFree(h_args);
    (void)hipFree(d_args);
  }
};

int ecsprt_hip_ensemble_create(ecsprt_hip_ctx *const *ctxs, int n, ecsprt_hip_ensemble **out) {
  using ecsprt::Error;
  if (!out) return ECSPRT_ERR_ARG;
  *out = nullptr;
  auto *e = new ecsprt_hip_ensemble();
  try {
    ECS_CHECK(ctxs && n > 0, ECSPRT_ERR_ARG, "ensemble: no contexts");
    int device = 0;
    ECS_HIP(hipGetDevice(&device));
    std::vector<ecsprt_params> pp;
    for (int q = 0; q < n; ++q) {
      ecsprt_hip_ctx *c = ctxs[q];
      ECS_CHECK(c != nullptr, ECSPRT_ERR_ARG, "ensemble: null context");
      for (int r = 0; r < q; ++r) ECS_CHECK(ctxs[r] != c, ECSPRT_ERR_ARG, "ensemble: a context appears twice");
      const ecsprt_options &o = c->o, &o0 = ctxs[0]->o;
      ECS_CHECK(c->p.ndim != 2, ECSPRT_ERR_ARG, "ensemble: three-dimensional runs only");
      ECS_CHECK(c->device == device, ECSPRT_ERR_ARG, "ensemble: every context on the current device");
      ECS_CHECK(o.collide_mode != ECSPRT_COLLIDE_PREDRAWN, ECSPRT_ERR_ARG,
                "ensemble: Philox collisions or none (pre-drawn records replay one run)");
      ECS_CHECK(o.pc_type == ECSPRT_PC_GMG && o.gmres_refine == 0 && o.initial_guess == 0, ECSPRT_ERR_ARG,
                "ensemble: multigrid preconditioning, classical Gram-Schmidt and a zero initial guess");
      ECS_CHECK(o.solve_precision == o0.solve_precision && o.ksp_restart == o0.ksp_restart &&
                    o.ksp_max_it == o0.ksp_max_it && o.ksp_rtol == o0.ksp_rtol && o.ksp_atol == o0.ksp_atol &&
                    o.mg_pre == o0.mg_pre && o.mg_post == o0.mg_post && o.inner_rtol == o0.inner_rtol,
                ECSPRT_ERR_ARG, "ensemble: the runs' solver options differ");
      pp.push_back(c->p);
      e->ctx.push_back(c);
    }
    e->solve = ecsprt::lockstep_create_runs(pp, ctxs[0]->o);
    e->st = ecsprt::lockstep_stream(*e->solve);
    const size_t bytes = ((ecsprt::tile_args_bytes() + 255) & ~(size_t)255) * n;
    ECS_HIP(hipHostMalloc(&e->h_args, bytes, hipHostMallocDefault));
    if (hipMalloc(&e->d_args, bytes) != hipSuccess) throw Error{ECSPRT_ERR_NOMEM, "ensemble: argument table"};
    e->op_set.assign(n, 0);
    for (hipEvent_t &v : e->ev) ECS_HIP(hipEventCreate(&v));
    ECS_HIP(hipEventCreateWithFlags(&e->rs.fork, hipEventDisableTiming));
    e->rs.done.assign(n, nullptr);
    for (hipEvent_t &v : e->rs.done) ECS_HIP(hipEventCreateWithFlags(&v, hipEventDisableTiming));
  } catch (const Error &x) {
    ecsprt::set_global_error(x.msg);
    delete e;
    return x.code;
  } catch (const std::bad_alloc &) {
    ecsprt::set_global_error("host allocation failed");
    delete e;
    return ECSPRT_ERR_NOMEM;
  }
  *out = e;
  return ECSPRT_OK;
}

int ecsprt_hip_ensemble_destroy(ecsprt_hip_ensemble *e) {
  delete e;
  return ECSPRT_OK;
}

const char *ecsprt_hip_ensemble_last_error(const ecsprt_hip_ensemble *e) { return e ? e->last_error.c_str() : ""; }

int ecsprt_hip_ensemble_traffic(const ecsprt_hip_ensemble *e, double *bytes) {
  if (!e || !bytes) return ECSPRT_ERR_ARG;
  *bytes = e->solve_bytes;
  return ECSPRT_OK;
}

int ecsprt_hip_ensemble_stage_times(const ecsprt_hip_ensemble *e, double *ms) {
  if (!e || !ms) return ECSPRT_ERR_ARG;
  if (!e->timed) return ECSPRT_ERR_STATE;
  const int pairs[4][2] = {{1, 2}, {0, 1}, {2, 3}, {3, 4}};
  for (int i = 0; i < 4; ++i) {
    float t = 0.f;
    if (hipEventElapsedTime(&t, e->ev[pairs[i][0]], e->ev[pairs[i][1]]) != hipSuccess) return ECSPRT_ERR_HIP;
    ms[i] = t;
  }
  return ECSPRT_OK;
}

// One timestep of every run in lockstep: the particle stages of all runs (one launch per
// stage; see run_fused_particles_runs), then each run's forcing, then one GMRES/multigrid
// solve of all runs with one launch per phase.  The contexts' streams are set to the
// ensemble's for the call.
int ecsprt_hip_ensemble_advance(ecsprt_hip_ensemble *e, double dt) {
  using ecsprt::Error;
  if (!e) return ECSPRT_ERR_ARG;
  const int n = (int)e->ctx.size();
  std::vector<hipStream_t> saved(n);
  for (int q = 0; q < n; ++q) saved[q] = e->ctx[q]->stream;
  std::vector<int> rc(n, ECSPRT_OK);
  int code = ECSPRT_OK;
  try {
    for (int q = 0; q < n; ++q) {
      ECS_HIP(hipStreamSynchronize(saved[q]));
      ECS_CHECK(e->ctx[q]->np > 0, ECSPRT_ERR_STATE, "ensemble: a run has no particles");
    }
    for (int q = 0; q < n; ++q) e->ctx[q]->stream = e->st;
    e->rs.st = saved;
    e->timed = false;
    ECS_HIP(hipEventRecord(e->ev[0], e->st));
    ecsprt::run_fused_particles_runs(e->ctx, e->st, dt, e->h_args, e->d_args, rc, e->ev[1], e->ev[2], &e->rs);
    ECS_HIP(hipEventRecord(e->ev[3], e->st));
    std::vector<char> take(n, 0);
    std::vector<const double *> dens(n, nullptr), rhs(n, nullptr);
    std::vector<double *> phi(n, nullptr);
    for (int q = 0; q < n; ++q) {
      ecsprt_hip_ctx *c = e->ctx[q];
      rhs[q] = c->d_rhs;
      phi[q] = c->d_phi;
      if (rc[q] != ECSPRT_OK) continue;
      take[q] = 1;
      // The operator is built at the first solve and kept (src/pic.c:192), or rebuilt
      // every solve in the rebuild mode: the batch copies it when the run's changes.
      if (!c->operator_built || c->o.operator_mode == ECSPRT_OPERATOR_REBUILD) {
        ecsprt::solver_build_operator(c);
        e->op_set[q] = 0;
      }
      if (!e->op_set[q]) dens[q] = ecsprt::solver_density(c);
      ecsprt::solver_rhs(c, c->d_rhs);
    }
    ecsprt::lockstep_solve_given(*e->solve, dens.data(), rhs.data(), phi.data(), take);
    ECS_HIP(hipEventRecord(e->ev[4], e->st));
    ECS_HIP(hipEventSynchronize(e->ev[4]));
    e->timed = true;
    std::vector<int> iters(n);
    std::vector<double> rnorm(n), bnorm(n);
    ecsprt::lockstep_stats(*e->solve, iters.data(), rnorm.data(), &e->solve_bytes);
    ecsprt::lockstep_bnorm(*e->solve, bnorm.data());
    const double rtol = e->ctx[0]->o.ksp_rtol > 0 ? e->ctx[0]->o.ksp_rtol : 1e-5;
    const double atol = e->ctx[0]->o.ksp_atol > 0 ? e->ctx[0]->o.ksp_atol : 1e-50;
    for (int q = 0; q < n; ++q) {
      if (!take[q]) continue;
      ecsprt_hip_ctx *c = e->ctx[q];
      if (dens[q]) e->op_set[q] = 1;
      c->last_iters = iters[q];
      c->last_rnorm = rnorm[q];
      c->last_bnorm = bnorm[q];
      if (!(rnorm[q] <= std::max(rtol * bnorm[q], atol))) {
        rc[q] = ECSPRT_ERR_NOT_CONVERGED;
        c->last_error = "KSP did not converge: ||r|| = " + std::to_string(rnorm[q]) + " after " +
                        std::to_string(iters[q]) + " iterations";
      }
    }
    for (int q = 0; q < n && code == ECSPRT_OK; ++q) code = rc[q];
    if (code != ECSPRT_OK) e->last_error = "ensemble: a run failed (see its context's last error)";
  } catch (const Error &x) {
    e->last_error = x.msg;
    code = x.code;
  } catch (const std::bad_alloc &) {
    e->last_error = "host allocation failed";
    code = ECSPRT_ERR_NOMEM;
  }
  for (int q = 0; q < n; ++q) e->ctx[q]->stream = saved[q];
  return code;
}

}  // extern "C"
