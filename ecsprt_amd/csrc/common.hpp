// Shared internals of libecsprt_hip: context layout, tile geometry, error plumbing.
#pragma once
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdio.h>
#include <string>
#include <vector>

#include "../../include/ecsprt_hip.h"

namespace ecsprt {

// Error raised inside the library; entry points convert it to an int code.
struct Error {
  int code;
  std::string msg;
};

#define ECS_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      throw ::ecsprt::Error{ECSPRT_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
    }                                                                                  \
  } while (0)

#define ECS_CHECK(cond, code, msg)                        \
  do {                                                    \
    if (!(cond)) throw ::ecsprt::Error{(code), (msg)};    \
  } while (0)

// FP64 particle storage in blocks of 64 ions (one wavefront): eight 512-byte planes per
// block, x, y, z, vx, vy, vz as double, then the 64 int32 ids (the last 256 bytes of the
// seventh plane and the eighth plane are unused).  A wave's field is one contiguous
// 512-byte run, and one address per ion reaches every field through immediate offsets,
// so a kernel carries one pointer per particle array instead of seven.
struct ParticlesF64 {
  double *p = nullptr;
  static constexpr long BLOCK = 512;  // doubles per 64-ion block
  __host__ __device__ static long at(long i) { return (i >> 6) * BLOCK + (i & 63); }
  __host__ __device__ double &x(long i) const { return p[at(i)]; }
  __host__ __device__ double &y(long i) const { return p[at(i) + 64]; }
  __host__ __device__ double &z(long i) const { return p[at(i) + 128]; }
  __host__ __device__ double &vx(long i) const { return p[at(i) + 192]; }
  __host__ __device__ double &vy(long i) const { return p[at(i) + 256]; }
  __host__ __device__ double &vz(long i) const { return p[at(i) + 320]; }
  __host__ __device__ int &id(long i) const {
    return reinterpret_cast<int *>(p + (i >> 6) * BLOCK + 384)[i & 63];
  }
  static size_t bytes(long n) { return (size_t)((n + 63) / 64) * BLOCK * sizeof(double); }
};

// Mixed-precision particle storage: a packed integer cell index (x in bits 0-10,
// y in bits 11-21, z in bits 22-31), the FP32 offset within the cell in units of the
// cell width, the FP32 velocity and the id, in blocks of 64 ions (one wavefront):
// eight 64-word planes per block.  A wave's fields are whole 256-byte lines and one
// address per ion reaches all eight fields through immediate offsets.
struct ParticlesMix {
  uint32_t *p = nullptr;
};

enum : int { MX_CELL = 0, MX_OX = 1, MX_OY = 2, MX_OZ = 3, MX_VX = 4, MX_VY = 5, MX_VZ = 6, MX_ID = 7 };

// Word index of field f of ion i.
__host__ __device__ inline long mx_word(long i, int f) { return (i >> 6) * 512 + f * 64 + (i & 63); }

constexpr int MX_MAX_NX = 2048, MX_MAX_NY = 2048, MX_MAX_NZ = 1024;

__host__ __device__ inline uint32_t mx_pack(int i, int j, int k) {
  return (uint32_t)i | ((uint32_t)j << 11) | ((uint32_t)k << 22);
}
__host__ __device__ inline int mx_i(uint32_t c) { return (int)(c & 2047u); }
__host__ __device__ inline int mx_j(uint32_t c) { return (int)((c >> 11) & 2047u); }
__host__ __device__ inline int mx_k(uint32_t c) { return (int)(c >> 22); }

// Geometry of the cell tiling.  A tile is a T^3 brick of cells (clipped at the upper
// grid faces).  Tile t's particles live in slots [base[t], base[t]+count[t]) of the
// particle arrays; capacity cap[t] leaves room for immigrants.
struct TileGeom {
  int Nx, Ny, Nz;
  int Tx, Ty, Tz;     // cells per tile along each axis
  int ntx, nty, ntz;
  int H;              // deposit halo (cells) of the accumulation region
  int RAx, RAy, RAz;  // accumulation region: T+1+2H nodes per axis, origin at tile origin - H
  int split;          // workgroups per tile
};

struct PushConsts {
  double dx, dy, dz;
  double t, s;
  double Escale[3];
  // When every cell width is a power of two, pos*inv equals pos/d exactly and the
  // kernels multiply instead of dividing.
  double inv[3];
  int pow2;
};

__host__ __device__ inline double cell_coord(double v, double d, double inv, int pow2) {
  return pow2 ? v * inv : v / d;
}

struct DriftConsts {
  double dt;
  double L[3], r0[3], r1[3];
};

// Bump allocators for solver workspace.  A context with an arena (the slots of a
// batch) carves its per-problem buffers from one device and one pinned-host block,
// so a batch allocates nothing between problems; reset() releases everything.
struct Arena {
  char *dev = nullptr, *host = nullptr;
  size_t dev_cap = 0, dev_used = 0, host_cap = 0, host_used = 0;
  void reset() { dev_used = host_used = 0; }
};

// Per-axis fold table: for global node i, up to FOLD_MAXP (tile index, local index) pairs.
constexpr int FOLD_MAXP = 4;

struct Solver;  // solver.hip

}  // namespace ecsprt

struct ecsprt_hip_ctx {
  ecsprt_params p;
  ecsprt_options o;
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;

  // tiling
  ecsprt::TileGeom g{};
  int ntiles = 0;
  long np = 0;
  long pcap = 0;
  std::vector<int> h_base, h_cap;
  int *d_base = nullptr, *d_cap = nullptr, *d_count = nullptr, *d_next = nullptr;
  ecsprt::ParticlesF64 A, B, E;  // current, next, emigrants
  ecsprt::ParticlesF64 S;        // spill for tiles that overflow their capacity
  ecsprt::ParticlesF64 D;        // ions deferred to the collision phase (tile layout of A)
  // Mixed-precision storage (particle_precision == ECSPRT_PARTICLES_MIXED): same roles.
  bool mixed = false;
  ecsprt::ParticlesMix Am, Bm, Em, Sm, Dm;
  long scap = 0;
  // counters: [0] emigrants, [1] spilled, [2] lost, [3] collision attempts,
  // [4] collision successes, [5] collision refusals, [6..8] collision pilot (attempts,
  // collisions, ions visited); CTR_N in all
  unsigned long long *d_ctr = nullptr;
  unsigned long long *h_ctr = nullptr;  // pinned mirror
  unsigned long long *d_pilot = nullptr;  // per-workgroup partial sums of a pilot pass
  long pilot_cap = 0;                     // workgroups they have room for

  // grid vectors
  long nnode = 0;
  double *d_phi = nullptr;      // [Nz][Ny][Nx]
  double *d_moments = nullptr;  // [Nz][Ny][Nx][4]
  double *d_fold = nullptr;     // per workgroup accumulation regions
  double *d_ovf = nullptr;      // [Nz][Ny][Nx][4] overflow deposits (global atomics)
  void *d_late = nullptr;       // per-tile lists of ions whose deposit is done by k_late (cap/4 entries each)
  int *d_late_cnt = nullptr;    // list lengths, per tile
  double *d_rhs = nullptr;      // [Nz][Ny][Nx]
  int *d_foldtab = nullptr;     // 3 axes x (count, pairs)
  int foldtab_off[3] = {0, 0, 0};
  size_t lds_bytes = 0;
  int block = 1024;  // threads per tile workgroup in effect
  int fx_bits = 40;  // FP64 path: fixed-point bits of the deposit regions (set at binning)

  // two-dimensional runs (p.ndim == 2): ions double[np][2] in the reference's layout,
  // fixed-point moment sums [Ny][Nx][3], per-cell counts plus their maximum
  double *d_pos2 = nullptr, *d_vel2 = nullptr;
  double *d_pos2b = nullptr, *d_vel2b = nullptr;  // the other buffer of the tile-order copy
  int *d_id2 = nullptr, *d_id2b = nullptr;        // index each ion was set with
  int *d_tile2 = nullptr;                          // each ion's tile, recorded by the fused push
  bool tiled2 = false;    // ions in the tile order and tile ranges of the last deposit, not moved since
  bool counted2 = false;  // per-tile counts of the current positions left by the fused push
  long cap2 = 0;
  unsigned long long *d_acc2 = nullptr;
  int *d_cnt2 = nullptr;
  int *d_perm2 = nullptr, *d_tiles2 = nullptr;  // ion indices binned by tile; tile counts and offsets
  long perm_cap2 = 0;
  int *d_seg2 = nullptr;  // colliding ions per list segment
  long seg_cap2 = 0;

  // staging (host<->device AoS)
  double *d_stage = nullptr;
  long stage_cap = 0;
  int *d_key = nullptr;
  long key_cap = 0;

  // collisions
  ecsprt_attempt *d_att = nullptr;
  long att_cap = 0, n_att = 0;
  int last_Nc = 0;
  long last_Ns = 0, last_Nf = 0, last_attempts = 0;
  double coll_lambda = -1.0;  // Philox mode: mean attempts per ion per step
  uint64_t step_index = 0;
  uint64_t lambda_step = ~0ull;  // step whose attempt rate coll_lambda was set for

  // solver
  ecsprt::Solver *solver = nullptr;
  ecsprt::Arena *arena = nullptr;  // null: solver buffers come from hipMalloc
  bool operator_built = false;
  bool rhs_sinusoidal = false;  // forcing of ComputeSinusoidalRHS3D (LoadFluidQuantities given no file)
  int last_iters = 0;
  double last_rnorm = 0.0, last_bnorm = 0.0;
  double solve_bytes = 0.0;  // algorithmic HBM bytes of the last potential solve (forcing included)

  // timing
  hipEvent_t ev[8] = {};
  double stage_ms[6] = {0, 0, 0, 0, 0, 0};
};

namespace ecsprt {

// particles.hip
void tiles_init(ecsprt_hip_ctx *c);
void tiles_free(ecsprt_hip_ctx *c);
void particles_bin_from_stage(ecsprt_hip_ctx *c, long np, bool ids_from_stage);
void particles_to_stage(ecsprt_hip_ctx *c);
void particles_init_device(ecsprt_hip_ctx *c, long np, uint64_t seed, double amp, int k);
void particles_init_distribution(ecsprt_hip_ctx *c, long np, int pdist);

// Two-dimensional ion stages (particles2d.hip).
void p2d_reserve(ecsprt_hip_ctx *c, long np);
void p2d_free(ecsprt_hip_ctx *c);
void p2d_set(ecsprt_hip_ctx *c, long np, const double *pos, const double *vel, hipMemcpyKind kind);
void p2d_get(ecsprt_hip_ctx *c, double *pos, double *vel);
void p2d_slot_of_id(ecsprt_hip_ctx *c, int *slot);
void p2d_init_device(ecsprt_hip_ctx *c, long np, uint64_t seed, double amp, int k);
void p2d_push(ecsprt_hip_ctx *c, double dt);
bool p2d_push_collide_drift(ecsprt_hip_ctx *c, double dt);
void p2d_drift(ecsprt_hip_ctx *c, double dt);
void p2d_apply_bc(ecsprt_hip_ctx *c, long np, double *pos);
void p2d_cell_counts(ecsprt_hip_ctx *c, int *d_counts);
void p2d_deposit(ecsprt_hip_ctx *c);
void p2d_collide_philox(ecsprt_hip_ctx *c, double dt);
void run_push(ecsprt_hip_ctx *c, double dt);
void run_drift_rebin(ecsprt_hip_ctx *c, double dt);
void run_deposit(ecsprt_hip_ctx *c);
void run_fused_particles(ecsprt_hip_ctx *c, double dt);
// Streams of the runs of a lockstep step, one per run, with an event each and one for the
// fork: the runs' small per-run launches (collision-rate passes, re-binning, late
// deposits, folds) go out on them, concurrently, between a fork from and a join into the
// step's stream.
struct RunStreams {
  std::vector<hipStream_t> st;
  std::vector<hipEvent_t> done;
  hipEvent_t fork = nullptr;
};
// Run r's launches go to its own stream (after everything queued on st so far) from
// leave(r) to back(r), when rs is given; join() makes st wait for the runs that left.
struct RunFork {
  const std::vector<ecsprt_hip_ctx *> &cs;
  hipStream_t st;
  const RunStreams *rs;
  std::vector<char> out;
  RunFork(const std::vector<ecsprt_hip_ctx *> &c, hipStream_t s, const RunStreams *r);
  void leave(int r);
  void back(int r);
  void join();
};
// One fused particle step of several 3-D runs on stream st, one launch per stage for all
// of them; the kernel arguments are staged through h_args (pinned) and d_args, each at
// least tile_args_bytes() rounded up to 256, per run.  With rs, the per-run launches run
// on the runs' own streams.
void run_fused_particles_runs(const std::vector<ecsprt_hip_ctx *> &cs, hipStream_t st, double dt, void *h_args,
                              void *d_args, std::vector<int> &rc, hipEvent_t ev_pass0, hipEvent_t ev_pass1,
                              const RunStreams *rs = nullptr);
size_t tile_args_bytes();
void run_cell_counts(ecsprt_hip_ctx *c, int *d_counts);
void run_apply_bc_aos(ecsprt_hip_ctx *c, long np, double *pos);
void ensure_stage(ecsprt_hip_ctx *c, long np);

// collide.hip
void run_collide_predrawn(ecsprt_hip_ctx *c, double dt);
void run_collide_philox_pilot(ecsprt_hip_ctx *c, double dt);
void collide_predict(ecsprt_hip_ctx *c, int Nc);
struct PredictState {
  bool began = false, active = false;
  int Nc = 0, stride = 1, it = 0;
  uint32_t tag = 0;
  double lam = 0.0;
};
void predict_begin(ecsprt_hip_ctx *c, int Nc, PredictState &s);
void predict_enqueue(ecsprt_hip_ctx *c, const PredictState &s);
bool predict_collect(ecsprt_hip_ctx *c, PredictState &s);
void predict_end(ecsprt_hip_ctx *c, const PredictState &s);
void run_collide_philox(ecsprt_hip_ctx *c, double dt);  // particles.hip
int collision_count(const ecsprt_params *p, long np, double dt);

// batch.hip: the synthetic density generator on stream st (fields addressed as
// base[q * stride]; see ecsprt_hip_lognormal_moments).
constexpr int LN_BLOCKS = 256;
hipError_t lognormal_launch(hipStream_t st, int Nx, int Ny, int Nz, uint64_t seed, double sigma, double smooth, int R,
                            double fs, double *gen, int sg, double *b1, int s1, double *b2, int s2, double *w,
                            double *partial, double *moments);

// bsolve.hip: the batched solves in lockstep (one launch per phase for the whole batch).
struct LockstepBatch;
LockstepBatch *lockstep_create(const ecsprt_params &p, const ecsprt_options &o, int nprob, const int *dims);
// One problem per params entry (its grid and physics), the shared solver options o.
LockstepBatch *lockstep_create_runs(const std::vector<ecsprt_params> &pp, const ecsprt_options &o);
void lockstep_destroy(LockstepBatch *b);
hipStream_t lockstep_stream(const LockstepBatch &b);
int lockstep_solve(LockstepBatch &b, double *const *moments, double *const *phi);
int lockstep_solve_given(LockstepBatch &b, const double *const *dens, const double *const *rhs, double *const *phi,
                         const std::vector<char> &take);
void lockstep_stats(const LockstepBatch &b, int *iters, double *rnorm, double *bytes);
void lockstep_bnorm(const LockstepBatch &b, double *bnorm);

// api.hip: the message ecsprt_hip_last_error(NULL) returns on this thread.
void set_global_error(const std::string &msg);

// solver.hip
void solver_create(ecsprt_hip_ctx *c);
void solver_destroy(ecsprt_hip_ctx *c);
void solver_build_operator(ecsprt_hip_ctx *c);
void solver_rhs(ecsprt_hip_ctx *c, double *b);
const double *solver_density(const ecsprt_hip_ctx *c);
void solver_matmult(ecsprt_hip_ctx *c, const double *x, double *y);
void solver_pcapply(ecsprt_hip_ctx *c, const double *r, double *z);
void solver_remove_mean(ecsprt_hip_ctx *c, double *v);
int solver_solve(ecsprt_hip_ctx *c, const double *b, double *x);
size_t solver_arena_bytes(const ecsprt_options &o, long n);
int solver_standalone(ecsprt_hip_ctx *c);

// Workgroups are dispatched round-robin over the 8 XCDs; with a grid of a multiple of 8
// each XCD takes a contiguous run of workgroup indices, so neighbouring tiles share an L2.
__device__ inline int xcd_order(int b, int nwg) { return (nwg & 7) == 0 ? (b & 7) * (nwg >> 3) + (b >> 3) : b; }

}  // namespace ecsprt
