// Elastic ion-neutral collisions, ScatterElastic (src/collisions.c:9-167), on the GPU.
//
// Pre-drawn mode reproduces the reference's serial loop exactly: attempts are
// processed in attempt order per particle, and the loop's stop at the Nc-th
// success becomes a global cut found by a scan over per-attempt outcomes.
//   1. ip of every attempt; group attempts by ip (counting sort, lists sorted by
//      attempt index).
//   2. pass 1: each particle replays its attempts in order and records each
//      attempt's outcome (rejected / collided / refused).
//   3. exclusive scans of the outcomes give the cut: the attempt after the Nc-th
//      success.  Ten refusals before the cut is the reference's Abort.
//   4. pass 2 replays the attempts below the cut and writes the velocities.
#include <algorithm>

#include "common.hpp"
#include "philox.hpp"
#include "pic_math.hpp"
#include "tiles.hpp"

namespace ecsprt {

CollideConsts collide_consts_of(const ecsprt_params &p);

// x86-64 double -> int32 (cvttsd2si): out of range or NaN gives INT_MIN.
static int x86_double_to_int(double v) {
  if (!(v > -2147483649.0 && v < 2147483648.0)) return (int)0x80000000u;
  return (int)v;
}

// Nc, src/collisions.c:48-63.
int collision_count(const ecsprt_params *p, long np, double dt) {
  double fc = p->nui * dt;
  double tmp = (p->vi0x - p->vn0x) * (p->vi0x - p->vn0x) + (p->vi0y - p->vn0y) * (p->vi0y - p->vn0y) +
               (p->vi0z - p->vn0z) * (p->vi0z - p->vn0z);
  double vrm = 4.0 * p->viT + sqrt(tmp);
  int Nc = x86_double_to_int(((double)np * fc * vrm) / ((p->vnT + p->vn0) * sqrt(p->mn / p->mi)));
  if (Nc > np) Nc = (int)np;
  else if (Nc < 0) Nc = 0;
  return Nc;
}

// ---- exclusive scan (int -> long long), 1024 elements per block --------------
__global__ void k_scan_local(const int *in, long long *out, long long *bsum, long n) {
  __shared__ long long wsum[16];
  const long base = (long)blockIdx.x * 1024;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  long long v[4], s = 0;
  for (int q = 0; q < 4; ++q) {
    long i = base + tid * 4 + q;
    v[q] = (i < n) ? in[i] : 0;
    s += v[q];
  }
  long long x = s;
  for (int o = 1; o < 64; o <<= 1) {
    long long y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  if (tid == 0) {
    long long acc = 0;
    for (int w = 0; w < (int)(blockDim.x / 64); ++w) {
      long long t = wsum[w];
      wsum[w] = acc;
      acc += t;
    }
    bsum[blockIdx.x] = acc;
  }
  __syncthreads();
  long long e = x - s + wsum[wid];
  for (int q = 0; q < 4; ++q) {
    long i = base + tid * 4 + q;
    if (i < n) out[i] = e;
    e += v[q];
  }
}

__global__ void k_scan_add(long long *out, const long long *boff, long n) {
  const long base = (long)blockIdx.x * 1024;
  const long long add = boff[blockIdx.x];
  for (int q = threadIdx.x; q < 1024; q += blockDim.x) {
    long i = base + q;
    if (i < n) out[i] += add;
  }
}

__global__ void k_i64_to_i32(const long long *in, int *out, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = (int)in[i];
}

// out[i] = sum_{j<i} in[j]; returns nothing, total available as out[n-1]+in[n-1].
static void scan_exclusive(ecsprt_hip_ctx *c, const int *in, long long *out, long n) {
  if (n <= 0) return;
  const long nb = (n + 1023) / 1024;
  long long *bsum = nullptr, *boff = nullptr;
  ECS_HIP(hipMallocAsync((void **)&bsum, sizeof(long long) * nb, c->stream));
  k_scan_local<<<nb, 256, 0, c->stream>>>(in, out, bsum, n);
  ECS_HIP(hipGetLastError());
  if (nb > 1) {
    ECS_HIP(hipMallocAsync((void **)&boff, sizeof(long long) * nb, c->stream));
    int *bsum32 = nullptr;
    ECS_HIP(hipMallocAsync((void **)&bsum32, sizeof(int) * nb, c->stream));
    k_i64_to_i32<<<(nb + 255) / 256, 256, 0, c->stream>>>(bsum, bsum32, nb);
    scan_exclusive(c, bsum32, boff, nb);
    k_scan_add<<<nb, 256, 0, c->stream>>>(out, boff, n);
    ECS_HIP(hipGetLastError());
    ECS_HIP(hipFreeAsync(bsum32, c->stream));
    ECS_HIP(hipFreeAsync(boff, c->stream));
  }
  ECS_HIP(hipFreeAsync(bsum, c->stream));
}

// ---- pre-drawn kernels --------------------------------------------------------
__global__ void k_att_count(const ecsprt_attempt *att, long na, long np, int *ipa, int *cnt) {
  for (long a = blockIdx.x * (long)blockDim.x + threadIdx.x; a < na; a += (long)gridDim.x * blockDim.x) {
    int ip = (int)pick_ip(np, att[a].u_ip);
    ipa[a] = ip;
    atomicAdd(&cnt[ip], 1);
  }
}

__global__ void k_att_place(long na, const int *ipa, const long long *off, int *fill, int *list) {
  for (long a = blockIdx.x * (long)blockDim.x + threadIdx.x; a < na; a += (long)gridDim.x * blockDim.x) {
    int ip = ipa[a];
    long long pos = off[ip] + atomicAdd(&fill[ip], 1);
    list[pos] = (int)a;
  }
}

__global__ void k_att_sort(long np, const int *cnt, const long long *off, int *list) {
  for (long ip = blockIdx.x * (long)blockDim.x + threadIdx.x; ip < np; ip += (long)gridDim.x * blockDim.x) {
    int n = cnt[ip];
    if (n < 2) continue;
    int *L = list + off[ip];
    for (int i = 1; i < n; ++i) {
      int v = L[i], j = i - 1;
      while (j >= 0 && L[j] > v) {
        L[j + 1] = L[j];
        --j;
      }
      L[j + 1] = v;
    }
  }
}

__global__ void k_slot_of_id(ParticlesF64 A, const int *base, const int *count, int *slot_of_id) {
  const int t = blockIdx.x;
  const long b = base[t], n = count[t];
  for (long q = threadIdx.x; q < n; q += blockDim.x) slot_of_id[A.id(b + q)] = (int)(b + q);
}

// mode 0: record outcomes; mode 1: apply attempts below the cut.
__global__ void k_att_replay(int mode, long np, const ecsprt_attempt *att, const int *cnt, const long long *off,
                             const int *list, const int *slot_of_id, ParticlesF64 A, CollideConsts k,
                             int *succ, int *fail, const long long *cut_ptr) {
  const long long cut = mode ? *cut_ptr : 0;
  for (long ip = blockIdx.x * (long)blockDim.x + threadIdx.x; ip < np; ip += (long)gridDim.x * blockDim.x) {
    const int n = cnt[ip];
    if (n == 0) continue;
    const int s = slot_of_id[ip];
    double v[3] = {A.vx(s), A.vy(s), A.vz(s)};
    const int *L = list + off[ip];
    for (int q = 0; q < n; ++q) {
      const int a = L[q];
      if (mode && a >= cut) break;
      const ecsprt_attempt r = att[a];
      double vf[3];
      const int res = collide_one<double>(k, r.g, r.u_acc1, r.u_acc2, r.u_z, r.u_phi, v, vf);
      if (res == 1) {
        v[0] = vf[0];
        v[1] = vf[1];
        v[2] = vf[2];
      }
      if (!mode) {
        succ[a] = res == 1;
        fail[a] = res == 2;
      }
    }
    if (mode) {
      A.vx(s) = v[0];
      A.vy(s) = v[1];
      A.vz(s) = v[2];
    }
  }
}

// The same replay on two-dimensional ions, double[np][2] velocities in ion order.
__global__ void k_att_replay2d(int mode, long np, const ecsprt_attempt *att, const int *cnt, const long long *off,
                               const int *list, const int *slot_of_id, double *vel, CollideConsts k, int *succ,
                               int *fail, const long long *cut_ptr) {
  const long long cut = mode ? *cut_ptr : 0;
  for (long ip = blockIdx.x * (long)blockDim.x + threadIdx.x; ip < np; ip += (long)gridDim.x * blockDim.x) {
    const int n = cnt[ip];
    if (n == 0) continue;
    const long s = slot_of_id[ip];
    double v[3] = {vel[2 * s], vel[2 * s + 1], 0.0};
    const int *L = list + off[ip];
    for (int q = 0; q < n; ++q) {
      const int a = L[q];
      if (mode && a >= cut) break;
      const ecsprt_attempt r = att[a];
      double vf[3];
      const int res = collide_one<double, 2>(k, r.g, r.u_acc1, r.u_acc2, r.u_z, r.u_phi, v, vf);
      if (res == 1) {
        v[0] = vf[0];
        v[1] = vf[1];
      }
      if (!mode) {
        succ[a] = res == 1;
        fail[a] = res == 2;
      }
    }
    if (mode) {
      vel[2 * s] = v[0];
      vel[2 * s + 1] = v[1];
    }
  }
}

// res[0] = cut, res[1] = successes in total, res[2] = refusals before the cut.
__global__ void k_att_cut(long na, int Nc, const int *succ, const long long *succ_ex, const int *fail,
                          const long long *fail_ex, long long *res) {
  for (long a = blockIdx.x * (long)blockDim.x + threadIdx.x; a < na; a += (long)gridDim.x * blockDim.x) {
    if (succ[a] && succ_ex[a] + 1 == Nc) res[0] = a + 1;
    if (a == na - 1) res[1] = succ_ex[a] + succ[a];
  }
  (void)fail;
  (void)fail_ex;
}

__global__ void k_att_fail_before(long na, const long long *cut_ptr, const int *fail, const long long *fail_ex,
                                  long long *res) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    long long cut = *cut_ptr;
    if (cut >= na) res[2] = fail_ex[na - 1] + fail[na - 1];
    else res[2] = fail_ex[cut];
  }
}

__global__ void k_set_i64(long long *p, long long v) { *p = v; }

void run_collide_predrawn(ecsprt_hip_ctx *c, double dt) {
  const long np = c->np, na = c->n_att;
  const int Nc = collision_count(&c->p, np, dt);
  c->last_Nc = Nc;
  c->last_Ns = 0;
  c->last_Nf = 0;
  c->last_attempts = 0;
  if (Nc == 0) return;
  ECS_CHECK(na > 0 && c->d_att, ECSPRT_ERR_ATTEMPTS, "no pre-drawn attempt records set");
  hipStream_t st = c->stream;
  int *cnt, *fill, *ipa, *list, *slot, *succ, *fail;
  long long *off, *succ_ex, *fail_ex, *res;
  ECS_HIP(hipMallocAsync((void **)&cnt, sizeof(int) * np, st));
  ECS_HIP(hipMallocAsync((void **)&fill, sizeof(int) * np, st));
  ECS_HIP(hipMallocAsync((void **)&slot, sizeof(int) * np, st));
  ECS_HIP(hipMallocAsync((void **)&off, sizeof(long long) * np, st));
  ECS_HIP(hipMallocAsync((void **)&ipa, sizeof(int) * na, st));
  ECS_HIP(hipMallocAsync((void **)&list, sizeof(int) * na, st));
  ECS_HIP(hipMallocAsync((void **)&succ, sizeof(int) * na, st));
  ECS_HIP(hipMallocAsync((void **)&fail, sizeof(int) * na, st));
  ECS_HIP(hipMallocAsync((void **)&succ_ex, sizeof(long long) * na, st));
  ECS_HIP(hipMallocAsync((void **)&fail_ex, sizeof(long long) * na, st));
  ECS_HIP(hipMallocAsync((void **)&res, sizeof(long long) * 3, st));
  ECS_HIP(hipMemsetAsync(cnt, 0, sizeof(int) * np, st));
  ECS_HIP(hipMemsetAsync(fill, 0, sizeof(int) * np, st));
  ECS_HIP(hipMemsetAsync(succ, 0, sizeof(int) * na, st));
  ECS_HIP(hipMemsetAsync(fail, 0, sizeof(int) * na, st));
  const int gb = (int)std::min<long>((std::max(na, np) + 255) / 256, 65536);
  k_att_count<<<gb, 256, 0, st>>>(c->d_att, na, np, ipa, cnt);
  scan_exclusive(c, cnt, off, np);
  k_att_place<<<gb, 256, 0, st>>>(na, ipa, off, fill, list);
  k_att_sort<<<gb, 256, 0, st>>>(np, cnt, off, list);
  const bool two_d = c->p.ndim == 2;
  if (two_d) p2d_slot_of_id(c, slot);
  else k_slot_of_id<<<c->ntiles, 256, 0, st>>>(c->A, c->d_base, c->d_count, slot);
  CollideConsts k = collide_consts_of(c->p);
  if (two_d)
    k_att_replay2d<<<gb, 256, 0, st>>>(0, np, c->d_att, cnt, off, list, slot, c->d_vel2, k, succ, fail, nullptr);
  else
    k_att_replay<<<gb, 256, 0, st>>>(0, np, c->d_att, cnt, off, list, slot, c->A, k, succ, fail, nullptr);
  scan_exclusive(c, succ, succ_ex, na);
  scan_exclusive(c, fail, fail_ex, na);
  k_set_i64<<<1, 1, 0, st>>>(res, (long long)na);
  k_att_cut<<<gb, 256, 0, st>>>(na, Nc, succ, succ_ex, fail, fail_ex, res);
  k_att_fail_before<<<1, 1, 0, st>>>(na, res, fail, fail_ex, res);
  long long h[3];
  ECS_HIP(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, st));
  ECS_HIP(hipStreamSynchronize(st));
  const long long cut = h[0], total_succ = h[1], fails = h[2];
  int rc = ECSPRT_OK;
  std::string msg;
  if (fails >= 10) {
    rc = ECSPRT_ERR_COLLISION;
    msg = "failed to collide " + std::to_string(fails) + " ion-neutral pairs";
  } else if (total_succ < Nc) {
    rc = ECSPRT_ERR_ATTEMPTS;
    msg = "pre-drawn records exhausted after " + std::to_string(total_succ) + " of " + std::to_string(Nc) +
          " collisions";
  } else {
    if (two_d)
      k_att_replay2d<<<gb, 256, 0, st>>>(1, np, c->d_att, cnt, off, list, slot, c->d_vel2, k, succ, fail, res);
    else
      k_att_replay<<<gb, 256, 0, st>>>(1, np, c->d_att, cnt, off, list, slot, c->A, k, succ, fail, res);
    ECS_HIP(hipGetLastError());
  }
  void *bufs[] = {cnt, fill, slot, off, ipa, list, succ, fail, succ_ex, fail_ex, res};
  for (void *b : bufs) ECS_HIP(hipFreeAsync(b, st));
  ECS_HIP(hipStreamSynchronize(st));
  c->last_Ns = std::min<long long>(total_succ, Nc);
  c->last_Nf = fails;
  c->last_attempts = rc == ECSPRT_OK ? cut : na;
  if (rc != ECSPRT_OK) throw Error{rc, msg};
}

// Acceptance estimate for the Philox mode: a dry run of the per-particle Poisson
// process (velocities are not written) on a stream independent of the real steps.
__device__ inline void pilot_ion(const ParticlesF64 &A, long s, double &vx, double &vy, double &vz, int &id) {
  vx = A.vx(s);
  vy = A.vy(s);
  vz = A.vz(s);
  id = A.id(s);
}

__device__ inline void pilot_ion(const ParticlesMix &A, long s, double &vx, double &vy, double &vz, int &id) {
  const uint32_t *b = A.p + mx_word(s, 0);
  vx = __uint_as_float(b[64 * MX_VX]);
  vy = __uint_as_float(b[64 * MX_VY]);
  vz = __uint_as_float(b[64 * MX_VZ]);
  id = (int)b[64 * MX_ID];
}

// Pilot pass `tag` with sampling stride `stride` (a power of two) visits every ion of one
// tile out of each run of `stride` consecutive tiles, picked by a hash of (run, tag): a
// stratified sample of whole tiles, so the result does not depend on the order of the
// ions inside a tile.  Each sampled tile is shared by gridDim.y blocks.
__device__ inline int pilot_tile(int j, uint32_t tag, int stride) {
  uint32_t h = (uint32_t)j * 0x9E3779B1u ^ tag * 0x85EBCA6Bu;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  return j * stride + (int)(h & (uint32_t)(stride - 1));
}

// Per-workgroup sums (attempts, collisions, ions visited) go to part[3 * workgroup]: the
// thousands of workgroups of one pass would otherwise end on the same three global
// atomics, serialised after the last of them.  The ions with attempts (about one in
// eight at 128^3 x 100 ions per cell) are queued in LDS and run their attempts with every
// lane busy, PILOT_CHUNK at a time.
constexpr int PILOT_CHUNK = 256, PILOT_QUEUE = 2 * PILOT_CHUNK;

template <class P>
__global__ void __launch_bounds__(PILOT_CHUNK) k_pilot(P A, const int *base, const int *count, int ntiles,
                                                       CollideConsts k, PoissonParams pt, uint32_t key0, uint32_t key1,
                                                       uint32_t tag, int stride, unsigned long long *part) {
  __shared__ double qv[3][PILOT_QUEUE];
  __shared__ int qid[PILOT_QUEUE], qk[PILOT_QUEUE], qn;
  __shared__ unsigned long long wsum[PILOT_CHUNK / 64][2];
  const int t = pilot_tile(blockIdx.x, tag, stride);
  const int tid = threadIdx.x;
  const long b = t < ntiles ? base[t] : 0, n = t < ntiles ? count[t] : 0;
  unsigned natt = 0, nsucc = 0, nfail = 0;
  auto run = [&](int e) {  // the attempts of queue entry e
    double vx = qv[0][e], vy = qv[1][e], vz = qv[2][e];
    collide_attempts<double>(k, key0, key1, tag, qid[e], 0, qk[e], vx, vy, vz, natt, nsucc, nfail);
  };
  if (tid == 0) qn = 0;
  __syncthreads();
  const long step = (long)PILOT_CHUNK * gridDim.y;
  for (long c0 = (long)blockIdx.y * PILOT_CHUNK; c0 < n; c0 += step) {
    const long q = c0 + tid;
    if (q < n) {
      double vx, vy, vz;
      int id;
      pilot_ion(A, b + q, vx, vy, vz, id);
      const int na = poisson_attempts(pt, key0, key1, tag, id);
      if (na > 0) {  // at most PILOT_CHUNK entries join a queue holding fewer than PILOT_CHUNK
        const int r = atomicAdd(&qn, 1);
        qv[0][r] = vx;
        qv[1][r] = vy;
        qv[2][r] = vz;
        qid[r] = id;
        qk[r] = na;
      }
    }
    __syncthreads();
    const int m = qn;  // every thread reads it before any changes it again
    __syncthreads();
    if (m >= PILOT_CHUNK) {
      run(m - PILOT_CHUNK + tid);
      __syncthreads();
      if (tid == 0) qn = m - PILOT_CHUNK;
      __syncthreads();
    }
  }
  if (tid < qn) run(tid);
  const unsigned long long na = wave_sum(natt), ns = wave_sum(nsucc);
  if ((tid & 63) == 0) {
    wsum[tid >> 6][0] = na;
    wsum[tid >> 6][1] = ns;
  }
  __syncthreads();
  if (tid == 0) {
    unsigned long long sa = 0, sc = 0;
    for (int w = 0; w < PILOT_CHUNK / 64; ++w) {
      sa += wsum[w][0];
      sc += wsum[w][1];
    }
    unsigned long long *o = part + 3 * ((long)blockIdx.y * gridDim.x + blockIdx.x);
    o[0] = sa;
    o[1] = sc;
    o[2] = blockIdx.y == 0 ? (unsigned long long)n : 0ull;
  }
}

// ctr[6] attempts, ctr[7] collisions, ctr[8] ions visited: the sums of the nwg partials.
__global__ void k_pilot_reduce(const unsigned long long *part, int nwg, unsigned long long *ctr) {
  __shared__ unsigned long long s[3][256];
  unsigned long long a = 0, c = 0, v = 0;
  for (int i = threadIdx.x; i < nwg; i += blockDim.x) {
    a += part[3 * i];
    c += part[3 * i + 1];
    v += part[3 * i + 2];
  }
  s[0][threadIdx.x] = a;
  s[1][threadIdx.x] = c;
  s[2][threadIdx.x] = v;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int j = 0; j < 3; ++j) s[j][threadIdx.x] += s[j][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < 3) ctr[6 + threadIdx.x] = s[threadIdx.x][0];
}

struct PilotResult {
  double collisions, ions;
};

// One dry run of the Poisson process at rate lam on the tiles of sample `tag`, queued on
// the context's stream; its sums reach the pinned counter mirror (pilot_result) once the
// stream has synchronised.
static void pilot_enqueue(ecsprt_hip_ctx *c, uint32_t tag, double lam, int stride) {
  CollideConsts k = collide_consts_of(c->p);
  const uint32_t k0 = (uint32_t)c->o.seed, k1 = (uint32_t)(c->o.seed >> 32);
  const PoissonParams pp = poisson_params(lam);
  // Blocks per sampled tile: enough for about 2048 blocks in all, at most one per 256
  // ions of an average tile, so that a sampled pass is not a few long serial blocks.
  const int ns = (c->ntiles + stride - 1) / stride;
  const long per_tile = c->np / std::max(1, c->ntiles);
  const int split = (int)std::max<long>(1, std::min<long>((2048 + ns - 1) / ns, per_tile / 256));
  const long nwg = (long)ns * split;
  if (nwg > c->pilot_cap) {
    (void)hipFree(c->d_pilot);
    c->d_pilot = nullptr;
    c->pilot_cap = 0;
    if (hipMalloc(&c->d_pilot, sizeof(unsigned long long) * 3 * nwg) != hipSuccess)
      throw Error{ECSPRT_ERR_NOMEM, "collision pilot partial sums"};
    c->pilot_cap = nwg;
  }
  const dim3 grid(ns, split);
  if (c->mixed)
    k_pilot<<<grid, PILOT_CHUNK, 0, c->stream>>>(c->Am, c->d_base, c->d_count, c->ntiles, k, pp, k0, k1, tag,
                                                 stride, c->d_pilot);
  else
    k_pilot<<<grid, PILOT_CHUNK, 0, c->stream>>>(c->A, c->d_base, c->d_count, c->ntiles, k, pp, k0, k1, tag,
                                                 stride, c->d_pilot);
  k_pilot_reduce<<<1, 256, 0, c->stream>>>(c->d_pilot, (int)nwg, c->d_ctr);
  ECS_HIP(hipGetLastError());
  ECS_HIP(hipMemcpyAsync(c->h_ctr, c->d_ctr, sizeof(unsigned long long) * CTR_N, hipMemcpyDeviceToHost, c->stream));
}

static PilotResult pilot_result(const ecsprt_hip_ctx *c) { return {(double)c->h_ctr[7], (double)c->h_ctr[8]}; }

static PilotResult pilot_pass(ecsprt_hip_ctx *c, uint32_t tag, double lam, int stride) {
  pilot_enqueue(c, tag, lam, stride);
  ECS_HIP(hipStreamSynchronize(c->stream));
  return pilot_result(c);
}

// Initial attempt rate: dry runs of the Poisson process over every ion, each rescaling
// lambda by Nc / (collisions it produced), starting from one attempt per collision.
void run_collide_philox_pilot(ecsprt_hip_ctx *c, double dt) {
  const int Nc = collision_count(&c->p, c->np, dt);
  double lam = c->np > 0 ? (double)Nc / (double)c->np : 0.0;
  for (int it = 0; it < 4 && Nc > 0 && lam > 0; ++it)
    lam = pilot_rescale(lam, Nc, pilot_pass(c, 0xFFFFFFF0u + it, std::min(lam, POISSON_LAMBDA_MAX), 1).collisions);
  pilot_check(lam);
  c->coll_lambda = lam;
}

// The attempt rate of this step, set from the current velocities before the step's
// draws: up to two dry-run passes on a sample of the tiles (at least 32 tiles and about
// 2 * 10^4 expected collisions, at least 1/64 of the ions), each rescaling lambda by
// Nc / (predicted collisions), on the same sample so that the second corrects the
// first's bias.  A rate carried over from the last step lags the state: in a run that
// cools on cold neutrals the acceptance falls every step and the collisions would fall
// short of Nc.  predict_begin sets up the passes, predict_enqueue queues one and
// predict_collect (after the stream has synchronised) applies it and says whether
// another is due, so that an ensemble can run the passes of all its runs together.
void predict_begin(ecsprt_hip_ctx *c, int Nc, PredictState &s) {
  s = PredictState{};
  if (Nc <= 0 || c->np <= 0 || c->coll_lambda <= 0.0) return;
  s.began = s.active = true;
  s.Nc = Nc;
  s.stride = 1;
  while (s.stride < 64 && (double)Nc / (2 * s.stride) >= 2.0e4 && c->ntiles / (2 * s.stride) >= 32) s.stride *= 2;
  s.tag = 0x5A17E000u ^ (uint32_t)c->step_index;
  s.lam = c->coll_lambda;
}

void predict_enqueue(ecsprt_hip_ctx *c, const PredictState &s) { pilot_enqueue(c, s.tag, s.lam, s.stride); }

bool predict_collect(ecsprt_hip_ctx *c, PredictState &s) {
  const PilotResult r = pilot_result(c);
  s.active = false;
  if (!(r.collisions > 0.0) || !(r.ions > 0.0)) return false;
  const double predicted = r.collisions * (double)c->np / r.ions;
  const double ratio = std::min(2.0, std::max(0.5, (double)s.Nc / predicted));
  s.lam = std::min(s.lam * ratio, POISSON_LAMBDA_MAX);
  // A rescale within 1% leaves a second-order error in the rescale itself, far below
  // the sample's own noise: the second pass is skipped.
  s.active = ++s.it < 2 && std::fabs(ratio - 1.0) > 0.01;
  return s.active;
}

void predict_end(ecsprt_hip_ctx *c, const PredictState &s) {
  if (s.began) c->coll_lambda = s.lam;
}

void collide_predict(ecsprt_hip_ctx *c, int Nc) {
  PredictState s;
  predict_begin(c, Nc, s);
  while (s.active) {
    predict_enqueue(c, s);
    ECS_HIP(hipStreamSynchronize(c->stream));
    predict_collect(c, s);
  }
  predict_end(c, s);
}

}  // namespace ecsprt
