// The allocation-free variant of the synthetic density generator, for callers that cannot
// call hipMalloc (code under HIP graph capture, tight serving loops, harnesses that forbid
// allocation).  It launches the kernels of ecsprt_hip_lognormal_moments on the caller's
// stream: the field and its smoothing passes live in the flux components of the caller's
// FluidNode output, and the block partial sums and smoothing weights in `scratch`, a
// fixed device array of this file.
#include <cmath>
#include <mutex>

#include "common.hpp"

namespace ecsprt {

// Workspace: LN_BLOCKS partial sums, then up to 2 LN_R_MAX + 1 smoothing weights (8 KB).
// Each device holds its own copy until the process ends.  Every call on a device shares
// it; each call writes all it reads before reading it, so nothing needs clearing, and
// calls that overlap on one device race.
constexpr int LN_R_MAX = 383;
constexpr int LN_SCRATCH = LN_BLOCKS + 2 * LN_R_MAX + 1;
static __device__ __align__(256) double scratch[LN_SCRATCH];

// Device address of `scratch` on the current device, looked up with that device current
// on the first call there; null when the lookup fails.  Host code uses only this address,
// never the array's name, which is a host-side stand-in.
static double *scratch_address() {
  static std::mutex mu;
  static double *addr[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    addr[dev] = static_cast<double *>(p);
  }
  return addr[dev];
}

}  // namespace ecsprt

extern "C" int ecsprt_hip_lognormal_moments_noalloc(int Nx, int Ny, int Nz, uint64_t seed, double sigma,
                                                    double smooth, double flux_scale, double *moments, void *stream) {
  if (Nx < 1 || Ny < 1 || Nz < 1 || !moments || !(smooth >= 0)) return ECSPRT_ERR_ARG;
  const int R = smooth > 0 ? (int)std::ceil(4.0 * smooth) : 0;
  if (R > ecsprt::LN_R_MAX) {
    ecsprt::set_global_error("ecsprt_hip_lognormal_moments_noalloc: smoothing radius ceil(4 smooth) = " +
                             std::to_string(R) + " exceeds the workspace limit of " +
                             std::to_string(ecsprt::LN_R_MAX) + " cells (smooth <= 95.75)");
    return ECSPRT_ERR_ARG;
  }
  double *ws = ecsprt::scratch_address();
  if (!ws) {
    ecsprt::set_global_error("ecsprt_hip_lognormal_moments_noalloc: device address of the workspace not found");
    return ECSPRT_ERR_HIP;
  }
  // Four doubles per node: the field is drawn into Gx and smoothed between the flux
  // components (x: Gx -> Gy, y: Gy -> Gz, z: Gz -> Gy); each node's thread then reads its
  // own Gy and overwrites its own four values.
  const hipError_t e = ecsprt::lognormal_launch(static_cast<hipStream_t>(stream), Nx, Ny, Nz, seed, sigma, smooth, R,
                                                flux_scale, moments + 1, 4, moments + 2, 4, moments + 3, 4,
                                                ws + ecsprt::LN_BLOCKS, ws, moments);
  if (e != hipSuccess) {
    ecsprt::set_global_error(std::string("ecsprt_hip_lognormal_moments_noalloc: ") + hipGetErrorString(e));
    return ECSPRT_ERR_HIP;
  }
  return ECSPRT_OK;
}
