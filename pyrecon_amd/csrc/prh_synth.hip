// Synthetic catalogs generated on the device, for load generation (the
// serving stream and benchmarks): the catalogs of a job are produced where
// they are consumed, from a seed, instead of being built on the host and
// copied.
//
// Every point i is a function of (seed, i) only (a counter-based generator),
// so a catalog is the same whatever the launch shape, stream or company.
//   * Uniform points: q uniform in the box, weight uniform in [wlo, whi).
//   * Clustered points: q uniform, then displaced by a Zel'dovich field
//     psi(q) = sum_m a_m khat_m sin(k_m . q + phase_m), a sum of random
//     longitudinal Fourier modes (Kraichnan's randomization method for a
//     Gaussian random field, the mode table drawn on the host from the power
//     spectrum), then shifted along the local line of sight by
//     f (psi . rhat) rhat and wrapped into the box.
#include "prh_common.hpp"

namespace prh {

constexpr int kSynthThreads = 256;

// splitmix64 finaliser: a bijective 64-bit mix.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Uniform double in [0, 1) from 53 random bits of stream `k` of point i.
__device__ __forceinline__ double uniform(uint64_t seed, uint64_t i, uint64_t k) {
  const uint64_t z = mix64(mix64(seed ^ (k * 0xD1B54A32D192ED03ull)) + i);
  return static_cast<double>(z >> 11) * 0x1.0p-53;
}

template <typename P>
__global__ __launch_bounds__(kSynthThreads) void k_synth(P* __restrict__ pos,
                                                         P* __restrict__ weights, size_t n,
                                                         uint64_t seed, double3 offset,
                                                         double3 box,
                                                         const double* __restrict__ modes,
                                                         int nmodes, double f, double wlo,
                                                         double whi) {
  const size_t i = blockIdx.x * static_cast<size_t>(kSynthThreads) + threadIdx.x;
  if (i >= n) return;
  double x = offset.x + box.x * uniform(seed, i, 0);
  double y = offset.y + box.y * uniform(seed, i, 1);
  double z = offset.z + box.z * uniform(seed, i, 2);
  if (nmodes > 0) {
    double sx = 0., sy = 0., sz = 0.;
    for (int m = 0; m < nmodes; m++) {
      const double* md = modes + 6 * m;  // kx, ky, kz, amplitude / |k|, phase, unused
      const double s = md[3] * sin(md[0] * x + md[1] * y + md[2] * z + md[4]);
      sx += s * md[0];
      sy += s * md[1];
      sz += s * md[2];
    }
    x += sx;
    y += sy;
    z += sz;
    const double r2 = x * x + y * y + z * z;
    if (r2 > 0.) {
      const double proj = f * (sx * x + sy * y + sz * z) / r2;
      x += proj * x;
      y += proj * y;
      z += proj * z;
    }
    // wrap into [offset, offset + box)
    x = offset.x + (x - offset.x - box.x * floor((x - offset.x) / box.x));
    y = offset.y + (y - offset.y - box.y * floor((y - offset.y) / box.y));
    z = offset.z + (z - offset.z - box.z * floor((z - offset.z) / box.z));
    if (x >= offset.x + box.x) x = offset.x;
    if (y >= offset.y + box.y) y = offset.y;
    if (z >= offset.z + box.z) z = offset.z;
  }
  pos[3 * i + 0] = static_cast<P>(x);
  pos[3 * i + 1] = static_cast<P>(y);
  pos[3 * i + 2] = static_cast<P>(z);
  if (weights) weights[i] = static_cast<P>(wlo + (whi - wlo) * uniform(seed, i, 3));
}

}  // namespace prh

using namespace prh;

extern "C" {

int prh_synth_points(void* positions, void* weights, int dtype, size_t n, uint64_t seed,
                     const double* boxsize, const double* boxcenter, const double* modes,
                     int nmodes, double f, double wlo, double whi, void* stream) {
  if (!positions || !boxsize || !boxcenter || nmodes < 0 || (nmodes > 0 && !modes) ||
      (dtype != PRH_F32 && dtype != PRH_F64)) {
    set_error("synth: bad arguments");
    return PRH_ERR_ARG;
  }
  if (n == 0) return PRH_OK;
  const double3 off = make_double3(boxcenter[0] - boxsize[0] / 2., boxcenter[1] - boxsize[1] / 2.,
                                   boxcenter[2] - boxsize[2] / 2.);
  const double3 box = make_double3(boxsize[0], boxsize[1], boxsize[2]);
  const unsigned grid = ceil_div(n, kSynthThreads);
  if (dtype == PRH_F64)
    k_synth<double><<<grid, kSynthThreads, 0, as_stream(stream)>>>(
        static_cast<double*>(positions), static_cast<double*>(weights), n, seed, off, box, modes,
        nmodes, f, wlo, whi);
  else
    k_synth<float><<<grid, kSynthThreads, 0, as_stream(stream)>>>(
        static_cast<float*>(positions), static_cast<float*>(weights), n, seed, off, box, modes,
        nmodes, f, wlo, whi);
  PRH_LAUNCHED();
  return PRH_OK;
}

}  // extern "C"
