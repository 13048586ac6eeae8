// Density contrast on the device.
//
// Replaces the arithmetic of BaseReconstruction.set_density_contrast
// (pyrecon/recon.py:227-247) and of OriginalMultiGridReconstruction's
// (pyrecon/multigrid.py:69-97).  The sums are deterministic FP64 reductions
// (the reference uses numpy's pairwise sums in the mesh precision); the scalar
// algebra is done by the caller with numpy's scalar types, and the per-cell
// rules by one kernel each.  The smoothing is in prh_smooth.hip.
#include "prh_reduce.hpp"

namespace prh {

constexpr int kDenThreads = 256;

template <typename R>
__global__ __launch_bounds__(kReduceThreads) void k_sums(const R* __restrict__ a,
                                                         const R* __restrict__ b, size_t n,
                                                         double* __restrict__ partial) {
  if (b) {
    block_reduce_range<2>(
        n,
        [&](size_t i, double* out) {
          out[0] = static_cast<double>(a[i]);
          out[1] = static_cast<double>(b[i]);
        },
        partial, kReduceBlocks);
  } else {
    block_reduce_range<1>(
        n, [&](size_t i, double* out) { out[0] = static_cast<double>(a[i]); }, partial,
        kReduceBlocks);
  }
}

// recon.py:229, 235, 243-244 (with randoms) or 246-247 (without; alpha = mean).
template <typename R>
__global__ __launch_bounds__(kDenThreads) void k_density(const R* __restrict__ data,
                                                         const R* __restrict__ randoms,
                                                         R* __restrict__ delta, size_t n, R alpha,
                                                         R thr, R bias_alpha, R bias) {
  const size_t i = blockIdx.x * static_cast<size_t>(kDenThreads) + threadIdx.x;
  if (i >= n) return;
  const R d = data[i];
  if (randoms) {
    const R r = randoms[i];
    const R dd = d - alpha * r;
    delta[i] = (r > thr) ? dd / (bias_alpha * r) : R(0);
  } else {
    delta[i] = (d / alpha - R(1)) / bias;
  }
}

// Martin White's density contrast (pyrecon/multigrid.py:79-83): the data
// summed over all cells except those that 0 < randoms < ran_min sets to zero,
// and the randoms summed where randoms >= ran_min.
template <typename R>
__global__ __launch_bounds__(kReduceThreads) void k_martin_sums(const R* __restrict__ data,
                                                                const R* __restrict__ randoms,
                                                                size_t n, R ran_min,
                                                                double* __restrict__ partial) {
  block_reduce_range<2>(
      n,
      [&](size_t i, double* out) {
        const R r = randoms[i];
        out[0] = (r > R(0) && r < ran_min) ? 0.0 : static_cast<double>(data[i]);
        out[1] = (r >= ran_min) ? static_cast<double>(r) : 0.0;
      },
      partial, kReduceBlocks);
}

// multigrid.py:84-95: inside the mask delta = (data / (alpha * randoms) - 1) / bias,
// outside 0 / bias; sums the non-zero delta and counts them for the mean removal.
template <typename R>
__global__ __launch_bounds__(kReduceThreads) void k_martin_contrast(
    const R* __restrict__ data, const R* __restrict__ randoms, R* __restrict__ delta, size_t n,
    R ran_min, R alpha, R bias, double* __restrict__ partial) {
  block_reduce_range<2>(
      n,
      [&](size_t i, double* out) {
        const R r = randoms[i];
        R x;
        if (r >= ran_min) {
          x = data[i] / (alpha * r);
          x = x - R(1);
          x = x / bias;
        } else {
          x = R(0) / bias;
        }
        delta[i] = x;
        const bool nz = x != R(0);
        out[0] = nz ? static_cast<double>(x) : 0.0;
        out[1] = nz ? 1.0 : 0.0;
      },
      partial, kReduceBlocks);
}

// multigrid.py:95-96: delta[delta != 0] -= mean.
template <typename R>
__global__ __launch_bounds__(kDenThreads) void k_subtract_nonzero(R* __restrict__ delta, size_t n,
                                                                  R mean) {
  const size_t i = blockIdx.x * static_cast<size_t>(kDenThreads) + threadIdx.x;
  if (i >= n) return;
  const R x = delta[i];
  if (x != R(0)) delta[i] = x - mean;
}

// Mesh-scalar arithmetic of RealMesh (pyrecon/mesh.py:156-183): out = a * s,
// a / s or s / a, elementwise in the mesh precision (numpy's result for a
// Python scalar, which takes the array's dtype).
template <typename R>
__global__ __launch_bounds__(kDenThreads) void k_scalar_op(const R* __restrict__ a,
                                                           R* __restrict__ out, size_t n, int op,
                                                           R s) {
  const size_t i = blockIdx.x * static_cast<size_t>(kDenThreads) + threadIdx.x;
  if (i >= n) return;
  const R x = a[i];
  out[i] = op == PRH_OP_MUL ? x * s : (op == PRH_OP_DIV ? x / s : s / x);
}

// Sums the nv partial rows left in ws->red by a block_reduce_range kernel and
// copies the nv results to sums (host); synchronizes the stream.
static int finish_sums(prh_workspace* ws, int nv, double* sums, hipStream_t s) {
  double* out = ws->red + 2 * kReduceBlocks;  // past the partials
  if (nv == 2) {
    k_reduce_final<2><<<1, 1024, 0, s>>>(ws->red, kReduceBlocks, kReduceBlocks, out);
  } else {
    k_reduce_final<1><<<1, 1024, 0, s>>>(ws->red, kReduceBlocks, kReduceBlocks, out);
  }
  PRH_LAUNCHED();
  PRH_HIP(hipMemcpyAsync(ws->red_host, out, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  PRH_HIP(hipStreamSynchronize(s));
  sums[0] = ws->red_host[0];
  sums[1] = nv == 2 ? ws->red_host[1] : 0.0;
  return PRH_OK;
}

template <typename R>
static int do_sums(prh_workspace* ws, const R* a, const R* b, size_t n, double* sums,
                   void* stream) {
  if (!ws || !a || !sums) return PRH_ERR_ARG;
  hipStream_t s = as_stream(stream);
  k_sums<R><<<kReduceBlocks, kReduceThreads, 0, s>>>(a, b, n, ws->red);
  PRH_LAUNCHED();
  return finish_sums(ws, b ? 2 : 1, sums, s);
}

template <typename R>
static int do_martin_sums(prh_workspace* ws, const R* data, const R* randoms, size_t n,
                          R ran_min, double* sums, void* stream) {
  if (!ws || !data || !randoms || !sums) return PRH_ERR_ARG;
  hipStream_t s = as_stream(stream);
  k_martin_sums<R><<<kReduceBlocks, kReduceThreads, 0, s>>>(data, randoms, n, ran_min, ws->red);
  PRH_LAUNCHED();
  return finish_sums(ws, 2, sums, s);
}

template <typename R>
static int do_martin_contrast(prh_workspace* ws, const R* data, const R* randoms, R* delta,
                              size_t n, R ran_min, R alpha, R bias, double* sums,
                              void* stream) {
  if (!ws || !data || !randoms || !delta || !sums) return PRH_ERR_ARG;
  hipStream_t s = as_stream(stream);
  k_martin_contrast<R><<<kReduceBlocks, kReduceThreads, 0, s>>>(data, randoms, delta, n,
                                                                 ran_min, alpha, bias, ws->red);
  PRH_LAUNCHED();
  return finish_sums(ws, 2, sums, s);
}

template <typename R>
static int do_subtract_nonzero(R* delta, size_t n, R mean, void* stream) {
  if (n && !delta) return PRH_ERR_ARG;
  if (n == 0) return PRH_OK;
  k_subtract_nonzero<R><<<ceil_div(n, kDenThreads), kDenThreads, 0, as_stream(stream)>>>(
      delta, n, mean);
  PRH_LAUNCHED();
  return PRH_OK;
}

template <typename R>
static int do_scalar_op(const R* a, R* out, size_t n, int op, double s, void* stream) {
  if (n && (!a || !out)) return PRH_ERR_ARG;
  if (op != PRH_OP_MUL && op != PRH_OP_DIV && op != PRH_OP_RDIV) {
    set_error("unknown mesh-scalar operation %d", op);
    return PRH_ERR_ARG;
  }
  if (n == 0) return PRH_OK;
  k_scalar_op<R><<<ceil_div(n, kDenThreads), kDenThreads, 0, as_stream(stream)>>>(
      a, out, n, op, static_cast<R>(s));
  PRH_LAUNCHED();
  return PRH_OK;
}

template <typename R>
static int do_density(const R* data, const R* randoms, R* delta, size_t n, R alpha, R thr, R ba,
                      R bias, void* stream) {
  if (n && (!data || !delta)) return PRH_ERR_ARG;
  if (n == 0) return PRH_OK;
  k_density<R><<<ceil_div(n, kDenThreads), kDenThreads, 0, as_stream(stream)>>>(
      data, randoms, delta, n, alpha, thr, ba, bias);
  PRH_LAUNCHED();
  return PRH_OK;
}

}  // namespace prh

using namespace prh;

extern "C" {

int prh_mesh_sums_f32(prh_workspace* ws, const float* a, const float* b, size_t size,
                      double* sums, void* stream) {
  return do_sums<float>(ws, a, b, size, sums, stream);
}
int prh_mesh_sums_f64(prh_workspace* ws, const double* a, const double* b, size_t size,
                      double* sums, void* stream) {
  return do_sums<double>(ws, a, b, size, sums, stream);
}

int prh_density_contrast_f32(const float* data, const float* randoms, float* delta, size_t size,
                             float alpha, float threshold, float bias_alpha, float bias,
                             void* stream) {
  return do_density<float>(data, randoms, delta, size, alpha, threshold, bias_alpha, bias, stream);
}
int prh_density_contrast_f64(const double* data, const double* randoms, double* delta,
                             size_t size, double alpha, double threshold, double bias_alpha,
                             double bias, void* stream) {
  return do_density<double>(data, randoms, delta, size, alpha, threshold, bias_alpha, bias,
                            stream);
}

int prh_martin_sums_f32(prh_workspace* ws, const float* data, const float* randoms, size_t size,
                        float ran_min, double* sums, void* stream) {
  return do_martin_sums<float>(ws, data, randoms, size, ran_min, sums, stream);
}
int prh_martin_sums_f64(prh_workspace* ws, const double* data, const double* randoms,
                        size_t size, double ran_min, double* sums, void* stream) {
  return do_martin_sums<double>(ws, data, randoms, size, ran_min, sums, stream);
}

int prh_martin_contrast_f32(prh_workspace* ws, const float* data, const float* randoms,
                            float* delta, size_t size, float ran_min, float alpha, float bias,
                            double* sums, void* stream) {
  return do_martin_contrast<float>(ws, data, randoms, delta, size, ran_min, alpha, bias, sums,
                                   stream);
}
int prh_martin_contrast_f64(prh_workspace* ws, const double* data, const double* randoms,
                            double* delta, size_t size, double ran_min, double alpha,
                            double bias, double* sums, void* stream) {
  return do_martin_contrast<double>(ws, data, randoms, delta, size, ran_min, alpha, bias, sums,
                                    stream);
}

int prh_subtract_nonzero_mean_f32(float* delta, size_t size, float mean, void* stream) {
  return do_subtract_nonzero<float>(delta, size, mean, stream);
}
int prh_subtract_nonzero_mean_f64(double* delta, size_t size, double mean, void* stream) {
  return do_subtract_nonzero<double>(delta, size, mean, stream);
}

int prh_mesh_scalar_op_f32(const float* a, float* out, size_t size, int op, double scalar,
                           void* stream) {
  return do_scalar_op<float>(a, out, size, op, scalar, stream);
}
int prh_mesh_scalar_op_f64(const double* a, double* out, size_t size, int op, double scalar,
                           void* stream) {
  return do_scalar_op<double>(a, out, size, op, scalar, stream);
}

}  // extern "C"
