// Gaussian smoothing and density contrast on the device.
//
// Smoothing replaces RealMesh.smooth_gaussian(method='fft')
// (pyrecon/mesh.py:655-680): the reference multiplies the full complex
// spectrum by exp(-sum_i (r k_i)^2 / 2) with k_i = 2 pi fftfreq(n_i, L_i/n_i),
// using numpy's complex-to-complex transforms.  The field is real, so the
// half spectrum of a real-to-complex transform carries the same information
// (the reference's own test_hermitian checks that both agree); hipFFT does
// the transforms and one kernel applies the Gaussian and numpy's 1/N
// normalisation of the inverse transform.
//
// The density contrast replaces the arithmetic of
// BaseReconstruction.set_density_contrast (pyrecon/recon.py:227-247).  The
// two sums are deterministic FP64 reductions (the reference uses numpy's
// pairwise sums in the mesh precision); the scalar algebra is done by the
// caller with numpy's scalar types, and the per-cell rule by one kernel.
#include "prh_reduce.hpp"

namespace prh {

constexpr int kDenThreads = 256;

template <typename R>
struct Cplx {
  R re, im;
};

template <typename R>
__global__ __launch_bounds__(kDenThreads) void k_gauss_multiply(Cplx<R>* __restrict__ spec, int n0,
                                                                int n1, int n2, int nh, double v0,
                                                                double v1, double v2, double r0,
                                                                double r1, double r2,
                                                                double norm) {
  const size_t total = static_cast<size_t>(n0) * n1 * nh;
  const size_t i = blockIdx.x * static_cast<size_t>(kDenThreads) + threadIdx.x;
  if (i >= total) return;
  const int iz = static_cast<int>(i % nh);
  const size_t t = i / nh;
  const int iy = static_cast<int>(t % n1), ix = static_cast<int>(t / n1);
  // numpy.fft.fftfreq ordering: i < (n-1)//2 + 1 -> i, else i - n
  const int fx = ix < (n0 - 1) / 2 + 1 ? ix : ix - n0;
  const int fy = iy < (n1 - 1) / 2 + 1 ? iy : iy - n1;
  const int fz = iz < (n2 - 1) / 2 + 1 ? iz : iz - n2;
  constexpr double twopi = 2 * 3.141592653589793;
  const double kx = twopi * (static_cast<double>(fx) * v0);
  const double ky = twopi * (static_cast<double>(fy) * v1);
  const double kz = twopi * (static_cast<double>(fz) * v2);
  const double ax = r0 * kx, ay = r1 * ky, az = r2 * kz;
  const double e = -0.5 * (ax * ax) + -0.5 * (ay * ay) + -0.5 * (az * az);
  const double g = exp(e) * norm;
  Cplx<R> c = spec[i];
  c.re = static_cast<R>(static_cast<double>(c.re) * g);
  c.im = static_cast<R>(static_cast<double>(c.im) * g);
  spec[i] = c;
}

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

static int fft_status(hipfftResult r, const char* what) {
  if (r != HIPFFT_SUCCESS) {
    set_error("%s failed with hipfftResult %d", what, static_cast<int>(r));
    return PRH_ERR_FFT;
  }
  return PRH_OK;
}

template <typename R>
static int exec_r2c(prh_workspace* ws, R* mesh);
template <>
int exec_r2c<double>(prh_workspace* ws, double* mesh) {
  return fft_status(hipfftExecD2Z(ws->fft.r2c, mesh,
                                  static_cast<hipfftDoubleComplex*>(ws->fft.spectrum)),
                    "hipfftExecD2Z");
}
template <>
int exec_r2c<float>(prh_workspace* ws, float* mesh) {
  return fft_status(
      hipfftExecR2C(ws->fft.r2c, mesh, static_cast<hipfftComplex*>(ws->fft.spectrum)),
      "hipfftExecR2C");
}
template <typename R>
static int exec_c2r(prh_workspace* ws, R* mesh);
template <>
int exec_c2r<double>(prh_workspace* ws, double* mesh) {
  return fft_status(
      hipfftExecZ2D(ws->fft.c2r, static_cast<hipfftDoubleComplex*>(ws->fft.spectrum), mesh),
      "hipfftExecZ2D");
}
template <>
int exec_c2r<float>(prh_workspace* ws, float* mesh) {
  return fft_status(hipfftExecC2R(ws->fft.c2r, static_cast<hipfftComplex*>(ws->fft.spectrum), mesh),
                    "hipfftExecC2R");
}

template <typename R>
static int do_smooth(prh_workspace* ws, R* mesh, const size_t* nmesh, const R* boxsize,
                     const double* radius, void* stream) {
  if (!ws || !mesh || !nmesh || !boxsize || !radius) return PRH_ERR_ARG;
  if (!(ws->flags & PRH_WS_FFT) || !ws->fft.have) {
    set_error("workspace has no FFT plans");
    return PRH_ERR_STATE;
  }
  if (ws->dtype != Tag<R>::code) {
    set_error("workspace precision does not match the mesh");
    return PRH_ERR_ARG;
  }
  for (int d = 0; d < 3; d++)
    if (ws->n[d] != nmesh[d]) {
      set_error("workspace mesh shape does not match nmesh");
      return PRH_ERR_CAPACITY;
    }
  hipStream_t s = as_stream(stream);
  PRH_TRY(fft_status(hipfftSetStream(ws->fft.r2c, s), "hipfftSetStream"));
  PRH_TRY(fft_status(hipfftSetStream(ws->fft.c2r, s), "hipfftSetStream"));
  PRH_TRY(exec_r2c<R>(ws, mesh));
  double val[3], rad[3];
  for (int d = 0; d < 3; d++) {
    // d_i = boxsize / nmesh (promoted to float64), val = 1 / (n d) as in numpy.fft.fftfreq
    const double cell = static_cast<double>(boxsize[d]) / static_cast<double>(static_cast<int>(nmesh[d]));
    val[d] = 1.0 / (static_cast<double>(nmesh[d]) * cell);
    rad[d] = static_cast<double>(static_cast<R>(radius[d]));  // radius_ is in the mesh dtype
  }
  const int n0 = static_cast<int>(nmesh[0]), n1 = static_cast<int>(nmesh[1]),
            n2 = static_cast<int>(nmesh[2]);
  const int nh = n2 / 2 + 1;
  const size_t total = static_cast<size_t>(n0) * n1 * nh;
  const double norm = 1.0 / (static_cast<double>(n0) * n1 * n2);
  k_gauss_multiply<R><<<ceil_div(total, kDenThreads), kDenThreads, 0, s>>>(
      static_cast<Cplx<R>*>(ws->fft.spectrum), n0, n1, n2, nh, val[0], val[1], val[2], rad[0],
      rad[1], rad[2], norm);
  PRH_LAUNCHED();
  PRH_TRY(exec_c2r<R>(ws, mesh));
  return PRH_OK;
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

int prh_smooth_gaussian_f32(prh_workspace* ws, float* mesh, const size_t* nmesh,
                            const float* boxsize, const double* radius, void* stream) {
  return do_smooth<float>(ws, mesh, nmesh, boxsize, radius, stream);
}
int prh_smooth_gaussian_f64(prh_workspace* ws, double* mesh, const size_t* nmesh,
                            const double* boxsize, const double* radius, void* stream) {
  return do_smooth<double>(ws, mesh, nmesh, boxsize, radius, stream);
}

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

}  // extern "C"
