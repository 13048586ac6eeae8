// Multigrid solve of the local/global line-of-sight RSD Poisson equation on
// the device: damped Jacobi, residual, full-weighting restriction, trilinear
// prolongation (fused with the V-cycle correction), V-cycle and full
// multigrid, plus the mixed-precision (FP32 V-cycles inside an FP64
// defect-correction loop) driver.
//
// Iteration semantics follow src/multigrid.c:270-328 exactly: same sweep
// counts, zero coarse-grid guesses, the FMG prolongation overwriting the
// iterate, recursion while every dimension is even and > 4.  Jacobi is done
// as one pass per sweep into a second iterate (ping-pong) instead of the
// reference's two passes through a temporary; the arithmetic per cell is the
// same, so results are bit-identical.
#include <utility>

#include "prh_reduce.hpp"
#include "prh_stencil.hpp"

namespace prh {

constexpr int kThreads = 256;

__device__ __forceinline__ void unflatten(size_t i, int n1, int n2, int& ix, int& iy, int& iz) {
  iz = static_cast<int>(i % n2);
  const size_t t = i / n2;
  iy = static_cast<int>(t % n1);
  ix = static_cast<int>(t / n1);
}

template <typename R, bool ZERO_IN>
__global__ __launch_bounds__(kThreads) void k_jacobi(const R* __restrict__ v,
                                                     const R* __restrict__ f, R* __restrict__ out,
                                                     LevelGeom<R> G, R omega, R keep) {
  const size_t total = static_cast<size_t>(G.n[0]) * G.n[1] * G.n[2];
  const size_t i = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x;
  if (i >= total) return;
  int ix, iy, iz;
  unflatten(i, G.n[1], G.n[2], ix, iy, iz);
  const Coef<R> k = cell_coef(G, ix, iy, iz);
  const Hood<R> h = ZERO_IN ? zero_hood<R>() : gather_hood(v, G.n[0], G.n[1], G.n[2], ix, iy, iz);
  out[i] = jacobi_update(k, h, f[i], omega, keep, G.local);
}

template <typename R, typename O>
__global__ __launch_bounds__(kThreads) void k_residual(const R* __restrict__ v,
                                                       const R* __restrict__ f,
                                                       O* __restrict__ r, LevelGeom<R> G) {
  const size_t total = static_cast<size_t>(G.n[0]) * G.n[1] * G.n[2];
  const size_t i = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x;
  if (i >= total) return;
  int ix, iy, iz;
  unflatten(i, G.n[1], G.n[2], ix, iy, iz);
  const Coef<R> k = cell_coef(G, ix, iy, iz);
  const Hood<R> h = gather_hood(v, G.n[0], G.n[1], G.n[2], ix, iy, iz);
  r[i] = static_cast<O>(residual_value(k, h, f[i], G.local));
}

// Residual with block partial sums of r^2 (deterministic) -- optionally also
// stored (cast to O) for the mixed-precision correction solve.
template <typename R, typename O, bool STORE>
__global__ __launch_bounds__(kReduceThreads) void k_residual_norm(const R* __restrict__ v,
                                                                  const R* __restrict__ f,
                                                                  O* __restrict__ r,
                                                                  LevelGeom<R> G,
                                                                  double* __restrict__ partial) {
  const size_t total = static_cast<size_t>(G.n[0]) * G.n[1] * G.n[2];
  block_reduce_range<1>(
      total,
      [&](size_t i, double* out) {
        int ix, iy, iz;
        unflatten(i, G.n[1], G.n[2], ix, iy, iz);
        const Coef<R> k = cell_coef(G, ix, iy, iz);
        const Hood<R> h = gather_hood(v, G.n[0], G.n[1], G.n[2], ix, iy, iz);
        const R res = residual_value(k, h, f[i], G.local);
        if (STORE) r[i] = static_cast<O>(res);
        out[0] = static_cast<double>(res) * static_cast<double>(res);
      },
      partial, 0);
}

// Full weighting (src/multigrid.c:236-262): centre x8, faces x4, edges x2,
// corners x1, summed in the reference's order, divided by 64.
template <typename R>
__global__ __launch_bounds__(kThreads) void k_restrict(const R* __restrict__ u,
                                                       R* __restrict__ c, int f0, int f1,
                                                       int f2) {
  const int c0 = f0 / 2, c1 = f1 / 2, c2 = f2 / 2;
  const size_t total = static_cast<size_t>(c0) * c1 * c2;
  const size_t j = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x;
  if (j >= total) return;
  int jx, jy, jz;
  unflatten(j, c1, c2, jx, jy, jz);
  const size_t s0 = static_cast<size_t>(f1) * f2, s1 = static_cast<size_t>(f2);
  const int fx = 2 * jx, fy = 2 * jy, fz = 2 * jz;
  const size_t x0 = fx * s0, xp = (fx + 1 == f0 ? 0 : fx + 1) * s0,
               xm = (fx == 0 ? f0 - 1 : fx - 1) * s0;
  const size_t y0 = fy * s1, yp = (fy + 1 == f1 ? 0 : fy + 1) * s1,
               ym = (fy == 0 ? f1 - 1 : fy - 1) * s1;
  const size_t z0 = fz, zp = (fz + 1 == f2 ? 0 : fz + 1), zm = (fz == 0 ? f2 - 1 : fz - 1);
  const R faces = u[xp + y0 + z0] + u[xm + y0 + z0] + u[x0 + yp + z0] + u[x0 + ym + z0] +
                  u[x0 + y0 + zp] + u[x0 + y0 + zm];
  const R edges = u[xp + yp + z0] + u[xm + yp + z0] + u[xp + ym + z0] + u[xm + ym + z0] +
                  u[xp + y0 + zp] + u[xm + y0 + zp] + u[xp + y0 + zm] + u[xm + y0 + zm] +
                  u[x0 + yp + zp] + u[x0 + ym + zp] + u[x0 + yp + zm] + u[x0 + ym + zm];
  R s = R(8) * u[x0 + y0 + z0] + R(4) * faces + R(2) * edges;
  s += u[xp + yp + zp];
  s += u[xm + yp + zp];
  s += u[xp + ym + zp];
  s += u[xm + ym + zp];
  s += u[xp + yp + zm];
  s += u[xm + yp + zm];
  s += u[xp + ym + zm];
  s += u[xm + ym + zm];
  // reference divides by the double 64.0; an exact power-of-two scaling
  c[j] = s * R(0.015625);
}

// Trilinear prolongation from the coarse mesh (c0, c1, c2) to the fine mesh,
// one thread per fine cell (src/multigrid.c:177-205), either overwriting the
// fine value or adding to it (the V-cycle correction, src/multigrid.c:296).
template <typename R, bool ADD>
__global__ __launch_bounds__(kThreads) void k_prolong(const R* __restrict__ u,
                                                      R* __restrict__ fine, int c0, int c1,
                                                      int c2) {
  const int F1 = 2 * c1, F2 = 2 * c2;
  const size_t total = static_cast<size_t>(2 * c0) * F1 * F2;
  const size_t i = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x;
  if (i >= total) return;
  int X, Y, Z;
  unflatten(i, F1, F2, X, Y, Z);
  const int jx = X >> 1, jy = Y >> 1, jz = Z >> 1;
  const size_t s0 = static_cast<size_t>(c1) * c2, s1 = static_cast<size_t>(c2);
  const size_t x0 = jx * s0, xp = (jx + 1 == c0 ? 0 : jx + 1) * s0;
  const size_t y0 = jy * s1, yp = (jy + 1 == c1 ? 0 : jy + 1) * s1;
  const size_t z0 = jz, zp = (jz + 1 == c2 ? 0 : jz + 1);
  const R a = u[x0 + y0 + z0];
  R val;
  switch (((X & 1) << 2) | ((Y & 1) << 1) | (Z & 1)) {
    case 0: val = a; break;
    case 4: val = (a + u[xp + y0 + z0]) / R(2); break;
    case 2: val = (a + u[x0 + yp + z0]) / R(2); break;
    case 1: val = (a + u[x0 + y0 + zp]) / R(2); break;
    case 6: val = (a + u[xp + y0 + z0] + u[x0 + yp + z0] + u[xp + yp + z0]) / R(4); break;
    case 3: val = (a + u[x0 + yp + z0] + u[x0 + y0 + zp] + u[x0 + yp + zp]) / R(4); break;
    case 5: val = (a + u[xp + y0 + z0] + u[x0 + y0 + zp] + u[xp + y0 + zp]) / R(4); break;
    default:
      val = (a + u[xp + y0 + z0] + u[x0 + yp + z0] + u[x0 + y0 + zp] + u[xp + yp + z0] +
             u[xp + y0 + zp] + u[x0 + yp + zp] + u[xp + yp + zp]) /
            R(8);
      break;
  }
  if (ADD)
    fine[i] += val;
  else
    fine[i] = val;
}

template <typename R>
__global__ __launch_bounds__(kThreads) void k_fill_zero(R* __restrict__ a, size_t n) {
  const size_t i = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x;
  if (i < n) a[i] = R(0);
}

__global__ __launch_bounds__(kThreads) void k_cast_f64_f32(const double* __restrict__ a,
                                                           float* __restrict__ b, size_t n) {
  const size_t i = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x;
  if (i < n) b[i] = static_cast<float>(a[i]);
}

__global__ __launch_bounds__(kThreads) void k_cast_f32_f64(const float* __restrict__ a,
                                                           double* __restrict__ b, size_t n) {
  const size_t i = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x;
  if (i < n) b[i] = static_cast<double>(a[i]);
}

__global__ __launch_bounds__(kThreads) void k_add_f32_f64(double* __restrict__ v,
                                                          const float* __restrict__ e, size_t n) {
  const size_t i = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x;
  if (i < n) v[i] += static_cast<double>(e[i]);
}

// ---------------------------------------------------------------------------
// Host-side drivers
// ---------------------------------------------------------------------------

static size_t ncells(const size_t* n) { return n[0] * n[1] * n[2]; }

template <typename R>
static int launch_jacobi(const R* v, const R* f, R* out, const LevelGeom<R>& G, R omega, R keep,
                         bool zero_in, hipStream_t s, int timer = -1) {
  const size_t total = static_cast<size_t>(G.n[0]) * G.n[1] * G.n[2];
  const dim3 grid(ceil_div(total, kThreads));
  timer_before(timer, s);
  if (zero_in)
    k_jacobi<R, true><<<grid, kThreads, 0, s>>>(v, f, out, G, omega, keep);
  else
    k_jacobi<R, false><<<grid, kThreads, 0, s>>>(v, f, out, G, omega, keep);
  // algorithmic traffic: read v (unless zero) and f, write the new iterate
  timer_after(timer, s, static_cast<double>(total) * sizeof(R) * (zero_in ? 2 : 3));
  PRH_LAUNCHED();
  return PRH_OK;
}

template <typename R, typename O>
static int launch_residual(const R* v, const R* f, O* r, const LevelGeom<R>& G, hipStream_t s) {
  const size_t total = static_cast<size_t>(G.n[0]) * G.n[1] * G.n[2];
  k_residual<R, O><<<ceil_div(total, kThreads), kThreads, 0, s>>>(v, f, r, G);
  PRH_LAUNCHED();
  return PRH_OK;
}

template <typename R>
static int launch_restrict(const R* fine, R* coarse, const size_t* nf, hipStream_t s) {
  const size_t total = (nf[0] / 2) * (nf[1] / 2) * (nf[2] / 2);
  k_restrict<R><<<ceil_div(total, kThreads), kThreads, 0, s>>>(
      fine, coarse, static_cast<int>(nf[0]), static_cast<int>(nf[1]), static_cast<int>(nf[2]));
  PRH_LAUNCHED();
  return PRH_OK;
}

template <typename R>
static int launch_prolong(const R* coarse, R* fine, const size_t* nc, bool add, hipStream_t s) {
  const size_t total = 8 * ncells(nc);
  const int c0 = static_cast<int>(nc[0]), c1 = static_cast<int>(nc[1]),
            c2 = static_cast<int>(nc[2]);
  if (add)
    k_prolong<R, true><<<ceil_div(total, kThreads), kThreads, 0, s>>>(coarse, fine, c0, c1, c2);
  else
    k_prolong<R, false><<<ceil_div(total, kThreads), kThreads, 0, s>>>(coarse, fine, c0, c1, c2);
  PRH_LAUNCHED();
  return PRH_OK;
}

template <typename R>
static int launch_fill_zero(R* a, size_t n, hipStream_t s) {
  k_fill_zero<R><<<ceil_div(n, kThreads), kThreads, 0, s>>>(a, n);
  PRH_LAUNCHED();
  return PRH_OK;
}

// State of one solve over a hierarchy.
template <typename R>
struct Solver {
  Hierarchy* H;
  LevelGeom<R> G[kMaxLevels];
  R omega, keep;
  int niter;
  hipStream_t s;

  R* v(int l) { return static_cast<R*>(H->v[l]); }
  R* f(int l) { return static_cast<R*>(H->f[l]); }
  R* tmp(int l) { return static_cast<R*>(H->tmp[l]); }
  size_t size(int l) const { return ncells(H->n[l]); }

  void setup(Hierarchy* h, const R* boxsize, const R* boxcenter, R beta, R damping, int nit,
             const R* los, hipStream_t stream) {
    H = h;
    for (int l = 0; l < H->nlev; l++) G[l] = LevelGeom<R>::make(H->n[l], boxsize, boxcenter, beta, los);
    omega = damping;
    keep = R(1) - damping;
    niter = nit;
    s = stream;
  }

  // niter damped-Jacobi sweeps; the iterate moves between cur and alt.
  int smooth(int l, R*& cur, R*& alt, const R* src, bool zero_start) {
    if (niter <= 0) {
      if (zero_start) PRH_TRY(launch_fill_zero(cur, size(l), s));
      return PRH_OK;
    }
    for (int it = 0; it < niter; it++) {
      PRH_TRY(launch_jacobi(cur, src, alt, G[l], omega, keep, zero_start && it == 0, s,
                            l == 0 ? PRH_TIMER_JACOBI_FINEST : -1));
      std::swap(cur, alt);
    }
    return PRH_OK;
  }

  // src/multigrid.c:270-300
  int vcycle(int l, R*& cur, R*& alt, const R* src, bool zero_start) {
    PRH_TRY(smooth(l, cur, alt, src, zero_start));
    if (l + 1 < H->nlev) {
      PRH_TRY((launch_residual<R, R>(cur, src, alt, G[l], s)));
      PRH_TRY(launch_restrict(alt, f(l + 1), H->n[l], s));
      R* cc = v(l + 1);
      R* ca = tmp(l + 1);
      PRH_TRY(vcycle(l + 1, cc, ca, f(l + 1), true));
      PRH_TRY(launch_prolong(cc, cur, H->n[l + 1], true, s));
    }
    return smooth(l, cur, alt, src, false);
  }

  // src/multigrid.c:303-328.  v0 is the caller's iterate (used as the start
  // only when no coarser level exists), f0 the caller's source.
  int fmg(const R* f0, R* v0, int ncycle) {
    const int nl = H->nlev;
    const R* src[kMaxLevels];
    src[0] = f0;
    for (int l = 1; l < nl; l++) {
      PRH_TRY(launch_restrict(src[l - 1], f(l), H->n[l - 1], s));
      src[l] = f(l);
    }
    R* cur;
    R* alt;
    if (nl == 1) {
      cur = v0;
      alt = tmp(0);
    } else {
      cur = v(nl - 1);
      alt = tmp(nl - 1);
      PRH_TRY(launch_fill_zero(cur, size(nl - 1), s));
    }
    for (int c = 0; c < ncycle; c++) PRH_TRY(vcycle(nl - 1, cur, alt, src[nl - 1], false));
    for (int l = nl - 2; l >= 0; l--) {
      R* dst = (l == 0) ? v0 : v(l);
      PRH_TRY(launch_prolong(cur, dst, H->n[l + 1], false, s));
      cur = dst;
      alt = tmp(l);
      for (int c = 0; c < ncycle; c++) PRH_TRY(vcycle(l, cur, alt, src[l], false));
    }
    if (cur != v0) PRH_HIP(hipMemcpyAsync(v0, cur, size(0) * sizeof(R), hipMemcpyDeviceToDevice, s));
    return PRH_OK;
  }
};

static int check_ws(const prh_workspace* ws, const size_t* nmesh, int dtype, unsigned need) {
  if (!ws || !nmesh) {
    set_error("workspace and nmesh are required");
    return PRH_ERR_ARG;
  }
  if ((ws->flags & need) != need) {
    set_error("workspace was not built with the required flags 0x%x", need);
    return PRH_ERR_STATE;
  }
  if (ws->dtype != dtype) {
    set_error("workspace precision %d does not match the call (%d)", ws->dtype, dtype);
    return PRH_ERR_ARG;
  }
  for (int d = 0; d < 3; d++)
    if (ws->n[d] != nmesh[d]) {
      set_error("workspace mesh (%zu, %zu, %zu) does not match nmesh (%zu, %zu, %zu)", ws->n[0],
                ws->n[1], ws->n[2], nmesh[0], nmesh[1], nmesh[2]);
      return PRH_ERR_CAPACITY;
    }
  return PRH_OK;
}

template <typename R>
static int do_jacobi(prh_workspace* ws, R* v, const R* f, const size_t* nmesh, const R* boxsize,
                     const R* boxcenter, R beta, R damping, int niter, const R* los, void* stream) {
  PRH_TRY(check_ws(ws, nmesh, Tag<R>::code, PRH_WS_MULTIGRID));
  if (!v || !f || !boxsize || !boxcenter) return PRH_ERR_ARG;
  Solver<R> S;
  S.setup(&ws->mg, boxsize, boxcenter, beta, damping, niter, los, as_stream(stream));
  R* cur = v;
  R* alt = S.tmp(0);
  PRH_TRY(S.smooth(0, cur, alt, f, false));
  if (cur != v)
    PRH_HIP(hipMemcpyAsync(v, cur, S.size(0) * sizeof(R), hipMemcpyDeviceToDevice, S.s));
  return PRH_OK;
}

template <typename R>
static int do_vcycle(prh_workspace* ws, R* v, const R* f, const size_t* nmesh, const R* boxsize,
                     const R* boxcenter, R beta, R damping, int niter, const R* los, void* stream) {
  PRH_TRY(check_ws(ws, nmesh, Tag<R>::code, PRH_WS_MULTIGRID));
  if (!v || !f || !boxsize || !boxcenter) return PRH_ERR_ARG;
  Solver<R> S;
  S.setup(&ws->mg, boxsize, boxcenter, beta, damping, niter, los, as_stream(stream));
  R* cur = v;
  R* alt = S.tmp(0);
  PRH_TRY(S.vcycle(0, cur, alt, f, false));
  if (cur != v)
    PRH_HIP(hipMemcpyAsync(v, cur, S.size(0) * sizeof(R), hipMemcpyDeviceToDevice, S.s));
  return PRH_OK;
}

template <typename R>
static int do_fmg(prh_workspace* ws, const R* f, R* v, const size_t* nmesh, const R* boxsize,
                  const R* boxcenter, R beta, R damping, int niter, int ncycle, const R* los,
                  void* stream) {
  PRH_TRY(check_ws(ws, nmesh, Tag<R>::code, PRH_WS_MULTIGRID));
  if (!v || !f || !boxsize || !boxcenter) return PRH_ERR_ARG;
  Solver<R> S;
  S.setup(&ws->mg, boxsize, boxcenter, beta, damping, niter, los, as_stream(stream));
  return S.fmg(f, v, ncycle);
}

template <typename R>
static int do_residual(const R* v, const R* f, R* r, const size_t* nmesh, const R* boxsize,
                       const R* boxcenter, R beta, const R* los, void* stream) {
  if (!v || !f || !r || !nmesh || !boxsize || !boxcenter) return PRH_ERR_ARG;
  const LevelGeom<R> G = LevelGeom<R>::make(nmesh, boxsize, boxcenter, beta, los);
  return launch_residual<R, R>(v, f, r, G, as_stream(stream));
}

// FP64 residual RMS over a mesh, deterministic; optionally stores r as float.
static int residual_rms(prh_workspace* ws, const double* v, const double* f, float* r32,
                        const LevelGeom<double>& G, double* rms, hipStream_t s) {
  const size_t total = static_cast<size_t>(G.n[0]) * G.n[1] * G.n[2];
  if (r32)
    k_residual_norm<double, float, true><<<kReduceBlocks, kReduceThreads, 0, s>>>(v, f, r32, G, ws->red);
  else
    k_residual_norm<double, float, false><<<kReduceBlocks, kReduceThreads, 0, s>>>(v, f, nullptr, G, ws->red);
  PRH_LAUNCHED();
  k_reduce_final<1><<<1, 1024, 0, s>>>(ws->red, kReduceBlocks, 0, ws->red + 2 * kReduceBlocks);
  PRH_LAUNCHED();
  PRH_HIP(hipMemcpyAsync(ws->red_host, ws->red + 2 * kReduceBlocks, sizeof(double),
                         hipMemcpyDeviceToHost, s));
  PRH_HIP(hipStreamSynchronize(s));
  *rms = sqrt(ws->red_host[0] / static_cast<double>(total));
  return PRH_OK;
}

static int do_fmg_mixed(prh_workspace* ws, const double* f, double* v, const size_t* nmesh,
                        const double* boxsize, const double* boxcenter, double beta, double damping,
                        int niter, int ncycle, const double* los, int nrefine, int vpr,
                        double* norms, void* stream) {
  PRH_TRY(check_ws(ws, nmesh, PRH_F64, PRH_WS_MULTIGRID | PRH_WS_MIXED));
  if (!v || !f || !boxsize || !boxcenter || nrefine < 0 || vpr < 0) return PRH_ERR_ARG;
  hipStream_t s = as_stream(stream);
  const size_t total = ncells(nmesh);
  const unsigned grid = ceil_div(total, kThreads);
  // FP32 operator: the geometry cast to float, as the reference's FP32 build sees it
  float bs32[3], bc32[3], los32[3];
  for (int d = 0; d < 3; d++) {
    bs32[d] = static_cast<float>(boxsize[d]);
    bc32[d] = static_cast<float>(boxcenter[d]);
    if (los) los32[d] = static_cast<float>(los[d]);
  }
  Solver<float> S;
  S.setup(&ws->mixed, bs32, bc32, static_cast<float>(beta), static_cast<float>(damping), niter,
          los ? los32 : nullptr, s);
  const LevelGeom<double> G64 = LevelGeom<double>::make(nmesh, boxsize, boxcenter, beta, los);
  float* f32 = S.f(0);
  float* e32 = S.v(0);
  // 1. FP32 full multigrid on the FP32 source
  k_cast_f64_f32<<<grid, kThreads, 0, s>>>(f, f32, total);
  PRH_LAUNCHED();
  PRH_TRY(launch_fill_zero(e32, total, s));
  PRH_TRY(S.fmg(f32, e32, ncycle));
  k_cast_f32_f64<<<grid, kThreads, 0, s>>>(e32, v, total);
  PRH_LAUNCHED();
  // 2. defect correction: r = f - A v in FP64 (stored as FP32), e = V-cycles(r), v += e
  for (int it = 0; it < nrefine; it++) {
    if (norms) {
      PRH_TRY(residual_rms(ws, v, f, f32, G64, &norms[it], s));
    } else {
      PRH_TRY((launch_residual<double, float>(v, f, f32, G64, s)));
    }
    float* cur = e32;
    float* alt = S.tmp(0);
    for (int c = 0; c < vpr; c++) PRH_TRY(S.vcycle(0, cur, alt, f32, c == 0));
    if (vpr == 0) PRH_TRY(launch_fill_zero(cur, total, s));
    k_add_f32_f64<<<grid, kThreads, 0, s>>>(v, cur, total);
    PRH_LAUNCHED();
  }
  if (norms) PRH_TRY(residual_rms(ws, v, f, nullptr, G64, &norms[nrefine], s));
  return PRH_OK;
}

}  // namespace prh

using namespace prh;

extern "C" {

int prh_jacobi_f32(prh_workspace* ws, float* v, const float* f, const size_t* nmesh,
                   const float* boxsize, const float* boxcenter, float beta, float damping_factor,
                   int niterations, const float* los, void* stream) {
  return do_jacobi<float>(ws, v, f, nmesh, boxsize, boxcenter, beta, damping_factor, niterations,
                          los, stream);
}
int prh_jacobi_f64(prh_workspace* ws, double* v, const double* f, const size_t* nmesh,
                   const double* boxsize, const double* boxcenter, double beta,
                   double damping_factor, int niterations, const double* los, void* stream) {
  return do_jacobi<double>(ws, v, f, nmesh, boxsize, boxcenter, beta, damping_factor, niterations,
                           los, stream);
}

int prh_residual_f32(const float* v, const float* f, float* r, const size_t* nmesh,
                     const float* boxsize, const float* boxcenter, float beta, const float* los,
                     void* stream) {
  return do_residual<float>(v, f, r, nmesh, boxsize, boxcenter, beta, los, stream);
}
int prh_residual_f64(const double* v, const double* f, double* r, const size_t* nmesh,
                     const double* boxsize, const double* boxcenter, double beta,
                     const double* los, void* stream) {
  return do_residual<double>(v, f, r, nmesh, boxsize, boxcenter, beta, los, stream);
}

int prh_prolong_f32(const float* v2h, float* v1h, const size_t* nmesh, void* stream) {
  if (!v2h || !v1h || !nmesh) return PRH_ERR_ARG;
  return launch_prolong<float>(v2h, v1h, nmesh, false, as_stream(stream));
}
int prh_prolong_f64(const double* v2h, double* v1h, const size_t* nmesh, void* stream) {
  if (!v2h || !v1h || !nmesh) return PRH_ERR_ARG;
  return launch_prolong<double>(v2h, v1h, nmesh, false, as_stream(stream));
}

int prh_reduce_f32(const float* v1h, float* v2h, const size_t* nmesh, void* stream) {
  if (!v1h || !v2h || !nmesh) return PRH_ERR_ARG;
  return launch_restrict<float>(v1h, v2h, nmesh, as_stream(stream));
}
int prh_reduce_f64(const double* v1h, double* v2h, const size_t* nmesh, void* stream) {
  if (!v1h || !v2h || !nmesh) return PRH_ERR_ARG;
  return launch_restrict<double>(v1h, v2h, nmesh, as_stream(stream));
}

int prh_vcycle_f32(prh_workspace* ws, float* v, const float* f, const size_t* nmesh,
                   const float* boxsize, const float* boxcenter, float beta, float damping_factor,
                   int niterations, const float* los, void* stream) {
  return do_vcycle<float>(ws, v, f, nmesh, boxsize, boxcenter, beta, damping_factor, niterations,
                          los, stream);
}
int prh_vcycle_f64(prh_workspace* ws, double* v, const double* f, const size_t* nmesh,
                   const double* boxsize, const double* boxcenter, double beta,
                   double damping_factor, int niterations, const double* los, void* stream) {
  return do_vcycle<double>(ws, v, f, nmesh, boxsize, boxcenter, beta, damping_factor, niterations,
                           los, stream);
}

int prh_fmg_f32(prh_workspace* ws, const float* f1h, float* v1h, const size_t* nmesh,
                const float* boxsize, const float* boxcenter, float beta,
                float jacobi_damping_factor, int jacobi_niterations, int vcycle_niterations,
                const float* los, void* stream) {
  return do_fmg<float>(ws, f1h, v1h, nmesh, boxsize, boxcenter, beta, jacobi_damping_factor,
                       jacobi_niterations, vcycle_niterations, los, stream);
}
int prh_fmg_f64(prh_workspace* ws, const double* f1h, double* v1h, const size_t* nmesh,
                const double* boxsize, const double* boxcenter, double beta,
                double jacobi_damping_factor, int jacobi_niterations, int vcycle_niterations,
                const double* los, void* stream) {
  return do_fmg<double>(ws, f1h, v1h, nmesh, boxsize, boxcenter, beta, jacobi_damping_factor,
                        jacobi_niterations, vcycle_niterations, los, stream);
}

int prh_fmg_mixed(prh_workspace* ws, const double* f1h, double* v1h, const size_t* nmesh,
                  const double* boxsize, const double* boxcenter, double beta,
                  double jacobi_damping_factor, int jacobi_niterations, int vcycle_niterations,
                  const double* los, int nrefine, int vcycles_per_refine, double* residual_norms,
                  void* stream) {
  return do_fmg_mixed(ws, f1h, v1h, nmesh, boxsize, boxcenter, beta, jacobi_damping_factor,
                      jacobi_niterations, vcycle_niterations, los, nrefine, vcycles_per_refine,
                      residual_norms, stream);
}

int prh_residual_norm_f64(prh_workspace* ws, const double* v, const double* f, const size_t* nmesh,
                          const double* boxsize, const double* boxcenter, double beta,
                          const double* los, double* rms, void* stream) {
  if (!ws || !v || !f || !nmesh || !boxsize || !boxcenter || !rms) return PRH_ERR_ARG;
  const LevelGeom<double> G = LevelGeom<double>::make(nmesh, boxsize, boxcenter, beta, los);
  return residual_rms(ws, v, f, nullptr, G, rms, as_stream(stream));
}

}  // extern "C"
