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
#include <atomic>
#include <mutex>
#include <type_traits>
#include <utility>

#include "prh_reduce.hpp"
#include "prh_stencil.hpp"

namespace prh {

constexpr int kThreads = 256;

// (ix, iy, iz) of the flat index i of an (n0, n1, n2) mesh.  A 32-bit index
// (the single-workgroup levels) takes shifts for power-of-two extents and
// 32-bit division otherwise.
template <typename I>
__device__ __forceinline__ void unflatten(I i, int n1, int n2, int& ix, int& iy, int& iz) {
  if constexpr (sizeof(I) == 4) {
    const unsigned u = static_cast<unsigned>(i), m1 = n1, m2 = n2;
    if ((m1 & (m1 - 1)) == 0 && (m2 & (m2 - 1)) == 0) {
      const int l1 = __builtin_ctz(m1), l2 = __builtin_ctz(m2);
      iz = static_cast<int>(u & (m2 - 1));
      iy = static_cast<int>((u >> l2) & (m1 - 1));
      ix = static_cast<int>(u >> (l1 + l2));
    } else {
      const unsigned t = u / m2;
      iz = static_cast<int>(u - t * m2);
      iy = static_cast<int>(t % m1);
      ix = static_cast<int>(t / m1);
    }
  } else {
    iz = static_cast<int>(i % n2);
    const size_t t = i / n2;
    iy = static_cast<int>(t % n1);
    ix = static_cast<int>(t / n1);
  }
}

// Generic kernels: one thread per cell on a 3D grid (x = blockIdx.z), for any
// shape; used for the coarse levels and for shapes the tiled kernel does not
// cover.
template <typename R, bool ZERO_IN>
__global__ __launch_bounds__(kThreads) void k_jacobi(const R* __restrict__ v,
                                                     const R* __restrict__ f, R* __restrict__ out,
                                                     LevelGeom<R> G, R omega, R keep) {
  const int iz = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y * blockDim.y + threadIdx.y;
  const int ix = blockIdx.z;
  if (iz >= G.n[2] || iy >= G.n[1]) return;
  const size_t i = (static_cast<size_t>(ix) * G.n[1] + iy) * G.n[2] + iz;
  const Coef<R> k = cell_coef(G, ix, iy, iz);
  const Hood<R> h = ZERO_IN ? zero_hood<R>() : gather_hood(v, G.n[0], G.n[1], G.n[2], ix, iy, iz);
  out[i] = jacobi_update(k, h, f[i], omega, keep, G.local);
}

// A source read with its mean removed: each value f becomes f - sum[0] / n
// (sum on the device, n the number of cells), the same FP64 subtraction as a
// centred copy of f would hold.  sum == nullptr: the source as it is.
struct Centre {
  const double* sum = nullptr;
  double n = 1.0;
};
template <typename R>
__device__ __forceinline__ R centre_mean(const Centre& C) {
  return C.sum ? static_cast<R>(*C.sum / C.n) : R(0);
}

template <typename R, typename O>
__global__ __launch_bounds__(kThreads) void k_residual(const R* __restrict__ v,
                                                       const R* __restrict__ f,
                                                       O* __restrict__ r, LevelGeom<R> G,
                                                       Centre C) {
  const int iz = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y * blockDim.y + threadIdx.y;
  const int ix = blockIdx.z;
  if (iz >= G.n[2] || iy >= G.n[1]) return;
  const size_t i = (static_cast<size_t>(ix) * G.n[1] + iy) * G.n[2] + iz;
  const Coef<R> k = cell_coef(G, ix, iy, iz);
  const Hood<R> h = gather_hood(v, G.n[0], G.n[1], G.n[2], ix, iy, iz);
  const R fv = C.sum ? f[i] - centre_mean<R>(C) : f[i];
  r[i] = static_cast<O>(residual_value(k, h, fv, G.local));
}

// 2.5D-blocked kernel for large levels (n2 % 64 == 0, n1 % TY == 0).  A block
// owns a TY x 64 tile of (y, z) and walks a run of x planes.  Planes x-1, x,
// x+1 of the input (with a one-cell halo in y and z, periodic) sit in a ring
// of 4 LDS slots; plane x+2 is fetched into registers while plane x is
// computed, then stored into the free slot, so each input value is read from
// HBM about once and every stencil read hits LDS.  256 threads: 64 along z,
// 4 along y, TY/4 rows each.  The arithmetic per cell is the generic
// kernels' (same bits).
// kOpJacobiAcc: a Jacobi sweep whose result is added (as double) into TileCoef::acc
// instead of being stored (the last sweep of a mixed-precision correction)
// kOpJacobiAfterZero: the first two sweeps from a zero iterate in one launch
// (FP32 rows of 512 cells): the first sweep of a zero iterate is pointwise in
// the source, so the second stages the first's value of each cell, computed
// from the source as the zero-start sweep computes it, instead of reading it
enum StencilOp { kOpJacobi = 0, kOpResidual = 1, kOpJacobiZero = 2, kOpJacobiAcc = 3, kOpJacobiAfterZero = 4 };
constexpr int kTileZ = 64;

// Coefficients of a level passed to the blocked kernels: the geometry for a
// local line of sight, the precomputed constants for a global one.
template <typename R>
struct TileCoef {
  LevelGeom<R> G;
  GlobalCoef<R> K;
  R dinv;  // 1 / K.d (fast variant only)
  // kOpJacobiAcc: each result is added (as double) to acc at its cell instead of
  // being stored (the last sweep of a mixed-precision correction, whose FP32
  // result is only needed in the FP64 update v += e)
  double* acc = nullptr;
  // kOpResidual: the source's mean to remove (the mixed path's centred source)
  Centre C;
};

// One cell of a blocked sweep: the coefficients at (px, py, pz) (A, B, C are
// h2 * p^2 along x, y, z) and the Jacobi update or the residual.
template <typename R, int OP, bool LOCAL, bool FAST>
__device__ __forceinline__ R stencil_cell(const TileCoef<R>& T, const Hood<R>& h, R fv, R px, R A,
                                          R py, R B, R pz, R C, R pypz, R omega, R keep) {
  const LevelGeom<R>& G = T.G;
  R g, g2, cx, cy, cz, d, pxpy, pxpz, res;
  if (LOCAL) {
    pxpy = px * py;
    pxpz = px * pz;
    if constexpr (FAST) {
      g = G.beta * rcp_fast(A + B + C);
      cx = fmaf(g * px, px, G.ih2[0]);
      cy = fmaf(g * py, py, G.ih2[1]);
      cz = fmaf(g * pz, pz, G.ih2[2]);
      g2 = g * R(0.5);
      d = R(2) * (cx + cy + cz);
    } else {
      g = G.beta / (A + B + C);
      cx = G.ih2[0] + g * px * px;
      cy = G.ih2[1] + g * py * py;
      cz = G.ih2[2] + g * pz * pz;
      g2 = g / R(2);
      d = R(2) * (cx + cy + cz);
    }
  } else {
    g = T.K.g;
    g2 = T.K.g2;
    cx = T.K.cx;
    cy = T.K.cy;
    cz = T.K.cz;
    d = T.K.d;
    pxpy = T.K.pxpy;
    pxpz = T.K.pxpz;
  }
  if constexpr (FAST) {
    if (OP != kOpResidual) {
      const R dinv = LOCAL ? rcp_fast(d) : T.dinv;
      res = jacobi_fast<LOCAL>(h, fv, px, py, pz, pxpy, pxpz, pypz, g, g2, cx, cy, cz,
                               dinv, omega, keep);
    } else {
      res = residual_fast<LOCAL>(h, fv, px, py, pz, pxpy, pxpz, pypz, g, g2, cx, cy, cz,
                                 d);
    }
  } else {
    if (OP != kOpResidual)
      res = jacobi_split<R, LOCAL>(h, fv, px, py, pz, pxpy, pxpz, pypz, g, g2, cx, cy,
                                   cz, d, omega, keep);
    else
      res = residual_split<R, LOCAL>(h, fv, px, py, pz, pxpy, pxpz, pypz, g, g2, cx,
                                     cy, cz, d);
  }
  return res;
}

// Minimum waves per SIMD requested for the blocked kernel (a register cap);
// 1 leaves the allocation to the compiler.  The FP64-to-FP32 residual keeps 4
// waves per SIMD (128 VGPRs) with its sources loaded a plane ahead.
#ifndef PRH_TILED_MIN_WAVES
#define PRH_TILED_MIN_WAVES 1
#endif
template <typename R, typename O>
constexpr int tiled_min_waves() {
  return std::is_same<R, double>::value && std::is_same<O, float>::value ? 4 : PRH_TILED_MIN_WAVES;
}

template <typename R, typename O, int OP, int TY, bool LOCAL, bool FAST>
__global__ __launch_bounds__(256, (tiled_min_waves<R, O>())) void k_stencil_tiled(const R* __restrict__ v,
                                                       const R* __restrict__ f,
                                                       O* __restrict__ out, TileCoef<R> T,
                                                       R omega, R keep, int xlen) {
  constexpr int W = kTileZ + 2, H = TY + 2, PLANE = W * H, KY = TY / 4;
  constexpr int PER = (PLANE + 255) / 256;
  __shared__ R lds[4 * PLANE];
  const LevelGeom<R>& G = T.G;
  const R fmean = OP == kOpResidual ? centre_mean<R>(T.C) : R(0);
  const int n0 = G.n[0], n1 = G.n[1], n2 = G.n[2];
  const int tid = threadIdx.x, tz = tid & 63, ty = tid >> 6;
  // XCD-aware tile order: the hardware deals blocks to the 8 XCDs round-robin;
  // logical tiles are numbered z-tile fastest, then y-tile, then x-run, and each
  // XCD gets a contiguous range of them, so tiles that share halo rows and
  // columns run together on one L2
  const int ztiles = n2 / kTileZ, ytiles = n1 / TY;
  const unsigned nb = gridDim.x, bid = blockIdx.x, per = nb / 8;
  const unsigned L = bid < per * 8 ? (bid % 8) * per + bid / 8 : bid;
  const int zt = static_cast<int>(L % ztiles), yt = static_cast<int>((L / ztiles) % ytiles);
  const int xr = static_cast<int>(L / (static_cast<unsigned>(ztiles) * ytiles));
  const int z0 = zt * kTileZ, y0 = yt * TY;
  const int xb = xr * xlen;
  const int xe = xb + xlen < n0 ? xb + xlen : n0;
  const size_t s0 = static_cast<size_t>(n1) * n2;
  int off[PER];
  bool has[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int e = tid + 256 * q;
    has[q] = e < PLANE;
    const int row = e / W, col = e - (e / W) * W;
    int gy = y0 - 1 + row;
    gy += gy < 0 ? n1 : 0;
    gy -= gy >= n1 ? n1 : 0;
    int gz = z0 - 1 + col;
    gz += gz < 0 ? n2 : 0;
    gz -= gz >= n2 ? n2 : 0;
    off[q] = gy * n2 + gz;
  }
  auto wrapx = [n0](int x) { return x < 0 ? x + n0 : (x >= n0 ? x - n0 : x); };
  // ring slot of plane x, relative to the run start (static inside the unrolled loop)
  auto slot = [xb](int x) { return (x - xb + 4) & 3; };
  if (OP != kOpJacobiZero)
  for (int p = -1; p <= 1; p++) {
    const R* src = v + static_cast<size_t>(wrapx(xb + p)) * s0;
    R* dst = lds + slot(xb + p) * PLANE;
#pragma unroll
    for (int q = 0; q < PER; q++)
      if (has[q]) dst[tid + 256 * q] = src[off[q]];
  }
  // per-thread constants: rows iy_k = y0 + ty + 4k, column iz
  const int iz = z0 + tz;
  int rowoff[KY], lo[KY];
  R py[KY], B[KY], pypz[KY];
  R pz = R(0), C = R(0);
  if (LOCAL) {
    pz = static_cast<R>(iz) + G.org[2];
    C = G.h2[2] * pz * pz;
  }
#pragma unroll
  for (int k = 0; k < KY; k++) {
    const int iy = y0 + ty + 4 * k;
    rowoff[k] = iy * n2 + iz;
    lo[k] = (ty + 4 * k + 1) * W + tz + 1;
    if (LOCAL) {
      py[k] = static_cast<R>(iy) + G.org[1];
      B[k] = G.h2[1] * py[k] * py[k];
      pypz[k] = py[k] * pz;
    } else {
      py[k] = R(0);
      B[k] = R(0);
      pypz[k] = T.K.pypz;
    }
  }
  // planes x + 2 and x + 3 are in flight in registers while plane x is computed:
  // set A holds the planes at an even distance from the run start, set B the odd
  R preA[PER], preB[PER];
  auto fetch = [&](R(&dst)[PER], int xp) {
    const R* src = v + static_cast<size_t>(wrapx(xp)) * s0;
#pragma unroll
    for (int q = 0; q < PER; q++)
      if (has[q]) dst[q] = src[off[q]];
  };
  auto stash = [&](const R(&src)[PER], int sl) {
    R* dst = lds + sl * PLANE;
#pragma unroll
    for (int q = 0; q < PER; q++)
      if (has[q]) dst[tid + 256 * q] = src[q];
  };
  // the source values of plane x, loaded one plane ahead (see k_stencil_rows)
  auto load_f = [&](R(&dst)[KY], int x) {
    const R* fx = f + static_cast<size_t>(x) * s0;
#pragma unroll
    for (int k = 0; k < KY; k++) dst[k] = fx[rowoff[k]];
  };
  auto compute = [&](int x, const R(&fvs)[KY], const R* M, const R* P, const R* N) {
    O* ox = out + static_cast<size_t>(x) * s0;
    R px = R(0), A = R(0);
    if (LOCAL) {
      px = static_cast<R>(x) + G.org[0];
      A = G.h2[0] * px * px;
    }
#pragma unroll
    for (int k = 0; k < KY; k++) {
      const int o = lo[k];
      Hood<R> h = zero_hood<R>();
      if (OP != kOpJacobiZero) {
      h.c = P[o];
      h.xp = N[o];
      h.xm = M[o];
      h.yp = P[o + W];
      h.ym = P[o - W];
      h.zp = P[o + 1];
      h.zm = P[o - 1];
      h.xpyp = N[o + W];
      h.xmym = M[o - W];
      h.xmyp = M[o + W];
      h.xpym = N[o - W];
      h.xpzp = N[o + 1];
      h.xmzm = M[o - 1];
      h.xmzp = M[o + 1];
      h.xpzm = N[o - 1];
      h.ypzp = P[o + W + 1];
      h.ymzm = P[o - W - 1];
      h.ymzp = P[o - W + 1];
      h.ypzm = P[o + W - 1];
      }
      const R fv = (OP == kOpResidual && T.C.sum) ? fvs[k] - fmean : fvs[k];
      const R res = stencil_cell<R, OP, LOCAL, FAST>(T, h, fv, px, A, py[k], B[k], pz, C, pypz[k],
                                                     omega, keep);
      if constexpr (OP == kOpJacobiAcc)
        T.acc[static_cast<size_t>(x) * s0 + rowoff[k]] += static_cast<double>(res);
      else
        ox[rowoff[k]] = static_cast<O>(res);
    }
  };
  // Loads in the order of k_stencil_rows: plane x's sources during plane x - 1,
  // after that plane's prefetch, every step issuing the same (clamped) loads.
  // The FP64 sweeps and residuals of the plain path keep the sources inside the
  // plane's step: the early loads cost them 4.6% at 1024^3 and gained nothing
  // at 512^3, while the FP64-to-FP32 residual gains 4-5% at both.
  constexpr bool kAhead = !(std::is_same<R, double>::value && std::is_same<O, double>::value);
  R fA[KY], fB[KY];
  if constexpr (!kAhead) {
    if (OP == kOpJacobiZero) {
      for (int x = xb; x < xe; x++) {
        load_f(fA, x);
        compute(x, fA, lds, lds, lds);
      }
      return;
    }
    if (xb + 1 < xe) fetch(preA, xb + 2);
    __syncthreads();
    for (int x0 = xb; x0 < xe; x0 += 4) {
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int x = x0 + u;
        if (x < xe) {
          if (x + 2 < xe) {
            if ((u + 3) & 1)
              fetch(preB, x + 3);
            else
              fetch(preA, x + 3);
          }
          load_f(fA, x);
          compute(x, fA, lds + ((u + 3) & 3) * PLANE, lds + (u & 3) * PLANE, lds + ((u + 1) & 3) * PLANE);
          if (x + 1 < xe) {
            if ((u + 2) & 1)
              stash(preB, (u + 2) & 3);
            else
              stash(preA, (u + 2) & 3);
          }
          __syncthreads();
        }
      }
    }
    return;
  }
  if (OP == kOpJacobiZero) {
    // zero iterate: no neighbourhood to stage, one pass over the run
    load_f(fA, xb);
    for (int x0 = xb; x0 < xe; x0 += 2) {
      load_f(fB, min(x0 + 1, xe - 1));
      compute(x0, fA, lds, lds, lds);
      load_f(fA, min(x0 + 2, xe - 1));
      if (x0 + 1 < xe) compute(x0 + 1, fB, lds, lds, lds);
    }
    return;
  }
  fetch(preA, min(xb + 2, xe));
  load_f(fA, xb);
  __syncthreads();
  for (int x0 = xb; x0 < xe; x0 += 4) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int x = x0 + u;
      if ((u + 3) & 1)
        fetch(preB, min(x + 3, xe));
      else
        fetch(preA, min(x + 3, xe));
      if ((u + 1) & 1)
        load_f(fB, min(x + 1, xe - 1));
      else
        load_f(fA, min(x + 1, xe - 1));
      __builtin_amdgcn_sched_barrier(0);
      if (x < xe) {
        compute(x, (u & 1) ? fB : fA, lds + ((u + 3) & 3) * PLANE, lds + (u & 3) * PLANE,
                lds + ((u + 1) & 3) * PLANE);
        if (x + 1 < xe) {
          if ((u + 2) & 1)
            stash(preB, (u + 2) & 3);
          else
            stash(preA, (u + 2) & 3);
        }
        __syncthreads();
      }
    }
  }
}
#ifndef PRH_TILE_Y
#define PRH_TILE_Y 8
#endif
constexpr int kTileY = PRH_TILE_Y;  // rows per 64-wide tile (2 per thread at 8)

// Full-row blocked kernel for FP32 levels with n2 = 64 * WPR * CZ and
// n1 % RY == 0.  A block owns RY complete rows (every z) of a run of x
// planes: WPR wavefronts per row, each lane CZ cells 64 * WPR apart.  Whole rows
// make the z neighbours wrap inside the staged row, so besides the tile only
// the two y-halo rows are read (the 64-wide tiles above also fetch a cache
// line on each side of every row).  Planes x-1, x, x+1 sit in a ring of 4
// LDS slots and planes x+2, x+3 are in flight in registers, as above; the
// arithmetic per cell is stencil_cell()'s.
constexpr int kRowsY = 8;
#ifndef PRH_ROWS_BLOCKS
#define PRH_ROWS_BLOCKS 512
#endif
#ifndef PRH_ROWS_MIN_XLEN
#define PRH_ROWS_MIN_XLEN 4  // shortest x run of a whole-row block
#endif
#ifndef PRH_ROWS_MIN_WAVES
#define PRH_ROWS_MIN_WAVES 4  // 128 VGPRs: two 512-thread blocks per CU on rows of 256 cells
#endif

template <typename R, typename O, int OP, int WPR, int CZ, bool LOCAL, bool FAST, int RY = kRowsY>
__global__ __launch_bounds__(64 * RY * WPR, PRH_ROWS_MIN_WAVES) void k_stencil_rows(const R* __restrict__ v,
                                                              const R* __restrict__ f,
                                                              O* __restrict__ out, TileCoef<R> T,
                                                              R omega, R keep, int xlen) {
  constexpr int N2 = 64 * WPR * CZ, PLANE = (RY + 2) * N2, NT = 64 * RY * WPR;
  constexpr int PER = (PLANE + NT - 1) / NT;
  // kOpJacobiAfterZero: v is the source, and the staged value of each cell is
  // the zero-start sweep's; the computation is then a plain sweep
  constexpr bool AZ = OP == kOpJacobiAfterZero;
  constexpr int OPC = AZ ? kOpJacobi : OP;
  static_assert(!AZ || FAST, "two sweeps from zero: the FP32 fast sweeps");
  __shared__ R lds[4 * PLANE];
  const LevelGeom<R>& G = T.G;
  const int n0 = G.n[0], n1 = G.n[1];
  const int tid = threadIdx.x, lane = tid & 63, ty = (tid >> 6) / WPR;
  const int zb = lane + 64 * ((tid >> 6) % WPR);  // first column of the lane; then every 64 * WPR
  // XCD-aware order (see k_stencil_tiled): y-tile fastest, then x-run
  const int ytiles = n1 / RY;
  const unsigned nb = gridDim.x, bid = blockIdx.x, per = nb / 8;
  const unsigned L = bid < per * 8 ? (bid % 8) * per + bid / 8 : bid;
  const int yt = static_cast<int>(L % ytiles), xr = static_cast<int>(L / ytiles);
  const int y0 = yt * RY;
  const int xb = xr * xlen;
  const int xe = xb + xlen < n0 ? xb + xlen : n0;
  const size_t s0 = static_cast<size_t>(n1) * N2;
  int off[PER];
  bool has[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int e = tid + NT * q;
    has[q] = (PLANE % NT == 0) || e < PLANE;
    const int row = e / N2, col = e % N2;
    int gy = y0 - 1 + row;
    gy += gy < 0 ? n1 : 0;
    gy -= gy >= n1 ? n1 : 0;
    off[q] = gy * N2 + col;
  }
  auto wrapx = [n0](int x) { return x < 0 ? x + n0 : (x >= n0 ? x - n0 : x); };
  auto slot = [xb](int x) { return (x - xb + 4) & 3; };
  // the value staged for element q of plane xp (a plane index in [0, n0))
  auto staged = [&](R val, int xp, int q) -> R {
    if constexpr (!AZ) {
      return val;
    } else {
      // the zero-start sweep at (xp, gy, gz), with the coefficients that
      // sweep's own kernel computes for that cell.  They are recomputed at each
      // plane: the empty asm hides that they do not change along x, which would
      // otherwise hold all of them in registers across the loop (and spill).
      int o = off[q];
      asm volatile("" : "+v"(o));
      const int gy = o / N2, gz = o - gy * N2;
      R px = R(0), A = R(0), py = R(0), B = R(0), pz = R(0), C = R(0), pypz;
      if (LOCAL) {
        px = static_cast<R>(xp) + G.org[0];
        A = G.h2[0] * px * px;
        py = static_cast<R>(gy) + G.org[1];
        B = G.h2[1] * py * py;
        pz = static_cast<R>(gz) + G.org[2];
        C = G.h2[2] * pz * pz;
        pypz = py * pz;
      } else {
        pypz = T.K.pypz;
      }
      return stencil_cell<R, kOpJacobiZero, LOCAL, FAST>(T, zero_hood<R>(), val, px, A, py, B, pz, C,
                                                         pypz, omega, keep);
    }
  };
  if (OP != kOpJacobiZero)
    for (int p = -1; p <= 1; p++) {
      const int xp = wrapx(xb + p);
      const R* src = v + static_cast<size_t>(xp) * s0;
      R* dst = lds + slot(xb + p) * PLANE;
#pragma unroll
      for (int q = 0; q < PER; q++)
        if (has[q]) dst[tid + NT * q] = staged(src[off[q]], xp, q);
    }
  // per-thread constants: row iy = y0 + ty, columns iz = zb + 64 WPR c
  const int iy = y0 + ty;
  const int rowoff = iy * N2;
  R py = R(0), B = R(0);
  R pz[CZ], C[CZ], pypz[CZ];
  if (LOCAL) {
    py = static_cast<R>(iy) + G.org[1];
    B = G.h2[1] * py * py;
  }
#pragma unroll
  for (int c = 0; c < CZ; c++) {
    if (LOCAL) {
      pz[c] = static_cast<R>(zb + 64 * WPR * c) + G.org[2];
      C[c] = G.h2[2] * pz[c] * pz[c];
      pypz[c] = py * pz[c];
    } else {
      pz[c] = R(0);
      C[c] = R(0);
      pypz[c] = T.K.pypz;
    }
  }
  R preA[PER], preB[PER];
  auto fetch = [&](R(&dst)[PER], int xp) {
    const R* src = v + static_cast<size_t>(wrapx(xp)) * s0;
#pragma unroll
    for (int q = 0; q < PER; q++)
      if (has[q]) dst[q] = src[off[q]];
  };
  auto stash = [&](const R(&src)[PER], int sl, int xp) {
    R* dst = lds + sl * PLANE;
#pragma unroll
    for (int q = 0; q < PER; q++)
      if (has[q]) dst[tid + NT * q] = src[q];
    if constexpr (AZ) {
      // the thread's own elements, in place, once the loaded values are out of
      // registers (kept apart from the sweep's own code, which holds the rest)
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int q = 0; q < PER; q++) {
        if (has[q]) dst[tid + NT * q] = staged(dst[tid + NT * q], xp, q);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  };
  const int rm = ty * N2, r0 = (ty + 1) * N2, rp = (ty + 2) * N2;
  // the source values of plane x (and, for kOpJacobiAcc, the FP64 values the
  // results are added to), loaded one plane ahead (see the loop below)
  struct Src {
    R f[CZ];
    double a[CZ];
  };
  auto load_f = [&](Src& dst, int x) {
    const R* fx = f + static_cast<size_t>(x) * s0 + rowoff;
#pragma unroll
    for (int c = 0; c < CZ; c++) dst.f[c] = fx[zb + 64 * WPR * c];
    if constexpr (OP == kOpJacobiAcc && N2 == 512) {
      const double* ax = T.acc + static_cast<size_t>(x) * s0 + rowoff;
#pragma unroll
      for (int c = 0; c < CZ; c++) dst.a[c] = ax[zb + 64 * WPR * c];
    }
  };
  auto compute = [&](int x, const Src& src, const R* M, const R* P, const R* N) {
    O* ox = out + static_cast<size_t>(x) * s0 + rowoff;
    R px = R(0), A = R(0);
    if (LOCAL) {
      px = static_cast<R>(x) + G.org[0];
      A = G.h2[0] * px * px;
    }
#pragma unroll
    for (int c = 0; c < CZ; c++) {
      const int zc = zb + 64 * WPR * c;
      const int zl = zc == 0 ? N2 - 1 : zc - 1;
      const int zr = zc == N2 - 1 ? 0 : zc + 1;
      Hood<R> h = zero_hood<R>();
      if (OP != kOpJacobiZero) {
        h.c = P[r0 + zc];
        h.xp = N[r0 + zc];
        h.xm = M[r0 + zc];
        h.yp = P[rp + zc];
        h.ym = P[rm + zc];
        h.zp = P[r0 + zr];
        h.zm = P[r0 + zl];
        h.xpyp = N[rp + zc];
        h.xmym = M[rm + zc];
        h.xmyp = M[rp + zc];
        h.xpym = N[rm + zc];
        h.xpzp = N[r0 + zr];
        h.xmzm = M[r0 + zl];
        h.xmzp = M[r0 + zr];
        h.xpzm = N[r0 + zl];
        h.ypzp = P[rp + zr];
        h.ymzm = P[rm + zl];
        h.ymzp = P[rm + zr];
        h.ypzm = P[rp + zl];
      }
      const R res = stencil_cell<R, OPC, LOCAL, FAST>(T, h, src.f[c], px, A, py, B, pz[c], C[c],
                                                     pypz[c], omega, keep);
      if constexpr (OP == kOpJacobiAcc) {
        double* a = T.acc + static_cast<size_t>(x) * s0 + rowoff + zc;
        *a = (N2 == 512 ? src.a[c] : *a) + static_cast<double>(res);
      } else
        ox[zc] = static_cast<O>(res);
    }
  };
  // Global loads retire in issue order (one vmcnt), so a load issued after the
  // prefetch of plane x + 3 could not be waited for without waiting for that
  // prefetch too.  Plane x's source values are therefore loaded during plane
  // x - 1, after that plane's prefetch: computing plane x waits only for loads
  // issued a plane earlier, and plane x + 3 and the sources of x + 1 stay in
  // flight meanwhile.  Every step issues the same loads, also past the end of
  // the run (plane indices clamped to it: such a load re-reads a plane that is
  // in L2 and its value is not used), so that the waits can count on them.
  // Only rows of 512 cells load ahead: their blocks walk long runs (64 planes
  // at 512^3; 0.322 -> 0.306 ms per sweep), while the short runs of the smaller
  // levels lose a few percent to the clamped loads at their ends (64^3: 5.15 ->
  // 5.35 us) and rows of 1024 cells would spill the registers in flight.
  constexpr bool kAhead = N2 == 512;
  Src fA, fB;
  if (OP == kOpJacobiZero) {
    load_f(fA, xb);
    for (int x0 = xb; x0 < xe; x0 += 2) {
      load_f(fB, min(x0 + 1, xe - 1));
      compute(x0, fA, lds, lds, lds);
      load_f(fA, min(x0 + 2, xe - 1));
      if (x0 + 1 < xe) compute(x0 + 1, fB, lds, lds, lds);
    }
    return;
  }
  if constexpr (!kAhead) {
    if (xb + 1 < xe) fetch(preA, xb + 2);
    __syncthreads();
    for (int x0 = xb; x0 < xe; x0 += 4) {
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int x = x0 + u;
        if (x < xe) {
          if (x + 2 < xe) {
            if ((u + 3) & 1)
              fetch(preB, x + 3);
            else
              fetch(preA, x + 3);
          }
          load_f(fA, x);
          compute(x, fA, lds + ((u + 3) & 3) * PLANE, lds + (u & 3) * PLANE, lds + ((u + 1) & 3) * PLANE);
          if (x + 1 < xe) {
            if ((u + 2) & 1)
              stash(preB, (u + 2) & 3, wrapx(x + 2));
            else
              stash(preA, (u + 2) & 3, wrapx(x + 2));
          }
          __syncthreads();
        }
      }
    }
    return;
  }
  fetch(preA, min(xb + 2, xe));
  load_f(fA, xb);
  __syncthreads();
  for (int x0 = xb; x0 < xe; x0 += 4) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int x = x0 + u;
      if ((u + 3) & 1)
        fetch(preB, min(x + 3, xe));
      else
        fetch(preA, min(x + 3, xe));
      if ((u + 1) & 1)
        load_f(fB, min(x + 1, xe - 1));
      else
        load_f(fA, min(x + 1, xe - 1));
      // keep the loads ahead of the computation (the scheduler would otherwise
      // sink them below the waits for this plane's sources)
      __builtin_amdgcn_sched_barrier(0);
      if (x < xe) {
        compute(x, (u & 1) ? fB : fA, lds + ((u + 3) & 3) * PLANE, lds + (u & 3) * PLANE,
                lds + ((u + 1) & 3) * PLANE);
        if (x + 1 < xe) {
          if ((u + 2) & 1)
            stash(preB, (u + 2) & 3, wrapx(x + 2));
          else
            stash(preA, (u + 2) & 3, wrapx(x + 2));
        }
        __syncthreads();
      }
    }
  }
}

// Two stencil operations fused along x (temporal blocking), FP32 fast kernels
// on levels with whole rows of N2 = 256 or 512 cells and n1 % 8 == 0: OP1
// (a Jacobi sweep, or one from zero) maps v to u and OP2 (a Jacobi sweep, or
// one added into the FP64 mesh acc) maps u to the output, and u never leaves
// LDS.  A block owns RY = 8 complete rows of a run of x planes.  For each plane
// p it computes u on rows y0 - 1 .. y0 + RY (the halo rows OP2 needs; the
// neighbouring blocks compute the same values) from the planes p - 1, p, p + 1
// of v on rows y0 - 2 .. y0 + RY + 1, held in a ring of 3 LDS slots, then the
// output at plane p - 1 from the ring of 3 u planes.  v plane p + 3 is in
// flight in registers meanwhile.  Per cell, the two sweeps move 12 bytes
// instead of 24 (8 from a zero start, 24 with the FP64 update instead of 40).
// The arithmetic per cell is stencil_cell()'s: the same bits as two launches.
constexpr int kPairRY = 8;
template <int OP1, int OP2, int NT, int N2, bool LOCAL>
__global__ __launch_bounds__(NT, 1) void k_stencil_pair(const float* __restrict__ v,
                                                       const float* __restrict__ f,
                                                       float* __restrict__ out, TileCoef<float> T,
                                                       float omega, float keep, int xlen) {
  constexpr int RY = kPairRY, VR = RY + 4, UR = RY + 2;
  constexpr bool ZERO = OP1 == kOpJacobiZero;
  constexpr int VPLANE = VR * N2, UPLANE = UR * N2, WPLANE = RY * N2;
  constexpr int VPER = (VPLANE + NT - 1) / NT, UPER = (UPLANE + NT - 1) / NT, WPER = WPLANE / NT;
  static_assert(WPLANE % NT == 0, "output rows split evenly over the threads");
  __shared__ float sv[ZERO ? 1 : 3 * VPLANE];
  __shared__ float su[3 * UPLANE];
  const LevelGeom<float>& G = T.G;
  const int n0 = G.n[0], n1 = G.n[1];
  const int tid = threadIdx.x;
  // XCD-aware order (see k_stencil_tiled): y-tile fastest, then x-run
  const int ytiles = n1 / RY;
  const unsigned nb = gridDim.x, bid = blockIdx.x, per = nb / 8;
  const unsigned L = bid < per * 8 ? (bid % 8) * per + bid / 8 : bid;
  const int yt = static_cast<int>(L % ytiles), xr = static_cast<int>(L / ytiles);
  const int y0 = yt * RY;
  const int xb = xr * xlen;
  const int xe = xb + xlen < n0 ? xb + xlen : n0;
  const size_t s0 = static_cast<size_t>(n1) * N2;
  auto wrapx = [n0](int x) { return x < 0 ? x + n0 : (x >= n0 ? x - n0 : x); };
  auto wrapy = [n1](int y) { return y < 0 ? y + n1 : (y >= n1 ? y - n1 : y); };
  auto slot = [xb](int p) { return (p - xb + 3) % 3; };
  // v rows y0 - 2 .. y0 + RY + 1 of a plane: element e = tid + NT q
  int voff[VPER];
  bool vhas[VPER];
  float pre[VPER];
#pragma unroll
  for (int q = 0; q < VPER; q++) {
    const int e = tid + NT * q;
    vhas[q] = !ZERO && e < VPLANE;
    voff[q] = wrapy(y0 - 2 + e / N2) * N2 + e % N2;
  }
  auto fetch = [&](int p) {
    const float* src = v + static_cast<size_t>(wrapx(p)) * s0;
#pragma unroll
    for (int q = 0; q < VPER; q++)
      if (vhas[q]) pre[q] = src[voff[q]];
  };
  auto stash = [&](int p) {
    float* dst = sv + slot(p) * VPLANE;
#pragma unroll
    for (int q = 0; q < VPER; q++)
      if (vhas[q]) dst[tid + NT * q] = pre[q];
  };
  // one cell of an operation: planes M, P, N (row stride N2), centre row rb / N2
  auto cell = [&](auto op, const float* M, const float* P, const float* N, int rb, int c, float fv,
                  int x, int iy) {
    constexpr int OPC = decltype(op)::value;
    const int cl = c == 0 ? N2 - 1 : c - 1;
    const int cr = c == N2 - 1 ? 0 : c + 1;
    Hood<float> h = zero_hood<float>();
    if (OPC != kOpJacobiZero) {
      const int rm = rb - N2, rp = rb + N2;
      h.c = P[rb + c];
      h.xp = N[rb + c];
      h.xm = M[rb + c];
      h.yp = P[rp + c];
      h.ym = P[rm + c];
      h.zp = P[rb + cr];
      h.zm = P[rb + cl];
      h.xpyp = N[rp + c];
      h.xmym = M[rm + c];
      h.xmyp = M[rp + c];
      h.xpym = N[rm + c];
      h.xpzp = N[rb + cr];
      h.xmzm = M[rb + cl];
      h.xmzp = M[rb + cr];
      h.xpzm = N[rb + cl];
      h.ypzp = P[rp + cr];
      h.ymzm = P[rm + cl];
      h.ymzp = P[rm + cr];
      h.ypzm = P[rp + cl];
    }
    float px = 0.f, A = 0.f, py = 0.f, B = 0.f, pz = 0.f, C = 0.f, pypz = T.K.pypz;
    if (LOCAL) {
      px = static_cast<float>(x) + G.org[0];
      A = G.h2[0] * px * px;
      py = static_cast<float>(iy) + G.org[1];
      B = G.h2[1] * py * py;
      pz = static_cast<float>(c) + G.org[2];
      C = G.h2[2] * pz * pz;
      pypz = py * pz;
    }
    return stencil_cell<float, OPC, LOCAL, true>(T, h, fv, px, A, py, B, pz, C, pypz, omega, keep);
  };
  if (!ZERO) {
    fetch(xb - 2);
    stash(xb - 2);
    fetch(xb - 1);
    stash(xb - 1);
    fetch(xb);
    stash(xb);
    fetch(xb + 1);
  }
  __syncthreads();
  for (int p = xb - 1; p <= xe; p++) {
    // u at plane p, rows y0 - 1 .. y0 + RY
    {
      const int xp = wrapx(p);
      const float* fx = f + static_cast<size_t>(xp) * s0;
      const float* Vm = sv + (ZERO ? 0 : slot(p - 1) * VPLANE);
      const float* Vp = sv + (ZERO ? 0 : slot(p) * VPLANE);
      const float* Vn = sv + (ZERO ? 0 : slot(p + 1) * VPLANE);
      float* U = su + slot(p) * UPLANE;
#pragma unroll
      for (int q = 0; q < UPER; q++) {
        const int e = tid + NT * q;
        if (e < UPLANE) {
          const int r = e / N2, c = e % N2;
          const int iy = wrapy(y0 - 1 + r);
          U[e] = cell(std::integral_constant<int, OP1>(), Vm, Vp, Vn, (r + 1) * N2, c,
                      fx[static_cast<size_t>(iy) * N2 + c], xp, iy);
        }
      }
    }
    __syncthreads();
    if (!ZERO && p + 2 <= xe + 1) {
      stash(p + 2);  // into the slot of plane p - 1, which no later u reads
      if (p + 3 <= xe + 1) fetch(p + 3);
    }
    // the output at plane x = p - 1, rows y0 .. y0 + RY - 1
    const int x = p - 1;
    if (x >= xb) {
      const float* fx = f + static_cast<size_t>(x) * s0;
      const float* Um = su + slot(x - 1) * UPLANE;
      const float* Up = su + slot(x) * UPLANE;
      const float* Un = su + slot(x + 1) * UPLANE;
#pragma unroll
      for (int q = 0; q < WPER; q++) {
        const int e = tid + NT * q;
        const int r = e / N2, c = e % N2;
        const int iy = y0 + r;
        const size_t gi = static_cast<size_t>(x) * s0 + static_cast<size_t>(iy) * N2 + c;
        const float res = cell(std::integral_constant<int, OP2>(), Um, Up, Un, (r + 1) * N2, c,
                               fx[static_cast<size_t>(iy) * N2 + c], x, iy);
        if constexpr (OP2 == kOpJacobiAcc)
          T.acc[gi] += static_cast<double>(res);
        else
          out[gi] = res;
      }
    }
    __syncthreads();
  }
}

// Rows per block of the full-row kernel: 8, or 4 for rows of 1024 cells (four
// wavefronts per row, four cells per lane, 96 KB of LDS for the ring of 4
// planes).  Measured on the 1024^3 FP32 sweep: 2.71 ms (59% of 8 TB/s) against
// 2.88 ms with 2 rows of 8 wavefronts and 2.99 ms with the 64-wide tiles.
constexpr int kRowsY1024 = 4;
inline int rows_ry(const int* n) { return n[2] == 1024 ? kRowsY1024 : kRowsY; }

// row length in units of 64 cells (1, 2, 4, 8 or 16), 0 where the kernel does not apply
inline int rows_cz(const int* n) {
  if (n[2] % 64 != 0 || n[1] % rows_ry(n) != 0 || n[0] < 1) return 0;
  const int cz = n[2] / 64;
  return (cz == 1 || cz == 2 || cz == 4 || cz == 8 || cz == 16) ? cz : 0;
}

// x-run per block so that the grid has roughly PRH_ROWS_BLOCKS blocks
inline int rows_xlen(const int* n) {
  const long tiles = n[1] / rows_ry(n);
  long chunks = (PRH_ROWS_BLOCKS + tiles - 1) / tiles;
  if (chunks > n[0]) chunks = n[0];
  int xlen = static_cast<int>((n[0] + chunks - 1) / chunks);
  return xlen < PRH_ROWS_MIN_XLEN ? (n[0] < PRH_ROWS_MIN_XLEN ? n[0] : PRH_ROWS_MIN_XLEN) : xlen;
}

inline bool tiled_ok(const int* n) { return n[2] % kTileZ == 0 && n[1] % kTileY == 0 && n[0] >= 1; }

// PRH_TUNE_PAIRED_SWEEPS, off by default: on MI355X the fused pair is slower
// than two single sweeps (0.69 against 2 x 0.31 ms at 512^3), because a sweep
// of this 19-point operator with per-cell coefficients is nearly as VALU-bound
// (about 60 instructions per cell) as it is HBM-bound, and the pair recomputes
// a quarter of the intermediate cells with one 132 KB block per CU.
static std::atomic<int> g_paired_sweeps{0};

inline bool pair_ok(const int* n) {
  return (n[2] == 256 || n[2] == 512) && n[1] % kPairRY == 0 && n[0] >= 2;
}
#ifndef PRH_PAIR_BLOCKS
#define PRH_PAIR_BLOCKS 512
#endif
// x-run per block of the fused pair so that the grid has about PRH_PAIR_BLOCKS blocks
inline int pair_xlen(const int* n) {
  const long tiles = n[1] / kPairRY;
  long chunks = (PRH_PAIR_BLOCKS + tiles - 1) / tiles;
  if (chunks > n[0]) chunks = n[0];
  return static_cast<int>((n[0] + chunks - 1) / chunks);
}

// Two fused FP32 fast operations (k_stencil_pair): OP1 from v (kOpJacobi or
// kOpJacobiZero), OP2 into out (kOpJacobi) or into acc (kOpJacobiAcc).
template <int OP1, int OP2>
static void launch_pair(const float* v, const float* f, float* out, const LevelGeom<float>& G,
                        float omega, float keep, hipStream_t s, double* acc = nullptr) {
  TileCoef<float> T;
  T.G = G;
  T.K = G.local ? GlobalCoef<float>() : GlobalCoef<float>::make(G);
  T.dinv = G.local ? 0.f : 1.f / T.K.d;
  T.acc = acc;
  const int xlen = pair_xlen(G.n);
  const unsigned grid = static_cast<unsigned>((G.n[1] / kPairRY) * ((G.n[0] + xlen - 1) / xlen));
  if (G.n[2] == 512) {
    if (G.local)
      k_stencil_pair<OP1, OP2, 1024, 512, true><<<grid, 1024, 0, s>>>(v, f, out, T, omega, keep, xlen);
    else
      k_stencil_pair<OP1, OP2, 1024, 512, false><<<grid, 1024, 0, s>>>(v, f, out, T, omega, keep, xlen);
  } else {
    if (G.local)
      k_stencil_pair<OP1, OP2, 512, 256, true><<<grid, 512, 0, s>>>(v, f, out, T, omega, keep, xlen);
    else
      k_stencil_pair<OP1, OP2, 512, 256, false><<<grid, 512, 0, s>>>(v, f, out, T, omega, keep, xlen);
  }
}

// x-run per block so that the grid has roughly 2048+ blocks
inline int tiled_xlen(const int* n) {
  const long tiles = static_cast<long>(n[2] / kTileZ) * (n[1] / kTileY);
  long chunks = (2048 + tiles - 1) / tiles;
  if (chunks > n[0]) chunks = n[0];
  int xlen = static_cast<int>((n[0] + chunks - 1) / chunks);
  return xlen < 4 ? (n[0] < 4 ? n[0] : 4) : xlen;
}

template <typename R, typename O, int OP, bool FAST, int WPR, int CZ, int RY = kRowsY>
static void launch_rows_cz(const R* v, const R* f, O* out, const TileCoef<R>& T, R omega, R keep,
                           hipStream_t s) {
  const int xlen = rows_xlen(T.G.n);
  const unsigned grid =
      static_cast<unsigned>((T.G.n[1] / RY) * ((T.G.n[0] + xlen - 1) / xlen));
  constexpr int nt = 64 * RY * WPR;
  if (T.G.local)
    k_stencil_rows<R, O, OP, WPR, CZ, true, FAST, RY><<<grid, nt, 0, s>>>(v, f, out, T, omega, keep, xlen);
  else
    k_stencil_rows<R, O, OP, WPR, CZ, false, FAST, RY><<<grid, nt, 0, s>>>(v, f, out, T, omega, keep, xlen);
}

// The full-row kernel where it applies (FP32 levels with whole rows of 64, 128,
// 256 or 512 cells); false leaves the level to the 64-wide tiles.
template <typename R, typename O, int OP, bool FAST>
static bool launch_rows(const R* v, const R* f, O* out, const TileCoef<R>& T, R omega, R keep,
                        hipStream_t s) {
  if constexpr (OP == kOpJacobiAfterZero) {
    const int cz = rows_cz(T.G.n);
    if (!FAST || !std::is_same<R, float>::value || PRH_ROWS_BLOCKS <= 0 || cz == 0) return false;
    if constexpr (FAST && std::is_same<R, float>::value) {
      switch (cz) {
        case 1: launch_rows_cz<R, O, OP, FAST, 1, 1>(v, f, out, T, omega, keep, s); break;
        case 2: launch_rows_cz<R, O, OP, FAST, 1, 2>(v, f, out, T, omega, keep, s); break;
        case 4: launch_rows_cz<R, O, OP, FAST, 1, 4>(v, f, out, T, omega, keep, s); break;
        case 8: launch_rows_cz<R, O, OP, FAST, 2, 4>(v, f, out, T, omega, keep, s); break;
        default: launch_rows_cz<R, O, OP, FAST, 4, 4, kRowsY1024>(v, f, out, T, omega, keep, s); break;
      }
    }
    return true;
  } else if constexpr (std::is_same<R, float>::value && PRH_ROWS_BLOCKS > 0) {
    switch (rows_cz(T.G.n)) {
      case 1: launch_rows_cz<R, O, OP, FAST, 1, 1>(v, f, out, T, omega, keep, s); return true;
      case 2: launch_rows_cz<R, O, OP, FAST, 1, 2>(v, f, out, T, omega, keep, s); return true;
      case 4: launch_rows_cz<R, O, OP, FAST, 1, 4>(v, f, out, T, omega, keep, s); return true;
      // 512-cell rows: two wavefronts per row keep the registers per lane near those of 256
      case 8: launch_rows_cz<R, O, OP, FAST, 2, 4>(v, f, out, T, omega, keep, s); return true;
      // 1024-cell rows: four wavefronts per row, 4 rows per block
      case 16: launch_rows_cz<R, O, OP, FAST, 4, 4, kRowsY1024>(v, f, out, T, omega, keep, s); return true;
      default: return false;
    }
  }
  return false;
}

template <typename R, typename O, int OP, bool FAST>
static void launch_tiled(const R* v, const R* f, O* out, const LevelGeom<R>& G, R omega, R keep,
                         hipStream_t s, double* acc = nullptr, Centre C = {}) {
  const int xlen = tiled_xlen(G.n);
  const unsigned grid = static_cast<unsigned>((G.n[2] / kTileZ) * (G.n[1] / kTileY) *
                                              ((G.n[0] + xlen - 1) / xlen));
  TileCoef<R> T;
  T.G = G;
  T.K = G.local ? GlobalCoef<R>() : GlobalCoef<R>::make(G);
  T.dinv = G.local ? R(0) : R(1) / T.K.d;
  T.acc = acc;
  T.C = C;
  if constexpr (OP == kOpJacobiAfterZero) {
    // only where after_zero_ok() holds: FP32 fast rows of 512 cells
    launch_rows<R, O, OP, FAST>(v, f, out, T, omega, keep, s);
  } else {
    if (!C.sum && launch_rows<R, O, OP, FAST>(v, f, out, T, omega, keep, s)) return;
    if (G.local)
      k_stencil_tiled<R, O, OP, kTileY, true, FAST><<<grid, 256, 0, s>>>(v, f, out, T, omega, keep, xlen);
    else
      k_stencil_tiled<R, O, OP, kTileY, false, FAST><<<grid, 256, 0, s>>>(v, f, out, T, omega, keep, xlen);
  }
}

// Block shape of the generic kernels: up to 64 threads along z, the rest along y.
inline void generic_shape(const int* n, dim3& block, dim3& grid) {
  int bz = 1;
  while (bz < n[2] && bz < 64) bz <<= 1;
  const int by = kThreads / bz;
  block = dim3(bz, by, 1);
  grid = dim3((n[2] + bz - 1) / bz, (n[1] + by - 1) / by, n[0]);
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
// The 27-point sum of the full weighting; at(a, b, c) is the fine value at
// offset (a - 1, b - 1, c - 1) from the centre, a, b, c in [0, 3).
template <typename R, typename A>
__device__ __forceinline__ R restrict_sum(A at) {
  const R faces = at(2, 1, 1) + at(0, 1, 1) + at(1, 2, 1) + at(1, 0, 1) + at(1, 1, 2) + at(1, 1, 0);
  const R edges = at(2, 2, 1) + at(0, 2, 1) + at(2, 0, 1) + at(0, 0, 1) + at(2, 1, 2) + at(0, 1, 2) +
                  at(2, 1, 0) + at(0, 1, 0) + at(1, 2, 2) + at(1, 0, 2) + at(1, 2, 0) + at(1, 0, 0);
  R s = R(8) * at(1, 1, 1) + R(4) * faces + R(2) * edges;
  s += at(2, 2, 2);
  s += at(0, 2, 2);
  s += at(2, 0, 2);
  s += at(0, 0, 2);
  s += at(2, 2, 0);
  s += at(0, 2, 0);
  s += at(2, 0, 0);
  s += at(0, 0, 0);
  // reference divides by the double 64.0; an exact power-of-two scaling
  return s * R(0.015625);
}

template <typename R>
__device__ __forceinline__ R restrict_value(const R* __restrict__ u, int f0, int f1, int f2,
                                            int jx, int jy, int jz, const Centre& C = {}) {
  const R m = centre_mean<R>(C);
  const size_t s0 = static_cast<size_t>(f1) * f2, s1 = static_cast<size_t>(f2);
  const int fx = 2 * jx, fy = 2 * jy, fz = 2 * jz;
  const size_t xo[3] = {(fx == 0 ? f0 - 1 : fx - 1) * s0, fx * s0, (fx + 1 == f0 ? 0 : fx + 1) * s0};
  const size_t yo[3] = {(fy == 0 ? f1 - 1 : fy - 1) * s1, fy * s1, (fy + 1 == f1 ? 0 : fy + 1) * s1};
  const size_t zo[3] = {static_cast<size_t>(fz == 0 ? f2 - 1 : fz - 1), static_cast<size_t>(fz),
                        static_cast<size_t>(fz + 1 == f2 ? 0 : fz + 1)};
  return restrict_sum<R>([&](int a, int b, int c) {
    const R x = u[xo[a] + yo[b] + zo[c]];
    return C.sum ? x - m : x;
  });
}

template <typename R, typename RO, typename I>
__device__ __forceinline__ void restrict_at(const R* __restrict__ u, RO* __restrict__ c, int f0,
                                            int f1, int f2, I j, const Centre& C = {}) {
  int jx, jy, jz;
  unflatten(j, f1 / 2, f2 / 2, jx, jy, jz);
  c[j] = static_cast<RO>(restrict_value(u, f0, f1, f2, jx, jy, jz, C));
}

template <typename R, typename RO>
__device__ __forceinline__ void restrict_body(const R* __restrict__ u, RO* __restrict__ c, int f0,
                                              int f1, int f2, const Centre& C) {
  const size_t total = static_cast<size_t>(f0 / 2) * (f1 / 2) * (f2 / 2);
  const size_t j = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x;
  if (j < total) restrict_at<R, RO>(u, c, f0, f1, f2, j, C);
}

template <typename R, typename RO>
__global__ __launch_bounds__(kThreads) void k_restrict(const R* __restrict__ u,
                                                       RO* __restrict__ c, int f0, int f1,
                                                       int f2, Centre C) {
  restrict_body<R, RO>(u, c, f0, f1, f2, C);
}

// FP32 transfers with two coarse cells per thread (16-byte fine accesses)
#ifndef PRH_TRANSFER_QUADS
#define PRH_TRANSFER_QUADS 1
#endif
template <typename R> struct Pair;
template <> struct Pair<float> { using type = float2; };
template <> struct Pair<double> { using type = double2; };

// The same restriction on a 3D grid (coarse z along the block's x, coarse y
// along its y, coarse x = blockIdx.z).  Each of the 9 fine rows a thread
// needs is read as one pair (fine z = 2 jz, 2 jz + 1); fine z = 2 jz - 1 is
// the previous lane's second value, read directly only by the first thread
// of each block row.  Requires the fine mesh on a pair boundary.
template <typename R, typename RO>
__global__ __launch_bounds__(kThreads) void k_restrict_rows(const R* __restrict__ u,
                                                            RO* __restrict__ c, int f0, int f1,
                                                            int f2, Centre C) {
  using V = typename Pair<R>::type;
  const int c1 = f1 / 2, c2 = f2 / 2;
  const int jz0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int jy0 = blockIdx.y * blockDim.y + threadIdx.y;
  const int jx = blockIdx.z;
  const bool in = jz0 < c2 && jy0 < c1;
  // lanes past the edge compute a clamped cell (every lane takes part in the shuffles)
  const int jz = jz0 < c2 ? jz0 : c2 - 1, jy = jy0 < c1 ? jy0 : c1 - 1;
  const size_t s0 = static_cast<size_t>(f1) * f2, s1 = static_cast<size_t>(f2);
  const int fx = 2 * jx, fy = 2 * jy, fz = 2 * jz;
  const size_t xo[3] = {(fx == 0 ? f0 - 1 : fx - 1) * s0, fx * s0, (fx + 1 == f0 ? 0 : fx + 1) * s0};
  const size_t yo[3] = {(fy == 0 ? f1 - 1 : fy - 1) * s1, fy * s1, (fy + 1 == f1 ? 0 : fy + 1) * s1};
  const size_t zm = static_cast<size_t>(fz == 0 ? f2 - 1 : fz - 1);
  const bool first = threadIdx.x == 0;
  R w[3][3][3];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      const R* row = u + xo[a] + yo[b];
      const V pr = *reinterpret_cast<const V*>(row + fz);
      const R left = __shfl_up(pr.y, 1);
      w[a][b][1] = pr.x;
      w[a][b][2] = pr.y;
      w[a][b][0] = first ? row[zm] : left;
    }
  if (C.sum) {
    const R m = centre_mean<R>(C);
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++)
#pragma unroll
        for (int q = 0; q < 3; q++) w[a][b][q] -= m;
  }
  // fine z = 2 jz + 1 wraps to 0 only when f2 == 2 jz + 1, which cannot happen (f2 even)
  if (in)
    c[(static_cast<size_t>(jx) * c1 + jy) * c2 + jz] =
        static_cast<RO>(restrict_sum<R>([&](int a, int b, int q) { return w[a][b][q]; }));
}

// FP32 restriction with two coarse z per thread (coarse z = 2 jq, 2 jq + 1):
// each of the 9 fine rows is read as one float4 (fine z = 4 jq .. 4 jq + 3),
// fine z = 4 jq - 1 comes from the previous lane's last value (read directly
// by the first thread of each block row), and the two coarse values are
// stored as one float2.  The sums are restrict_sum()'s, so the bits are the
// pair kernel's.  Requires f2 % 4 == 0, the fine mesh on a 16-byte and the
// coarse mesh on an 8-byte boundary.
template <typename RO>
__global__ __launch_bounds__(kThreads) void k_restrict_quads(const float* __restrict__ u,
                                                             RO* __restrict__ c, int f0, int f1,
                                                             int f2) {
  using VO = typename Pair<RO>::type;
  const int c1 = f1 / 2, c2 = f2 / 2, q2 = c2 / 2;
  const int jq0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int jy0 = blockIdx.y * blockDim.y + threadIdx.y;
  const int jx = blockIdx.z;
  const bool in = jq0 < q2 && jy0 < c1;
  // lanes past the edge compute a clamped cell (every lane takes part in the shuffles)
  const int jq = jq0 < q2 ? jq0 : q2 - 1, jy = jy0 < c1 ? jy0 : c1 - 1;
  const size_t s0 = static_cast<size_t>(f1) * f2, s1 = static_cast<size_t>(f2);
  const int fx = 2 * jx, fy = 2 * jy, fz = 4 * jq;
  const size_t xo[3] = {(fx == 0 ? f0 - 1 : fx - 1) * s0, fx * s0, (fx + 1 == f0 ? 0 : fx + 1) * s0};
  const size_t yo[3] = {(fy == 0 ? f1 - 1 : fy - 1) * s1, fy * s1, (fy + 1 == f1 ? 0 : fy + 1) * s1};
  const size_t zm = static_cast<size_t>(fz == 0 ? f2 - 1 : fz - 1);
  const bool first = threadIdx.x == 0;
  float w0[3][3][3], w1[3][3][3];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      const float* row = u + xo[a] + yo[b];
      const float4 q = *reinterpret_cast<const float4*>(row + fz);
      const float left = __shfl_up(q.w, 1);
      w0[a][b][0] = first ? row[zm] : left;
      w0[a][b][1] = q.x;
      w0[a][b][2] = q.y;
      w1[a][b][0] = q.y;
      w1[a][b][1] = q.z;
      w1[a][b][2] = q.w;
    }
  // fine z = 4 jq + 3 is the last of the quad and never wraps (f2 % 4 == 0)
  if (in) {
    VO out;
    out.x = static_cast<RO>(restrict_sum<float>([&](int a, int b, int k) { return w0[a][b][k]; }));
    out.y = static_cast<RO>(restrict_sum<float>([&](int a, int b, int k) { return w1[a][b][k]; }));
    *reinterpret_cast<VO*>(c + (static_cast<size_t>(jx) * c1 + jy) * c2 + 2 * jq) = out;
  }
}

// Trilinear prolongation from the coarse mesh (c0, c1, c2) to the fine mesh,
// one thread per fine cell (src/multigrid.c:177-205), either overwriting the
// fine value or adding to it (the V-cycle correction, src/multigrid.c:296).
template <typename RC, typename R, bool ADD, typename I>
__device__ __forceinline__ void prolong_at(const RC* __restrict__ uc, R* __restrict__ fine,
                                           int c0, int c1, int c2, I i) {
  const int F1 = 2 * c1, F2 = 2 * c2;
  int X, Y, Z;
  unflatten(i, F1, F2, X, Y, Z);
  const int jx = X >> 1, jy = Y >> 1, jz = Z >> 1;
  const size_t s0 = static_cast<size_t>(c1) * c2, s1 = static_cast<size_t>(c2);
  const size_t x0 = jx * s0, xp = (jx + 1 == c0 ? 0 : jx + 1) * s0;
  const size_t y0 = jy * s1, yp = (jy + 1 == c1 ? 0 : jy + 1) * s1;
  const size_t z0 = jz, zp = (jz + 1 == c2 ? 0 : jz + 1);
  auto u = [uc](size_t idx) { return static_cast<R>(uc[idx]); };
  const R a = u(x0 + y0 + z0);
  R val;
  switch (((X & 1) << 2) | ((Y & 1) << 1) | (Z & 1)) {
    case 0: val = a; break;
    case 4: val = (a + u(xp + y0 + z0)) / R(2); break;
    case 2: val = (a + u(x0 + yp + z0)) / R(2); break;
    case 1: val = (a + u(x0 + y0 + zp)) / R(2); break;
    case 6: val = (a + u(xp + y0 + z0) + u(x0 + yp + z0) + u(xp + yp + z0)) / R(4); break;
    case 3: val = (a + u(x0 + yp + z0) + u(x0 + y0 + zp) + u(x0 + yp + zp)) / R(4); break;
    case 5: val = (a + u(xp + y0 + z0) + u(x0 + y0 + zp) + u(xp + y0 + zp)) / R(4); break;
    default:
      val = (a + u(xp + y0 + z0) + u(x0 + yp + z0) + u(x0 + y0 + zp) + u(xp + yp + z0) +
             u(xp + y0 + zp) + u(x0 + yp + zp) + u(xp + yp + zp)) /
            R(8);
      break;
  }
  if (ADD)
    fine[i] += val;
  else
    fine[i] = val;
}

template <typename RC, typename R, bool ADD>
__device__ __forceinline__ void prolong_body(const RC* __restrict__ uc, R* __restrict__ fine,
                                             int c0, int c1, int c2) {
  const size_t total = static_cast<size_t>(2 * c0) * (2 * c1) * (2 * c2);
  const size_t i = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x;
  if (i < total) prolong_at<RC, R, ADD>(uc, fine, c0, c1, c2, i);
}

template <typename RC, typename R, bool ADD>
__global__ __launch_bounds__(kThreads) void k_prolong(const RC* __restrict__ uc,
                                                      R* __restrict__ fine, int c0, int c1,
                                                      int c2) {
  prolong_body<RC, R, ADD>(uc, fine, c0, c1, c2);
}

// The same prolongation with one thread per fine (X, Y) row and coarse z:
// the thread writes fine cells Z = 2 jz and 2 jz + 1 as one vector (fine
// rows have even length), from the coarse values that both share.  The
// parities of X and Y are uniform across a wavefront when c2 >= 64; each
// value is formed with the operands and order of prolong_at().
template <typename RC, typename R, bool ADD>
__global__ __launch_bounds__(kThreads) void k_prolong_rows(const RC* __restrict__ uc,
                                                           R* __restrict__ fine, int c0, int c1,
                                                           int c2) {
  const int F1 = 2 * c1, F2 = 2 * c2;
  const int jz = blockIdx.x * blockDim.x + threadIdx.x;
  const int Y = blockIdx.y * blockDim.y + threadIdx.y;
  const int X = blockIdx.z;
  if (jz >= c2 || Y >= F1) return;
  const int jx = X >> 1, jy = Y >> 1;
  const size_t s0 = static_cast<size_t>(c1) * c2, s1 = static_cast<size_t>(c2);
  const size_t x0 = jx * s0, xp = (jx + 1 == c0 ? 0 : jx + 1) * s0;
  const size_t y0 = jy * s1, yp = (jy + 1 == c1 ? 0 : jy + 1) * s1;
  const size_t z0 = jz, zp = (jz + 1 == c2 ? 0 : jz + 1);
  auto u = [uc](size_t idx) { return static_cast<R>(uc[idx]); };
  const R a = u(x0 + y0 + z0), az = u(x0 + y0 + zp);
  R even, odd;
  if (X & 1) {
    const R ax = u(xp + y0 + z0), axz = u(xp + y0 + zp);
    if (Y & 1) {
      const R ay = u(x0 + yp + z0), ayz = u(x0 + yp + zp);
      const R axy = u(xp + yp + z0), axyz = u(xp + yp + zp);
      even = (a + ax + ay + axy) / R(4);
      odd = (a + ax + ay + az + axy + axz + ayz + axyz) / R(8);
    } else {
      even = (a + ax) / R(2);
      odd = (a + ax + az + axz) / R(4);
    }
  } else if (Y & 1) {
    const R ay = u(x0 + yp + z0), ayz = u(x0 + yp + zp);
    even = (a + ay) / R(2);
    odd = (a + ay + az + ayz) / R(4);
  } else {
    even = a;
    odd = (a + az) / R(2);
  }
  using V = typename Pair<R>::type;
  V* dst = reinterpret_cast<V*>(fine + (static_cast<size_t>(X) * F1 + Y) * F2 + 2 * jz);
  V w;
  if (ADD) {
    w = *dst;
    w.x += even;
    w.y += odd;
  } else {
    w.x = even;
    w.y = odd;
  }
  *dst = w;
}

// FP32 prolongation with two coarse z per thread: fine z = 4 jq .. 4 jq + 3
// (the even and odd cells of coarse z = 2 jq and 2 jq + 1) written as one
// float4; each value is formed with the operands and order of prolong_at().
// Requires c2 even and the fine mesh on a 16-byte boundary.
template <typename RC, bool ADD>
__global__ __launch_bounds__(kThreads) void k_prolong_quads(const RC* __restrict__ uc,
                                                            float* __restrict__ fine, int c0,
                                                            int c1, int c2) {
  using R = float;
  const int F1 = 2 * c1, F2 = 2 * c2;
  const int jq = blockIdx.x * blockDim.x + threadIdx.x;
  const int Y = blockIdx.y * blockDim.y + threadIdx.y;
  const int X = blockIdx.z;
  if (2 * jq >= c2 || Y >= F1) return;
  const int jx = X >> 1, jy = Y >> 1;
  const size_t s0 = static_cast<size_t>(c1) * c2, s1 = static_cast<size_t>(c2);
  const size_t x0 = jx * s0, xp = (jx + 1 == c0 ? 0 : jx + 1) * s0;
  const size_t y0 = jy * s1, yp = (jy + 1 == c1 ? 0 : jy + 1) * s1;
  const size_t z0 = 2 * jq, z1 = z0 + 1, z2 = (z1 + 1 == static_cast<size_t>(c2) ? 0 : z1 + 1);
  auto u = [uc](size_t idx) { return static_cast<R>(uc[idx]); };
  // even and odd fine values of coarse z = zc, whose next coarse z is zn
  auto cell = [&](size_t zc, size_t zn, R& even, R& odd) {
    const R a = u(x0 + y0 + zc), az = u(x0 + y0 + zn);
    if (X & 1) {
      const R ax = u(xp + y0 + zc), axz = u(xp + y0 + zn);
      if (Y & 1) {
        const R ay = u(x0 + yp + zc), ayz = u(x0 + yp + zn);
        const R axy = u(xp + yp + zc), axyz = u(xp + yp + zn);
        even = (a + ax + ay + axy) / R(4);
        odd = (a + ax + ay + az + axy + axz + ayz + axyz) / R(8);
      } else {
        even = (a + ax) / R(2);
        odd = (a + ax + az + axz) / R(4);
      }
    } else if (Y & 1) {
      const R ay = u(x0 + yp + zc), ayz = u(x0 + yp + zn);
      even = (a + ay) / R(2);
      odd = (a + ay + az + ayz) / R(4);
    } else {
      even = a;
      odd = (a + az) / R(2);
    }
  };
  R e0, o0, e1, o1;
  cell(z0, z1, e0, o0);
  cell(z1, z2, e1, o1);
  float4* dst = reinterpret_cast<float4*>(fine + (static_cast<size_t>(X) * F1 + Y) * F2 + 4 * jq);
  float4 w;
  if (ADD) {
    w = *dst;
    w.x += e0;
    w.y += o0;
    w.z += e1;
    w.w += o1;
  } else {
    w = make_float4(e0, o0, e1, o1);
  }
  *dst = w;
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

// Block partial sums of an FP64 mesh (deterministic, see prh_reduce.hpp).
__global__ __launch_bounds__(kReduceThreads) void k_sum_f64(const double* __restrict__ a,
                                                             size_t n, double* __restrict__ partial) {
  block_reduce_range<1>(n, [&](size_t i, double* out) { out[0] = a[i]; }, partial, 0);
}

// ---------------------------------------------------------------------------
// Grouped launches: one launch applies the same operation at the same level
// shape to up to kMaxGroup independent problems (serving several
// reconstructions at once), each with its own buffers and geometry.  The job
// index is folded into the grid; per cell the arithmetic is the single-job
// kernels', so each job's bits do not depend on its company.
// ---------------------------------------------------------------------------
constexpr int kMaxGroup = 8;

template <typename RI, typename RO, typename RG>
struct GroupArgs {
  const RI* in[kMaxGroup];
  const RI* src[kMaxGroup];
  RO* out[kMaxGroup];
  LevelGeom<RG> G[kMaxGroup];
};

template <typename R, bool ZERO_IN>
__global__ __launch_bounds__(kThreads) void k_jacobi_group(GroupArgs<R, R, R> A, R omega, R keep,
                                                           int n0) {
  const int job = blockIdx.z / n0;
  const int ix = blockIdx.z - job * n0;
  const LevelGeom<R>& G = A.G[job];
  const int iz = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y * blockDim.y + threadIdx.y;
  if (iz >= G.n[2] || iy >= G.n[1]) return;
  const size_t i = (static_cast<size_t>(ix) * G.n[1] + iy) * G.n[2] + iz;
  const Coef<R> k = cell_coef(G, ix, iy, iz);
  const Hood<R> h =
      ZERO_IN ? zero_hood<R>() : gather_hood(A.in[job], G.n[0], G.n[1], G.n[2], ix, iy, iz);
  A.out[job][i] = jacobi_update(k, h, A.src[job][i], omega, keep, G.local);
}

template <typename R>
__global__ __launch_bounds__(kThreads) void k_residual_group(GroupArgs<R, R, R> A, int n0) {
  const int job = blockIdx.z / n0;
  const int ix = blockIdx.z - job * n0;
  const LevelGeom<R>& G = A.G[job];
  const int iz = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y * blockDim.y + threadIdx.y;
  if (iz >= G.n[2] || iy >= G.n[1]) return;
  const size_t i = (static_cast<size_t>(ix) * G.n[1] + iy) * G.n[2] + iz;
  const Coef<R> k = cell_coef(G, ix, iy, iz);
  const Hood<R> h = gather_hood(A.in[job], G.n[0], G.n[1], G.n[2], ix, iy, iz);
  A.out[job][i] = residual_value(k, h, A.src[job][i], G.local);
}

// restriction / prolongation / zero fill of every job: blockIdx.y is the job
template <typename R>
__global__ __launch_bounds__(kThreads) void k_restrict_group(GroupArgs<R, R, R> A, int f0, int f1,
                                                             int f2) {
  restrict_body<R, R>(A.in[blockIdx.y], A.out[blockIdx.y], f0, f1, f2, Centre{});
}

template <typename R, bool ADD>
__global__ __launch_bounds__(kThreads) void k_prolong_group(GroupArgs<R, R, R> A, int c0, int c1,
                                                            int c2) {
  prolong_body<R, R, ADD>(A.in[blockIdx.y], A.out[blockIdx.y], c0, c1, c2);
}

template <typename R>
__global__ __launch_bounds__(kThreads) void k_fill_zero_group(GroupArgs<R, R, R> A, size_t n) {
  const size_t i = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x;
  if (i < n) A.out[blockIdx.y][i] = R(0);
}

// ---------------------------------------------------------------------------
// Coarse levels in one workgroup.  Below kCoarseCells cells the launches of a
// level are latency, not work: a V-cycle from such a level down to the
// coarsest runs in one 1024-thread workgroup per problem, every step sweeping
// all cells of its level and ending with a barrier (the levels sit in L2).
// Per cell the arithmetic is the generic kernels', so the bits are the same.
// ---------------------------------------------------------------------------
constexpr int kCoarseThreads = 1024;
constexpr size_t kCoarseCells = 4096;
constexpr int kMaxCoarse = 4;
constexpr int kCoarseGroup = 4;

template <typename R>
struct CoarseLevel {
  LevelGeom<R> G;
  R* v;
  R* f;
  R* tmp;
};

template <typename R>
struct CoarseArgs {
  int nlev, niter, zero_start;
  R omega, keep;
  CoarseLevel<R> L[kCoarseGroup][kMaxCoarse];  // relative level 0 is the V-cycle's top level
  R* cur0[kCoarseGroup];
  R* alt0[kCoarseGroup];
  const R* src0[kCoarseGroup];
};

// All single-workgroup levels live in LDS during the call: iterate, its
// ping-pong partner, source and the per-cell g of each level (at most
// 4096 (1 + 1/8 + 1/64 + 1/512) cells: 150 KB in FP64).
constexpr int kCoarseLds = 4682;

// The cells of one level that a thread visits (i = threadIdx.x + kCoarseThreads u),
// with their coordinates and the offsets of their periodic neighbours, worked
// out once per visit of the level rather than once per sweep.  Offsets are
// below the level's size in magnitude, so they pack in pairs of 16 bits.
constexpr int kCoarsePer = static_cast<int>((kCoarseCells + kCoarseThreads - 1) / kCoarseThreads);
static_assert(kCoarseCells <= 32767, "neighbour offsets are packed in 16 bits");
struct CoarseCells {
  int total;
  int xy[kCoarsePer];  // ix | iy << 16
  int z[kCoarsePer];
  int ox[kCoarsePer];  // offset of x + 1 in the low half, of x - 1 in the high half
  int oy[kCoarsePer];
  int oz[kCoarsePer];
};

__device__ __forceinline__ int pack16(int lo, int hi) { return (lo & 0xffff) | (hi << 16); }
__device__ __forceinline__ int lo16(int p) { return static_cast<int>(static_cast<int16_t>(p & 0xffff)); }
__device__ __forceinline__ int hi16(int p) { return p >> 16; }

template <typename R>
__device__ __forceinline__ CoarseCells coarse_cells(const LevelGeom<R>& G) {
  const int n0 = G.n[0], n1 = G.n[1], n2 = G.n[2], s0 = n1 * n2;
  CoarseCells C;
  C.total = n0 * s0;
#pragma unroll
  for (int u = 0; u < kCoarsePer; u++) {
    const int i = threadIdx.x + kCoarseThreads * u;
    int ix = 0, iy = 0, iz = 0;
    if (i < C.total) unflatten(i, n1, n2, ix, iy, iz);
    const int xp = ix + 1 == n0 ? 0 : ix + 1, xm = ix == 0 ? n0 - 1 : ix - 1;
    const int yp = iy + 1 == n1 ? 0 : iy + 1, ym = iy == 0 ? n1 - 1 : iy - 1;
    const int zp = iz + 1 == n2 ? 0 : iz + 1, zm = iz == 0 ? n2 - 1 : iz - 1;
    C.xy[u] = ix | (iy << 16);
    C.z[u] = iz;
    C.ox[u] = pack16((xp - ix) * s0, (xm - ix) * s0);
    C.oy[u] = pack16((yp - iy) * n2, (ym - iy) * n2);
    C.oz[u] = pack16(zp - iz, zm - iz);
  }
  return C;
}

// gather_hood() of cell u of C (the same neighbours)
template <typename R>
__device__ __forceinline__ Hood<R> coarse_hood(const R* v, const CoarseCells& C, int u, int i) {
  const int xp = lo16(C.ox[u]), xm = hi16(C.ox[u]), yp = lo16(C.oy[u]), ym = hi16(C.oy[u]);
  const int zp = lo16(C.oz[u]), zm = hi16(C.oz[u]);
  const R* c = v + i;
  Hood<R> h;
  h.c = c[0];
  h.xp = c[xp];
  h.xm = c[xm];
  h.yp = c[yp];
  h.ym = c[ym];
  h.zp = c[zp];
  h.zm = c[zm];
  h.xpyp = c[xp + yp];
  h.xmym = c[xm + ym];
  h.xmyp = c[xm + yp];
  h.xpym = c[xp + ym];
  h.xpzp = c[xp + zp];
  h.xmzm = c[xm + zm];
  h.xmzp = c[xm + zp];
  h.xpzm = c[xp + zm];
  h.ypzp = c[yp + zp];
  h.ymzm = c[ym + zm];
  h.ymzp = c[ym + zp];
  h.ypzm = c[yp + zm];
  return h;
}

template <typename R>
__device__ __forceinline__ Coef<R> coarse_coef(const LevelGeom<R>& G, const CoarseCells& C, int u,
                                               const R* gtab, int i) {
  return cell_coef_g(G, C.xy[u] & 0xffff, C.xy[u] >> 16, C.z[u], gtab[i]);
}

template <typename R>
__device__ __forceinline__ void cw_jacobi(const LevelGeom<R>& G, const CoarseCells& C, const R* in,
                                          const R* src, R* out, const R* gtab, R omega, R keep,
                                          bool zero_in) {
#pragma unroll
  for (int u = 0; u < kCoarsePer; u++) {
    const int i = threadIdx.x + kCoarseThreads * u;
    if (i < C.total) {
      const Coef<R> k = coarse_coef(G, C, u, gtab, i);
      const Hood<R> h = zero_in ? zero_hood<R>() : coarse_hood(in, C, u, i);
      out[i] = jacobi_update(k, h, src[i], omega, keep, G.local);
    }
  }
  __syncthreads();
}

template <typename R>
__device__ __forceinline__ void cw_smooth(const LevelGeom<R>& G, const CoarseCells& C, R*& cur,
                                          R*& alt, const R* src, const R* gtab, int niter,
                                          R omega, R keep, bool zero_start) {
  if (niter <= 0) {
    if (zero_start) {
      for (int i = threadIdx.x; i < C.total; i += kCoarseThreads) cur[i] = R(0);
      __syncthreads();
    }
    return;
  }
  for (int it = 0; it < niter; it++) {
    cw_jacobi(G, C, cur, src, alt, gtab, omega, keep, zero_start && it == 0);
    R* t = cur;
    cur = alt;
    alt = t;
  }
}

// src/multigrid.c:270-300 from relative level 0 to the coarsest, unrolled over
// the (at most kMaxCoarse) levels.  Level 0's iterate and source are read from
// global memory at the start and its iterate written back at the end; every
// other step stays in LDS.  Per cell the arithmetic is the generic kernels'
// (g computed once per call by cell_coef() itself), so the bits are the same.
template <typename R>
__global__ __launch_bounds__(kCoarseThreads) void k_vcycle_coarse(CoarseArgs<R> A) {
  __shared__ R sv[kCoarseLds], st[kCoarseLds], sf[kCoarseLds], sg[kCoarseLds];
  const int job = blockIdx.x;
  const int nl = A.nlev;
  int off[kMaxCoarse];
  {
    int o = 0;
#pragma unroll
    for (int m = 0; m < kMaxCoarse; m++) {
      off[m] = o;
      if (m < nl) {
        const LevelGeom<R>& G = A.L[job][m].G;
        o += G.n[0] * G.n[1] * G.n[2];
      }
    }
  }
  {
    const LevelGeom<R>& G = A.L[job][0].G;
    const int total = G.n[0] * G.n[1] * G.n[2];
    const R* c0 = A.cur0[job];
    const R* f0 = A.src0[job];
    for (int i = threadIdx.x; i < total; i += kCoarseThreads) {
      sf[i] = f0[i];
      if (!A.zero_start) sv[i] = c0[i];
    }
  }
#pragma unroll
  for (int m = 0; m < kMaxCoarse; m++) {
    if (m >= nl) break;
    const LevelGeom<R>& G = A.L[job][m].G;
    const int total = G.n[0] * G.n[1] * G.n[2];
    for (int i = threadIdx.x; i < total; i += kCoarseThreads) {
      int ix, iy, iz;
      unflatten(i, G.n[1], G.n[2], ix, iy, iz);
      sg[off[m] + i] = cell_coef(G, ix, iy, iz).g;
    }
  }
  __syncthreads();
  R* cur[kMaxCoarse];
  R* alt[kMaxCoarse];
#pragma unroll
  for (int m = 0; m < kMaxCoarse; m++) {
    cur[m] = sv + off[m];
    alt[m] = st + off[m];
  }
#pragma unroll
  for (int m = 0; m < kMaxCoarse; m++) {
    if (m >= nl) break;
    const LevelGeom<R>& G = A.L[job][m].G;
    const R* src = sf + off[m];
    const R* gt = sg + off[m];
    const CoarseCells C = coarse_cells(G);
    cw_smooth(G, C, cur[m], alt[m], src, gt, A.niter, A.omega, A.keep, m > 0 || A.zero_start);
    if (m + 1 < nl) {
#pragma unroll
      for (int u = 0; u < kCoarsePer; u++) {
        const int i = threadIdx.x + kCoarseThreads * u;
        if (i < C.total) {
          const Coef<R> k = coarse_coef(G, C, u, gt, i);
          const Hood<R> h = coarse_hood(cur[m], C, u, i);
          alt[m][i] = residual_value(k, h, src[i], G.local);
        }
      }
      __syncthreads();
      const int ctotal = (G.n[0] / 2) * (G.n[1] / 2) * (G.n[2] / 2);
      R* fc = sf + off[m + 1];
      for (int j = threadIdx.x; j < ctotal; j += kCoarseThreads)
        restrict_at<R, R>(alt[m], fc, G.n[0], G.n[1], G.n[2], j);
      __syncthreads();
    }
  }
#pragma unroll
  for (int m = kMaxCoarse - 1; m >= 0; m--) {
    if (m >= nl) continue;
    const LevelGeom<R>& G = A.L[job][m].G;
    if (m + 1 < nl) {
      const LevelGeom<R>& Gc = A.L[job][m + 1].G;
      const int ftotal = 8 * Gc.n[0] * Gc.n[1] * Gc.n[2];
      for (int i = threadIdx.x; i < ftotal; i += kCoarseThreads)
        prolong_at<R, R, true>(cur[m + 1], cur[m], Gc.n[0], Gc.n[1], Gc.n[2], i);
      __syncthreads();
    }
    cw_smooth(G, coarse_cells(G), cur[m], alt[m], sf + off[m], sg + off[m], A.niter, A.omega, A.keep,
              false);
  }
  {
    const LevelGeom<R>& G = A.L[job][0].G;
    const int total = G.n[0] * G.n[1] * G.n[2];
    R* c0 = A.cur0[job];
    for (int i = threadIdx.x; i < total; i += kCoarseThreads) c0[i] = cur[0][i];
  }
}

// ---------------------------------------------------------------------------
// Host-side drivers
// ---------------------------------------------------------------------------

static size_t ncells(const size_t* n) { return n[0] * n[1] * n[2]; }

template <typename R>
// acc (FP32 fast sweeps on blocked levels only): add the result into this FP64
// mesh instead of storing it; *fused tells whether that was done.
static int launch_jacobi(const R* v, const R* f, R* out, const LevelGeom<R>& G, R omega, R keep,
                         bool zero_in, hipStream_t s, int timer = -1, bool fast = false,
                         double* acc = nullptr, bool* fused = nullptr) {
  const size_t total = static_cast<size_t>(G.n[0]) * G.n[1] * G.n[2];
  dim3 block, grid;
  generic_shape(G.n, block, grid);
  if (fused) *fused = false;
  // a sweep added into the FP64 iterate is another kernel (kOpJacobiAcc): not in the timer
  const bool into_acc = acc && !zero_in && tiled_ok(G.n) && fast && std::is_same<R, float>::value;
  if (into_acc) timer = -1;
  timer_before(timer, s);
  if (into_acc) {
    launch_tiled<R, R, kOpJacobiAcc, std::is_same<R, float>::value>(v, f, out, G, omega, keep, s, acc);
    if (fused) *fused = true;
  } else if (zero_in && tiled_ok(G.n) && fast)
    launch_tiled<R, R, kOpJacobiZero, std::is_same<R, float>::value>(v, f, out, G, omega, keep, s);
  else if (zero_in && tiled_ok(G.n))
    launch_tiled<R, R, kOpJacobiZero, false>(v, f, out, G, omega, keep, s);
  else if (zero_in)
    k_jacobi<R, true><<<grid, block, 0, s>>>(v, f, out, G, omega, keep);
  else if (tiled_ok(G.n) && fast)
    launch_tiled<R, R, kOpJacobi, std::is_same<R, float>::value>(v, f, out, G, omega, keep, s);
  else if (tiled_ok(G.n))
    launch_tiled<R, R, kOpJacobi, false>(v, f, out, G, omega, keep, s);
  else
    k_jacobi<R, false><<<grid, block, 0, s>>>(v, f, out, G, omega, keep);
  // algorithmic traffic: read v (unless zero) and f, write the new iterate
  timer_after(timer, s, static_cast<double>(total) * sizeof(R) * (zero_in ? 2 : 3));
  PRH_LAUNCHED();
  return PRH_OK;
}

template <typename R, typename O>
static int launch_residual(const R* v, const R* f, O* r, const LevelGeom<R>& G, hipStream_t s,
                           bool fast = false, Centre C = {}) {
  if (tiled_ok(G.n) && fast) {
    launch_tiled<R, O, kOpResidual, std::is_same<R, float>::value && std::is_same<O, float>::value>(
        v, f, r, G, R(0), R(0), s, nullptr, C);
  } else if (tiled_ok(G.n)) {
    launch_tiled<R, O, kOpResidual, false>(v, f, r, G, R(0), R(0), s, nullptr, C);
  } else {
    dim3 block, grid;
    generic_shape(G.n, block, grid);
    k_residual<R, O><<<grid, block, 0, s>>>(v, f, r, G, C);
  }
  PRH_LAUNCHED();
  return PRH_OK;
}

template <typename R, typename RO>
static int launch_restrict(const R* fine, RO* coarse, const size_t* nf, hipStream_t s, Centre C = {}) {
  if (nf[2] % 2 != 0 || reinterpret_cast<uintptr_t>(fine) % (2 * sizeof(R)) != 0) {
    // rows that do not start on pair boundaries: one thread per coarse cell
    const size_t total = (nf[0] / 2) * (nf[1] / 2) * (nf[2] / 2);
    k_restrict<R, RO><<<ceil_div(total, kThreads), kThreads, 0, s>>>(
        fine, coarse, static_cast<int>(nf[0]), static_cast<int>(nf[1]), static_cast<int>(nf[2]), C);
    PRH_LAUNCHED();
    return PRH_OK;
  }
  const int nc[3] = {static_cast<int>(nf[0] / 2), static_cast<int>(nf[1] / 2),
                     static_cast<int>(nf[2] / 2)};
  dim3 block, grid;
  if constexpr (std::is_same<R, float>::value) {
    if (!C.sum && PRH_TRANSFER_QUADS && nf[2] % 4 == 0 && reinterpret_cast<uintptr_t>(fine) % 16 == 0 &&
        reinterpret_cast<uintptr_t>(coarse) % (2 * sizeof(RO)) == 0) {
      const int nq[3] = {nc[0], nc[1], nc[2] / 2};
      generic_shape(nq, block, grid);
      k_restrict_quads<RO><<<grid, block, 0, s>>>(fine, coarse, static_cast<int>(nf[0]),
                                                  static_cast<int>(nf[1]), static_cast<int>(nf[2]));
      PRH_LAUNCHED();
      return PRH_OK;
    }
  }
  generic_shape(nc, block, grid);
  k_restrict_rows<R, RO><<<grid, block, 0, s>>>(fine, coarse, static_cast<int>(nf[0]),
                                                static_cast<int>(nf[1]), static_cast<int>(nf[2]), C);
  PRH_LAUNCHED();
  return PRH_OK;
}

template <typename RC, typename R>
static int launch_prolong(const RC* coarse, R* fine, const size_t* nc, bool add, hipStream_t s) {
  const int c0 = static_cast<int>(nc[0]), c1 = static_cast<int>(nc[1]),
            c2 = static_cast<int>(nc[2]);
  if (reinterpret_cast<uintptr_t>(fine) % (2 * sizeof(R)) != 0) {
    // a fine mesh that does not start on a pair boundary: one thread per cell
    const size_t total = 8 * ncells(nc);
    if (add)
      k_prolong<RC, R, true><<<ceil_div(total, kThreads), kThreads, 0, s>>>(coarse, fine, c0, c1, c2);
    else
      k_prolong<RC, R, false><<<ceil_div(total, kThreads), kThreads, 0, s>>>(coarse, fine, c0, c1, c2);
    PRH_LAUNCHED();
    return PRH_OK;
  }
  dim3 block, grid;
  if constexpr (std::is_same<R, float>::value) {
    if (PRH_TRANSFER_QUADS && c2 % 2 == 0 && reinterpret_cast<uintptr_t>(fine) % 16 == 0) {
      // threads over (coarse z pair, fine y, fine x)
      const int shape[3] = {2 * c0, 2 * c1, c2 / 2};
      generic_shape(shape, block, grid);
      if (add)
        k_prolong_quads<RC, true><<<grid, block, 0, s>>>(coarse, fine, c0, c1, c2);
      else
        k_prolong_quads<RC, false><<<grid, block, 0, s>>>(coarse, fine, c0, c1, c2);
      PRH_LAUNCHED();
      return PRH_OK;
    }
  }
  // threads over (coarse z, fine y, fine x)
  const int shape[3] = {2 * c0, 2 * c1, c2};
  generic_shape(shape, block, grid);
  if (add)
    k_prolong_rows<RC, R, true><<<grid, block, 0, s>>>(coarse, fine, c0, c1, c2);
  else
    k_prolong_rows<RC, R, false><<<grid, block, 0, s>>>(coarse, fine, c0, c1, c2);
  PRH_LAUNCHED();
  return PRH_OK;
}

template <typename R>
static int launch_fill_zero(R* a, size_t n, hipStream_t s) {
  k_fill_zero<R><<<ceil_div(n, kThreads), kThreads, 0, s>>>(a, n);
  PRH_LAUNCHED();
  return PRH_OK;
}

// Per-job buffers of a (grouped) solve.
template <typename R>
struct Ptrs {
  R* p[kMaxGroup];
  R*& operator[](int j) { return p[j]; }
  R* operator[](int j) const { return p[j]; }
};

template <typename R>
static void swap_all(Ptrs<R>& a, Ptrs<R>& b, int count) {
  for (int j = 0; j < count; j++) std::swap(a.p[j], b.p[j]);
}

// State of one solve over `count` hierarchies of the same shape (count = 1
// for the reference-shaped entry points).  Every level runs the schedule of
// src/multigrid.c:270-328 for all jobs in lock step.  Levels that the blocked
// kernels cover are launched once per job (they fill the device alone); the
// smaller levels, which the generic kernels handle and whose launches are
// latency-bound, are launched once for the whole group.
template <typename R>
struct Solver {
  int count = 1;
  Hierarchy* H[kMaxGroup];
  LevelGeom<R> G[kMaxGroup][kMaxLevels];
  R omega, keep;
  int niter;
  hipStream_t s;
  bool fast = false;  // fast FP32 kernels on the blocked levels (mixed-precision corrections)
  bool pair_enabled = true;  // with fast: two sweeps per launch where the level allows
  // When fuse_top is set, the last post-smoothing sweep of level 0 adds its result into the FP64
  // mesh acc_top[j] instead of storing it; fused[j] tells whether it could (blocked levels).
  bool fuse_top = false;
  double* acc_top[kMaxGroup] = {};
  bool fused[kMaxGroup] = {};

  int nlev() const { return H[0]->nlev; }
  const size_t* n(int l) const { return H[0]->n[l]; }
  size_t size(int l) const { return ncells(H[0]->n[l]); }
  Ptrs<R> v(int l) const { return level_ptrs(&Hierarchy::v, l); }
  Ptrs<R> f(int l) const { return level_ptrs(&Hierarchy::f, l); }
  Ptrs<R> tmp(int l) const { return level_ptrs(&Hierarchy::tmp, l); }
  Ptrs<R> level_ptrs(void* (Hierarchy::*m)[kMaxLevels], int l) const {
    Ptrs<R> r;
    for (int j = 0; j < count; j++) r.p[j] = static_cast<R*>((H[j]->*m)[l]);
    return r;
  }
  bool grouped(int l) const {
    int nn[3] = {static_cast<int>(n(l)[0]), static_cast<int>(n(l)[1]), static_cast<int>(n(l)[2])};
    return count > 1 && !tiled_ok(nn);
  }

  void setup(Hierarchy* h, const R* boxsize, const R* boxcenter, R beta, R damping, int nit,
             const R* los, hipStream_t stream) {
    count = 1;
    H[0] = h;
    geometry(0, boxsize, boxcenter, beta, los);
    params(damping, nit, stream);
  }
  void geometry(int j, const R* boxsize, const R* boxcenter, R beta, const R* los) {
    for (int l = 0; l < H[j]->nlev; l++)
      G[j][l] = LevelGeom<R>::make(H[j]->n[l], boxsize, boxcenter, beta, los);
  }
  void params(R damping, int nit, hipStream_t stream) {
    omega = damping;
    keep = R(1) - damping;
    niter = nit;
    s = stream;
  }

  GroupArgs<R, R, R> args(int l, const Ptrs<R>& in, const Ptrs<R>& src, const Ptrs<R>& out) const {
    GroupArgs<R, R, R> A;
    for (int j = 0; j < kMaxGroup; j++) {
      const int q = j < count ? j : 0;
      A.in[j] = in.p[q];
      A.src[j] = src.p[q];
      A.out[j] = out.p[q];
      A.G[j] = G[q][l];
    }
    return A;
  }

  int jacobi(int l, const Ptrs<R>& in, const Ptrs<R>& src, const Ptrs<R>& out, bool zero_in,
             bool fuse = false) {
    if (fuse)
      for (int j = 0; j < count; j++) fused[j] = false;
    if (grouped(l)) {
      const int nn[3] = {static_cast<int>(n(l)[0]), static_cast<int>(n(l)[1]),
                         static_cast<int>(n(l)[2])};
      dim3 block, grid;
      generic_shape(nn, block, grid);
      grid.z *= count;
      if (zero_in)
        k_jacobi_group<R, true><<<grid, block, 0, s>>>(args(l, in, src, out), omega, keep, nn[0]);
      else
        k_jacobi_group<R, false><<<grid, block, 0, s>>>(args(l, in, src, out), omega, keep, nn[0]);
      PRH_LAUNCHED();
      return PRH_OK;
    }
    for (int j = 0; j < count; j++)
      PRH_TRY(launch_jacobi(in[j], src[j], out[j], G[j][l], omega, keep, zero_in, s,
                            l == 0 ? PRH_TIMER_JACOBI_FINEST : -1, fast, fuse ? acc_top[j] : nullptr,
                            fuse ? &fused[j] : nullptr));
    return PRH_OK;
  }

  int residual(int l, const Ptrs<R>& in, const Ptrs<R>& src, const Ptrs<R>& out) {
    if (grouped(l)) {
      const int nn[3] = {static_cast<int>(n(l)[0]), static_cast<int>(n(l)[1]),
                         static_cast<int>(n(l)[2])};
      dim3 block, grid;
      generic_shape(nn, block, grid);
      grid.z *= count;
      k_residual_group<R><<<grid, block, 0, s>>>(args(l, in, src, out), nn[0]);
      PRH_LAUNCHED();
      return PRH_OK;
    }
    for (int j = 0; j < count; j++)
      PRH_TRY((launch_residual<R, R>(in[j], src[j], out[j], G[j][l], s, fast)));
    return PRH_OK;
  }

  // fine level l -> coarse level l + 1
  int restrict_to(int l, const Ptrs<R>& fine, const Ptrs<R>& coarse) {
    if (grouped(l + 1)) {
      const size_t total = size(l + 1);
      const dim3 grid(ceil_div(total, kThreads), count);
      k_restrict_group<R><<<grid, kThreads, 0, s>>>(
          args(l + 1, fine, fine, coarse), static_cast<int>(n(l)[0]), static_cast<int>(n(l)[1]),
          static_cast<int>(n(l)[2]));
      PRH_LAUNCHED();
      return PRH_OK;
    }
    for (int j = 0; j < count; j++) PRH_TRY((launch_restrict<R, R>(fine[j], coarse[j], n(l), s)));
    return PRH_OK;
  }

  // coarse level l + 1 -> fine level l, added to or overwriting the fine values
  int prolong_to(int l, const Ptrs<R>& coarse, const Ptrs<R>& fine, bool add) {
    if (grouped(l + 1)) {
      const size_t total = 8 * size(l + 1);
      const dim3 grid(ceil_div(total, kThreads), count);
      const int c0 = static_cast<int>(n(l + 1)[0]), c1 = static_cast<int>(n(l + 1)[1]),
                c2 = static_cast<int>(n(l + 1)[2]);
      if (add)
        k_prolong_group<R, true><<<grid, kThreads, 0, s>>>(args(l + 1, coarse, coarse, fine), c0, c1, c2);
      else
        k_prolong_group<R, false><<<grid, kThreads, 0, s>>>(args(l + 1, coarse, coarse, fine), c0, c1, c2);
      PRH_LAUNCHED();
      return PRH_OK;
    }
    for (int j = 0; j < count; j++) PRH_TRY((launch_prolong<R, R>(coarse[j], fine[j], n(l + 1), add, s)));
    return PRH_OK;
  }

  int fill_zero(int l, const Ptrs<R>& a) {
    if (grouped(l)) {
      const dim3 grid(ceil_div(size(l), kThreads), count);
      k_fill_zero_group<R><<<grid, kThreads, 0, s>>>(args(l, a, a, a), size(l));
      PRH_LAUNCHED();
      return PRH_OK;
    }
    for (int j = 0; j < count; j++) PRH_TRY(launch_fill_zero(a[j], size(l), s));
    return PRH_OK;
  }

  // two sweeps in one launch (k_stencil_pair): FP32 fast kernels on levels with
  // whole rows of 256 or 512 cells
  bool pairable(int l) const {
    if constexpr (std::is_same<R, float>::value) {
      const int nn[3] = {static_cast<int>(n(l)[0]), static_cast<int>(n(l)[1]), static_cast<int>(n(l)[2])};
      return fast && pair_enabled && pair_ok(nn);
    }
    return false;
  }
  int jacobi_pair(int l, const Ptrs<R>& in, const Ptrs<R>& src, const Ptrs<R>& out, bool zero_in, bool acc) {
    if constexpr (std::is_same<R, float>::value) {
      const double total = static_cast<double>(size(l));
      // algorithmic traffic: v (unless zero) and f read, the result written (or added into FP64)
      const double bytes = total * (acc ? 24.0 : zero_in ? 8.0 : 12.0);
      for (int j = 0; j < count; j++) {
        timer_before(l == 0 ? PRH_TIMER_PAIR_FINEST : -1, s);
        if (acc) {
          launch_pair<kOpJacobi, kOpJacobiAcc>(in[j], src[j], out[j], G[j][l], omega, keep, s, acc_top[j]);
          fused[j] = true;
        } else if (zero_in) {
          launch_pair<kOpJacobiZero, kOpJacobi>(in[j], src[j], out[j], G[j][l], omega, keep, s);
        } else {
          launch_pair<kOpJacobi, kOpJacobi>(in[j], src[j], out[j], G[j][l], omega, keep, s);
        }
        timer_after(l == 0 ? PRH_TIMER_PAIR_FINEST : -1, s, bytes);
        PRH_LAUNCHED();
      }
      return PRH_OK;
    }
    return PRH_ERR_STATE;
  }

  // the first two sweeps from a zero iterate in one launch (kOpJacobiAfterZero)
  bool after_zero_ok(int l) const {
    if (!fast || !std::is_same<R, float>::value || grouped(l) || PRH_ROWS_BLOCKS <= 0) return false;
    const int nn[3] = {static_cast<int>(n(l)[0]), static_cast<int>(n(l)[1]), static_cast<int>(n(l)[2])};
    return tiled_ok(nn) && rows_cz(nn) != 0;
  }
  // both sweeps' result goes to out (where two single sweeps would leave it)
  int jacobi_after_zero(int l, const Ptrs<R>& src, const Ptrs<R>& out) {
    if constexpr (std::is_same<R, float>::value) {
      for (int j = 0; j < count; j++) {
        launch_tiled<R, R, kOpJacobiAfterZero, true>(src[j], src[j], out[j], G[j][l], omega, keep, s);
        PRH_LAUNCHED();
      }
      return PRH_OK;
    }
    return PRH_ERR_STATE;
  }

  // niter damped-Jacobi sweeps; the iterate moves between cur and alt.  Where
  // the level allows, the sweeps go in fused pairs (the last one alone when
  // niter is odd), and a zero start's first two sweeps go in one launch.
  int smooth(int l, Ptrs<R>& cur, Ptrs<R>& alt, const Ptrs<R>& src, bool zero_start, bool post = false) {
    if (niter <= 0) {
      if (zero_start) PRH_TRY(fill_zero(l, cur));
      return PRH_OK;
    }
    const bool last_fused = post && fuse_top && l == 0;
    for (int it = 0; it < niter;) {
      if (zero_start && it == 0 && niter >= 2 && !(pairable(l)) && !(last_fused && niter == 2) &&
          after_zero_ok(l)) {
        // two sweeps: the iterate ends in cur, as after two single sweeps (two swaps)
        PRH_TRY(jacobi_after_zero(l, src, cur));
        it += 2;
        continue;
      }
      if (pairable(l) && niter - it >= 2) {
        PRH_TRY(jacobi_pair(l, cur, src, alt, zero_start && it == 0, last_fused && it + 2 == niter));
        swap_all(cur, alt, count);
        it += 2;
        continue;
      }
      PRH_TRY(jacobi(l, cur, src, alt, zero_start && it == 0, last_fused && it == niter - 1));
      swap_all(cur, alt, count);
      it++;
    }
    return PRH_OK;
  }

  // first level handled by the single-workgroup coarse kernel
  int coarse_level() const {
    for (int l = 0; l < nlev(); l++)
      if (size(l) <= kCoarseCells && nlev() - l <= kMaxCoarse) return l;
    return nlev();
  }

  // V-cycle from level l (>= coarse_level()) to the coarsest in one launch per
  // kCoarseGroup problems; cur and alt end where they started (2 niter swaps
  // per level).
  int vcycle_coarse(int l, const Ptrs<R>& cur, const Ptrs<R>& alt, const Ptrs<R>& src,
                    bool zero_start) {
    const int nl = nlev() - l;
    for (int j0 = 0; j0 < count; j0 += kCoarseGroup) {
      const int nj = count - j0 < kCoarseGroup ? count - j0 : kCoarseGroup;
      CoarseArgs<R> A;
      A.nlev = nl;
      A.niter = niter;
      A.zero_start = zero_start ? 1 : 0;
      A.omega = omega;
      A.keep = keep;
      for (int q = 0; q < kCoarseGroup; q++) {
        const int j = j0 + (q < nj ? q : 0);
        A.cur0[q] = cur[j];
        A.alt0[q] = alt[j];
        A.src0[q] = src[j];
        for (int m = 0; m < kMaxCoarse; m++) {
          const int lm = l + (m < nl ? m : 0);
          A.L[q][m].G = G[j][lm];
          A.L[q][m].v = static_cast<R*>(H[j]->v[lm]);
          A.L[q][m].f = static_cast<R*>(H[j]->f[lm]);
          A.L[q][m].tmp = static_cast<R*>(H[j]->tmp[lm]);
        }
      }
      k_vcycle_coarse<R><<<nj, kCoarseThreads, 0, s>>>(A);
      PRH_LAUNCHED();
    }
    return PRH_OK;
  }

  // src/multigrid.c:270-300
  int vcycle(int l, Ptrs<R>& cur, Ptrs<R>& alt, const Ptrs<R>& src, bool zero_start) {
    if (l >= coarse_level()) return vcycle_coarse(l, cur, alt, src, zero_start);
    PRH_TRY(smooth(l, cur, alt, src, zero_start));
    if (l + 1 < nlev()) {
      PRH_TRY(residual(l, cur, src, alt));
      const Ptrs<R> fc = f(l + 1);
      PRH_TRY(restrict_to(l, alt, fc));
      Ptrs<R> cc = v(l + 1);
      Ptrs<R> ca = tmp(l + 1);
      PRH_TRY(vcycle(l + 1, cc, ca, fc, true));
      PRH_TRY(prolong_to(l, cc, cur, true));
    }
    return smooth(l, cur, alt, src, false, true);
  }

  // src/multigrid.c:303-328, with level `top` as the finest level.  v0 is the
  // caller's iterate (used as the start only when no coarser level exists),
  // f0 the caller's source.
  int fmg(const Ptrs<R>& f0, const Ptrs<R>& v0, int ncycle, int top = 0) {
    const int nl = nlev();
    Ptrs<R> src[kMaxLevels];
    src[top] = f0;
    for (int l = top + 1; l < nl; l++) {
      src[l] = f(l);
      PRH_TRY(restrict_to(l - 1, src[l - 1], src[l]));
    }
    Ptrs<R> cur, alt;
    if (nl == top + 1) {
      cur = v0;
      alt = tmp(top);
    } else {
      cur = v(nl - 1);
      alt = tmp(nl - 1);
      PRH_TRY(fill_zero(nl - 1, cur));
    }
    for (int c = 0; c < ncycle; c++) PRH_TRY(vcycle(nl - 1, cur, alt, src[nl - 1], false));
    for (int l = nl - 2; l >= top; l--) {
      const Ptrs<R> dst = (l == top) ? v0 : v(l);
      PRH_TRY(prolong_to(l, cur, dst, false));
      cur = dst;
      alt = tmp(l);
      for (int c = 0; c < ncycle; c++) PRH_TRY(vcycle(l, cur, alt, src[l], false));
    }
    for (int j = 0; j < count; j++)
      if (cur[j] != v0[j])
        PRH_HIP(hipMemcpyAsync(v0[j], cur[j], size(top) * sizeof(R), hipMemcpyDeviceToDevice, s));
    return PRH_OK;
  }
};

template <typename R>
static Ptrs<R> one(R* p) {
  Ptrs<R> r;
  r.p[0] = p;
  return r;
}

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
  Ptrs<R> cur = one(v), alt = S.tmp(0);
  PRH_TRY(S.smooth(0, cur, alt, one(const_cast<R*>(f)), false));
  if (cur[0] != v)
    PRH_HIP(hipMemcpyAsync(v, cur[0], S.size(0) * sizeof(R), hipMemcpyDeviceToDevice, S.s));
  return PRH_OK;
}

// The mixed path's FP32 smoother on its own: niter fast damped-Jacobi sweeps
// (two per launch where the level allows), from zero if zero_start.
static int do_jacobi_fast(prh_workspace* ws, float* v, const float* f, const size_t* nmesh,
                          const float* boxsize, const float* boxcenter, float beta, float damping,
                          int niter, const float* los, int zero_start, void* stream) {
  PRH_TRY(check_ws(ws, nmesh, PRH_F32, PRH_WS_MULTIGRID));
  if (!v || !f || !boxsize || !boxcenter) return PRH_ERR_ARG;
  Solver<float> S;
  S.setup(&ws->mg, boxsize, boxcenter, beta, damping, niter, los, as_stream(stream));
  S.fast = true;
  S.pair_enabled = g_paired_sweeps.load() != 0;
  Ptrs<float> cur = one(v), alt = S.tmp(0);
  PRH_TRY(S.smooth(0, cur, alt, one(const_cast<float*>(f)), zero_start != 0));
  if (cur[0] != v)
    PRH_HIP(hipMemcpyAsync(v, cur[0], S.size(0) * sizeof(float), hipMemcpyDeviceToDevice, S.s));
  return PRH_OK;
}

template <typename R>
static int do_vcycle(prh_workspace* ws, R* v, const R* f, const size_t* nmesh, const R* boxsize,
                     const R* boxcenter, R beta, R damping, int niter, const R* los, void* stream) {
  PRH_TRY(check_ws(ws, nmesh, Tag<R>::code, PRH_WS_MULTIGRID));
  if (!v || !f || !boxsize || !boxcenter) return PRH_ERR_ARG;
  Solver<R> S;
  S.setup(&ws->mg, boxsize, boxcenter, beta, damping, niter, los, as_stream(stream));
  Ptrs<R> cur = one(v), alt = S.tmp(0);
  PRH_TRY(S.vcycle(0, cur, alt, one(const_cast<R*>(f)), false));
  if (cur[0] != v)
    PRH_HIP(hipMemcpyAsync(v, cur[0], S.size(0) * sizeof(R), hipMemcpyDeviceToDevice, S.s));
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
  return S.fmg(one(const_cast<R*>(f)), one(v), ncycle);
}

// ---------------------------------------------------------------------------
// Full multigrid without a workspace (prh_fmg_noalloc_f64), for callers that
// cannot allocate device memory (HIP graph capture, tight serving loops): the
// hierarchy of meshes of at most kNoAllocCells cells lives in a file-scope
// device array.  Every GPU holds its own copy of the array until the process
// ends; its address is looked up once per device (hipGetSymbolAddress) and
// only that address is passed to kernels.  All calls on one GPU share the
// array: calls that overlap on one GPU (other streams, other threads) race.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocCells = size_t(1) << 18;  // 64^3
constexpr size_t kNoAllocAlign = 32;                // doubles (256 bytes) per sub-buffer boundary
// level-0 scratch iterate, plus v, f and the scratch iterate of the coarser levels (each level has
// at most 1/8 of the previous one's cells), each rounded up to kNoAllocAlign
constexpr size_t kNoAllocDoubles =
    kNoAllocCells + 3 * (kNoAllocCells / 7 + 1) + (1 + 3 * kMaxLevels) * kNoAllocAlign;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kNoAllocDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kNoAllocDevices] = {};

// The current device's address of `scratch`, looked up on the first call on that device.
static int noalloc_scratch(double** out) {
  int dev = 0;
  PRH_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocDevices) {
    set_error("device %d: at most %d devices", dev, kNoAllocDevices);
    return PRH_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("hipGetSymbolAddress of the multigrid scratch failed: %s", hipGetErrorString(e));
      return PRH_ERR_HIP;
    }
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return PRH_OK;
}

static int do_fmg_noalloc(const double* f, double* v, const size_t* nmesh, const double* boxsize,
                          const double* boxcenter, double beta, double damping, int niter,
                          int ncycle, const double* los, void* stream) {
  if (!f || !v || !nmesh || !boxsize || !boxcenter) return PRH_ERR_ARG;
  const size_t cells = nmesh[0] * nmesh[1] * nmesh[2];
  if (cells == 0 || cells > kNoAllocCells) {
    set_error("prh_fmg_noalloc_f64 takes meshes of 1 to %zu cells, got %zu", kNoAllocCells, cells);
    return PRH_ERR_CAPACITY;
  }
  double* base = nullptr;
  PRH_TRY(noalloc_scratch(&base));
  Hierarchy H;
  size_t m[3] = {nmesh[0], nmesh[1], nmesh[2]};
  size_t off = 0;
  auto take = [&](size_t count) {
    double* p = base + off;
    off += (count + kNoAllocAlign - 1) / kNoAllocAlign * kNoAllocAlign;
    return static_cast<void*>(p);
  };
  for (;;) {
    const int l = H.nlev;
    for (int d = 0; d < 3; d++) H.n[l][d] = m[d];
    const size_t c = m[0] * m[1] * m[2];
    H.tmp[l] = take(c);
    if (l > 0) {
      H.v[l] = take(c);
      H.f[l] = take(c);
    }
    H.nlev++;
    if (!coarsenable(m) || H.nlev >= kMaxLevels) break;
    for (int d = 0; d < 3; d++) m[d] /= 2;
  }
  if (off > kNoAllocDoubles) {
    set_error("multigrid scratch overflow (%zu > %zu doubles)", off, kNoAllocDoubles);
    return PRH_ERR_CAPACITY;
  }
  Solver<double> S;
  S.setup(&H, boxsize, boxcenter, beta, damping, niter, los, as_stream(stream));
  return S.fmg(one(const_cast<double*>(f)), one(v), ncycle);
}

// Checks a group of jobs: 1..kMaxGroup workspaces of the same shape and
// precision, each with its buffers and geometry.
template <typename T>
static int check_group(prh_workspace* const* ws, int count, const size_t* nmesh, int dtype,
                       unsigned need, const T* const* f, T* const* v, const T* boxsize,
                       const T* boxcenter, const T* beta) {
  if (!ws || !f || !v || !nmesh || !boxsize || !boxcenter || !beta) {
    set_error("group solve: ws, f, v, nmesh, boxsize, boxcenter and beta are required");
    return PRH_ERR_ARG;
  }
  if (count < 1 || count > kMaxGroup) {
    set_error("group solve: count %d out of [1, %d]", count, kMaxGroup);
    return PRH_ERR_ARG;
  }
  for (int j = 0; j < count; j++) {
    PRH_TRY(check_ws(ws[j], nmesh, dtype, need));
    if (!f[j] || !v[j]) {
      set_error("group solve: job %d has no source or iterate", j);
      return PRH_ERR_ARG;
    }
    for (int q = 0; q < j; q++)
      if (ws[q] == ws[j]) {
        set_error("group solve: jobs %d and %d share a workspace", q, j);
        return PRH_ERR_ARG;
      }
  }
  return PRH_OK;
}

template <typename R>
static int do_fmg_group(prh_workspace* const* ws, int count, const R* const* f, R* const* v,
                        const size_t* nmesh, const R* boxsize, const R* boxcenter, const R* beta,
                        R damping, int niter, int ncycle, const R* const* los, void* stream) {
  PRH_TRY(check_group<R>(ws, count, nmesh, Tag<R>::code, PRH_WS_MULTIGRID, f, v, boxsize,
                         boxcenter, beta));
  Solver<R> S;
  S.count = count;
  Ptrs<R> f0, v0;
  for (int j = 0; j < count; j++) {
    S.H[j] = &ws[j]->mg;
    S.geometry(j, boxsize + 3 * j, boxcenter + 3 * j, beta[j], los ? los[j] : nullptr);
    f0[j] = const_cast<R*>(f[j]);
    v0[j] = v[j];
  }
  S.params(damping, niter, as_stream(stream));
  return S.fmg(f0, v0, ncycle);
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

// Mixed precision.  Every V-cycle of the multigrid is affine in (v, f), so
// V(v, f) = v + V(0, f - A v): the FP64 iterate of the reference's full
// multigrid can be advanced by an FP64 residual, a V-cycle on it in FP32 from
// a zero start, and an FP64 update.  The coarse FMG stages (levels >= 1)
// run entirely in FP32 (their errors are damped by the finest-level cycles);
// the ncycle finest-level cycles run as FP64 defect corrections with FP32
// V-cycles (fast kernels on the blocked levels); nextra further corrections
// may follow.  The iterate tracks the FP64 multigrid's to FP32 rounding of
// the corrections.  `count` problems of one shape advance in lock step (the
// FP32 levels are grouped as in Solver); norms only for a single problem.
// True when the operator's diagonal D is the same in every cell of every level:
// a global line of sight, or a local one with equal cell sizes on the three axes
// (then sum_i c_i = (3 + beta) / h^2).
static bool centrable(const LevelGeom<double>& G) {
  return !G.local || (G.h2[0] == G.h2[1] && G.h2[1] == G.h2[2]);
}

static int do_fmg_mixed_group(prh_workspace* const* ws, int count, const double* const* f,
                              double* const* v, const size_t* nmesh, const double* boxsize,
                              const double* boxcenter, const double* beta, double damping,
                              int niter, int ncycle, const double* const* los, int nextra, int vpr,
                              double* norms, void* stream) {
  PRH_TRY(check_group<double>(ws, count, nmesh, PRH_F64, PRH_WS_MULTIGRID | PRH_WS_MIXED, f, v,
                              boxsize, boxcenter, beta));
  if (nextra < 0 || vpr < 1 || ncycle < 0 || (norms && count != 1)) return PRH_ERR_ARG;
  hipStream_t s = as_stream(stream);
  const size_t total = ncells(nmesh);
  const unsigned grid = ceil_div(total, kThreads);
  Solver<float> S;
  S.count = count;
  LevelGeom<double> G64[kMaxGroup];
  for (int j = 0; j < count; j++) {
    const double* lj = los ? los[j] : nullptr;
    float bs32[3], bc32[3], los32[3];
    for (int d = 0; d < 3; d++) {
      bs32[d] = static_cast<float>(boxsize[3 * j + d]);
      bc32[d] = static_cast<float>(boxcenter[3 * j + d]);
      if (lj) los32[d] = static_cast<float>(lj[d]);
    }
    S.H[j] = &ws[j]->mixed;
    S.geometry(j, bs32, bc32, static_cast<float>(beta[j]), lj ? los32 : nullptr);
    G64[j] = LevelGeom<double>::make(nmesh, boxsize + 3 * j, boxcenter + 3 * j, beta[j], lj);
  }
  S.params(static_cast<float>(damping), niter, s);
  S.fast = true;
  S.pair_enabled = g_paired_sweeps.load() != 0;
  const Ptrs<float> r32 = S.f(0), e32 = S.v(0);
  // 0. the centred source.  The mean of f is a null mode of the operator: the
  // iterate drifts by a constant per sweep, and in an FP32 correction that
  // constant grows with the coarse cell sizes until its rounding swamps the
  // correction's gradients.  Where every level's diagonal D is the same in
  // every cell (global line of sight, or equal cell sizes), a sweep, residual,
  // restriction or prolongation maps a constant to a constant, so each cycle
  // maps f - m to its result on f minus a constant: solving with f - m (m the
  // mean of f) changes the potential by a constant only, and its gradient not
  // at all.  The kernels that read f (the restriction to level 1 and the FP64
  // residuals) subtract m as they load it (Centre).
  Centre cen[kMaxGroup];
  for (int j = 0; j < count; j++) {
    if (!centrable(G64[j]) || !ws[j]->mg.tmp[0]) continue;
    double* sum = ws[j]->red + 2 * kReduceBlocks + 2;
    k_sum_f64<<<kReduceBlocks, kReduceThreads, 0, s>>>(f[j], total, ws[j]->red);
    PRH_LAUNCHED();
    k_reduce_final<1><<<1, 1024, 0, s>>>(ws[j]->red, kReduceBlocks, 0, sum);
    PRH_LAUNCHED();
    cen[j].sum = sum;
    cen[j].n = static_cast<double>(total);
  }
  // 1. coarse FMG stages in FP32, prolonged into the FP64 iterate.  The full
  // multigrid starts its coarsest level from zero; when level 1 is that level,
  // Solver::fmg starts from the iterate it is given, so that is zeroed first.
  if (S.nlev() > 1) {
    const Ptrs<float> f1 = S.f(1), v1 = S.v(1);
    for (int j = 0; j < count; j++) PRH_TRY((launch_restrict<double, float>(f[j], f1[j], nmesh, s, cen[j])));
    if (S.nlev() == 2) PRH_TRY(S.fill_zero(1, v1));
    PRH_TRY(S.fmg(f1, v1, ncycle, 1));
    for (int j = 0; j < count; j++)
      PRH_TRY((launch_prolong<float, double>(v1[j], v[j], S.n(1), false, s)));
  }
  // 2. finest-level cycles as FP64 defect corrections
  const int ncorr = ncycle + nextra;
  for (int it = 0; it < ncorr; it++) {
    for (int j = 0; j < count; j++) {
      if (norms) {
        // the norm is of the residual of the true source, f - A v
        PRH_TRY(residual_rms(ws[j], v[j], f[j], cen[j].sum ? nullptr : r32[j], G64[j], &norms[it], s));
        if (cen[j].sum) PRH_TRY((launch_residual<double, float>(v[j], f[j], r32[j], G64[j], s, false, cen[j])));
      } else {
        PRH_TRY((launch_residual<double, float>(v[j], f[j], r32[j], G64[j], s, false, cen[j])));
      }
    }
    Ptrs<float> cur = e32, alt = S.tmp(0);
    // the last cycle's last sweep adds its correction into the FP64 iterate where it can
    for (int j = 0; j < count; j++) S.acc_top[j] = v[j];
    for (int c = 0; c < vpr; c++) {
      S.fuse_top = c == vpr - 1;
      for (int j = 0; j < count; j++) S.fused[j] = false;
      PRH_TRY(S.vcycle(0, cur, alt, r32, c == 0));
    }
    S.fuse_top = false;
    for (int j = 0; j < count; j++) {
      if (S.fused[j]) continue;
      k_add_f32_f64<<<grid, kThreads, 0, s>>>(v[j], cur[j], total);
      PRH_LAUNCHED();
    }
  }
  if (norms) PRH_TRY(residual_rms(ws[0], v[0], f[0], nullptr, G64[0], &norms[ncorr], s));
  return PRH_OK;
}

// One problem: the reference-shaped entry point of the mixed-precision solve.
static int do_fmg_mixed(prh_workspace* ws, const double* f, double* v, const size_t* nmesh,
                        const double* boxsize, const double* boxcenter, double beta, double damping,
                        int niter, int ncycle, const double* los, int nextra, int vpr,
                        double* norms, void* stream) {
  if (!v || !f || !boxsize || !boxcenter) return PRH_ERR_ARG;
  return do_fmg_mixed_group(&ws, 1, &f, &v, nmesh, boxsize, boxcenter, &beta, damping, niter,
                            ncycle, &los, nextra, vpr, norms, stream);
}

}  // namespace prh

using namespace prh;

extern "C" {

int prh_set_tuning(int option, int value) {
  if (option == PRH_TUNE_PAIRED_SWEEPS) return g_paired_sweeps.exchange(value ? 1 : 0);
  set_error("unknown tuning option %d", option);
  return PRH_ERR_ARG;
}


int prh_jacobi_f32(prh_workspace* ws, float* v, const float* f, const size_t* nmesh,
                   const float* boxsize, const float* boxcenter, float beta, float damping_factor,
                   int niterations, const float* los, void* stream) {
  return do_jacobi<float>(ws, v, f, nmesh, boxsize, boxcenter, beta, damping_factor, niterations,
                          los, stream);
}
int prh_jacobi_fast_f32(prh_workspace* ws, float* v, const float* f, const size_t* nmesh,
                        const float* boxsize, const float* boxcenter, float beta, float damping_factor,
                        int niterations, const float* los, int zero_start, void* stream) {
  return do_jacobi_fast(ws, v, f, nmesh, boxsize, boxcenter, beta, damping_factor, niterations, los,
                        zero_start, stream);
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
  return launch_prolong<float, float>(v2h, v1h, nmesh, false, as_stream(stream));
}
int prh_prolong_f64(const double* v2h, double* v1h, const size_t* nmesh, void* stream) {
  if (!v2h || !v1h || !nmesh) return PRH_ERR_ARG;
  return launch_prolong<double, double>(v2h, v1h, nmesh, false, as_stream(stream));
}

int prh_reduce_f32(const float* v1h, float* v2h, const size_t* nmesh, void* stream) {
  if (!v1h || !v2h || !nmesh) return PRH_ERR_ARG;
  return launch_restrict<float, float>(v1h, v2h, nmesh, as_stream(stream));
}
int prh_reduce_f64(const double* v1h, double* v2h, const size_t* nmesh, void* stream) {
  if (!v1h || !v2h || !nmesh) return PRH_ERR_ARG;
  return launch_restrict<double, double>(v1h, v2h, nmesh, as_stream(stream));
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

int prh_fmg_noalloc_f64(const double* f1h, double* v1h, const size_t* nmesh, const double* boxsize,
                        const double* boxcenter, double beta, double jacobi_damping_factor,
                        int jacobi_niterations, int vcycle_niterations, const double* los,
                        void* stream) {
  return do_fmg_noalloc(f1h, v1h, nmesh, boxsize, boxcenter, beta, jacobi_damping_factor,
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

int prh_fmg_group_f32(prh_workspace* const* ws, int count, const float* const* f1h,
                      float* const* v1h, const size_t* nmesh, const float* boxsize,
                      const float* boxcenter, const float* beta, float jacobi_damping_factor,
                      int jacobi_niterations, int vcycle_niterations, const float* const* los,
                      void* stream) {
  return do_fmg_group<float>(ws, count, f1h, v1h, nmesh, boxsize, boxcenter, beta,
                             jacobi_damping_factor, jacobi_niterations, vcycle_niterations, los,
                             stream);
}
int prh_fmg_group_f64(prh_workspace* const* ws, int count, const double* const* f1h,
                      double* const* v1h, const size_t* nmesh, const double* boxsize,
                      const double* boxcenter, const double* beta, double jacobi_damping_factor,
                      int jacobi_niterations, int vcycle_niterations, const double* const* los,
                      void* stream) {
  return do_fmg_group<double>(ws, count, f1h, v1h, nmesh, boxsize, boxcenter, beta,
                              jacobi_damping_factor, jacobi_niterations, vcycle_niterations, los,
                              stream);
}
int prh_fmg_mixed_group(prh_workspace* const* ws, int count, const double* const* f1h,
                        double* const* v1h, const size_t* nmesh, const double* boxsize,
                        const double* boxcenter, const double* beta, double jacobi_damping_factor,
                        int jacobi_niterations, int vcycle_niterations, const double* const* los,
                        int nrefine, int vcycles_per_refine, void* stream) {
  return do_fmg_mixed_group(ws, count, f1h, v1h, nmesh, boxsize, boxcenter, beta,
                            jacobi_damping_factor, jacobi_niterations, vcycle_niterations, los,
                            nrefine, vcycles_per_refine, nullptr, stream);
}

int prh_residual_norm_f64(prh_workspace* ws, const double* v, const double* f, const size_t* nmesh,
                          const double* boxsize, const double* boxcenter, double beta,
                          const double* los, double* rms, void* stream) {
  if (!ws || !v || !f || !nmesh || !boxsize || !boxcenter || !rms) return PRH_ERR_ARG;
  const LevelGeom<double> G = LevelGeom<double>::make(nmesh, boxsize, boxcenter, beta, los);
  return residual_rms(ws, v, f, nullptr, G, rms, as_stream(stream));
}

}  // extern "C"
