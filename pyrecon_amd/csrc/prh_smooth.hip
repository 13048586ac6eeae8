// Gaussian smoothing of a real mesh in Fourier space, on the device, without
// a 3-D FFT library.
//
// The reference (RealMesh.smooth_gaussian(method='fft'), pyrecon/mesh.py:
// 655-680) multiplies the 3-D spectrum by exp(-sum_i (r_i k_i)^2 / 2) with
// k_i = 2 pi fftfreq(n_i, L_i / n_i) and transforms back.  The filter is the
// product of one factor per axis, and the 3-D DFT is the product of 1-D DFTs
// along the axes, so the same field results from filtering along each axis in
// turn: for every line along the axis, forward DFT, multiply by
// exp(-(r k)^2 / 2) / n, inverse DFT.  Each pass reads and writes the mesh
// once.  The mesh is real and the 1-D filter is real and even, so two real
// lines are filtered at once as the real and imaginary parts of one complex
// line.
//
// The 1-D transforms are mixed-radix Stockham FFTs (radix 2, 3, 4, 5, 7) on
// lines held in LDS, ping-ponging between two buffers; twiddles come from a
// per-axis table exp(-2 pi i k / n) built on the host in double precision.
#include <cmath>
#include <cstdint>
#include <type_traits>
#include <vector>

#include "prh_workspace.hpp"

namespace prh {

constexpr int kSmoothThreads = 512;
#ifndef PRH_SMOOTH_REG512
#define PRH_SMOOTH_REG512 1
#endif
constexpr int kMaxFly = 2;  // butterflies per thread and pass (held in registers)
constexpr size_t kSmoothLds = 81920;      // target LDS bytes of a block (2 blocks per CU)
constexpr size_t kSmoothLdsMax = 163840;  // upper limit (one line per block)

template <typename R>
struct Cx {
  R re, im;
};

template <typename R>
__device__ __forceinline__ Cx<R> cmul(Cx<R> a, Cx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// Parameters of one Stockham pass: radix, span Ns of the sub-transforms done
// so far (a power of two except after a second odd radix), butterflies per
// line m = n / radix (looped over a power-of-two range 2^lm >= m, so that the
// (line, butterfly) split of a thread index is a shift), twiddle step.
struct Pass {
  int rad, ns, ns_shift;  // ns_shift = log2(ns), or -1 when ns is not a power of two
  int m, lm, span;
};

struct AxisPlan {
  int n, npass, ln;  // 2^ln >= n: padded element range of the load/store loops
  Pass pass[24];
};

// LDS position of element `idx` of the line buffers.  The early Stockham
// passes write with a stride of `radix` elements; without a remap those
// writes fall on a few banks.  Default: the low 3 bits of the index are
// XOR-ed with bits 3-5 (a permutation within each group of 8 elements, no
// extra space; the buffers are rounded up to whole groups).  Otherwise: one
// pad slot every 16 elements.
#ifndef PRH_SMOOTH_SWIZZLE
#define PRH_SMOOTH_SWIZZLE 1
#endif
#if PRH_SMOOTH_SWIZZLE
__device__ __forceinline__ int lds_pad(int idx) { return idx ^ ((idx >> 3) & 7); }
__host__ __device__ constexpr size_t padded_len(size_t n) { return (n + 7) & ~size_t(7); }
#else
__device__ __forceinline__ int lds_pad(int idx) { return idx + (idx >> 4); }
__host__ __device__ constexpr size_t padded_len(size_t n) { return n + n / 16 + 1; }
#endif

// 4-point DFT (forward: exp(-2 pi i / 4) = -i; inverse: +i).
template <typename R>
__device__ __forceinline__ void dft4(const Cx<R> c[4], Cx<R> y[4], bool inverse) {
  const Cx<R> s02 = {c[0].re + c[2].re, c[0].im + c[2].im};
  const Cx<R> d02 = {c[0].re - c[2].re, c[0].im - c[2].im};
  const Cx<R> s13 = {c[1].re + c[3].re, c[1].im + c[3].im};
  const Cx<R> d13 = {c[1].re - c[3].re, c[1].im - c[3].im};
  const Cx<R> rot = inverse ? Cx<R>{-d13.im, d13.re} : Cx<R>{d13.im, -d13.re};
  y[0] = {s02.re + s13.re, s02.im + s13.im};
  y[2] = {s02.re - s13.re, s02.im - s13.im};
  y[1] = {d02.re + rot.re, d02.im + rot.im};
  y[3] = {d02.re - rot.re, d02.im - rot.im};
}

// One Stockham pass of radix RAD over LP = 2^llp lines of length n, in place:
// every thread reads and transforms its (at most kMaxFly) butterflies into
// registers, the block synchronises, then the results are written over the
// same buffer.  The host keeps LP * n / RAD <= kMaxFly * kSmoothThreads.
template <typename R, int RAD>
__device__ __forceinline__ void stockham_pass(Cx<R>* __restrict__ a, int n, int llp, const Pass& P,
                                              const Cx<R>* __restrict__ tw, bool inverse) {
  const int m = P.m, Ns = P.ns;
  const int rstep = n / RAD;  // index step of the RAD-th roots of unity
  const int total = 1 << (llp + P.lm);
  Cx<R> o[kMaxFly][RAD];
  int dst[kMaxFly];
#pragma unroll
  for (int fly = 0; fly < kMaxFly; fly++) {
    const int t = threadIdx.x + fly * kSmoothThreads;
    const int l = t >> P.lm;
    const int j = t & ((1 << P.lm) - 1);
    dst[fly] = -1;
    if (t >= total || j >= m) continue;
    int k, jq;
    if (P.ns_shift >= 0) {
      k = j & (Ns - 1);
      jq = j >> P.ns_shift;
    } else {
      jq = j / Ns;
      k = j - jq * Ns;
    }
    const int la = l * n + j;
    Cx<R> v[RAD];
#pragma unroll
    for (int r = 0; r < RAD; r++) v[r] = a[lds_pad(la + r * m)];
    const int tstep = k * P.span;  // r * tstep < n
#pragma unroll
    for (int r = 1; r < RAD; r++) {
      Cx<R> w = tw[r * tstep];
      if (inverse) w.im = -w.im;
      v[r] = cmul(v[r], w);
    }
    Cx<R>* out = o[fly];
    if constexpr (RAD == 2) {
      out[0] = {v[0].re + v[1].re, v[0].im + v[1].im};
      out[1] = {v[0].re - v[1].re, v[0].im - v[1].im};
    } else if constexpr (RAD == 8) {
      // 2 x 4 split: a_k = v_k + v_k+4, b_k = (v_k - v_k+4) w8^k, then a 4-point DFT of
      // each; X_2m from a, X_2m+1 from b
      constexpr R h = R(0.70710678118654752440);
      Cx<R> aa[4], bb[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        aa[q] = {v[q].re + v[q + 4].re, v[q].im + v[q + 4].im};
        bb[q] = {v[q].re - v[q + 4].re, v[q].im - v[q + 4].im};
      }
      bb[1] = cmul(bb[1], Cx<R>{h, inverse ? h : -h});
      bb[2] = inverse ? Cx<R>{-bb[2].im, bb[2].re} : Cx<R>{bb[2].im, -bb[2].re};
      bb[3] = cmul(bb[3], Cx<R>{-h, inverse ? h : -h});
      Cx<R> ya[4], yb[4];
      dft4(aa, ya, inverse);
      dft4(bb, yb, inverse);
#pragma unroll
      for (int q = 0; q < 4; q++) {
        out[2 * q] = ya[q];
        out[2 * q + 1] = yb[q];
      }
    } else if constexpr (RAD == 4) {
      dft4(v, out, inverse);
    } else {
      // direct DFT with the RAD-th roots of unity taken from the table
#pragma unroll
      for (int q = 0; q < RAD; q++) {
        Cx<R> acc = v[0];
#pragma unroll
        for (int r = 1; r < RAD; r++) {
          Cx<R> w = tw[((q * r) % RAD) * rstep];
          if (inverse) w.im = -w.im;
          const Cx<R> p = cmul(v[r], w);
          acc.re += p.re;
          acc.im += p.im;
        }
        out[q] = acc;
      }
    }
    dst[fly] = l * n + jq * Ns * RAD + k;
  }
  __syncthreads();
#pragma unroll
  for (int fly = 0; fly < kMaxFly; fly++)
    if (dst[fly] >= 0) {
#pragma unroll
      for (int r = 0; r < RAD; r++) a[lds_pad(dst[fly] + r * Ns)] = o[fly][r];
    }
  __syncthreads();
}

// Runs the passes of `P` on the 2^llp lines in `buf` (in place, natural order
// in and out).
template <typename R>
__device__ void fft_lines(Cx<R>* buf, int llp, const AxisPlan& P, const Cx<R>* __restrict__ tw,
                          bool inverse) {
  for (int p = 0; p < P.npass; p++) {
    const Pass& S = P.pass[p];
    switch (S.rad) {
      case 2: stockham_pass<R, 2>(buf, P.n, llp, S, tw, inverse); break;
      case 3: stockham_pass<R, 3>(buf, P.n, llp, S, tw, inverse); break;
      case 4: stockham_pass<R, 4>(buf, P.n, llp, S, tw, inverse); break;
      case 8: stockham_pass<R, 8>(buf, P.n, llp, S, tw, inverse); break;
      case 5: stockham_pass<R, 5>(buf, P.n, llp, S, tw, inverse); break;
      default: stockham_pass<R, 7>(buf, P.n, llp, S, tw, inverse); break;
    }
  }
}

// Real lines of one axis: line q starts at (q / I) * ostride + (q % I) * istride
// and has n elements `es` apart.  Block b filters the 2^llp complex lines
// starting at complex line b * 2^llp (remapped so that consecutive bundles run
// on one XCD); complex line p is real line 2p (+ i real line 2p + 1).
// CONTIG: the elements of a line are contiguous (the z axis); otherwise
// consecutive real lines are (loads iterate over lines fastest).
template <typename R, bool CONTIG>
__global__ __launch_bounds__(kSmoothThreads) void k_smooth_axis(
    R* __restrict__ mesh, AxisPlan P, long long nreal, int I, long long ostride,
    long long istride, long long es, int llp, const Cx<R>* __restrict__ tw_global,
    const R* __restrict__ gain) {
  extern __shared__ double smem[];
  const int LP = 1 << llp;
  const size_t buf = padded_len(static_cast<size_t>(LP) * P.n);
  Cx<R>* b0 = reinterpret_cast<Cx<R>*>(smem);
  Cx<R>* stw = b0 + buf;  // twiddle table, staged once per block
  const int n = P.n;
  for (int k = threadIdx.x; k < n; k += kSmoothThreads) stw[k] = tw_global[k];
  // XCD-aware order: the 8 XCDs take blocks round-robin; give consecutive
  // bundles (which share cache lines in the strided passes) to one XCD
  const unsigned nb = gridDim.x, bid = blockIdx.x;
  const unsigned per = nb / 8;
  const unsigned blk = bid < per * 8 ? (bid % 8) * per + bid / 8 : bid;
  const long long q0 = 2LL * blk * LP;  // first real line of the block
  auto base = [&](long long q) { return (q / I) * ostride + (q % I) * istride; };
  const int lpl = llp + 1;  // log2 of the real lines of the block
  const int nel = CONTIG ? (2 * LP) << P.ln : (2 * LP) * n;
  // loads in batches of kBatch per thread, all issued before any LDS store, so
  // that a thread keeps kBatch reads in flight
  constexpr int kBatch = 8;
  for (int t0 = threadIdx.x; t0 < nel; t0 += kBatch * kSmoothThreads) {
    R val[kBatch];
    int dst[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; u++) {
      const int t = t0 + u * kSmoothThreads;
      int lp, e;
      if (CONTIG) {
        lp = t >> P.ln;
        e = t & ((1 << P.ln) - 1);
      } else {
        e = t >> lpl;
        lp = t & ((1 << lpl) - 1);
      }
      const long long q = q0 + lp;
      const bool ok = t < nel && e < n;
      val[u] = ok && q < nreal ? mesh[base(q) + e * es] : R(0);
      dst[u] = ok ? 2 * lds_pad((lp >> 1) * n + e) + (lp & 1) : -1;
    }
#pragma unroll
    for (int u = 0; u < kBatch; u++)
      if (dst[u] >= 0) reinterpret_cast<R*>(b0)[dst[u]] = val[u];
  }
  __syncthreads();
  fft_lines<R>(b0, llp, P, stw, false);
  Cx<R>* spec = b0;
  for (int t = threadIdx.x; t < (LP << P.ln); t += kSmoothThreads) {
    const int e = t & ((1 << P.ln) - 1);
    if (e >= n) continue;
    const int l = t >> P.ln;
    const R g = gain[e];
    const int q = lds_pad(l * n + e);
    spec[q].re *= g;
    spec[q].im *= g;
  }
  __syncthreads();
  fft_lines<R>(b0, llp, P, stw, true);
  const Cx<R>* res = b0;
  for (int t = threadIdx.x; t < nel; t += kSmoothThreads) {
    int lp, e;
    if (CONTIG) {
      lp = t >> P.ln;
      e = t & ((1 << P.ln) - 1);
      if (e >= n) continue;
    } else {
      e = t >> lpl;
      lp = t & ((1 << lpl) - 1);
    }
    const long long q = q0 + lp;
    if (q < nreal) {
      const R* slot = reinterpret_cast<const R*>(&res[lds_pad((lp >> 1) * n + e)]);
      mesh[base(q) + e * es] = slot[lp & 1];
    }
  }
}

// Lines of 512 cells along z (contiguous), with the transforms in registers:
// one wavefront per complex line (two real rows), lane L holding elements
// L + 64 r, r = 0..7.  A 512-point DFT is 8 x 64 and the 64 is 8 x 8:
//   1. DFT-8 over r in each lane, times W512^(L k1);
//   2. transpose through LDS to lanes (j1, k1) = j1 + 8 k1 holding the
//      values of lanes j1 + 8 j2, j2 = 0..7; DFT-8 over j2, times W64^(j1 m1);
//   3. transpose to lanes (k1, m1) = k1 + 8 m1; DFT-8 over j1.
// Lane k1 + 8 m1 then holds X[k1 + 8 m1 + 64 m2], m2 = 0..7: the same layout
// as the input, so the gain is applied in place and the inverse transform is
// the same three steps with conjugate roots.  Each transpose is one LDS write
// and one read of 8 values per lane; nothing else leaves registers.
template <typename R>
__device__ __forceinline__ void dft8(Cx<R> v[8], bool inverse) {
  constexpr R h = R(0.70710678118654752440);
  Cx<R> aa[4], bb[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    aa[q] = {v[q].re + v[q + 4].re, v[q].im + v[q + 4].im};
    bb[q] = {v[q].re - v[q + 4].re, v[q].im - v[q + 4].im};
  }
  bb[1] = cmul(bb[1], Cx<R>{h, inverse ? h : -h});
  bb[2] = inverse ? Cx<R>{-bb[2].im, bb[2].re} : Cx<R>{bb[2].im, -bb[2].re};
  bb[3] = cmul(bb[3], Cx<R>{-h, inverse ? h : -h});
  Cx<R> ya[4], yb[4];
  dft4(aa, ya, inverse);
  dft4(bb, yb, inverse);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    v[2 * q] = ya[q];
    v[2 * q + 1] = yb[q];
  }
}

constexpr int kReg512Waves = 4;  // lines in flight per workgroup

// exchange slot of value k of lane L (XOR-swizzled within each lane's 8)
__device__ __forceinline__ int xslot(int lane, int k) { return lane * 8 + (k ^ (lane & 7)); }

// tw holds W_{512 S}^k: S = 1 for 512-cell lines, 2 for the halves of 1024-cell lines.
template <typename R, int S = 1>
__device__ __forceinline__ void fft512_lane(Cx<R> v[8], Cx<R>* __restrict__ xb,
                                            const Cx<R>* __restrict__ tw, int lane, bool inverse) {
  auto root = [&](int k) {  // W512^k (conjugate for the inverse)
    Cx<R> w = tw[(k & 511) * S];
    if (inverse) w.im = -w.im;
    return w;
  };
  // 1. over r (elements lane + 64 r)
  dft8(v, inverse);
#pragma unroll
  for (int k = 1; k < 8; k++) v[k] = cmul(v[k], root(lane * k));
  // 2. lane (j1, k1) gathers the values k1 of lanes j1 + 8 j2
#pragma unroll
  for (int k = 0; k < 8; k++) xb[xslot(lane, k)] = v[k];
  __syncthreads();
  const int j1 = lane & 7, k1 = lane >> 3;
#pragma unroll
  for (int j2 = 0; j2 < 8; j2++) v[j2] = xb[xslot(j1 + 8 * j2, k1)];
  dft8(v, inverse);
#pragma unroll
  for (int m1 = 1; m1 < 8; m1++) v[m1] = cmul(v[m1], root(8 * j1 * m1));
  __syncthreads();
  // 3. lane (k1, m1) gathers the values m1 of lanes j1 + 8 k1
#pragma unroll
  for (int m1 = 0; m1 < 8; m1++) xb[xslot(lane, m1)] = v[m1];
  __syncthreads();
  const int kk1 = lane & 7, m1 = lane >> 3;
#pragma unroll
  for (int jj = 0; jj < 8; jj++) v[jj] = xb[xslot(jj + 8 * kk1, m1)];
  dft8(v, inverse);
  __syncthreads();
}

// `nrows` rows of 512 contiguous cells (row q at q * 512); wave w of block b
// filters complex line p = b * kReg512Waves + w (rows 2p, 2p + 1), then moves
// on by the grid's number of lines.
template <typename R>
__global__ __launch_bounds__(64 * kReg512Waves) void k_smooth_reg512(
    R* __restrict__ mesh, long long nrows, const Cx<R>* __restrict__ tw_global,
    const R* __restrict__ gain) {
  __shared__ Cx<R> stw[512];
  __shared__ Cx<R> xbuf[kReg512Waves][512];
  for (int k = threadIdx.x; k < 512; k += 64 * kReg512Waves) stw[k] = tw_global[k];
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  Cx<R>* xb = xbuf[w];
  const long long ncx = (nrows + 1) / 2;
  const long long step = static_cast<long long>(gridDim.x) * kReg512Waves;
  // every wave of the block runs the same number of iterations (barriers inside)
  for (long long p0 = static_cast<long long>(blockIdx.x) * kReg512Waves; p0 < ncx; p0 += step) {
    const long long p = p0 + w;
    const bool live = p < ncx;
    const long long qa = 2 * p, qb = 2 * p + 1;
    const R* ra = mesh + (live ? qa : 0) * 512;
    const bool hasb = live && qb < nrows;
    const R* rb = mesh + (hasb ? qb : 0) * 512;
    Cx<R> v[8];
#pragma unroll
    for (int r = 0; r < 8; r++) {
      const int e = lane + 64 * r;
      v[r] = {live ? ra[e] : R(0), hasb ? rb[e] : R(0)};
    }
    fft512_lane(v, xb, stw, lane, false);
    // lane holds X[lane + 64 m2]
#pragma unroll
    for (int m2 = 0; m2 < 8; m2++) {
      const R g = gain[lane + 64 * m2];
      v[m2].re *= g;
      v[m2].im *= g;
    }
    fft512_lane(v, xb, stw, lane, true);
    if (live) {
      R* wa = mesh + qa * 512;
#pragma unroll
      for (int r = 0; r < 8; r++) wa[lane + 64 * r] = v[r].re;
      if (hasb) {
        R* wb = mesh + qb * 512;
#pragma unroll
        for (int r = 0; r < 8; r++) wb[lane + 64 * r] = v[r].im;
      }
    }
  }
}

// Lines of 512 cells along x or y (stride `es`), transforms as in
// k_smooth_reg512.  A block takes 16 adjacent real lines (consecutive z, one
// 128-byte row of FP64 per element index): coalesced loads into an LDS tile
// [element][line] (row stride 17 against bank conflicts), each of the 8
// wavefronts reads its complex line (real lines 2w, 2w + 1) into registers,
// the tile memory then serves as the transposition buffers, and the results
// go back through the tile to coalesced stores.  Requires I % 16 == 0 and
// consecutive lines adjacent (istride 1).
constexpr int kReg512SWaves = 8;
constexpr int kTileRow = 17;  // LDS row stride of the [element][16 lines] tile

template <typename R>
__global__ __launch_bounds__(64 * kReg512SWaves) void k_smooth_reg512s(
    R* __restrict__ mesh, long long nreal, int I, long long ostride, long long es,
    const Cx<R>* __restrict__ tw_global, const R* __restrict__ gain) {
  constexpr int NT = 64 * kReg512SWaves;
  __shared__ Cx<R> stw[512];
  __shared__ R tile[512 * kTileRow];  // also 8 transposition buffers of 512 Cx<R>
  static_assert(sizeof(R) * 512 * kTileRow >= sizeof(Cx<R>) * 512 * kReg512SWaves, "tile");
  for (int k = threadIdx.x; k < 512; k += NT) stw[k] = tw_global[k];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  Cx<R>* xb = reinterpret_cast<Cx<R>*>(tile) + 512 * w;
  const int c = threadIdx.x & 15, e0 = threadIdx.x >> 4;  // load/store: line c, elements e0 + 32 k
  const long long ngroups = nreal / 16;
  for (long long g = blockIdx.x; g < ngroups; g += gridDim.x) {
    const long long q0 = 16 * g;
    R* base = mesh + (q0 / I) * ostride + (q0 % I);
    __syncthreads();  // the tile is free (previous group stored, twiddles staged)
#pragma unroll 4
    for (int k = 0; k < 16; k++) {
      const int e = e0 + 32 * k;
      tile[e * kTileRow + c] = base[static_cast<long long>(e) * es + c];
    }
    __syncthreads();
    Cx<R> v[8];
#pragma unroll
    for (int r = 0; r < 8; r++) {
      const R* t = tile + (lane + 64 * r) * kTileRow + 2 * w;
      v[r] = {t[0], t[1]};
    }
    __syncthreads();  // the tile becomes the transposition buffers
    fft512_lane(v, xb, stw, lane, false);
#pragma unroll
    for (int m2 = 0; m2 < 8; m2++) {
      const R gg = gain[lane + 64 * m2];
      v[m2].re *= gg;
      v[m2].im *= gg;
    }
    fft512_lane(v, xb, stw, lane, true);  // ends with a barrier
#pragma unroll
    for (int r = 0; r < 8; r++) {
      R* t = tile + (lane + 64 * r) * kTileRow + 2 * w;
      t[0] = v[r].re;
      t[1] = v[r].im;
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < 16; k++) {
      const int e = e0 + 32 * k;
      base[static_cast<long long>(e) * es + c] = tile[e * kTileRow + c];
    }
  }
}

// Lines of 1024 cells in registers: lane L holds elements L + 64 r, r = 0..15.
// Forward: one radix-2 step in the lane (a = x_n + x_{n+512}, b = (x_n -
// x_{n+512}) W1024^n for n = L + 64 r, r < 8), then the 512-point transforms of
// a and b (fft512_lane with every second root of the 1024-point table): lane
// L then holds X[2 (L + 64 m)] in a[m] and X[2 (L + 64 m) + 1] in b[m].  The
// inverse runs the two 512-point inverse transforms and combines x_n = a'_n +
// W1024^-n b'_n, x_{n+512} = a'_n - W1024^-n b'_n.
template <typename R>
__device__ __forceinline__ void filter1024_lane(Cx<R> v[16], Cx<R>* __restrict__ xb,
                                                const Cx<R>* __restrict__ tw,
                                                const R* __restrict__ gain, int lane) {
  Cx<R> a[8], b[8];
#pragma unroll
  for (int r = 0; r < 8; r++) {
    a[r] = {v[r].re + v[r + 8].re, v[r].im + v[r + 8].im};
    b[r] = cmul(Cx<R>{v[r].re - v[r + 8].re, v[r].im - v[r + 8].im}, tw[lane + 64 * r]);
  }
  fft512_lane<R, 2>(a, xb, tw, lane, false);
  fft512_lane<R, 2>(b, xb, tw, lane, false);
#pragma unroll
  for (int m = 0; m < 8; m++) {
    const R ga = gain[2 * (lane + 64 * m)], gb = gain[2 * (lane + 64 * m) + 1];
    a[m].re *= ga;
    a[m].im *= ga;
    b[m].re *= gb;
    b[m].im *= gb;
  }
  fft512_lane<R, 2>(a, xb, tw, lane, true);
  fft512_lane<R, 2>(b, xb, tw, lane, true);  // ends with a barrier
#pragma unroll
  for (int r = 0; r < 8; r++) {
    Cx<R> w = tw[lane + 64 * r];
    w.im = -w.im;
    const Cx<R> t = cmul(b[r], w);
    v[r] = {a[r].re + t.re, a[r].im + t.im};
    v[r + 8] = {a[r].re - t.re, a[r].im - t.im};
  }
}

// `nrows` rows of 1024 contiguous cells; as k_smooth_reg512 with one complex
// line (two real rows) per wavefront.
template <typename R>
__global__ __launch_bounds__(64 * kReg512Waves) void k_smooth_reg1024(
    R* __restrict__ mesh, long long nrows, const Cx<R>* __restrict__ tw_global,
    const R* __restrict__ gain) {
  __shared__ Cx<R> stw[1024];
  __shared__ Cx<R> xbuf[kReg512Waves][512];
  for (int k = threadIdx.x; k < 1024; k += 64 * kReg512Waves) stw[k] = tw_global[k];
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  Cx<R>* xb = xbuf[w];
  const long long ncx = (nrows + 1) / 2;
  const long long step = static_cast<long long>(gridDim.x) * kReg512Waves;
  // every wave of the block runs the same number of iterations (barriers inside)
  for (long long p0 = static_cast<long long>(blockIdx.x) * kReg512Waves; p0 < ncx; p0 += step) {
    const long long p = p0 + w;
    const bool live = p < ncx;
    const long long qa = 2 * p, qb = 2 * p + 1;
    const R* ra = mesh + (live ? qa : 0) * 1024;
    const bool hasb = live && qb < nrows;
    const R* rb = mesh + (hasb ? qb : 0) * 1024;
    Cx<R> v[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int e = lane + 64 * r;
      v[r] = {live ? ra[e] : R(0), hasb ? rb[e] : R(0)};
    }
    filter1024_lane<R>(v, xb, stw, gain, lane);
    if (live) {
      R* wa = mesh + qa * 1024;
#pragma unroll
      for (int r = 0; r < 16; r++) wa[lane + 64 * r] = v[r].re;
      if (hasb) {
        R* wb = mesh + qb * 1024;
#pragma unroll
        for (int r = 0; r < 16; r++) wb[lane + 64 * r] = v[r].im;
      }
    }
  }
}

// Lines of 1024 cells along x or y, as k_smooth_reg512s: 16 adjacent real
// lines staged through an LDS tile [element][line] (139 KB in FP64), which
// then serves as the 8 wavefronts' transposition buffers.
template <typename R>
__global__ __launch_bounds__(64 * kReg512SWaves) void k_smooth_reg1024s(
    R* __restrict__ mesh, long long nreal, int I, long long ostride, long long es,
    const Cx<R>* __restrict__ tw_global, const R* __restrict__ gain) {
  constexpr int NT = 64 * kReg512SWaves;
  __shared__ Cx<R> stw[1024];
  __shared__ R tile[1024 * kTileRow];
  static_assert(sizeof(R) * 1024 * kTileRow >= sizeof(Cx<R>) * 512 * kReg512SWaves, "tile");
  for (int k = threadIdx.x; k < 1024; k += NT) stw[k] = tw_global[k];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  Cx<R>* xb = reinterpret_cast<Cx<R>*>(tile) + 512 * w;
  const int c = threadIdx.x & 15, e0 = threadIdx.x >> 4;  // load/store: line c, elements e0 + 32 k
  const long long ngroups = nreal / 16;
  for (long long g = blockIdx.x; g < ngroups; g += gridDim.x) {
    const long long q0 = 16 * g;
    R* base = mesh + (q0 / I) * ostride + (q0 % I);
    __syncthreads();  // the tile is free (previous group stored, twiddles staged)
#pragma unroll 4
    for (int k = 0; k < 32; k++) {
      const int e = e0 + 32 * k;
      tile[e * kTileRow + c] = base[static_cast<long long>(e) * es + c];
    }
    __syncthreads();
    Cx<R> v[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const R* t = tile + (lane + 64 * r) * kTileRow + 2 * w;
      v[r] = {t[0], t[1]};
    }
    __syncthreads();  // the tile becomes the transposition buffers
    filter1024_lane<R>(v, xb, stw, gain, lane);  // ends with a barrier
#pragma unroll
    for (int r = 0; r < 16; r++) {
      R* t = tile + (lane + 64 * r) * kTileRow + 2 * w;
      t[0] = v[r].re;
      t[1] = v[r].im;
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < 32; k++) {
      const int e = e0 + 32 * k;
      base[static_cast<long long>(e) * es + c] = tile[e * kTileRow + c];
    }
  }
}

// Direct DFT along one axis, for lengths the radix passes do not cover (a prime
// factor above 7, or a line too long for them): X[k] = sum_m x[m] w^(k m) with
// w = exp(-2 pi i / n) (conjugated for the inverse), accumulated in double from
// a double twiddle table, (k m) mod n kept by increments.  One block per line,
// staged in LDS as double; O(n^2) per line.  Lines are addressed as in
// k_smooth_axis.  OP: kDftFwd / kDftInv (IO as k_fft_axis), or kDftSmooth (a
// real line in place: forward, times gain[k], inverse, real part).
enum FftIO { kFftComplex = 0, kFftRealIn = 1, kFftRealOut = 2 };
constexpr int kDirectMax = 4096;  // two lines of double complex in LDS
constexpr int kDftThreads = 256;
enum DftOp { kDftFwd = 0, kDftInv = 1, kDftSmooth = 2 };

__device__ __forceinline__ void dft_line(const Cx<double>* __restrict__ in, Cx<double>* __restrict__ out,
                                         int n, const Cx<double>* __restrict__ tw, bool inverse) {
  for (int k = threadIdx.x; k < n; k += kDftThreads) {
    double re = 0.0, im = 0.0;
    int idx = 0;
    for (int m = 0; m < n; m++) {
      const Cx<double> x = in[m], w = tw[idx];
      const double wi = inverse ? -w.im : w.im;
      re += x.re * w.re - x.im * wi;
      im += x.re * wi + x.im * w.re;
      idx += k;
      if (idx >= n) idx -= n;
    }
    out[k] = Cx<double>{re, im};
  }
}

template <typename R, int OP, int IO>
__global__ __launch_bounds__(kDftThreads) void k_dft_axis(const void* src, void* dst, int n, long long nlines,
                                                          int I, long long ostride, long long istride,
                                                          long long es, const Cx<double>* __restrict__ tw,
                                                          const R* __restrict__ gain) {
  extern __shared__ double smem[];
  Cx<double>* a = reinterpret_cast<Cx<double>*>(smem);
  Cx<double>* b = a + n;
  const long long q = blockIdx.x;
  if (q >= nlines) return;
  const long long base = (q / I) * ostride + (q % I) * istride;
  for (int e = threadIdx.x; e < n; e += kDftThreads) {
    const long long at = base + e * es;
    if (OP == kDftSmooth || IO == kFftRealIn) {
      a[e] = Cx<double>{static_cast<double>(static_cast<const R*>(src)[at]), 0.0};
    } else {
      const Cx<R> v = static_cast<const Cx<R>*>(src)[at];
      a[e] = Cx<double>{static_cast<double>(v.re), static_cast<double>(v.im)};
    }
  }
  __syncthreads();
  dft_line(a, b, n, tw, OP == kDftInv);
  __syncthreads();
  if (OP == kDftSmooth) {
    // the gain in the mesh precision (with its 1 / n), applied as the radix kernels do
    for (int k = threadIdx.x; k < n; k += kDftThreads) {
      const R g = gain[k];
      const Cx<R> v{static_cast<R>(b[k].re) * g, static_cast<R>(b[k].im) * g};
      b[k] = Cx<double>{static_cast<double>(v.re), static_cast<double>(v.im)};
    }
    __syncthreads();
    dft_line(b, a, n, tw, true);
    __syncthreads();
    for (int e = threadIdx.x; e < n; e += kDftThreads)
      static_cast<R*>(dst)[base + e * es] = static_cast<R>(a[e].re);
    return;
  }
  for (int e = threadIdx.x; e < n; e += kDftThreads) {
    const long long at = base + e * es;
    if (IO == kFftRealOut)
      static_cast<R*>(dst)[at] = static_cast<R>(b[e].re);
    else
      static_cast<Cx<R>*>(dst)[at] = Cx<R>{static_cast<R>(b[e].re), static_cast<R>(b[e].im)};
  }
}

template <typename R, int OP, int IO>
static int launch_dft_axis(const void* src, void* dst, int n, long long nlines, int I, long long ostride,
                           long long istride, long long es, const double* tw, const R* gain,
                           hipStream_t s) {
  const size_t lds = 2 * sizeof(Cx<double>) * static_cast<size_t>(n);
  const void* fn = reinterpret_cast<const void*>(&k_dft_axis<R, OP, IO>);
  if (lds > 65536)
    PRH_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
  k_dft_axis<R, OP, IO><<<static_cast<unsigned>(nlines), kDftThreads, lds, s>>>(
      src, dst, n, nlines, I, ostride, istride, es, reinterpret_cast<const Cx<double>*>(tw), gain);
  PRH_LAUNCHED();
  return PRH_OK;
}

// gain[k] = exp(-(r k_phys)^2 / 2) / n with k_phys = 2 pi fftfreq(n, L / n)[k]
// (numpy's ordering: k < (n - 1) / 2 + 1 -> k, else k - n), in double.
template <typename R>
__global__ void k_axis_gain(R* __restrict__ gain, int n, double val, double rad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int f = k < (n - 1) / 2 + 1 ? k : k - n;
  constexpr double twopi = 2 * 3.141592653589793;
  const double kp = twopi * (static_cast<double>(f) * val);
  const double a = rad * kp;
  gain[k] = static_cast<R>(exp(-0.5 * (a * a)) / static_cast<double>(n));
}

static int ceil_log2(int v) {
  int l = 0;
  while ((1 << l) < v) l++;
  return l;
}

// Radix sequence 8, ..., 8, (4), (2), then the odd factors 3, 5, 7, so that the
// span of every power-of-two pass, and of the first odd one, is a power of two.
static bool factor_plan(int n, AxisPlan& P) {
  P.n = n;
  P.npass = 0;
  P.ln = ceil_log2(n);
  int rads[24], nr = 0, m = n;
  while (m % 8 == 0) {
    rads[nr++] = 8;
    m /= 8;
  }
  while (m % 4 == 0) {
    rads[nr++] = 4;
    m /= 4;
  }
  while (m % 2 == 0) {
    rads[nr++] = 2;
    m /= 2;
  }
  for (int q : {3, 5, 7})
    while (m % q == 0 && nr < 24) {
      rads[nr++] = q;
      m /= q;
    }
  if (m != 1) return false;
  int ns = 1;
  for (int i = 0; i < nr; i++) {
    Pass& S = P.pass[P.npass++];
    S.rad = rads[i];
    S.ns = ns;
    S.ns_shift = (ns & (ns - 1)) == 0 ? ceil_log2(ns) : -1;
    S.m = n / rads[i];
    S.lm = ceil_log2(S.m);
    S.span = n / (ns * rads[i]);
    ns *= rads[i];
  }
  return true;
}

// log2 of the butterflies a block can hold in one pass
static int fly_log2() { return ceil_log2(kMaxFly * kSmoothThreads); }

// largest log2 butterfly range of the passes of a plan
static int plan_max_lm(const AxisPlan& P) {
  int lm = 0;
  for (int i = 0; i < P.npass; i++) lm = P.pass[i].lm > lm ? P.pass[i].lm : lm;
  return lm;
}

int smooth_setup(prh_workspace* ws) {
  SmoothState& S = ws->smooth;
  S.have = true;
  const size_t w = static_cast<size_t>(ws->dtype);
  for (int d = 0; d < 3; d++) {
    const int n = static_cast<int>(ws->n[d]);
    AxisPlan P;
    S.direct[d] = !factor_plan(n, P) || (padded_len(n) + n) * 2 * w > kSmoothLdsMax ||
                  plan_max_lm(P) > fly_log2();
    if (S.direct[d] && n > kDirectMax) {
      S.have = false;  // transforms refused at call time for this shape
      continue;
    }
    PRH_TRY(ws_alloc(ws, &S.tw[d], 2 * w * n));
    PRH_TRY(ws_alloc(ws, &S.gain[d], w * n));
    std::vector<double> tw(2 * static_cast<size_t>(n));
    for (int k = 0; k < n; k++) {
      const double ang = -2.0 * M_PI * static_cast<double>(k) / static_cast<double>(n);
      tw[2 * k] = std::cos(ang);
      tw[2 * k + 1] = std::sin(ang);
    }
    if (S.direct[d]) {
      void* p;
      PRH_TRY(ws_alloc(ws, &p, 2 * sizeof(double) * n));
      S.tw64[d] = static_cast<double*>(p);
      PRH_HIP(hipMemcpy(S.tw64[d], tw.data(), 2 * sizeof(double) * n, hipMemcpyHostToDevice));
    }
    if (ws->dtype == PRH_F64) {
      PRH_HIP(hipMemcpy(S.tw[d], tw.data(), 2 * w * n, hipMemcpyHostToDevice));
    } else {
      std::vector<float> tf(tw.begin(), tw.end());
      PRH_HIP(hipMemcpy(S.tw[d], tf.data(), 2 * w * n, hipMemcpyHostToDevice));
    }
  }
  return PRH_OK;
}

size_t smooth_bytes(const size_t* n, int dtype) {
  size_t b = 0;
  for (int d = 0; d < 3; d++) {
    b += 3 * static_cast<size_t>(dtype) * n[d];
    AxisPlan P;
    const bool direct = !factor_plan(static_cast<int>(n[d]), P) ||
                        (padded_len(n[d]) + n[d]) * 2 * static_cast<size_t>(dtype) > kSmoothLdsMax ||
                        plan_max_lm(P) > fly_log2();
    if (direct) b += 2 * sizeof(double) * n[d];
  }
  return b;
}

template <typename R>
static int do_smooth(prh_workspace* ws, R* mesh, const size_t* nmesh, const R* boxsize,
                     const double* radius, void* stream) {
  if (!ws || !mesh || !nmesh || !boxsize || !radius) return PRH_ERR_ARG;
  if (!(ws->flags & PRH_WS_FFT)) {
    set_error("workspace was built without smoothing tables (PRH_WS_FFT)");
    return PRH_ERR_STATE;
  }
  if (!ws->smooth.have) {
    set_error("smoothing supports dimensions of at most %d cells when they have a prime factor above 7 "
              "or exceed the radix passes", kDirectMax);
    return PRH_ERR_ARG;
  }
  if (ws->dtype != Tag<R>::code) {
    set_error("workspace precision does not match the mesh");
    return PRH_ERR_ARG;
  }
  for (int d = 0; d < 3; d++)
    if (ws->n[d] != nmesh[d]) {
      set_error("workspace mesh does not match nmesh");
      return PRH_ERR_CAPACITY;
    }
  hipStream_t s = as_stream(stream);
  const long long n0 = static_cast<long long>(nmesh[0]), n1 = static_cast<long long>(nmesh[1]),
                  n2 = static_cast<long long>(nmesh[2]);
  for (int d = 0; d < 3; d++) {
    // d_i = boxsize / nmesh (float64), val = 1 / (n d) as in numpy.fft.fftfreq; the radius is
    // held in the mesh dtype (mesh.py:664-666)
    const double cell = static_cast<double>(boxsize[d]) / static_cast<double>(static_cast<int>(nmesh[d]));
    const double val = 1.0 / (static_cast<double>(nmesh[d]) * cell);
    const double rad = static_cast<double>(static_cast<R>(radius[d]));
    const int n = static_cast<int>(nmesh[d]);
    k_axis_gain<R><<<ceil_div(n, 256), 256, 0, s>>>(static_cast<R*>(ws->smooth.gain[d]), n, val, rad);
    PRH_LAUNCHED();
  }
  // axis order z, y, x: the z pass reads whole rows
  for (int d = 2; d >= 0; d--) {
    AxisPlan P;
    factor_plan(static_cast<int>(nmesh[d]), P);
    const int n = P.n;
    // lines per block: the largest power of two whose buffer fits kSmoothLds and whose
    // butterflies fit the block's registers in every pass
    int llp = 0;
    auto bytes = [&](int l) {
      return (padded_len((size_t(1) << l) * static_cast<size_t>(n)) + n) * 2 * sizeof(R);
    };
    const int lm = plan_max_lm(P);
    while (llp < 5 && bytes(llp + 1) <= kSmoothLds && llp + 1 + lm <= fly_log2()) llp++;
    const int LP = 1 << llp;
    long long nreal, ostride, istride, es;
    int I;
    if (d == 2) {
      nreal = n0 * n1; I = static_cast<int>(n1); ostride = n1 * n2; istride = n2; es = 1;
    } else if (d == 1) {
      nreal = n0 * n2; I = static_cast<int>(n2); ostride = n1 * n2; istride = 1; es = n2;
    } else {
      nreal = n1 * n2; I = static_cast<int>(n2); ostride = n2; istride = 1; es = n1 * n2;
    }
    if (ws->smooth.direct[d]) {
      PRH_TRY((launch_dft_axis<R, kDftSmooth, 0>(mesh, mesh, n, nreal, I, ostride, istride, es, ws->smooth.tw64[d],
                                                 static_cast<const R*>(ws->smooth.gain[d]), s)));
      continue;
    }
    const long long ncomplex = (nreal + 1) / 2;
    const unsigned grid = static_cast<unsigned>((ncomplex + LP - 1) / LP);
    const size_t lds = bytes(llp);
    const Cx<R>* tw = static_cast<const Cx<R>*>(ws->smooth.tw[d]);
    const R* gain = static_cast<const R*>(ws->smooth.gain[d]);
    if (lds > 65536) {
      PRH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_smooth_axis<R, true>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
      PRH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_smooth_axis<R, false>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    }
    if (d == 2 && n == 1024 && PRH_SMOOTH_REG512) {
      const long long ncx = (nreal + 1) / 2;
      const long long nb = (ncx + kReg512Waves - 1) / kReg512Waves;
      const unsigned g1024 = static_cast<unsigned>(nb < 8192 ? nb : 8192);
      k_smooth_reg1024<R><<<g1024, 64 * kReg512Waves, 0, s>>>(mesh, nreal, tw, gain);
    } else if (n == 1024 && PRH_SMOOTH_REG512 && I % 16 == 0 && istride == 1) {
      const long long ngroups = nreal / 16;
      const unsigned g1024 = static_cast<unsigned>(ngroups < 8192 ? ngroups : 8192);
      k_smooth_reg1024s<R><<<g1024, 64 * kReg512SWaves, 0, s>>>(mesh, nreal, I, ostride, es, tw, gain);
    } else if (d == 2 && n == 512 && PRH_SMOOTH_REG512) {
      const long long ncx = (nreal + 1) / 2;
      const long long nb = (ncx + kReg512Waves - 1) / kReg512Waves;
      const unsigned g512 = static_cast<unsigned>(nb < 8192 ? nb : 8192);
      k_smooth_reg512<R><<<g512, 64 * kReg512Waves, 0, s>>>(mesh, nreal, tw, gain);
    } else if (d == 2) {
      k_smooth_axis<R, true><<<grid, kSmoothThreads, lds, s>>>(mesh, P, nreal, I, ostride, istride,
                                                               es, llp, tw, gain);
    } else if (n == 512 && PRH_SMOOTH_REG512 && I % 16 == 0 && istride == 1) {
      const long long ngroups = nreal / 16;
      const unsigned g512 = static_cast<unsigned>(ngroups < 8192 ? ngroups : 8192);
      k_smooth_reg512s<R><<<g512, 64 * kReg512SWaves, 0, s>>>(mesh, nreal, I, ostride, es, tw, gain);
    }
    else
      k_smooth_axis<R, false><<<grid, kSmoothThreads, lds, s>>>(mesh, P, nreal, I, ostride,
                                                                istride, es, llp, tw, gain);
    PRH_LAUNCHED();
  }
  return PRH_OK;
}


// ---------------------------------------------------------------------------
// 3-D complex transforms and Fourier-space operators of the FFT
// reconstructions (pyrecon/plane_parallel_fft.py, pyrecon/iterative_fft.py).
//
// The reference transforms with numpy's c2c fftn / ifftn (RealMesh.to_complex
// and ComplexMesh.to_real, pyrecon/mesh.py:697-705, 827-844, 974-988).  Here a
// 3-D transform is three passes of 1-D transforms, one per axis, over an
// interleaved complex mesh of the same (n0, n1, n2) shape, with the Stockham
// machinery of the smoothing.  The forward pass along z may read a real mesh
// (imaginary part 0), and the last inverse pass may write the real part only,
// so no separate real <-> complex copies are made.  No pass normalises: the
// 1 / N of the inverse is folded into the Fourier-space operator that precedes
// it.
// ---------------------------------------------------------------------------

// The lines of one axis (same addressing as k_smooth_axis, in elements of the
// mesh); block b takes complex lines [b LP, (b + 1) LP), remapped so that
// consecutive bundles run on one XCD.  With kFftComplex, src and dst may be the
// same buffer: a block reads all its lines before it writes any, and blocks
// own disjoint lines.
template <typename R, bool CONTIG, int IO>
__global__ __launch_bounds__(kSmoothThreads) void k_fft_axis(const void* src, void* dst, AxisPlan P,
                                                             long long nlines, int I, long long ostride,
                                                             long long istride, long long es, int llp,
                                                             const Cx<R>* __restrict__ tw_global,
                                                             int inverse) {
  extern __shared__ double smem[];
  const int LP = 1 << llp;
  const int n = P.n;
  Cx<R>* b0 = reinterpret_cast<Cx<R>*>(smem);
  Cx<R>* stw = b0 + padded_len(static_cast<size_t>(LP) * n);
  for (int k = threadIdx.x; k < n; k += kSmoothThreads) stw[k] = tw_global[k];
  const unsigned nb = gridDim.x, bid = blockIdx.x, per = nb / 8;
  const unsigned blk = bid < per * 8 ? (bid % 8) * per + bid / 8 : bid;
  const long long l0 = static_cast<long long>(blk) * LP;
  auto base = [&](long long q) { return (q / I) * ostride + (q % I) * istride; };
  const int nel = CONTIG ? LP << P.ln : LP * n;
  constexpr int kBatch = 8;
  for (int t0 = threadIdx.x; t0 < nel; t0 += kBatch * kSmoothThreads) {
    Cx<R> val[kBatch];
    int slot[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; u++) {
      const int t = t0 + u * kSmoothThreads;
      int lp, e;
      if (CONTIG) {
        lp = t >> P.ln;
        e = t & ((1 << P.ln) - 1);
      } else {
        e = t >> llp;
        lp = t & (LP - 1);
      }
      const long long q = l0 + lp;
      const bool ok = t < nel && e < n;
      val[u] = Cx<R>{R(0), R(0)};
      if (ok && q < nlines) {
        const long long at = base(q) + e * es;
        if (IO == kFftRealIn)
          val[u].re = static_cast<const R*>(src)[at];
        else
          val[u] = static_cast<const Cx<R>*>(src)[at];
      }
      slot[u] = ok ? lds_pad(lp * n + e) : -1;
    }
#pragma unroll
    for (int u = 0; u < kBatch; u++)
      if (slot[u] >= 0) b0[slot[u]] = val[u];
  }
  __syncthreads();
  fft_lines<R>(b0, llp, P, stw, inverse != 0);
  for (int t = threadIdx.x; t < nel; t += kSmoothThreads) {
    int lp, e;
    if (CONTIG) {
      lp = t >> P.ln;
      e = t & ((1 << P.ln) - 1);
      if (e >= n) continue;
    } else {
      e = t >> llp;
      lp = t & (LP - 1);
    }
    const long long q = l0 + lp;
    if (q >= nlines) continue;
    const long long at = base(q) + e * es;
    const Cx<R> v = b0[lds_pad(lp * n + e)];
    if (IO == kFftRealOut)
      static_cast<R*>(dst)[at] = v.re;
    else
      static_cast<Cx<R>*>(dst)[at] = v;
  }
}

template <typename R, bool CONTIG, int IO>
static int launch_fft_axis(const AxisPlan& P, const void* src, void* dst, long long nlines, int I,
                           long long ostride, long long istride, long long es, int llp, size_t lds,
                           const Cx<R>* tw, int inverse, hipStream_t s) {
  const void* fn = reinterpret_cast<const void*>(&k_fft_axis<R, CONTIG, IO>);
  if (lds > 65536)
    PRH_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
  const long long LP = 1LL << llp;
  const unsigned grid = static_cast<unsigned>((nlines + LP - 1) / LP);
  k_fft_axis<R, CONTIG, IO><<<grid, kSmoothThreads, lds, s>>>(src, dst, P, nlines, I, ostride, istride,
                                                              es, llp, tw, inverse);
  PRH_LAUNCHED();
  return PRH_OK;
}

// One pass of 1-D transforms along axis d of an (n0, n1, n2) mesh.
template <typename R>
static int fft_pass(prh_workspace* ws, int d, const void* src, void* dst, int io, int inverse,
                    hipStream_t s) {
  AxisPlan P;
  factor_plan(static_cast<int>(ws->n[d]), P);
  const int n = P.n;
  const long long n0 = static_cast<long long>(ws->n[0]), n1 = static_cast<long long>(ws->n[1]),
                  n2 = static_cast<long long>(ws->n[2]);
  auto bytes = [&](int l) {
    return (padded_len((size_t(1) << l) * static_cast<size_t>(n)) + n) * 2 * sizeof(R);
  };
  // lines per block: as many as fit kSmoothLds and the block's butterflies (at most 32)
  const int lm = plan_max_lm(P);
  int llp = 0;
  while (llp < 5 && bytes(llp + 1) <= kSmoothLds && llp + 1 + lm <= fly_log2()) llp++;
  long long nlines, ostride, istride, es;
  int I;
  if (d == 2) {
    nlines = n0 * n1; I = static_cast<int>(n1); ostride = n1 * n2; istride = n2; es = 1;
  } else if (d == 1) {
    nlines = n0 * n2; I = static_cast<int>(n2); ostride = n1 * n2; istride = 1; es = n2;
  } else {
    nlines = n1 * n2; I = static_cast<int>(n2); ostride = n2; istride = 1; es = n1 * n2;
  }
  const Cx<R>* tw = static_cast<const Cx<R>*>(ws->smooth.tw[d]);
  if (ws->smooth.direct[d]) {
    const double* t64 = ws->smooth.tw64[d];
    if (inverse) {
      if (io == kFftRealOut)
        return launch_dft_axis<R, kDftInv, kFftRealOut>(src, dst, n, nlines, I, ostride, istride, es, t64, nullptr, s);
      return launch_dft_axis<R, kDftInv, kFftComplex>(src, dst, n, nlines, I, ostride, istride, es, t64, nullptr, s);
    }
    if (io == kFftRealIn)
      return launch_dft_axis<R, kDftFwd, kFftRealIn>(src, dst, n, nlines, I, ostride, istride, es, t64, nullptr, s);
    return launch_dft_axis<R, kDftFwd, kFftComplex>(src, dst, n, nlines, I, ostride, istride, es, t64, nullptr, s);
  }
  const size_t lds = bytes(llp);
  if (d == 2) {
    if (io == kFftRealIn)
      return launch_fft_axis<R, true, kFftRealIn>(P, src, dst, nlines, I, ostride, istride, es, llp, lds,
                                                  tw, inverse, s);
    if (io == kFftRealOut)
      return launch_fft_axis<R, true, kFftRealOut>(P, src, dst, nlines, I, ostride, istride, es, llp, lds,
                                                   tw, inverse, s);
    return launch_fft_axis<R, true, kFftComplex>(P, src, dst, nlines, I, ostride, istride, es, llp, lds,
                                                 tw, inverse, s);
  }
  if (io == kFftRealIn)
    return launch_fft_axis<R, false, kFftRealIn>(P, src, dst, nlines, I, ostride, istride, es, llp, lds,
                                                 tw, inverse, s);
  if (io == kFftRealOut)
    return launch_fft_axis<R, false, kFftRealOut>(P, src, dst, nlines, I, ostride, istride, es, llp, lds,
                                                  tw, inverse, s);
  return launch_fft_axis<R, false, kFftComplex>(P, src, dst, nlines, I, ostride, istride, es, llp, lds,
                                                tw, inverse, s);
}

static int check_transform_ws(prh_workspace* ws, const size_t* nmesh, int code) {
  if (!(ws->flags & PRH_WS_FFT)) {
    set_error("workspace was built without Fourier tables (PRH_WS_FFT)");
    return PRH_ERR_STATE;
  }
  if (!ws->smooth.have) {
    set_error("Fourier transforms support dimensions of at most %d cells when they have a prime factor "
              "above 7 or exceed the radix passes", kDirectMax);
    return PRH_ERR_ARG;
  }
  if (ws->dtype != code) {
    set_error("workspace precision does not match the mesh");
    return PRH_ERR_ARG;
  }
  for (int d = 0; d < 3; d++)
    if (ws->n[d] != nmesh[d]) {
      set_error("workspace mesh does not match nmesh");
      return PRH_ERR_CAPACITY;
    }
  return PRH_OK;
}

// Forward: src real (PRH_FFT_REAL_IN) or complex -> dst complex, axes z, y, x.
// Inverse: src complex (transformed in place along x and y) -> dst complex, or
// its real part (PRH_FFT_REAL_OUT), along z last.
// One pass of a 3-D transform, in the PRH_TIMER_FFT_PASS timer: algorithmic bytes 4 n^3 W for a
// complex pass, 3 n^3 W for the real-in and real-out passes.
template <typename R>
static int timed_pass(prh_workspace* ws, int d, const void* src, void* dst, int io, int inverse,
                      hipStream_t s) {
  timer_before(PRH_TIMER_FFT_PASS, s);
  PRH_TRY(fft_pass<R>(ws, d, src, dst, io, inverse, s));
  const double cells = static_cast<double>(ws->n[0]) * ws->n[1] * ws->n[2];
  timer_after(PRH_TIMER_FFT_PASS, s, cells * sizeof(R) * (io == kFftComplex ? 4 : 3));
  return PRH_OK;
}

template <typename R>
static int do_fft3d(prh_workspace* ws, const void* src, void* dst, const size_t* nmesh, int inverse,
                    int io, void* stream) {
  if (!ws || !src || !dst || !nmesh || io < 0 || io > 2) return PRH_ERR_ARG;
  if ((io == kFftRealIn && inverse) || (io == kFftRealOut && !inverse)) {
    set_error("a real input goes with the forward transform, a real output with the inverse");
    return PRH_ERR_ARG;
  }
  PRH_TRY(check_transform_ws(ws, nmesh, Tag<R>::code));
  hipStream_t s = as_stream(stream);
  if (!inverse) {
    PRH_TRY(timed_pass<R>(ws, 2, src, dst, io, 0, s));
    PRH_TRY(timed_pass<R>(ws, 1, dst, dst, kFftComplex, 0, s));
    PRH_TRY(timed_pass<R>(ws, 0, dst, dst, kFftComplex, 0, s));
  } else {
    void* work = const_cast<void*>(src);
    PRH_TRY(timed_pass<R>(ws, 0, work, work, kFftComplex, 1, s));
    PRH_TRY(timed_pass<R>(ws, 1, work, work, kFftComplex, 1, s));
    PRH_TRY(timed_pass<R>(ws, 2, work, dst, io, 1, s));
  }
  return PRH_OK;
}

// Fourier-space operators.  k_i = 2 pi fftfreq(n_i, L_i / n_i) (ComplexMesh.
// coords, pyrecon/mesh.py:812-820) is formed per cell from its index exactly
// as numpy forms it: (2 pi) * (f * val) with val = 1 / (n_i (L_i / n_i)).
// The arithmetic is double (the reference's wavenumbers are float64 arrays);
// values the reference stores in the mesh precision are rounded to R where it
// stores them.
struct SpecArgs {
  double val[3];
  double los[3];
  double beta;
  double norm;  // factor on the output (1 / N before an inverse transform)
  int n[3];
  int a, b;     // axes of the operator
};

__device__ __forceinline__ double wavenumber(int i, int n, double val) {
  const int f = i < (n - 1) / 2 + 1 ? i : i - n;
  constexpr double twopi = 2 * 3.141592653589793;
  return twopi * (static_cast<double>(f) * val);
}

enum SpecOp {
  kSpecPlaneParallel = 0,  // dst = i k_a (src / k^2) / (1 + beta mu^2) (plane_parallel_fft.py:45-53)
  kSpecInvK2 = 1,          // dst = src / (k_y^2 + k_x^2 + k_z^2) in R, 0 at k = 0 (iterative_fft.py:37-39)
  kSpecLosSq = 2,          // dst = (k . los)^2 src (iterative_fft.py:44)
  kSpecKaKb = 3,           // dst = k_a k_b src (iterative_fft.py:58)
  kSpecGrad = 4,           // dst = i k_a (src / k^2), 0 at k = 0 (iterative_fft.py:75-81)
  kSpecIKa = 5,            // dst = (src i) k_a (iterative_fft_particle.py:112)
};

template <typename R, int OP>
__global__ __launch_bounds__(256) void k_spectral(const Cx<R>* src, Cx<R>* dst, SpecArgs A) {
  const int iz = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y;
  const int ix = blockIdx.z;
  if (iz >= A.n[2]) return;
  const size_t i = (static_cast<size_t>(ix) * A.n[1] + iy) * A.n[2] + iz;
  const double k[3] = {wavenumber(ix, A.n[0], A.val[0]), wavenumber(iy, A.n[1], A.val[1]),
                       wavenumber(iz, A.n[2], A.val[2])};
  const Cx<R> c = src[i];
  const bool zero = (ix | iy | iz) == 0;
  Cx<R> out;
  if constexpr (OP == kSpecInvK2) {
    // RealMesh.prod_sum(exp=-1) (src/mesh.c:327-365): the squared wavenumbers are cast to R and
    // summed in R as (y + x) + z; complex parts are divided separately
    const R k2 = (static_cast<R>(k[1] * k[1]) + static_cast<R>(k[0] * k[0])) + static_cast<R>(k[2] * k[2]);
    out = zero ? Cx<R>{R(0), R(0)} : Cx<R>{c.re / k2, c.im / k2};
  } else if constexpr (OP == kSpecLosSq) {
    const double kl = ((0. + k[0] * A.los[0]) + k[1] * A.los[1]) + k[2] * A.los[2];
    const double m = kl * kl;
    out = {static_cast<R>(m * static_cast<double>(c.re) * A.norm),
           static_cast<R>(m * static_cast<double>(c.im) * A.norm)};
  } else if constexpr (OP == kSpecIKa) {
    // (src * 1j) = (-im, re), then times k_a
    const double ka = k[A.a];
    out = {static_cast<R>(-(static_cast<double>(c.im) * ka) * A.norm),
           static_cast<R>(static_cast<double>(c.re) * ka * A.norm)};
  } else if constexpr (OP == kSpecKaKb) {
    const double m = k[A.a] * k[A.b];
    out = {static_cast<R>(m * static_cast<double>(c.re) * A.norm),
           static_cast<R>(m * static_cast<double>(c.im) * A.norm)};
  } else {
    // k^2 as the Python sum of the broadcast squares, 1 at k = 0; the spectrum divided in place
    // (rounded to the mesh precision)
    const double k2 = zero ? 1. : ((0. + k[0] * k[0]) + k[1] * k[1]) + k[2] * k[2];
    const double dre = static_cast<double>(static_cast<R>(static_cast<double>(c.re) / k2));
    const double dim = static_cast<double>(static_cast<R>(static_cast<double>(c.im) / k2));
    const double ka = k[A.a];
    // (i k_a) d = (-k_a d.im, k_a d.re)
    double re = -(ka * dim), im = ka * dre;
    if constexpr (OP == kSpecPlaneParallel) {
      const double kl = ((0. + k[0] * A.los[0]) + k[1] * A.los[1]) + k[2] * A.los[2];
      const double den = 1. + A.beta * (kl * kl / k2);
      re /= den;
      im /= den;
      out = {static_cast<R>(re * A.norm), static_cast<R>(im * A.norm)};
    } else {
      out = zero ? Cx<R>{R(0), R(0)} : Cx<R>{static_cast<R>(re * A.norm), static_cast<R>(im * A.norm)};
    }
  }
  dst[i] = out;
}

template <typename R>
static int do_spectral(prh_workspace* ws, int op, const void* src, void* dst, const size_t* nmesh,
                       const R* boxsize, const double* los, double beta, double norm, int a, int b,
                       void* stream) {
  if (!ws || !src || !dst || !nmesh || !boxsize || op < 0 || op > 5 || a < 0 || a > 2 || b < 0 ||
      b > 2)
    return PRH_ERR_ARG;
  if ((op == kSpecPlaneParallel || op == kSpecLosSq) && !los) return PRH_ERR_ARG;
  PRH_TRY(check_transform_ws(ws, nmesh, Tag<R>::code));
  SpecArgs A;
  for (int d = 0; d < 3; d++) {
    A.n[d] = static_cast<int>(nmesh[d]);
    // d = boxsize / nmesh (float64, as numpy divides the mesh-dtype boxsize by the int nmesh)
    const double cell = static_cast<double>(boxsize[d]) / static_cast<double>(A.n[d]);
    A.val[d] = 1.0 / (static_cast<double>(A.n[d]) * cell);
    A.los[d] = los ? los[d] : 0.;
  }
  A.beta = beta;
  A.norm = norm;
  A.a = a;
  A.b = b;
  const dim3 grid(ceil_div(nmesh[2], 256), A.n[1], A.n[0]);
  const Cx<R>* in = static_cast<const Cx<R>*>(src);
  Cx<R>* out = static_cast<Cx<R>*>(dst);
  hipStream_t s = as_stream(stream);
  switch (op) {
    case kSpecPlaneParallel: k_spectral<R, kSpecPlaneParallel><<<grid, 256, 0, s>>>(in, out, A); break;
    case kSpecInvK2: k_spectral<R, kSpecInvK2><<<grid, 256, 0, s>>>(in, out, A); break;
    case kSpecLosSq: k_spectral<R, kSpecLosSq><<<grid, 256, 0, s>>>(in, out, A); break;
    case kSpecKaKb: k_spectral<R, kSpecKaKb><<<grid, 256, 0, s>>>(in, out, A); break;
    case kSpecIKa: k_spectral<R, kSpecIKa><<<grid, 256, 0, s>>>(in, out, A); break;
    default: k_spectral<R, kSpecGrad><<<grid, 256, 0, s>>>(in, out, A); break;
  }
  PRH_LAUNCHED();
  return PRH_OK;
}

// Real-space RSD removal of one iteration (iterative_fft.py:41-67): with a
// global line of sight (a < 0) delta_real = delta_real - factor * t; with a
// local one delta_real = delta_real - factor * ((x_a x_b) t / |x|^2), where
// x_i = offset_i + (L_i / n_i) j_i are the mesh coordinates (RealMesh.coords,
// pyrecon/mesh.py:522-527) in float64.
template <typename R>
__global__ __launch_bounds__(256) void k_rsd_update(R* __restrict__ delta, const R* __restrict__ t,
                                                    int n0, int n1, int n2, double o0, double o1,
                                                    double o2, double c0, double c1, double c2,
                                                    double factor, int a, int b) {
  const int iz = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y;
  const int ix = blockIdx.z;
  if (iz >= n2) return;
  const size_t i = (static_cast<size_t>(ix) * n1 + iy) * n2 + iz;
  double rsd = static_cast<double>(t[i]);
  if (a >= 0) {
    const double x[3] = {o0 + c0 * static_cast<double>(ix), o1 + c1 * static_cast<double>(iy),
                         o2 + c2 * static_cast<double>(iz)};
    const double x2 = ((0. + x[0] * x[0]) + x[1] * x[1]) + x[2] * x[2];
    rsd = ((x[a] * x[b]) * rsd) / x2;
  }
  delta[i] = static_cast<R>(static_cast<double>(delta[i]) - factor * rsd);
}

template <typename R>
static int do_rsd_update(R* delta, const R* t, const size_t* nmesh, const R* boxsize,
                         const R* boxcenter, double factor, int a, int b, void* stream) {
  if (!delta || !t || !nmesh || !boxsize || !boxcenter || a > 2 || b > 2 || (a >= 0 && b < 0))
    return PRH_ERR_ARG;
  double o[3], c[3];
  for (int d = 0; d < 3; d++) {
    // MeshInfo.offset in the mesh dtype (mesh.py:462-465); cellsize in float64
    o[d] = static_cast<double>(static_cast<R>(boxcenter[d] - boxsize[d] / R(2)));
    c[d] = static_cast<double>(boxsize[d]) / static_cast<double>(static_cast<int>(nmesh[d]));
  }
  const int n0 = static_cast<int>(nmesh[0]), n1 = static_cast<int>(nmesh[1]),
            n2 = static_cast<int>(nmesh[2]);
  const dim3 grid(ceil_div(nmesh[2], 256), n1, n0);
  k_rsd_update<R><<<grid, 256, 0, as_stream(stream)>>>(delta, t, n0, n1, n2, o[0], o[1], o[2], c[0],
                                                       c[1], c[2], factor, a, b);
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

int prh_fft3d_f32(prh_workspace* ws, const void* src, void* dst, const size_t* nmesh, int inverse,
                  int io, void* stream) {
  return do_fft3d<float>(ws, src, dst, nmesh, inverse, io, stream);
}
int prh_fft3d_f64(prh_workspace* ws, const void* src, void* dst, const size_t* nmesh, int inverse,
                  int io, void* stream) {
  return do_fft3d<double>(ws, src, dst, nmesh, inverse, io, stream);
}

int prh_spectral_f32(prh_workspace* ws, int op, const void* src, void* dst, const size_t* nmesh,
                     const float* boxsize, const double* los, double beta, double norm, int a, int b,
                     void* stream) {
  return do_spectral<float>(ws, op, src, dst, nmesh, boxsize, los, beta, norm, a, b, stream);
}
int prh_spectral_f64(prh_workspace* ws, int op, const void* src, void* dst, const size_t* nmesh,
                     const double* boxsize, const double* los, double beta, double norm, int a, int b,
                     void* stream) {
  return do_spectral<double>(ws, op, src, dst, nmesh, boxsize, los, beta, norm, a, b, stream);
}

int prh_rsd_update_f32(float* delta, const float* t, const size_t* nmesh, const float* boxsize,
                       const float* boxcenter, double factor, int a, int b, void* stream) {
  return do_rsd_update<float>(delta, t, nmesh, boxsize, boxcenter, factor, a, b, stream);
}
int prh_rsd_update_f64(double* delta, const double* t, const size_t* nmesh, const double* boxsize,
                       const double* boxcenter, double factor, int a, int b, void* stream) {
  return do_rsd_update<double>(delta, t, nmesh, boxsize, boxcenter, factor, a, b, stream);
}

}  // extern "C"
