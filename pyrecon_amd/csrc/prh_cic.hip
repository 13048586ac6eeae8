// Cloud-in-cell paint and reads on the device.
//
// Paint (replaces assign_cic, src/mesh.c:26-63, whose loop is serial).  The
// reference adds each point's 8 corner weights to the mesh in point order, so
// every cell receives its contributions in increasing point index.  Here,
// per chunk of points:
//   1. key:     base cell b(i) and fractional offsets of every point, with
//               the Cartesian -> mesh-unit transform, the wrap or range check
//               and the finite check of pyrecon/mesh.py:544-556 fused in;
//   2. sort:    a stable LSD radix sort of the (base cell, point index)
//               pairs (digits of up to 9 bits, one histogram, scan and
//               ranked scatter per digit), so that point indices come out
//               grouped by base cell and increasing within each group;
//   3. starts:  each cell's first position in the sorted order, from the
//               sorted keys;
//   4. permute: the payload (offsets, weight) into the sorted order;
//   5. gather:  each cell merges the 8 groups of the base cells that touch
//               it by point index and accumulates the corner weights into its
//               current value in that order.
// The arithmetic per contribution is the reference's ((wx*wy)*wz)*w, so the
// mesh is bit-identical to the reference's serial accumulation, and repeated
// runs give identical bits (no step depends on the timing of atomics).
// Chunks are applied in order, which preserves the point order across chunk
// boundaries.
#include <cstring>
#include <type_traits>
#include <utility>

#include "prh_reduce.hpp"

namespace prh {

constexpr int kCicThreads = 256;
constexpr uint32_t kErrRange = 1u, kErrNonFinite = 2u;

// Workgroups are dealt to the 8 XCDs (each with its own L2) round-robin by
// block index.  Kernels whose neighbouring work items share reads take their
// item as xcd_block(per) instead, on a grid of kXcds * per blocks: block b
// runs on XCD b % kXcds and takes item (b % kXcds) * per + b / kXcds, so each
// XCD covers a contiguous range of items; blocks past the last item return.
#ifndef PRH_XCD_MAP
#define PRH_XCD_MAP 1
#endif
constexpr int kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t per) {
  return PRH_XCD_MAP ? (blockIdx.x % kXcds) * per + blockIdx.x / kXcds : blockIdx.x;
}
static inline uint32_t xcd_per(size_t nitems) {
  return static_cast<uint32_t>((nitems + kXcds - 1) / kXcds);
}

template <typename A, typename B>
using promote_t = typename std::conditional<(sizeof(A) >= sizeof(B)), A, B>::type;

template <typename R>
struct Payload {
  R t[3];
  R w;
};

// numpy's float remainder (npy_divmod): fmod, moved to the sign of b.
template <typename T>
__device__ __forceinline__ T np_mod(T a, T b) {
  T m = fmod(a, b);
  if (m != T(0)) {
    if ((b < T(0)) != (m < T(0))) m += b;
  } else {
    m = copysign(T(0), b);
  }
  return m;
}

// Geometry of the Cartesian -> mesh-unit transform, in the mesh precision R.
template <typename R>
struct Frame {
  R bs[3], bc[3], off[3], lo[3], hi[3];
  int n[3];
  double nd[3];
};

template <typename R>
static Frame<R> make_frame(const size_t* nmesh, const R* boxsize, const R* boxcenter) {
  Frame<R> F;
  for (int d = 0; d < 3; d++) {
    F.n[d] = static_cast<int>(nmesh[d]);
    F.nd[d] = static_cast<double>(static_cast<int>(nmesh[d]));
    F.bs[d] = boxsize[d];
    F.bc[d] = boxcenter[d];
    // MeshInfo.offset and the range-check bounds, in R (mesh.py:462-465, 548)
    F.off[d] = boxcenter[d] - boxsize[d] / R(2);
    F.lo[d] = boxcenter[d] - boxsize[d] / R(2);
    F.hi[d] = boxcenter[d] + boxsize[d] / R(2);
  }
  return F;
}

// Transform of one Cartesian coordinate triple (positions dtype P) to mesh
// units in R with numpy's promotion: Q = promote(P, R) for the box arithmetic,
// float64 for the product with the int32 nmesh (mesh.py:554-555).  Sets
// error bits for out-of-box (no wrap) and non-finite coordinates.
template <typename R, typename P>
__device__ __forceinline__ void to_mesh_units(const P x[3], const Frame<R>& F, int wrap,
                                              R out[3], uint32_t& err) {
  using Q = promote_t<P, R>;
  for (int d = 0; d < 3; d++) {
    Q p;
    if (wrap) {
      p = np_mod(static_cast<Q>(x[d]) - static_cast<Q>(F.off[d]), static_cast<Q>(F.bs[d])) +
          static_cast<Q>(F.off[d]);
    } else {
      p = static_cast<Q>(x[d]);
      if (!(static_cast<Q>(x[d]) >= static_cast<Q>(F.lo[d]) &&
            static_cast<Q>(x[d]) <= static_cast<Q>(F.hi[d])))
        err |= kErrRange;
    }
    if (!isfinite(p)) err |= kErrNonFinite;
    Q a = p - static_cast<Q>(F.bc[d]);
    a = a / static_cast<Q>(F.bs[d]);
    a = a + static_cast<Q>(0.5);
    out[d] = static_cast<R>(static_cast<double>(a) * F.nd[d]);
  }
}

// ((int) p) % n with the reference's conversion of the int to size_t.
__device__ __forceinline__ int base_index(int ip, int n) {
  // the common cases without an integer division: in the mesh, or one period above
  if (ip >= 0 && ip < n) return ip;
  if (ip >= n && ip - n < n) return ip - n;
  if (ip >= 0) return ip % n;
  const unsigned long long u = static_cast<unsigned long long>(static_cast<long long>(ip));
  return static_cast<int>(u % static_cast<unsigned long long>(n));
}

// 64-bit hash of a catalog's position bytes: the wrapping sum over points of a
// mix of (index, x, y, z) bits (splitmix64's finaliser).  Integer sums do not
// depend on the summation order, so block partials may be combined in any
// order; the index in the mix makes a permutation of the points hash
// differently.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z ^= z >> 30;
  z *= 0xbf58476d1ce4e5b9ull;
  z ^= z >> 27;
  z *= 0x94d049bb133111ebull;
  z ^= z >> 31;
  return z;
}
__device__ __forceinline__ uint64_t word_bits(double x) { return __double_as_longlong(x); }
__device__ __forceinline__ uint64_t word_bits(float x) { return static_cast<uint32_t>(__float_as_int(x)); }
template <typename P>
__device__ __forceinline__ uint64_t point_hash(size_t i, const P x[3]) {
  uint64_t h = mix64(static_cast<uint64_t>(i) + 0x9e3779b97f4a7c15ull);
  for (int d = 0; d < 3; d++) h = mix64(h ^ word_bits(x[d]));
  return h;
}

// Sum of one value per thread over a block of kCicThreads threads, written to
// partial[blockIdx.x] (wrapping 64-bit integer sum).
__device__ __forceinline__ void block_hash_partial(uint64_t h, uint64_t* __restrict__ partial) {
  __shared__ uint64_t s_h[kCicThreads / 64];
  for (int off = 32; off > 0; off >>= 1) h += __shfl_xor(h, off);
  if ((threadIdx.x & 63) == 0) s_h[threadIdx.x >> 6] = h;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int w = 0; w < kCicThreads / 64; w++) t += s_h[w];
    partial[blockIdx.x] = t;
  }
}

// Wave-level OR of a paint kernel's error bits into misc[0], one atomic per wave.
__device__ __forceinline__ void flag_paint_errors(uint32_t err, uint32_t* misc) {
  const uint64_t bad = __ballot(err != 0);
  if (bad) {
    uint32_t e = err;
    for (int off = 32; off > 0; off >>= 1) e |= __shfl_xor(e, off);
    if ((threadIdx.x & 63) == 0) atomicOr(&misc[0], e);
  }
}

// out[0] = sum of partial[0, n) (one block of 1024 threads).
__global__ __launch_bounds__(1024) void k_hash_final(const uint64_t* __restrict__ partial, size_t n,
                                                     uint64_t* __restrict__ out) {
  __shared__ uint64_t s_h[16];
  uint64_t h = 0;
  for (size_t i = threadIdx.x; i < n; i += 1024) h += partial[i];
  for (int off = 32; off > 0; off >>= 1) h += __shfl_xor(h, off);
  if ((threadIdx.x & 63) == 0) s_h[threadIdx.x >> 6] = h;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int w = 0; w < 16; w++) t += s_h[w];
    out[0] = t;
  }
}

// Partial hashes of a catalog (the read's side of the comparison): one
// kCicThreads block per kCicThreads * kHashItems points.
constexpr int kHashItems = 16;
template <typename P>
__global__ __launch_bounds__(kCicThreads) void k_pos_hash(const P* __restrict__ pos, size_t m,
                                                          uint64_t* __restrict__ partial) {
  const size_t base = blockIdx.x * static_cast<size_t>(kCicThreads) * kHashItems + threadIdx.x;
  uint64_t h = 0;
#pragma unroll
  for (int j = 0; j < kHashItems; j++) {
    const size_t i = base + static_cast<size_t>(j) * kCicThreads;
    if (i < m) {
      const P x[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
      h += point_hash<P>(i, x);
    }
  }
  block_hash_partial(h, partial);
}

// ---------------------------------------------------------------------------
// Paint step 1 (MODE 0: positions already in mesh units, no checks;
// MODE 1: Cartesian positions, range check; MODE 2: Cartesian, wrapped)
// ---------------------------------------------------------------------------
template <typename R, typename P, typename W, int MODE>
__global__ __launch_bounds__(kCicThreads) void k_paint_key(
    const P* __restrict__ pos, const W* __restrict__ w, size_t m, Frame<R> F,
    uint32_t* __restrict__ key, Payload<R>* __restrict__ payload, uint32_t* __restrict__ misc,
    int check_only) {
  const size_t i = blockIdx.x * static_cast<size_t>(kCicThreads) + threadIdx.x;
  uint32_t err = 0;
  if (i < m) {
    const P x[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    R p[3];
    if (MODE == 0) {
      for (int d = 0; d < 3; d++) p[d] = static_cast<R>(x[d]);
    } else {
      to_mesh_units<R, P>(x, F, MODE == 2, p, err);
      if (w && !isfinite(w[i])) err |= kErrNonFinite;
    }
    if (!check_only) {
      int b[3];
      Payload<R> q;
      for (int d = 0; d < 3; d++) {
        b[d] = base_index(static_cast<int>(p[d]), F.n[d]);
        q.t[d] = p[d] - static_cast<R>(b[d]);
      }
      q.w = w ? static_cast<R>(w[i]) : R(1);
      const uint32_t k = (static_cast<uint32_t>(b[0]) * F.n[1] + b[1]) * F.n[2] + b[2];
      key[i] = k;
      payload[i] = q;
    }
  }
  // one atomic per wavefront with a non-zero error
  const uint64_t bad = __ballot(err != 0);
  if (bad) {
    uint32_t e = err;
    for (int off = 32; off > 0; off >>= 1) e |= __shfl_xor(e, off);
    if ((threadIdx.x & 63) == 0) atomicOr(&misc[0], e);
  }
}

// Exclusive scan of count[0, n) into start[0, n] (start[n] = total), zeroing
// count: the digit offsets of the radix sort and the tile starts of the read
// binning.
constexpr int kScanItems = 32;
constexpr int kScanBlock = kCicThreads * kScanItems;  // 8192 cells per block

__global__ __launch_bounds__(kCicThreads) void k_scan_reduce(const uint32_t* __restrict__ count,
                                                             size_t n,
                                                             uint32_t* __restrict__ partial) {
  __shared__ uint32_t lds[kCicThreads];
  const size_t base = blockIdx.x * static_cast<size_t>(kScanBlock);
  uint32_t acc = 0;
  for (int k = 0; k < kScanItems; k++) {
    const size_t i = base + k * kCicThreads + threadIdx.x;
    if (i < n) acc += count[i];
  }
  lds[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kCicThreads / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) lds[threadIdx.x] += lds[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = lds[0];
}

// Block-wide exclusive scan of one value per thread (256 threads).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds,
                                                         uint32_t& total) {
  lds[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < kCicThreads; off <<= 1) {
    uint32_t t = threadIdx.x >= static_cast<unsigned>(off) ? lds[threadIdx.x - off] : 0u;
    __syncthreads();
    lds[threadIdx.x] += t;
    __syncthreads();
  }
  total = lds[kCicThreads - 1];
  const uint32_t incl = lds[threadIdx.x];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(kCicThreads) void k_scan_partials(uint32_t* __restrict__ partial,
                                                               int np) {
  __shared__ uint32_t lds[kCicThreads];
  uint32_t carry = 0;
  for (int base = 0; base < np; base += kCicThreads) {
    const int i = base + threadIdx.x;
    const uint32_t v = i < np ? partial[i] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan(v, lds, total);
    if (i < np) partial[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) partial[np] = carry;
}

// Coalesced loads into LDS (padded by one word every 32 so that each thread's
// 32 consecutive cells sit in distinct banks), a per-thread serial scan of
// its 32 cells, a block scan of the thread totals, coalesced stores.
__global__ __launch_bounds__(kCicThreads) void k_scan_final(uint32_t* __restrict__ count,
                                                            size_t n,
                                                            const uint32_t* __restrict__ partial,
                                                            uint32_t* __restrict__ start) {
  __shared__ uint32_t tile[kScanBlock + kScanBlock / 32];
  __shared__ uint32_t lds[kCicThreads];
  const size_t base = blockIdx.x * static_cast<size_t>(kScanBlock);
  auto pad = [](int e) { return e + (e >> 5); };
  for (int k = 0; k < kScanItems; k++) {
    const int e = k * kCicThreads + threadIdx.x;
    const size_t i = base + e;
    tile[pad(e)] = i < n ? count[i] : 0u;
  }
  __syncthreads();
  uint32_t acc = 0;
  const int mine = threadIdx.x * kScanItems;
  for (int k = 0; k < kScanItems; k++) acc += tile[pad(mine + k)];
  uint32_t total;
  uint32_t run = block_exclusive_scan(acc, lds, total) + partial[blockIdx.x];
  for (int k = 0; k < kScanItems; k++) {
    const uint32_t c = tile[pad(mine + k)];
    tile[pad(mine + k)] = run;
    run += c;
  }
  __syncthreads();
  for (int k = 0; k < kScanItems; k++) {
    const int e = k * kCicThreads + threadIdx.x;
    const size_t i = base + e;
    if (i < n) {
      start[i] = tile[pad(e)];
      count[i] = 0u;
    }
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) start[n] = partial[gridDim.x];
}

// Counting-sort scatter of the read binning: index i to start[key[i]] + rank[i].
__global__ __launch_bounds__(kCicThreads) void k_group_scatter(
    const uint32_t* __restrict__ key, const uint32_t* __restrict__ rank, size_t m,
    const uint32_t* __restrict__ start, uint32_t* __restrict__ sorted) {
  const size_t i = blockIdx.x * static_cast<size_t>(kCicThreads) + threadIdx.x;
  if (i < m) sorted[start[key[i]] + rank[i]] = static_cast<uint32_t>(i);
}

// Paint step 2: stable LSD radix sort of (key, index) pairs, one digit of
// `bits` (<= 9) bits per pass.  A pass splits the pairs into tiles of
// kRadixTile; k_radix_count writes each tile's digit counts digit-major
// (hist[d * ntiles + tile]), their exclusive scan gives every (digit, tile)
// its first output position, and k_radix_scatter ranks each tile's pairs
// within their digit in input order and writes them there.  Tiles are dealt to
// the 8 XCDs in contiguous ranges, so that the counts and the output runs of
// neighbouring tiles meet in one L2.
constexpr int kRadixThreads = 256;
constexpr int kRadixWaves = kRadixThreads / 64;
constexpr int kRadixItems = 16;                             // per thread
constexpr int kRadixTile = kRadixThreads * kRadixItems;     // 4096 pairs
constexpr int kRadixWaveItems = 64 * kRadixItems;           // 1024 pairs per wave
constexpr int kRadixMaxBits = 9;
constexpr int kRadixMaxBins = 1 << kRadixMaxBits;

__global__ __launch_bounds__(kRadixThreads) void k_radix_count(const uint32_t* __restrict__ key,
                                                               size_t m, int shift, int bits,
                                                               uint32_t ntiles, uint32_t per,
                                                               uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[kRadixMaxBins];
  const uint32_t tile = xcd_block(per);
  if (tile >= ntiles) return;
  const int nb = 1 << bits;
  const uint32_t mask = static_cast<uint32_t>(nb - 1);
  for (int d = threadIdx.x; d < nb; d += kRadixThreads) h[d] = 0u;
  __syncthreads();
  const size_t base = static_cast<size_t>(tile) * kRadixTile;
  uint32_t k[kRadixItems];
#pragma unroll
  for (int j = 0; j < kRadixItems; j++) {
    const size_t i = base + j * kRadixThreads + threadIdx.x;
    k[j] = i < m ? key[i] : 0u;
  }
#pragma unroll
  for (int j = 0; j < kRadixItems; j++) {
    const size_t i = base + j * kRadixThreads + threadIdx.x;
    if (i < m) atomicAdd(&h[(k[j] >> shift) & mask], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < nb; d += kRadixThreads)
    hist[static_cast<size_t>(d) * ntiles + tile] = h[d];
}

// Wave-wide inclusive prefix sum (64 lanes).
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t u = __shfl_up(v, off);
    if (lane >= off) v += u;
  }
  return v;
}

// One pass of the sort.  Wave w of a tile holds the tile's pairs
// [w * 1024, (w + 1) * 1024) in 16 rounds of 64 consecutive pairs, so
// (wave, round, lane) is the input order.  In each round the lanes with equal
// digits find each other by `bits` ballots; a pair's rank is the number of
// equal-digit pairs before it in the wave (the wave's running count in LDS
// plus the equal lanes below it).  The tile then stages its pairs in LDS in
// (digit, rank) order and writes each digit's run to its scanned offset.
// `vin` null: the values are the input positions (first pass).
__global__ __launch_bounds__(kRadixThreads) void k_radix_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, size_t m, int shift,
    int bits, uint32_t ntiles, uint32_t per, const uint32_t* __restrict__ goff,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout) {
  __shared__ uint32_t s_wcnt[kRadixWaves][kRadixMaxBins];  // per-wave counts, then prefixes
  __shared__ uint32_t s_lbeg[kRadixMaxBins];               // digit's first slot in the tile
  __shared__ uint32_t s_gbeg[kRadixMaxBins];               // digit's first output position
  __shared__ uint32_t s_wsum[kRadixWaves];
  __shared__ uint32_t s_key[kRadixTile];
  __shared__ uint32_t s_val[kRadixTile];
  const uint32_t tile = xcd_block(per);
  if (tile >= ntiles) return;
  const int nb = 1 << bits;
  const uint32_t mask = static_cast<uint32_t>(nb - 1);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int e = threadIdx.x; e < kRadixWaves * kRadixMaxBins; e += kRadixThreads)
    (&s_wcnt[0][0])[e] = 0u;
  const size_t tbase = static_cast<size_t>(tile) * kRadixTile;
  const size_t wbase = tbase + static_cast<size_t>(w) * kRadixWaveItems;
  uint32_t k[kRadixItems], v[kRadixItems], r[kRadixItems];
#pragma unroll
  for (int j = 0; j < kRadixItems; j++) {
    const size_t i = wbase + j * 64 + lane;
    k[j] = i < m ? kin[i] : 0u;
    v[j] = i < m ? (vin ? vin[i] : static_cast<uint32_t>(i)) : 0u;
  }
  __syncthreads();
  const uint64_t below = (uint64_t(1) << lane) - 1u;
#pragma unroll
  for (int j = 0; j < kRadixItems; j++) {
    const bool ok = wbase + j * 64 + lane < m;
    const uint32_t d = (k[j] >> shift) & mask;
    uint64_t peers = __ballot(ok);
    for (int b = 0; b < bits; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t set = __ballot(bit);
      peers &= bit ? set : ~set;
    }
    const uint32_t before = static_cast<uint32_t>(__popcll(peers & below));
    const uint32_t prior = ok ? s_wcnt[w][d] : 0u;
    r[j] = prior + before;
    // the lowest lane of each digit advances the wave's count
    if (ok && before == 0u) s_wcnt[w][d] = prior + static_cast<uint32_t>(__popcll(peers));
  }
  __syncthreads();
  // per digit: exclusive prefix over the waves, and the tile's count
  uint32_t tc[2] = {0u, 0u};
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int d = 2 * threadIdx.x + q;
    uint32_t run = 0;
#pragma unroll
    for (int u = 0; u < kRadixWaves; u++) {
      const uint32_t c = s_wcnt[u][d];
      s_wcnt[u][d] = run;
      run += c;
    }
    tc[q] = run;
  }
  // exclusive scan of the tile counts over the digits (two digits per thread)
  const uint32_t pair = tc[0] + tc[1];
  const uint32_t incl = wave_inclusive_sum(pair);
  if (lane == 63) s_wsum[w] = incl;
  __syncthreads();
  uint32_t wpre = 0;
  for (int u = 0; u < w; u++) wpre += s_wsum[u];
  const uint32_t ex = wpre + incl - pair;
  {
    const int d0 = 2 * threadIdx.x;
    s_lbeg[d0] = ex;
    s_lbeg[d0 + 1] = ex + tc[0];
    if (d0 < nb) s_gbeg[d0] = goff[static_cast<size_t>(d0) * ntiles + tile];
    if (d0 + 1 < nb) s_gbeg[d0 + 1] = goff[static_cast<size_t>(d0 + 1) * ntiles + tile];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kRadixItems; j++) {
    if (wbase + j * 64 + lane < m) {
      const uint32_t d = (k[j] >> shift) & mask;
      const uint32_t slot = s_lbeg[d] + s_wcnt[w][d] + r[j];
      if (slot < static_cast<uint32_t>(kRadixTile)) {
        s_key[slot] = k[j];
        s_val[slot] = v[j];
      }
    }
  }
  __syncthreads();
  const size_t rem = m - tbase;
  const int cnt = rem < static_cast<size_t>(kRadixTile) ? static_cast<int>(rem) : kRadixTile;
  for (int e = threadIdx.x; e < cnt; e += kRadixThreads) {
    const uint32_t kk = s_key[e];
    const uint32_t d = (kk >> shift) & mask;
    const size_t g = static_cast<size_t>(s_gbeg[d]) + (static_cast<uint32_t>(e) - s_lbeg[d]);
    if (g < m) {
      kout[g] = kk;
      vout[g] = s_val[e];
    }
  }
}

// Paint step 3: cell starts from the sorted keys, kStartCells cells per block.
// bnd[b] is the first sorted position whose key is >= b * kStartCells; a block
// marks the first position of each of its non-empty cells and fills the empty
// ones with the next non-empty cell's start by a suffix minimum.
constexpr int kStartItems = 32;
constexpr int kStartCells = kCicThreads * kStartItems;  // 8192

// The keys' cell is key >> shift (the tile paint sorts by tile id, shift 9).
__global__ __launch_bounds__(kCicThreads) void k_start_bounds(const uint32_t* __restrict__ skey,
                                                              size_t m, uint32_t nsb,
                                                              uint32_t* __restrict__ bnd, int shift) {
  const size_t p = blockIdx.x * static_cast<size_t>(kCicThreads) + threadIdx.x;
  if (p > m) return;
  // blocks b with key[p - 1] < b * kStartCells <= key[p] start at p (key[m] = infinity)
  const int64_t lo = p > 0 ? static_cast<int64_t>((skey[p - 1] >> shift) / kStartCells) : -1;
  const int64_t hi = p < m ? static_cast<int64_t>((skey[p] >> shift) / kStartCells) : static_cast<int64_t>(nsb);
  for (int64_t b = lo + 1; b <= hi && b <= static_cast<int64_t>(nsb); b++)
    bnd[b] = static_cast<uint32_t>(p);
}

__global__ __launch_bounds__(kCicThreads) void k_cell_start(const uint32_t* __restrict__ skey,
                                                            const uint32_t* __restrict__ bnd,
                                                            size_t ncell, size_t m,
                                                            uint32_t* __restrict__ start, int shift) {
  __shared__ uint32_t s[kStartCells + kStartCells / 32];
  __shared__ uint32_t s_wmin[kCicThreads / 64];
  auto pad = [](int e) { return e + (e >> 5); };
  const size_t c0 = blockIdx.x * static_cast<size_t>(kStartCells);
  const size_t rem = ncell - c0;
  const int cnt = rem < static_cast<size_t>(kStartCells) ? static_cast<int>(rem) : kStartCells;
  // (clamped into [0, m]: every position read below stays inside the sorted keys)
  const uint32_t hi = min(bnd[blockIdx.x + 1], static_cast<uint32_t>(m));
  const uint32_t lo = min(bnd[blockIdx.x], hi);
  for (int e = threadIdx.x; e < kStartCells; e += kCicThreads) s[pad(e)] = hi;
  __syncthreads();
  for (uint32_t p = lo + threadIdx.x; p < hi; p += kCicThreads) {
    const uint32_t k = skey[p] >> shift;
    if (p == 0u || (skey[p - 1] >> shift) != k) {
      const size_t e = static_cast<size_t>(k) - c0;
      if (e < static_cast<size_t>(cnt)) s[pad(static_cast<int>(e))] = p;
    }
  }
  __syncthreads();
  // suffix minimum: thread t owns cells [32 t, 32 t + 32)
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int e0 = threadIdx.x * kStartItems;
  uint32_t seg = hi;
  for (int q = 0; q < kStartItems; q++) seg = min(seg, s[pad(e0 + q)]);
  uint32_t x = seg;  // minimum over this lane and the lanes above it
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_down(x, off);
    if (lane + off < 64) x = min(x, y);
  }
  uint32_t after = __shfl_down(x, 1);
  if (lane == 63) after = hi;
  if (lane == 0) s_wmin[w] = x;
  __syncthreads();
  for (int u = w + 1; u < kCicThreads / 64; u++) after = min(after, s_wmin[u]);
  uint32_t run = after;
  for (int q = kStartItems - 1; q >= 0; q--) {
    run = min(run, s[pad(e0 + q)]);
    s[pad(e0 + q)] = run;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < cnt; e += kCicThreads) start[c0 + e] = s[pad(e)];
  if (c0 + cnt == ncell && threadIdx.x == 0) start[ncell] = static_cast<uint32_t>(m);
}

// Paint step 4: payload into cell-grouped order, so that step 5 reads each
// group contiguously.
template <typename R>
__global__ __launch_bounds__(kCicThreads) void k_paint_permute(const uint32_t* __restrict__ sorted,
                                                               const Payload<R>* __restrict__ payload,
                                                               Payload<R>* __restrict__ spay, size_t m) {
  const size_t s = blockIdx.x * static_cast<size_t>(kCicThreads) + threadIdx.x;
  if (s < m) spay[s] = payload[sorted[s]];
}

// CIC weight of a point at corner k = (a, b, c) of its base cell, in the
// reference's product order (src/mesh.c:45-60).
template <typename R>
__device__ __forceinline__ R corner_weight(const Payload<R>& q, int k) {
  const R fx = (k & 4) ? q.t[0] : R(1) - q.t[0];
  const R fy = (k & 2) ? q.t[1] : R(1) - q.t[1];
  const R fz = (k & 1) ? q.t[2] : R(1) - q.t[2];
  return fx * fy * fz * q.w;
}

// Merge of the (at most 8) groups of one target cell in point-index order,
// accumulating onto `acc`; groups are [beg[k], end[k]) of `idx` / `pay`
// (in LDS or global memory).  Returns whether any point contributed.  Each
// step picks the group with the smallest next index by a select chain and
// issues one index read (the chosen group's next entry) and one payload read.
template <typename R>
__device__ __forceinline__ bool merge_groups(uint32_t beg[8], const uint32_t end[8],
                                             const uint32_t* __restrict__ idx,
                                             const Payload<R>* __restrict__ pay, R& acc) {
  constexpr uint32_t kNone = 0xffffffffu;
  uint32_t cur[8];
  bool any = false;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    cur[k] = beg[k] < end[k] ? idx[beg[k]] : kNone;
    any |= beg[k] < end[k];
  }
  if (!any) return false;
  for (;;) {
    uint32_t best = cur[0], pos = beg[0], lim = end[0];
    int bk = 0;
#pragma unroll
    for (int k = 1; k < 8; k++) {
      const bool lt = cur[k] < best;
      best = lt ? cur[k] : best;
      pos = lt ? beg[k] : pos;
      lim = lt ? end[k] : lim;
      bk = lt ? k : bk;
    }
    if (best == kNone) break;
    const uint32_t nxt = pos + 1;
    const uint32_t nv = nxt < lim ? idx[nxt] : kNone;
    const Payload<R> q = pay[pos];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      beg[k] = bk == k ? nxt : beg[k];
      cur[k] = bk == k ? nv : cur[k];
    }
    acc += corner_weight(q, bk);
  }
  return true;
}

// Paint step 5, tiled: a workgroup owns an 8^3 tile of target cells; the
// groups of the 9^3 base cells that reach it (the tile and its lower
// neighbours) are staged in LDS (indices and payloads, read once, coalesced),
// and each thread merges its cell's 8 groups from LDS.  A tile whose groups
// exceed the LDS capacity merges from global memory instead.  Source k =
// (a, b, c) of a target cell is the base cell at (x - a, y - b, z - c).
constexpr int kGT = 8;                      // tile edge
constexpr int kGE = kGT + 1;                // staged region edge
constexpr int kGCells = kGE * kGE * kGE;    // 729 staged base cells
constexpr int kGThreads = kGT * kGT * kGT;  // one thread per target cell
template <typename R>
struct GatherCap {
  static constexpr int value = sizeof(R) == 8 ? 1792 : 3072;
};

template <typename R>
__global__ __launch_bounds__(kGThreads) void k_paint_gather_tiled(
    R* __restrict__ mesh, int n0, int n1, int n2, int tiles1, int tiles2, uint32_t ntiles,
    uint32_t per, const uint32_t* __restrict__ start, const uint32_t* __restrict__ sorted,
    const Payload<R>* __restrict__ spay, const uint32_t* __restrict__ misc) {
  constexpr int kCap = GatherCap<R>::value;
  __shared__ uint32_t s_gbeg[kGCells];  // global start of each staged base cell
  __shared__ uint32_t s_cnt[kGCells];   // its number of points
  __shared__ uint32_t s_off[kGCells];   // its first slot in the staged arrays
  __shared__ uint32_t s_row[kGE * kGE];
  __shared__ uint32_t s_idx[kCap];
  __shared__ uint16_t s_map[kCap];  // staged slot -> staged cell
  __shared__ Payload<R> s_pay[kCap];
  if (misc[0] != 0u) return;  // invalid input: leave the mesh untouched
  const int tid = threadIdx.x;
  const uint32_t tile = xcd_block(per);
  if (tile >= ntiles) return;
  const int bz = tile % tiles2;
  const int by = (tile / tiles2) % tiles1;
  const int bx = tile / (tiles2 * tiles1);
  const int x0 = bx * kGT, y0 = by * kGT, z0 = bz * kGT;
  auto wrap = [](int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); };
  // staged cell e = (ex, ey, ez) is the base cell at (x0 - 1 + ex, ...) and serves the target
  // planes ex - 1 and ex; it is staged only if the first of them that exists lies inside the
  // mesh (x0 + max(ex - 1, 0) < n0, ...).  Its coordinate then lies in [-1, n), so one period of
  // wrap maps it into the mesh.  Other cells (past the end of a partial tile, where a mesh
  // extent is below the tile edge) stage nothing.
  auto needed = [](int o, int ei, int n) { return o + (ei > 0 ? ei - 1 : 0) < n; };
  for (int e = tid; e < kGCells; e += kGThreads) {
    const int ex = e / (kGE * kGE), ey = (e / kGE) % kGE, ez = e % kGE;
    uint32_t g0 = 0u, cnt = 0u;
    if (needed(x0, ex, n0) && needed(y0, ey, n1) && needed(z0, ez, n2)) {
      const size_t bc = (static_cast<size_t>(wrap(x0 - 1 + ex, n0)) * n1 + wrap(y0 - 1 + ey, n1)) * n2 +
                        wrap(z0 - 1 + ez, n2);
      g0 = start[bc];
      cnt = start[bc + 1] - g0;
    }
    s_gbeg[e] = g0;
    s_cnt[e] = cnt;
  }
  __syncthreads();
  // offsets: row sums, an exclusive scan over the 81 rows (first wave, by
  // shuffles), then within the rows
  if (tid < kGE * kGE) {
    uint32_t sum = 0;
    for (int q = 0; q < kGE; q++) sum += s_cnt[tid * kGE + q];
    s_row[tid] = sum;
  }
  __syncthreads();
  if (tid < 64) {
    uint32_t v0 = s_row[tid];
    uint32_t v1 = tid + 64 < kGE * kGE ? s_row[tid + 64] : 0u;
    uint32_t i0 = v0, i1 = v1;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t u0 = __shfl_up(i0, off), u1 = __shfl_up(i1, off);
      if (tid >= off) {
        i0 += u0;
        i1 += u1;
      }
    }
    const uint32_t t0 = __shfl(i0, 63);
    s_row[tid] = i0 - v0;
    if (tid + 64 < kGE * kGE) s_row[tid + 64] = t0 + i1 - v1;
    if (tid == 63) s_off[kGCells - 1] = t0 + __shfl(i1, 63);  // total, rewritten below
  }
  __syncthreads();
  const uint32_t total = s_off[kGCells - 1];
  __syncthreads();
  if (total == 0) return;
  if (tid < kGE * kGE) {
    uint32_t run = s_row[tid];
    for (int q = 0; q < kGE; q++) {
      s_off[tid * kGE + q] = run;
      run += s_cnt[tid * kGE + q];
    }
  }
  __syncthreads();
  // Staged in one part (the whole region) when it fits; otherwise in two halves
  // along x: target planes [0, 4) need staged planes [0, 5), target planes
  // [4, 8) staged planes [4, 9).  The cells of consecutive planes hold a
  // contiguous range of slots.  A part that still does not fit merges from
  // global memory.
  const int lz = tid % kGT, ly = (tid / kGT) % kGT, lx = tid / (kGT * kGT);
  const int x = x0 + lx, y = y0 + ly, z = z0 + lz;
  const bool valid = x < n0 && y < n1 && z < n2;
  // the target value is read now, so that its latency overlaps the staging
  const size_t c = valid ? (static_cast<size_t>(x) * n1 + y) * n2 + z : 0;
  R acc = valid ? mesh[c] : R(0);
  const int nparts = total <= static_cast<uint32_t>(kCap) ? 1 : 2;
  constexpr int kPlane = kGE * kGE;
  for (int part = 0; part < nparts; part++) {
    const int ex0 = nparts == 1 ? 0 : 4 * part;
    const int ex1 = nparts == 1 ? kGE : 4 * part + 5;
    const uint32_t base = s_off[ex0 * kPlane];
    const uint32_t lim = ex1 == kGE ? total : s_off[ex1 * kPlane];
    const uint32_t cnt = lim - base;
    const bool staged = cnt <= static_cast<uint32_t>(kCap);
    if (staged && cnt > 0) {
      // slot -> staged cell map, then the copy over slots in batches whose reads are
      // all issued before the LDS stores (several reads in flight per thread)
      for (int e = ex0 * kPlane + tid; e < ex1 * kPlane; e += kGThreads) {
        const uint32_t o0 = s_off[e] - base, c = s_cnt[e];
        for (uint32_t j = 0; j < c; j++) s_map[o0 + j] = static_cast<uint16_t>(e);
      }
      __syncthreads();
      constexpr int kBatch = 4;
      for (uint32_t p0 = tid; p0 < cnt; p0 += kBatch * kGThreads) {
        uint32_t ix[kBatch];
        Payload<R> pl[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
          // slots past the end re-read the last slot (in bounds; not stored)
          const uint32_t p = min(p0 + u * kGThreads, cnt - 1);
          const int e = s_map[p];
          const uint32_t g = s_gbeg[e] + (p + base - s_off[e]);
          ix[u] = sorted[g];
          pl[u] = spay[g];
        }
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
          const uint32_t p = p0 + u * kGThreads;
          if (p < cnt) {
            s_idx[p] = ix[u];
            s_pay[p] = pl[u];
          }
        }
      }
      __syncthreads();
    }
    const bool mine = nparts == 1 || (lx >= 4 * part && lx < 4 * part + 4);
    if (valid && mine) {
      uint32_t beg[8], end[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const int e = (lx + 1 - ((k >> 2) & 1)) * kPlane + (ly + 1 - ((k >> 1) & 1)) * kGE +
                      (lz + 1 - (k & 1));
        beg[k] = staged ? s_off[e] - base : s_gbeg[e];
        end[k] = beg[k] + s_cnt[e];
      }
      const bool any = staged ? merge_groups<R>(beg, end, s_idx, s_pay, acc)
                              : merge_groups<R>(beg, end, sorted, spay, acc);
      if (any) mesh[c] = acc;
    }
    if (part + 1 < nparts) __syncthreads();  // the next part restages the LDS arrays
  }
}

// Paint step 5 without staging: one thread per target cell, the cells of an
// 8^3 tile per workgroup (the tile order of k_paint_gather_tiled), merging
// straight from global memory.  A tile's groups and their neighbours' bounds
// are shared by its threads through L1/L2; with no LDS the occupancy is set
// by registers alone.
template <typename R>
__global__ __launch_bounds__(kGThreads) void k_paint_gather_direct(
    R* __restrict__ mesh, int n0, int n1, int n2, int tiles1, int tiles2, uint32_t ntiles,
    uint32_t per, const uint32_t* __restrict__ start, const uint32_t* __restrict__ sorted,
    const Payload<R>* __restrict__ spay, const uint32_t* __restrict__ misc) {
  if (misc[0] != 0u) return;  // invalid input: leave the mesh untouched
  const int tid = threadIdx.x;
  const uint32_t tile = xcd_block(per);
  if (tile >= ntiles) return;
  const int bz = tile % tiles2;
  const int by = (tile / tiles2) % tiles1;
  const int bx = tile / (tiles2 * tiles1);
  const int x = bx * kGT + tid / (kGT * kGT), y = by * kGT + (tid / kGT) % kGT, z = bz * kGT + tid % kGT;
  if (x >= n0 || y >= n1 || z >= n2) return;
  const int xs[2] = {x, x == 0 ? n0 - 1 : x - 1};
  const int ys[2] = {y, y == 0 ? n1 - 1 : y - 1};
  const int zs[2] = {z, z == 0 ? n2 - 1 : z - 1};
  uint32_t beg[8], end[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    // source k = (a, b, c): the base cell at (x - a, y - b, z - c)
    const size_t bc = (static_cast<size_t>(xs[(k >> 2) & 1]) * n1 + ys[(k >> 1) & 1]) * n2 + zs[k & 1];
    beg[k] = start[bc];
    end[k] = start[bc + 1];
  }
  const size_t c = (static_cast<size_t>(x) * n1 + y) * n2 + z;
  R acc = mesh[c];
  if (merge_groups<R>(beg, end, sorted, spay, acc)) mesh[c] = acc;
}

// Below this many points per 100 target cells the gather merges straight from
// global memory; above it the tiles stage their groups in LDS.  Measured at
// 512^3 (FP64): 2e7 shell points take 3.4 ms direct and 4.8 ms staged, 1e8
// take 10.4 ms direct and 8.2 ms staged (crossover near 0.37 per cell).
#ifndef PRH_GATHER_DIRECT_PER100
#define PRH_GATHER_DIRECT_PER100 37
#endif

__global__ void k_zero_misc(uint32_t* misc) {
  if (threadIdx.x < 4) misc[threadIdx.x] = 0u;
}

// ---------------------------------------------------------------------------
// Paint with exact accumulation per tile (the mixed-precision path)
// ---------------------------------------------------------------------------
// The paint above reproduces the reference's serial sum bit for bit.  The
// mixed-precision reconstruction instead sums the FP64 corner weights in fixed
// point per tile: a cell's k contributions are each rounded by at most 2^-62 B
// (B the tile's weight bound, below), against about 2^-53 of the cell's value
// for each term of the serial FP64 sum, so the two meshes agree to the serial
// sum's own rounding (a cell lands on the other side of the density rule's
// threshold only if it lies within that rounding of it), and every repeated run
// gives the same bits.  Points are sorted by
// kPT^3-cell tile alone (kPT = 16; the key's low 12 bits carry the cell inside
// the tile and ride along), so a 2-digit radix sort replaces the cell sort and
// the point-order merge.  One workgroup per tile:
//   - loads its points (positions in tile order, their FP64 weights),
//   - bounds B = sum |w| prod_d max(|1 - d_d|, |d_d|) over them (an FP64
//     reduction in a fixed order) and takes the scale 2^(62 - e), B < 2^e,
//   - rounds each FP64 contribution ((wx*wy)*wz)*w times the scale to an
//     integer and adds it into a 17^3 array of 64-bit integers in LDS: integer
//     sums do not depend on their order, and the scaled sum stays below 2^62,
//   - converts to FP64 (the scale is a power of two) and adds the cells that
//     only this tile reaches (local indices 1 .. extent - 1 on every axis)
//     into the mesh; the others go to a per-tile boundary array.
// A second pass adds, for each cell with a local index 0, its own tile's value
// and the upper halos of the (up to 7) lower neighbour tiles in a fixed order.
// The contributions are rounded to within 2^-62 B (B the tile's own weight
// bound), about the rounding of the FP64 products themselves.
constexpr int kPTB = kPaintTileBits;            // log2 of the tile edge
constexpr int kPT = kPaintTile;                  // tile edge (cells)
constexpr int kPTH = kPT + 1;                    // with the upper halo
constexpr int kPTCells = kPTH * kPTH * kPTH;     // accumulation cells per tile
constexpr int kPTShift = 3 * kPTB;               // key = tile << kPTShift | ((cx * kPT + cy) * kPT + cz)
static_assert(kPTCells == kPaintTileCells, "tile record size");

struct TileGrid {
  int n[3];  // mesh shape
  int t[3];  // tiles per axis
};

// A point's position (in the positions dtype) and FP64 weight, packed by the key
// step into one aligned 32-byte record so that the permute gathers a single record per point.
template <typename P>
struct alignas(32) TileRec {
  P x[3];
  double w;
};
static_assert(sizeof(TileRec<double>) == 32 && sizeof(TileRec<float>) == 32, "tile records are 32 bytes");

template <typename P, typename W, int MODE>
__global__ __launch_bounds__(kCicThreads) void k_tile_key(const P* __restrict__ pos,
                                                          const W* __restrict__ w, size_t m,
                                                          Frame<double> F, TileGrid TG,
                                                          uint32_t* __restrict__ key,
                                                          TileRec<P>* __restrict__ rec,
                                                          uint32_t* __restrict__ misc,
                                                          uint64_t* __restrict__ hpart) {
  const size_t i = blockIdx.x * static_cast<size_t>(kCicThreads) + threadIdx.x;
  uint32_t err = 0;
  uint64_t hash = 0;
  if (i < m) {
    const P x[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    if (hpart) hash = point_hash<P>(i, x);
    double p[3];
    to_mesh_units<double, P>(x, F, MODE == 2, p, err);
    if (w && !isfinite(static_cast<double>(w[i]))) err |= kErrNonFinite;
    int b[3];
    for (int d = 0; d < 3; d++) b[d] = base_index(static_cast<int>(p[d]), F.n[d]);
    const uint32_t tile =
        (static_cast<uint32_t>(b[0] >> kPTB) * TG.t[1] + (b[1] >> kPTB)) * TG.t[2] + (b[2] >> kPTB);
    key[i] = (tile << kPTShift) |
             (((b[0] & (kPT - 1)) * kPT + (b[1] & (kPT - 1))) * kPT + (b[2] & (kPT - 1)));
    TileRec<P> r;
    r.x[0] = x[0];
    r.x[1] = x[1];
    r.x[2] = x[2];
    r.w = w ? static_cast<double>(w[i]) : 1.0;
    rec[i] = r;
  }
  flag_paint_errors(err, misc);
  if (hpart) block_hash_partial(hash, hpart);
}

// start[t] = first sorted position whose tile (key >> kPTShift) is >= t, for
// t in [0, ntiles]; one thread per tile, by binary search over the sorted keys.
__global__ __launch_bounds__(kCicThreads) void k_tile_starts(const uint32_t* __restrict__ skey, size_t m,
                                                             uint32_t ntiles, uint32_t* __restrict__ start) {
  const size_t t = blockIdx.x * static_cast<size_t>(kCicThreads) + threadIdx.x;
  if (t > ntiles) return;
  size_t lo = 0, hi = m;  // first position in [lo, hi) with tile >= t
  while (lo < hi) {
    const size_t mid = lo + (hi - lo) / 2;
    if ((skey[mid] >> kPTShift) < t)
      lo = mid + 1;
    else
      hi = mid;
  }
  start[t] = static_cast<uint32_t>(lo);
}

// Positions and weights in tile order.
template <typename P>
__global__ __launch_bounds__(kCicThreads) void k_tile_permute(const uint32_t* __restrict__ sorted,
                                                              const TileRec<P>* __restrict__ rec, size_t m,
                                                              P* __restrict__ spos,
                                                              double* __restrict__ sw) {
  const size_t s = blockIdx.x * static_cast<size_t>(kCicThreads) + threadIdx.x;
  if (s >= m) return;
  const TileRec<P> r = rec[sorted[s]];
  sw[s] = r.w;
  spos[3 * s] = r.x[0];
  spos[3 * s + 1] = r.x[1];
  spos[3 * s + 2] = r.x[2];
}

// Offsets of a point from its base cell (recomputed from its position as
// k_tile_key computed the cell).
template <typename P>
__device__ __forceinline__ void tile_offsets(const P* __restrict__ spos, size_t s,
                                             const Frame<double>& F, int wrap, double d[3]) {
  const P x[3] = {spos[3 * s], spos[3 * s + 1], spos[3 * s + 2]};
  double p[3];
  uint32_t err = 0;
  to_mesh_units<double, P>(x, F, wrap, p, err);
  for (int k = 0; k < 3; k++) {
    const int b = base_index(static_cast<int>(p[k]), F.n[k]);
    d[k] = p[k] - static_cast<double>(b);
  }
}

struct TilePos {
  int o[3], e[3];  // origin and extent (cells) on each axis
};
__device__ __forceinline__ TilePos tile_pos(uint32_t t, const TileGrid& TG) {
  TilePos T;
  const int c[3] = {static_cast<int>(t / (static_cast<uint32_t>(TG.t[1]) * TG.t[2])),
                    static_cast<int>((t / TG.t[2]) % TG.t[1]), static_cast<int>(t % TG.t[2])};
  for (int d = 0; d < 3; d++) {
    T.o[d] = kPT * c[d];
    T.e[d] = TG.n[d] - T.o[d] < kPT ? TG.n[d] - T.o[d] : kPT;
  }
  return T;
}

constexpr int kAccThreads = 512;  // threads per tile (a 16^3 tile holds about 3000 randoms at 512^3)
constexpr int kAccKeep = 4;  // points per thread whose offsets stay in registers between the passes
template <typename P>
__global__ __launch_bounds__(kAccThreads) void k_tile_accumulate(
    double* __restrict__ mesh, Frame<double> F, TileGrid TG, int wrap, uint32_t ntiles,
    uint32_t per, const uint32_t* __restrict__ start, const uint32_t* __restrict__ skey,
    const P* __restrict__ spos, const double* __restrict__ sw, double* __restrict__ bnd,
    const uint32_t* __restrict__ misc) {
  __shared__ unsigned long long acc[kPTCells];
  __shared__ double s_red[kAccThreads / 64];
  if (misc[0] != 0u) return;  // invalid input: leave the mesh untouched
  const uint32_t t = xcd_block(per);
  if (t >= ntiles) return;
  const uint32_t beg = start[t], end = start[t + 1];
  if (beg >= end) return;  // empty: the boundary pass skips it
  const TilePos T = tile_pos(t, TG);
  for (int c = threadIdx.x; c < kPTCells; c += kAccThreads) acc[c] = 0ull;
  // weight bound of the tile (fixed order: strided per thread, then a fixed tree).  The offsets
  // and weights of a thread's first kAccKeep points stay in registers for the second pass, so
  // that tiles of up to kAccKeep x kAccThreads points read their positions and weights once.
  double bsum = 0.0;
  double dk[kAccKeep][3], wk[kAccKeep];
#pragma unroll
  for (int j = 0; j < kAccKeep; j++) {
    const uint32_t q = beg + threadIdx.x + j * kAccThreads;
    dk[j][0] = dk[j][1] = dk[j][2] = 0.0;
    wk[j] = 0.0;
    if (q < end) {
      tile_offsets<P>(spos, q, F, wrap, dk[j]);
      wk[j] = sw[q];
      double pw = fabs(wk[j]);
      for (int k = 0; k < 3; k++) pw *= fmax(fabs(1.0 - dk[j][k]), fabs(dk[j][k]));
      bsum += pw;
    }
  }
  for (uint32_t q = beg + threadIdx.x + kAccKeep * kAccThreads; q < end; q += kAccThreads) {
    double d[3];
    tile_offsets<P>(spos, q, F, wrap, d);
    double pw = fabs(sw[q]);
    for (int k = 0; k < 3; k++) pw *= fmax(fabs(1.0 - d[k]), fabs(d[k]));
    bsum += pw;
  }
  for (int off = 32; off > 0; off >>= 1) bsum += __shfl_xor(bsum, off);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = bsum;
  __syncthreads();
  double B = 0.0;
  for (int u = 0; u < kAccThreads / 64; u++) B += s_red[u];
  int ex = 0;
  if (B > 0.0 && isfinite(B)) frexp(B, &ex);  // B < 2^ex
  const double scale = ldexp(1.0, 62 - ex), inv = ldexp(1.0, ex - 62);
  auto add_point = [&](const double d[3], double wq, uint32_t key) {
    const uint32_t cell = key & ((1u << kPTShift) - 1u);
    const int cx = static_cast<int>(cell >> (2 * kPTB)), cy = static_cast<int>((cell >> kPTB) & (kPT - 1u)),
              cz = static_cast<int>(cell & (kPT - 1u));
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int a = (k >> 2) & 1, b = (k >> 1) & 1, c = k & 1;
      const double wx = a ? d[0] : 1.0 - d[0];
      const double wy = b ? d[1] : 1.0 - d[1];
      const double wz = c ? d[2] : 1.0 - d[2];
      const long long fixed = __double2ll_rn(((wx * wy) * wz) * wq * scale);
      atomicAdd(&acc[((cx + a) * kPTH + (cy + b)) * kPTH + (cz + c)], static_cast<unsigned long long>(fixed));
    }
  };
#pragma unroll
  for (int j = 0; j < kAccKeep; j++) {
    const uint32_t q = beg + threadIdx.x + j * kAccThreads;
    if (q < end) add_point(dk[j], wk[j], skey[q]);
  }
  for (uint32_t q = beg + threadIdx.x + kAccKeep * kAccThreads; q < end; q += kAccThreads) {
    double d[3];
    tile_offsets<P>(spos, q, F, wrap, d);
    add_point(d, sw[q], skey[q]);
  }
  __syncthreads();
  const size_t n1 = static_cast<size_t>(TG.n[1]), n2 = static_cast<size_t>(TG.n[2]);
  for (int c = threadIdx.x; c < kPTCells; c += kAccThreads) {
    const int i = c / (kPTH * kPTH), j = (c / kPTH) % kPTH, k = c % kPTH;
    if (i > T.e[0] || j > T.e[1] || k > T.e[2]) continue;
    const double v = static_cast<double>(static_cast<long long>(acc[c])) * inv;
    const bool own = i >= 1 && i < T.e[0] && j >= 1 && j < T.e[1] && k >= 1 && k < T.e[2];
    if (own)
      mesh[(static_cast<size_t>(T.o[0] + i) * n1 + (T.o[1] + j)) * n2 + (T.o[2] + k)] += v;
    else
      bnd[static_cast<size_t>(t) * kPTCells + c] = v;
  }
}

// Cells with a local index 0: own value plus the lower neighbours' halos.
__global__ __launch_bounds__(kCicThreads) void k_tile_boundary(double* __restrict__ mesh,
                                                               TileGrid TG, uint32_t ntiles,
                                                               uint32_t per,
                                                               const uint32_t* __restrict__ start,
                                                               const double* __restrict__ bnd,
                                                               const uint32_t* __restrict__ misc) {
  if (misc[0] != 0u) return;  // invalid input: leave the mesh untouched
  const uint32_t t = xcd_block(per);
  if (t >= ntiles) return;
  const TilePos T = tile_pos(t, TG);
  const int c[3] = {T.o[0] / kPT, T.o[1] / kPT, T.o[2] / kPT};
  // lower neighbours: subset s of the axes (bit 2: x, bit 1: y, bit 0: z), s = 0 the tile itself
  uint32_t nb[8];
  bool full[8];
  int em[3];  // extent of the lower neighbour along each axis (its halo index)
  for (int d = 0; d < 3; d++) {
    const int cm = c[d] == 0 ? TG.t[d] - 1 : c[d] - 1;
    em[d] = TG.n[d] - kPT * cm < kPT ? TG.n[d] - kPT * cm : kPT;
  }
  for (int sub = 0; sub < 8; sub++) {
    int cc[3];
    for (int d = 0; d < 3; d++) {
      const bool lo = (sub >> (2 - d)) & 1;
      cc[d] = lo ? (c[d] == 0 ? TG.t[d] - 1 : c[d] - 1) : c[d];
    }
    nb[sub] = (static_cast<uint32_t>(cc[0]) * TG.t[1] + cc[1]) * TG.t[2] + cc[2];
    full[sub] = start[nb[sub]] < start[nb[sub] + 1];
  }
  const size_t n1 = static_cast<size_t>(TG.n[1]), n2 = static_cast<size_t>(TG.n[2]);
  // the cells with a local index 0, listed face by face: i = 0 (kPT^2 cells),
  // then i > 0 and j = 0 ((kPT - 1) kPT), then i, j > 0 and k = 0 ((kPT - 1)^2)
  constexpr int kFaceI = kPT * kPT, kFaceJ = (kPT - 1) * kPT, kFaceK = (kPT - 1) * (kPT - 1);
  for (int q = threadIdx.x; q < kFaceI + kFaceJ + kFaceK; q += kCicThreads) {
    int i, j, k;
    if (q < kFaceI) {
      i = 0;
      j = q >> kPTB;
      k = q & (kPT - 1);
    } else if (q < kFaceI + kFaceJ) {
      const int r = q - kFaceI;
      i = 1 + (r >> kPTB);
      j = 0;
      k = r & (kPT - 1);
    } else {
      const int r = q - kFaceI - kFaceJ;
      i = 1 + r / (kPT - 1);
      j = 1 + r % (kPT - 1);
      k = 0;
    }
    if (i >= T.e[0] || j >= T.e[1] || k >= T.e[2]) continue;
    const int zero = (i == 0 ? 4 : 0) | (j == 0 ? 2 : 0) | (k == 0 ? 1 : 0);
    bool any = false;
    double sum = 0.0;
    if (full[0]) {
      sum = bnd[static_cast<size_t>(nb[0]) * kPTCells + (i * kPTH + j) * kPTH + k];
      any = true;
    }
    for (int sub = 1; sub < 8; sub++) {
      if ((sub & zero) != sub || !full[sub]) continue;
      const int ii = (sub & 4) ? em[0] : i, jj = (sub & 2) ? em[1] : j, kk = (sub & 1) ? em[2] : k;
      sum += bnd[static_cast<size_t>(nb[sub]) * kPTCells + (ii * kPTH + jj) * kPTH + kk];
      any = true;
    }
    if (any) mesh[(static_cast<size_t>(T.o[0] + i) * n1 + (T.o[1] + j)) * n2 + (T.o[2] + k)] += sum;
  }
}

// ---------------------------------------------------------------------------
// Reads
// ---------------------------------------------------------------------------

// Base cell and offsets of a position in mesh units (src/mesh.c:205-214, 277-283).
template <typename R>
__device__ __forceinline__ void cic_base(const R p[3], const int nn[3], int b[3], R t[3]) {
  for (int d = 0; d < 3; d++) {
    const int ip = static_cast<int>(p[d]);
    t[d] = p[d] - static_cast<R>(ip);
    b[d] = base_index(ip, nn[d]);
  }
}

// src/mesh.c:277-306.  at(a, b, c) is the value at base + (a, b, c), a, b, c in {0, 1}.
template <typename R, typename A>
__device__ __forceinline__ R cic_core(A at, const R t[3]) {
  R acc = R(0);
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int a = (k >> 2) & 1, bb = (k >> 1) & 1, e = k & 1;
    const R wx = a ? t[0] : R(1) - t[0];
    const R wy = bb ? t[1] : R(1) - t[1];
    const R wz = e ? t[2] : R(1) - t[2];
    const R term = at(a, bb, e) * wx * wy * wz;
    acc = (k == 0) ? term : acc + term;
  }
  return acc;
}

// src/mesh.c:205-262: central differences at the 8 corners, corner order
// 000, 100, 010, 001, 110, 101, 011, 111; divided by cell = 2 L / n.
// at(i, j, k) is the value at base + (i, j, k) - 1, i, j, k in [0, 4).
template <typename R, typename A>
__device__ __forceinline__ void fd_core(A at, const R t[3], const R cell[3], R g[3]) {
  constexpr int corner[8] = {0, 4, 2, 1, 6, 5, 3, 7};  // bit 2: x, bit 1: y, bit 0: z
  R gx = R(0), gy = R(0), gz = R(0);
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int a = (corner[q] >> 2) & 1, bb = (corner[q] >> 1) & 1, e = corner[q] & 1;
    const R wx = a ? t[0] : R(1) - t[0];
    const R wy = bb ? t[1] : R(1) - t[1];
    const R wz = e ? t[2] : R(1) - t[2];
    const R wt = wx * wy * wz;
    const R dx = (at(2 + a, 1 + bb, 1 + e) - at(a, 1 + bb, 1 + e)) * wt;
    const R dy = (at(1 + a, 2 + bb, 1 + e) - at(1 + a, bb, 1 + e)) * wt;
    const R dz = (at(1 + a, 1 + bb, 2 + e) - at(1 + a, 1 + bb, e)) * wt;
    if (q == 0) {
      gx = dx;
      gy = dy;
      gz = dz;
    } else {
      gx += dx;
      gy += dy;
      gz += dz;
    }
  }
  g[0] = gx / cell[0];
  g[1] = gy / cell[1];
  g[2] = gz / cell[2];
}

template <typename R>
__device__ __forceinline__ R cic_value(const R* __restrict__ mesh, int n0, int n1, int n2,
                                       const R p[3]) {
  R t[3];
  int b[3];
  const int nn[3] = {n0, n1, n2};
  cic_base(p, nn, b, t);
  const size_t s0 = static_cast<size_t>(n1) * n2, s1 = static_cast<size_t>(n2);
  const size_t X[2] = {b[0] * s0, (b[0] + 1 == n0 ? 0 : b[0] + 1) * s0};
  const size_t Y[2] = {b[1] * s1, (b[1] + 1 == n1 ? 0 : b[1] + 1) * s1};
  const size_t Z[2] = {static_cast<size_t>(b[2]), static_cast<size_t>(b[2] + 1 == n2 ? 0 : b[2] + 1)};
  return cic_core<R>([&](int a, int bb, int e) { return mesh[X[a] + Y[bb] + Z[e]]; }, t);
}

template <typename R>
__device__ __forceinline__ void fd_gradient(const R* __restrict__ mesh, int n0, int n1, int n2,
                                            const R cell[3], const R p[3], R g[3]) {
  R t[3];
  int b[3];
  const int nn[3] = {n0, n1, n2};
  cic_base(p, nn, b, t);
  const size_t s0 = static_cast<size_t>(n1) * n2, s1 = static_cast<size_t>(n2);
  // (b - 1 + k) mod n: b lies in [0, n), so for n >= 2 one correction by n is enough
  auto wrapk = [](int b, int k, int n) {
    if (n < 2) return (b + n - 1 + k) % n;
    int v = b - 1 + k;
    v += v < 0 ? n : 0;
    v -= v >= n ? n : 0;
    return v;
  };
  size_t X[4], Y[4], Z[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    X[k] = static_cast<size_t>(wrapk(b[0], k, n0)) * s0;
    Y[k] = static_cast<size_t>(wrapk(b[1], k, n1)) * s1;
    Z[k] = static_cast<size_t>(wrapk(b[2], k, n2));
  }
  fd_core<R>([&](int i, int j, int k) { return mesh[X[i] + Y[j] + Z[k]]; }, t, cell, g);
}

template <typename R>
__global__ __launch_bounds__(kCicThreads) void k_read_cic(const R* __restrict__ mesh, int n0,
                                                          int n1, int n2,
                                                          const R* __restrict__ pos,
                                                          R* __restrict__ out, size_t m) {
  const size_t i = blockIdx.x * static_cast<size_t>(kCicThreads) + threadIdx.x;
  if (i >= m) return;
  const R p[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
  out[i] = cic_value(mesh, n0, n1, n2, p);
}

template <typename R>
__global__ __launch_bounds__(kCicThreads) void k_read_fd(const R* __restrict__ mesh, int n0,
                                                         int n1, int n2, R c0, R c1, R c2,
                                                         const R* __restrict__ pos,
                                                         R* __restrict__ out, size_t m) {
  const size_t i = blockIdx.x * static_cast<size_t>(kCicThreads) + threadIdx.x;
  if (i >= m) return;
  const R p[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
  const R cell[3] = {c0, c1, c2};
  R g[3];
  fd_gradient(mesh, n0, n1, n2, cell, p, g);
  out[3 * i] = g[0];
  out[3 * i + 1] = g[1];
  out[3 * i + 2] = g[2];
}

// The potential as a read sees it: global memory with periodic indices.
template <typename R>
struct GlobalMesh {
  const R* __restrict__ phi;
  int n0, n1, n2;
  __device__ __forceinline__ R value(const R p[3]) const { return cic_value(phi, n0, n1, n2, p); }
  __device__ __forceinline__ void gradient(const R cell[3], const R p[3], R g[3]) const {
    fd_gradient(phi, n0, n1, n2, cell, p, g);
  }
};

// The (kPT + 3)^3 cells o - 1 .. o + kPT + 1 of a tile with origin o, staged in LDS
// (periodic indices): every cell the value (base .. base + 1) or the central
// differences (base - 1 .. base + 2) of a point whose base cell lies in the tile
// reads; other points read global memory.  Same values, same arithmetic: the
// results are GlobalMesh's bit for bit.
constexpr int kTS = kPT + 3;  // staged cells per axis
template <typename R>
struct TileMesh {
  const R* lds;
  int o[3];
  GlobalMesh<R> g;
  __device__ __forceinline__ bool local(const R p[3], int u[3], R t[3]) const {
    int b[3];
    const int nn[3] = {g.n0, g.n1, g.n2};
    cic_base(p, nn, b, t);
    bool in = true;
    for (int d = 0; d < 3; d++) {
      u[d] = b[d] - o[d];
      in = in && u[d] >= 0 && u[d] < kPT;
    }
    return in;
  }
  __device__ __forceinline__ R value(const R p[3]) const {
    int u[3];
    R t[3];
    if (!local(p, u, t)) return g.value(p);
    const R* c = lds + ((u[0] + 1) * kTS + (u[1] + 1)) * kTS + (u[2] + 1);
    return cic_core<R>([&](int a, int bb, int e) { return c[(a * kTS + bb) * kTS + e]; }, t);
  }
  __device__ __forceinline__ void gradient(const R cell[3], const R p[3], R gr[3]) const {
    int u[3];
    R t[3];
    if (!local(p, u, t)) return g.gradient(cell, p, gr);
    const R* c = lds + (u[0] * kTS + u[1]) * kTS + u[2];
    fd_core<R>([&](int i, int j, int k) { return c[(i * kTS + j) * kTS + k]; }, t, cell, gr);
  }
};

// Parameters of a fused read shared by all points.
template <typename R>
struct ReadArgs {
  Frame<R> F;
  R cell[3];
  double fgrowth;
  R los[3];
  int local, field, wrap, shifted;
};

constexpr int kRT = 8;  // read binning tile edge (cells)

// Fused read of the shifts and shifted positions (multigrid.py:137-158,
// recon.py:304-325), with numpy's dtypes: the gradient is cast to the
// positions dtype P; the local line of sight x/|x| is in P, computed from
// the unwrapped input; with a global line of sight (mesh dtype R) the RSD term
// is in Q = promote(P, R).  O is the output dtype.
// SPEC: the shifted positions with the displacement and RSD terms along the
// local line of sight, without wrap (RecSym's read, the common case), fixed at
// compile time so that the registers of the other cases are not held.
// The shifts (or shifted positions) of point i from its displacement s, with numpy's dtypes and
// order (recon.py:255-325): s and the local line of sight x/|x| in P, a global line of sight's
// RSD term in Q = promote(P, R); written at out[3 i .. 3 i + 2].
template <typename R, typename P, typename O>
__device__ __forceinline__ void shift_out(size_t i, const P x[3], const P s[3], const ReadArgs<R>& A, int field,
                                          int shifted, int local, int wrap, O* __restrict__ out) {
  using Q = promote_t<P, R>;
  const Frame<R>& F = A.F;
  // shifts in P (disp, disp+rsd, local rsd) or in Q (global rsd)
  O res[3];
  if (field == PRH_FIELD_DISP) {
    for (int d = 0; d < 3; d++) res[d] = static_cast<O>(s[d]);
    if (shifted)
      for (int d = 0; d < 3; d++) res[d] = static_cast<O>(x[d] - s[d]);
  } else if (local) {
    const P dist = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    const P l[3] = {x[0] / dist, x[1] / dist, x[2] / dist};
    // numpy's sum along the last axis starts from +0 (a sum of three -0 is +0)
    const P dot = ((P(0) + s[0] * l[0]) + s[1] * l[1]) + s[2] * l[2];
    const P fd = static_cast<P>(A.fgrowth) * dot;
    P sh[3];
    for (int d = 0; d < 3; d++) {
      const P rsd = fd * l[d];
      sh[d] = (field == PRH_FIELD_RSD) ? rsd : static_cast<P>(s[d] + rsd);
    }
    for (int d = 0; d < 3; d++)
      res[d] = static_cast<O>(shifted ? static_cast<P>(x[d] - sh[d]) : sh[d]);
  } else {
    const Q l[3] = {static_cast<Q>(A.los[0]), static_cast<Q>(A.los[1]), static_cast<Q>(A.los[2])};
    const Q dot = ((Q(0) + static_cast<Q>(s[0]) * l[0]) + static_cast<Q>(s[1]) * l[1]) +
                  static_cast<Q>(s[2]) * l[2];
    const Q fd = static_cast<Q>(A.fgrowth) * dot;
    for (int d = 0; d < 3; d++) {
      const Q rsd = fd * l[d];
      if (field == PRH_FIELD_RSD) {
        res[d] = static_cast<O>(shifted ? static_cast<Q>(static_cast<Q>(x[d]) - rsd) : rsd);
      } else {
        const P sh = static_cast<P>(static_cast<Q>(s[d]) + rsd);
        res[d] = static_cast<O>(shifted ? static_cast<P>(x[d] - sh) : sh);
      }
    }
  }
  if (shifted && wrap) {
    using QO = promote_t<O, R>;
    for (int d = 0; d < 3; d++)
      res[d] = static_cast<O>(np_mod(static_cast<QO>(res[d]) - static_cast<QO>(F.off[d]),
                                     static_cast<QO>(F.bs[d])) +
                              static_cast<QO>(F.off[d]));
  }
  out[3 * i] = res[0];
  out[3 * i + 1] = res[1];
  out[3 * i + 2] = res[2];
}

template <typename R, typename P, typename O, typename M, bool SPEC = false>
__device__ __forceinline__ uint32_t read_value(size_t i, const P x[3], const M& mesh,
                                               const ReadArgs<R>& A, O* __restrict__ out) {
  const Frame<R>& F = A.F;
  const int wrap = SPEC ? 0 : A.wrap, local = SPEC ? 1 : A.local;
  uint32_t err = 0;
  R p[3];
  to_mesh_units<R, P>(x, F, wrap, p, err);
  if (!SPEC && A.field == PRH_FIELD_VALUE) {
    // RealMesh.read_cic (mesh.py:570-610): the CIC value, cast to the positions dtype
    out[i] = static_cast<O>(static_cast<P>(mesh.value(p)));
    return err;
  }
  R g[3];
  mesh.gradient(A.cell, p, g);
  const P s[3] = {static_cast<P>(g[0]), static_cast<P>(g[1]), static_cast<P>(g[2])};
  const int field = SPEC ? PRH_FIELD_DISP_RSD : A.field, shifted = SPEC ? 1 : A.shifted;
  shift_out<R, P, O>(i, x, s, A, field, shifted, local, wrap, out);
  return err;
}

// The point i of `pos`, written at out[i].
template <typename R, typename P, typename O>
__device__ __forceinline__ uint32_t read_point(size_t i, const R* __restrict__ phi,
                                               const ReadArgs<R>& A, const P* __restrict__ pos,
                                               O* __restrict__ out) {
  const P x[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
  return read_value<R, P, O>(i, x, GlobalMesh<R>{phi, A.F.n[0], A.F.n[1], A.F.n[2]}, A, out);
}

// Wave-level OR of the status flags, one atomic per wave.
__device__ __forceinline__ void flag_errors(uint32_t err, uint32_t* misc) {
  const uint64_t bad = __ballot(err != 0);
  if (bad) {
    uint32_t e = err;
    for (int off = 32; off > 0; off >>= 1) e |= __shfl_xor(e, off);
    if ((threadIdx.x & 63) == 0) atomicOr(&misc[0], e);
  }
}

// Points in input order, reading the mesh from global memory.
template <typename R, typename P, typename O>
__global__ __launch_bounds__(kCicThreads) void k_read_shifts(const R* __restrict__ phi,
                                                             ReadArgs<R> A,
                                                             const P* __restrict__ pos, size_t m,
                                                             O* __restrict__ out,
                                                             uint32_t* __restrict__ misc) {
  const size_t i = blockIdx.x * static_cast<size_t>(kCicThreads) + threadIdx.x;
  uint32_t err = 0;
  if (i < m) err = read_point<R, P, O>(i, phi, A, pos, out);
  flag_errors(err, misc);
}

// Points in binned order (`order`: point indices grouped by cell or by 8^3
// tile), so that the mesh reads of a wave fall in a small neighbourhood and hit
// in cache.  (Dealing contiguous block ranges to the XCDs measured slower
// here: 8.8 against 8.6 ms for 1e8 points at 512^3.)
template <typename R, typename P, typename O>
__global__ __launch_bounds__(kCicThreads) void k_read_binned(const R* __restrict__ phi,
                                                             ReadArgs<R> A,
                                                             const P* __restrict__ pos, size_t m,
                                                             O* __restrict__ out,
                                                             uint32_t* __restrict__ misc,
                                                             const uint32_t* __restrict__ order) {
  const size_t t = blockIdx.x * static_cast<size_t>(kCicThreads) + threadIdx.x;
  uint32_t err = 0;
  if (t < m) err = read_point<R, P, O>(order[t], phi, A, pos, out);
  flag_errors(err, misc);
}

// Points of the catalog painted last in its cell order, streaming the paint's
// copy of the positions (spos[t] is the position of point order[t]); each
// result is written at its point's own index.
template <typename R, typename P, typename O>
__global__ __launch_bounds__(kCicThreads) void k_read_sorted(const R* __restrict__ phi,
                                                             ReadArgs<R> A,
                                                             const P* __restrict__ spos, size_t m,
                                                             O* __restrict__ out,
                                                             uint32_t* __restrict__ misc,
                                                             const uint32_t* __restrict__ order) {
  const size_t t = blockIdx.x * static_cast<size_t>(kCicThreads) + threadIdx.x;
  uint32_t err = 0;
  if (t < m) {
    const P x[3] = {spos[3 * t], spos[3 * t + 1], spos[3 * t + 2]};
    err = read_value<R, P, O>(order[t], x, GlobalMesh<R>{phi, A.F.n[0], A.F.n[1], A.F.n[2]}, A, out);
  }
  flag_errors(err, misc);
}

// The painted positions tile by tile (one workgroup per tile of the tile
// paint, tiles dealt to the XCDs in contiguous ranges): the tile's potential is
// staged in LDS once and each point's 32 mesh reads come from there; results
// go to the points' input positions (order).
#ifndef PRH_READ_TILE_THREADS
#define PRH_READ_TILE_THREADS 512
#endif
constexpr int kReadTileThreads = PRH_READ_TILE_THREADS;
template <typename R, typename P, typename O, bool SPEC>
__global__ __launch_bounds__(kReadTileThreads, 4) void k_read_tiles(const R* __restrict__ phi, ReadArgs<R> A,
                                                            TileGrid TG, uint32_t ntiles, uint32_t per,
                                                            const uint32_t* __restrict__ tstart,
                                                            const P* __restrict__ spos,
                                                            const uint32_t* __restrict__ order,
                                                            O* __restrict__ out, uint32_t* __restrict__ misc) {
  __shared__ R s_m[kTS * kTS * kTS];
  __shared__ size_t s_off[3 * kTS];  // periodic offsets of the staged planes, rows and columns
  const uint32_t t = xcd_block(per);
  if (t >= ntiles) return;
  const uint32_t beg = tstart[t], end = tstart[t + 1];
  if (beg >= end) return;
  const TilePos T = tile_pos(t, TG);
  const int n0 = A.F.n[0], n1 = A.F.n[1], n2 = A.F.n[2];
  if (threadIdx.x < 3 * kTS) {
    const int d = threadIdx.x / kTS, u = threadIdx.x % kTS;
    const int n = d == 0 ? n0 : (d == 1 ? n1 : n2);
    const int o = d == 0 ? T.o[0] : (d == 1 ? T.o[1] : T.o[2]);  // no dynamic index into T
    const int g = ((o - 1 + u) % n + n) % n;
    s_off[threadIdx.x] = static_cast<size_t>(g) * (d == 0 ? static_cast<size_t>(n1) * n2 : (d == 1 ? n2 : 1));
  }
  // the loads of this thread's first point are issued before the staging, to overlap with it
  const uint32_t q0 = beg + threadIdx.x;
  P x0[3] = {P(0), P(0), P(0)};
  uint32_t i0 = 0;
  if (q0 < end) {
    x0[0] = spos[3 * static_cast<size_t>(q0)];
    x0[1] = spos[3 * static_cast<size_t>(q0) + 1];
    x0[2] = spos[3 * static_cast<size_t>(q0) + 2];
    i0 = order[q0];
  }
  __syncthreads();
  for (int c = threadIdx.x; c < kTS * kTS * kTS; c += kReadTileThreads) {
    const int u = c / (kTS * kTS), v = (c / kTS) % kTS, w = c % kTS;
    s_m[c] = phi[s_off[u] + s_off[kTS + v] + s_off[2 * kTS + w]];
  }
  __syncthreads();
  uint32_t err = 0;
  const TileMesh<R> M{s_m, {T.o[0], T.o[1], T.o[2]}, GlobalMesh<R>{phi, n0, n1, n2}};
  if constexpr (SPEC) {
    // the registers the general case leaves free hold the next point's loads while
    // the current point is computed
    for (uint32_t q = q0; q < end; q += kReadTileThreads) {
      const uint32_t qn = q + kReadTileThreads;
      P xn[3] = {P(0), P(0), P(0)};
      uint32_t in = 0;
      if (qn < end) {
        xn[0] = spos[3 * static_cast<size_t>(qn)];
        xn[1] = spos[3 * static_cast<size_t>(qn) + 1];
        xn[2] = spos[3 * static_cast<size_t>(qn) + 2];
        in = order[qn];
      }
      err |= read_value<R, P, O, TileMesh<R>, true>(i0, x0, M, A, out);
      x0[0] = xn[0];
      x0[1] = xn[1];
      x0[2] = xn[2];
      i0 = in;
    }
  } else {
    if (q0 < end) err = read_value<R, P, O>(i0, x0, M, A, out);
    for (uint32_t q = q0 + kReadTileThreads; q < end; q += kReadTileThreads) {
      const P x[3] = {spos[3 * static_cast<size_t>(q)], spos[3 * static_cast<size_t>(q) + 1],
                      spos[3 * static_cast<size_t>(q) + 2]};
      err |= read_value<R, P, O>(order[q], x, M, A, out);
    }
  }
  flag_errors(err, misc);
}

// Read binning: a counting sort of the points by 8^3-cell tile (the order
// inside a tile is irrelevant: every point is evaluated independently and
// written at its own index), so that the reads of one wavefront touch a small
// neighbourhood of the potential and hit in cache.
template <typename R, typename P>
__global__ __launch_bounds__(kCicThreads) void k_bin_key(const P* __restrict__ pos, size_t m,
                                                         Frame<R> F, int wrap, int t1, int t2,
                                                         uint32_t* __restrict__ key,
                                                         uint32_t* __restrict__ rank,
                                                         uint32_t* __restrict__ count) {
  const size_t i = blockIdx.x * static_cast<size_t>(kCicThreads) + threadIdx.x;
  if (i >= m) return;
  const P x[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
  R p[3];
  uint32_t err = 0;
  to_mesh_units<R, P>(x, F, wrap, p, err);
  int t[3];
  for (int d = 0; d < 3; d++) {
    int ip = static_cast<int>(p[d]);
    ip = ip < 0 ? 0 : (ip >= F.n[d] ? F.n[d] - 1 : ip);
    t[d] = ip >> 3;
  }
  const uint32_t k = (static_cast<uint32_t>(t[0]) * t1 + t[1]) * t2 + t[2];
  key[i] = k;
  rank[i] = atomicAdd(&count[k], 1u);
}

// ---------------------------------------------------------------------------
// Host drivers
// ---------------------------------------------------------------------------

// Digit widths of the radix sort of keys in [0, ncell): the fewest passes of
// at most kRadixMaxBits bits, as even as possible.
struct RadixPlan {
  int npass = 0;
  int bits[4] = {};
  int shift[4] = {};
};
static RadixPlan radix_plan(size_t ncell) {
  int nbits = 1;
  while (nbits < 32 && (size_t(1) << nbits) < ncell) nbits++;
  RadixPlan P;
  P.npass = (nbits + kRadixMaxBits - 1) / kRadixMaxBits;
  int sh = 0;
  for (int p = 0; p < P.npass; p++) {
    P.bits[p] = nbits / P.npass + (p < nbits % P.npass ? 1 : 0);
    P.shift[p] = sh;
    sh += P.bits[p];
  }
  return P;
}

template <typename R, typename P, typename W, int MODE>
static int paint_chunk(prh_workspace* ws, R* mesh, const Frame<R>& F, const P* pos, const W* w,
                       size_t m, bool check_only, hipStream_t s) {
  PaintScratch& S = ws->paint;
  S.bins = PaintScratch::Bins();  // the payload memory is overwritten
  const unsigned grid = ceil_div(m, kCicThreads);
  k_paint_key<R, P, W, MODE><<<grid, kCicThreads, 0, s>>>(
      pos, w, m, F, S.key, static_cast<Payload<R>*>(S.payload), S.misc, check_only ? 1 : 0);
  PRH_LAUNCHED();
  if (check_only) return PRH_OK;
  const size_t ncell = S.ncell;
  // sort: keys alternate between key and rank, values between vtmp and sorted (the last
  // pass writes sorted); the first pass takes the input positions as values
  const RadixPlan RP = radix_plan(ncell);
  const uint32_t ntiles = static_cast<uint32_t>(ceil_div(m, kRadixTile));
  const uint32_t per = static_cast<uint32_t>(ceil_div(ntiles, kXcds));
  const unsigned rgrid = per * kXcds;
  uint32_t* kbuf[2] = {S.key, S.rank};
  uint32_t* vbuf[2] = {S.sorted, S.vtmp};
  const uint32_t* vin = nullptr;
  for (int p = 0; p < RP.npass; p++) {
    const size_t nh = (size_t(1) << RP.bits[p]) * ntiles;
    const unsigned nscan = ceil_div(nh, kScanBlock);
    k_radix_count<<<rgrid, kRadixThreads, 0, s>>>(kbuf[p & 1], m, RP.shift[p], RP.bits[p], ntiles,
                                                  per, S.hist);
    PRH_LAUNCHED();
    k_scan_reduce<<<nscan, kCicThreads, 0, s>>>(S.hist, nh, S.partial);
    PRH_LAUNCHED();
    k_scan_partials<<<1, kCicThreads, 0, s>>>(S.partial, static_cast<int>(nscan));
    PRH_LAUNCHED();
    k_scan_final<<<nscan, kCicThreads, 0, s>>>(S.hist, nh, S.partial, S.goff);
    PRH_LAUNCHED();
    uint32_t* vout = vbuf[(RP.npass - 1 - p) & 1];
    k_radix_scatter<<<rgrid, kRadixThreads, 0, s>>>(kbuf[p & 1], vin, m, RP.shift[p], RP.bits[p],
                                                    ntiles, per, S.goff, kbuf[(p + 1) & 1], vout);
    PRH_LAUNCHED();
    vin = vout;
  }
  const uint32_t* skey = kbuf[RP.npass & 1];
  const uint32_t nsb = static_cast<uint32_t>(ceil_div(ncell, kStartCells));
  k_start_bounds<<<ceil_div(m + 1, kCicThreads), kCicThreads, 0, s>>>(skey, m, nsb, S.bnd, 0);
  PRH_LAUNCHED();
  k_cell_start<<<nsb, kCicThreads, 0, s>>>(skey, S.bnd, ncell, m, S.start, 0);
  PRH_LAUNCHED();
  k_paint_permute<R><<<grid, kCicThreads, 0, s>>>(S.sorted,
                                                  static_cast<const Payload<R>*>(S.payload),
                                                  static_cast<Payload<R>*>(S.spay), m);
  PRH_LAUNCHED();
  timer_before(PRH_TIMER_PAINT_GATHER, s);
  {
    const int t0 = (F.n[0] + kGT - 1) / kGT, t1 = (F.n[1] + kGT - 1) / kGT,
              t2 = (F.n[2] + kGT - 1) / kGT;
    const unsigned gtiles = static_cast<unsigned>(static_cast<size_t>(t0) * t1 * t2);
    const uint32_t per = xcd_per(gtiles);
    if (100 * static_cast<double>(m) < PRH_GATHER_DIRECT_PER100 * static_cast<double>(ncell))
      k_paint_gather_direct<R><<<per * kXcds, kGThreads, 0, s>>>(
          mesh, F.n[0], F.n[1], F.n[2], t1, t2, gtiles, per, S.start, S.sorted,
          static_cast<const Payload<R>*>(S.spay), S.misc);
    else
      k_paint_gather_tiled<R><<<per * kXcds, kGThreads, 0, s>>>(
          mesh, F.n[0], F.n[1], F.n[2], t1, t2, gtiles, per, S.start, S.sorted,
          static_cast<const Payload<R>*>(S.spay), S.misc);
  }
  // algorithmic: cell starts + 8 contributions' payload per point + mesh read/write
  timer_after(PRH_TIMER_PAINT_GATHER, s,
              static_cast<double>(ncell) * (4 + 2 * sizeof(R)) +
                  static_cast<double>(m) * (4 + sizeof(Payload<R>)));
  PRH_LAUNCHED();
  return PRH_OK;
}

// Status flags of the last paint; with `hash`, also the hash of its positions.
static int read_flags(prh_workspace* ws, hipStream_t s, uint32_t* flags, uint64_t* hash = nullptr) {
  PRH_HIP(hipMemcpyAsync(ws->flag_host, ws->paint.misc, 2 * sizeof(uint32_t),
                         hipMemcpyDeviceToHost, s));
  if (hash)
    PRH_HIP(hipMemcpyAsync(ws->red_host, ws->paint.hash, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  PRH_HIP(hipStreamSynchronize(s));
  *flags = ws->flag_host[0];
  if (hash) std::memcpy(hash, ws->red_host, sizeof(uint64_t));
  return PRH_OK;
}

static int flags_to_status(uint32_t flags) {
  if (flags & kErrRange) {
    set_error("positions not in box range");
    return PRH_ERR_RANGE;
  }
  if (flags & kErrNonFinite) {
    set_error("Some positions or weights are NaN/inf");
    return PRH_ERR_NONFINITE;
  }
  return PRH_OK;
}

template <typename R, typename P, typename W, int MODE>
static int paint_all(prh_workspace* ws, R* mesh, const Frame<R>& F, const P* pos, const W* w,
                     size_t n, hipStream_t s) {
  PaintScratch& S = ws->paint;
  // this paint's order goes to the older of the two order arrays
  std::swap(S.sorted, S.sorted_prev);
  std::swap(S.spos, S.spos_prev);
  std::swap(S.tstart, S.tstart_prev);
  S.order_prev = S.order;
  S.order = PaintScratch::Order();
  k_zero_misc<<<1, 64, 0, s>>>(S.misc);
  PRH_LAUNCHED();
  if (n == 0) return PRH_OK;
  const bool multi = n > S.chunk;
  if (multi && MODE != 0) {
    // validate every point before the first chunk touches the mesh
    for (size_t o = 0; o < n; o += S.chunk) {
      const size_t m = n - o < S.chunk ? n - o : S.chunk;
      PRH_TRY((paint_chunk<R, P, W, MODE>(ws, mesh, F, pos + 3 * o, w ? w + o : nullptr, m, true, s)));
    }
    uint32_t flags;
    PRH_TRY(read_flags(ws, s, &flags));
    PRH_TRY(flags_to_status(flags));
  }
  for (size_t o = 0; o < n; o += S.chunk) {
    const size_t m = n - o < S.chunk ? n - o : S.chunk;
    if (o > 0) {
      k_zero_misc<<<1, 64, 0, s>>>(S.misc);
      PRH_LAUNCHED();
    }
    // no copy of the positions here: gathering them into cell order costs this paint more than
    // streaming them saves a later read (its cell order is kept and reused instead)
    PRH_TRY((paint_chunk<R, P, W, MODE>(ws, mesh, F, pos + 3 * o, w ? w + o : nullptr, m, false, s)));
  }
  if (MODE == 0 || multi) return PRH_OK;
  uint32_t flags;
  PRH_TRY(read_flags(ws, s, &flags));
  PRH_TRY(flags_to_status(flags));
  S.order.pos = pos;
  S.order.n = n;
  S.order.pdt = Tag<P>::code;
  return PRH_OK;
}

// The tile paint of one chunk (see k_tile_accumulate).  `keep`: also hash the
// positions into S.hash[0], so that a later read of the same catalog may stream
// the positions this paint lays out in tile order (S.spos).
template <typename P, typename W, int MODE>
static int tile_paint_chunk(prh_workspace* ws, double* mesh, const Frame<double>& F,
                            const TileGrid& TG, const P* pos, const W* w, size_t m, bool check_only,
                            hipStream_t s, bool keep) {
  PaintScratch& S = ws->paint;
  S.bins = PaintScratch::Bins();  // the payload memory is overwritten
  const unsigned grid = ceil_div(m, kCicThreads);
  keep = keep && !check_only && grid <= S.nhpart;
  // the packed records go to the paint payload memory (32 bytes per point at most; chunk x 4 x 8 held)
  TileRec<P>* rec = static_cast<TileRec<P>*>(S.payload);
  k_tile_key<P, W, MODE><<<grid, kCicThreads, 0, s>>>(pos, w, m, F, TG, S.key, rec, S.misc,
                                                      keep ? S.hpart : nullptr);
  PRH_LAUNCHED();
  if (check_only) return PRH_OK;
  if (keep) {
    k_hash_final<<<1, 1024, 0, s>>>(S.hpart, grid, S.hash);
    PRH_LAUNCHED();
  }
  const size_t ntiles = static_cast<size_t>(TG.t[0]) * TG.t[1] * TG.t[2];
  // sort the (key, index) pairs by tile: digits above the 9 cell bits
  RadixPlan RP = radix_plan(ntiles);
  for (int p = 0; p < RP.npass; p++) RP.shift[p] += kPTShift;
  const uint32_t rtiles = static_cast<uint32_t>(ceil_div(m, kRadixTile));
  const uint32_t rper = static_cast<uint32_t>(ceil_div(rtiles, kXcds));
  uint32_t* kbuf[2] = {S.key, S.rank};
  uint32_t* vbuf[2] = {S.sorted, S.vtmp};
  const uint32_t* vin = nullptr;
  for (int p = 0; p < RP.npass; p++) {
    const size_t nh = (size_t(1) << RP.bits[p]) * rtiles;
    const unsigned nscan = ceil_div(nh, kScanBlock);
    k_radix_count<<<rper * kXcds, kRadixThreads, 0, s>>>(kbuf[p & 1], m, RP.shift[p], RP.bits[p], rtiles,
                                                         rper, S.hist);
    PRH_LAUNCHED();
    k_scan_reduce<<<nscan, kCicThreads, 0, s>>>(S.hist, nh, S.partial);
    PRH_LAUNCHED();
    k_scan_partials<<<1, kCicThreads, 0, s>>>(S.partial, static_cast<int>(nscan));
    PRH_LAUNCHED();
    k_scan_final<<<nscan, kCicThreads, 0, s>>>(S.hist, nh, S.partial, S.goff);
    PRH_LAUNCHED();
    uint32_t* vout = vbuf[(RP.npass - 1 - p) & 1];
    k_radix_scatter<<<rper * kXcds, kRadixThreads, 0, s>>>(kbuf[p & 1], vin, m, RP.shift[p], RP.bits[p],
                                                           rtiles, rper, S.goff, kbuf[(p + 1) & 1], vout);
    PRH_LAUNCHED();
    vin = vout;
  }
  const uint32_t* skey = kbuf[RP.npass & 1];
  k_tile_starts<<<ceil_div(ntiles + 1, kCicThreads), kCicThreads, 0, s>>>(skey, m, static_cast<uint32_t>(ntiles),
                                                                         S.tstart);
  PRH_LAUNCHED();
  // the weights in tile order go to the sorted-payload memory (free in this paint)
  double* sw = static_cast<double*>(S.spay);
  k_tile_permute<P><<<grid, kCicThreads, 0, s>>>(S.sorted, rec, m, static_cast<P*>(S.spos), sw);
  PRH_LAUNCHED();
  const uint32_t per = xcd_per(ntiles);
  timer_before(PRH_TIMER_PAINT_GATHER, s);
  k_tile_accumulate<P><<<per * kXcds, kAccThreads, 0, s>>>(mesh, F, TG, MODE == 2, static_cast<uint32_t>(ntiles),
                                                           per, S.tstart, skey, static_cast<const P*>(S.spos), sw,
                                                           S.tbnd, S.misc);
  // algorithmic: tile starts, key, position and weight per point, the mesh read and written
  timer_after(PRH_TIMER_PAINT_GATHER, s,
              static_cast<double>(ntiles) * 4 + static_cast<double>(m) * (4 + 3 * sizeof(P) + 8) +
                  static_cast<double>(S.ncell) * 16);
  PRH_LAUNCHED();
  k_tile_boundary<<<per * kXcds, kCicThreads, 0, s>>>(mesh, TG, static_cast<uint32_t>(ntiles), per, S.tstart,
                                                      S.tbnd, S.misc);
  PRH_LAUNCHED();
  return PRH_OK;
}

template <typename P, typename W, int MODE>
static int tile_paint_all(prh_workspace* ws, double* mesh, const Frame<double>& F, const P* pos,
                          const W* w, size_t n, hipStream_t s) {
  PaintScratch& S = ws->paint;
  TileGrid TG;
  for (int d = 0; d < 3; d++) {
    TG.n[d] = F.n[d];
    TG.t[d] = (F.n[d] + kPT - 1) / kPT;
  }
  std::swap(S.sorted, S.sorted_prev);
  std::swap(S.spos, S.spos_prev);
  std::swap(S.tstart, S.tstart_prev);
  S.order_prev = S.order;
  S.order = PaintScratch::Order();
  k_zero_misc<<<1, 64, 0, s>>>(S.misc);
  PRH_LAUNCHED();
  if (n == 0) return PRH_OK;
  const bool multi = n > S.chunk;
  if (multi) {
    // validate every point before the first chunk touches the mesh
    for (size_t o = 0; o < n; o += S.chunk) {
      const size_t m = n - o < S.chunk ? n - o : S.chunk;
      PRH_TRY((tile_paint_chunk<P, W, MODE>(ws, mesh, F, TG, pos + 3 * o, w ? w + o : nullptr, m, true, s,
                                            false)));
    }
    uint32_t flags;
    PRH_TRY(read_flags(ws, s, &flags));
    PRH_TRY(flags_to_status(flags));
  }
  for (size_t o = 0; o < n; o += S.chunk) {
    const size_t m = n - o < S.chunk ? n - o : S.chunk;
    if (o > 0) {
      k_zero_misc<<<1, 64, 0, s>>>(S.misc);
      PRH_LAUNCHED();
    }
    PRH_TRY((tile_paint_chunk<P, W, MODE>(ws, mesh, F, TG, pos + 3 * o, w ? w + o : nullptr, m, false, s,
                                          !multi)));
  }
  if (multi) return PRH_OK;
  uint32_t flags;
  uint64_t hash = 0;
  PRH_TRY(read_flags(ws, s, &flags, &hash));
  PRH_TRY(flags_to_status(flags));
  S.order.pos = pos;
  S.order.n = n;
  S.order.pdt = Tag<P>::code;
  S.order.copy = ceil_div(n, kCicThreads) <= S.nhpart;
  S.order.tiles = S.order.copy;
  S.order.hash = hash;
  return PRH_OK;
}

template <typename P>
static int tile_paint_dispatch(prh_workspace* ws, double* mesh, const Frame<double>& F, const P* pos,
                               const void* w, int wdt, size_t n, int wrap, hipStream_t s) {
  if (!w || wdt == PRH_F64) {
    const double* wd = static_cast<const double*>(w);
    return wrap ? tile_paint_all<P, double, 2>(ws, mesh, F, pos, wd, n, s)
                : tile_paint_all<P, double, 1>(ws, mesh, F, pos, wd, n, s);
  }
  if (wdt == PRH_F32) {
    const float* wf = static_cast<const float*>(w);
    return wrap ? tile_paint_all<P, float, 2>(ws, mesh, F, pos, wf, n, s)
                : tile_paint_all<P, float, 1>(ws, mesh, F, pos, wf, n, s);
  }
  set_error("weight dtype must be PRH_F32 or PRH_F64");
  return PRH_ERR_ARG;
}

static int check_paint_ws(const prh_workspace* ws, const size_t* nmesh, int dtype) {
  if (!ws || !nmesh) {
    set_error("workspace and nmesh are required");
    return PRH_ERR_ARG;
  }
  if (!(ws->flags & PRH_WS_PAINT)) {
    set_error("workspace has no paint scratch");
    return PRH_ERR_STATE;
  }
  if (ws->dtype != dtype) {
    set_error("workspace precision does not match the mesh");
    return PRH_ERR_ARG;
  }
  for (int d = 0; d < 3; d++)
    if (ws->n[d] != nmesh[d]) {
      set_error("workspace mesh shape does not match nmesh");
      return PRH_ERR_CAPACITY;
    }
  return PRH_OK;
}

template <typename R>
static int do_assign_cic(prh_workspace* ws, R* mesh, const size_t* nmesh, const R* pos,
                         const R* w, size_t n, void* stream) {
  PRH_TRY(check_paint_ws(ws, nmesh, Tag<R>::code));
  if (n && (!mesh || !pos || !w)) return PRH_ERR_ARG;
  const R one[3] = {R(1), R(1), R(1)}, zero[3] = {R(0), R(0), R(0)};
  const Frame<R> F = make_frame<R>(nmesh, one, zero);
  return paint_all<R, R, R, 0>(ws, mesh, F, pos, w, n, as_stream(stream));
}

template <typename R, typename P, typename W>
static int paint_dispatch_w(prh_workspace* ws, R* mesh, const Frame<R>& F, const P* pos,
                            const W* w, size_t n, int wrap, hipStream_t s) {
  if (wrap) return paint_all<R, P, W, 2>(ws, mesh, F, pos, w, n, s);
  return paint_all<R, P, W, 1>(ws, mesh, F, pos, w, n, s);
}

template <typename R, typename P>
static int paint_dispatch_p(prh_workspace* ws, R* mesh, const Frame<R>& F, const P* pos,
                            const void* w, int wdt, size_t n, int wrap, hipStream_t s) {
  if (!w || wdt == PRH_F64)
    return paint_dispatch_w<R, P, double>(ws, mesh, F, pos, static_cast<const double*>(w), n,
                                          wrap, s);
  if (wdt == PRH_F32)
    return paint_dispatch_w<R, P, float>(ws, mesh, F, pos, static_cast<const float*>(w), n, wrap,
                                         s);
  set_error("weight dtype must be PRH_F32 or PRH_F64");
  return PRH_ERR_ARG;
}

template <typename R>
static int do_paint(prh_workspace* ws, R* mesh, const size_t* nmesh, const R* boxsize,
                    const R* boxcenter, const void* pos, int pdt, const void* w, int wdt, size_t n,
                    int wrap, void* stream) {
  PRH_TRY(check_paint_ws(ws, nmesh, Tag<R>::code));
  if (!boxsize || !boxcenter || (n && (!mesh || !pos))) return PRH_ERR_ARG;
  const Frame<R> F = make_frame<R>(nmesh, boxsize, boxcenter);
  hipStream_t s = as_stream(stream);
  if (pdt == PRH_F64)
    return paint_dispatch_p<R, double>(ws, mesh, F, static_cast<const double*>(pos), w, wdt, n,
                                       wrap, s);
  if (pdt == PRH_F32)
    return paint_dispatch_p<R, float>(ws, mesh, F, static_cast<const float*>(pos), w, wdt, n,
                                      wrap, s);
  set_error("positions dtype must be PRH_F32 or PRH_F64");
  return PRH_ERR_ARG;
}

static int do_paint_mixed(prh_workspace* ws, double* mesh, const size_t* nmesh, const double* boxsize,
                          const double* boxcenter, const void* pos, int pdt, const void* w, int wdt,
                          size_t n, int wrap, void* stream) {
  PRH_TRY(check_paint_ws(ws, nmesh, PRH_F64));
  if (!boxsize || !boxcenter || (n && (!mesh || !pos))) return PRH_ERR_ARG;
  if (!ws->paint.tbnd) {
    set_error("the tile paint needs a workspace with PRH_WS_PAINT | PRH_WS_MIXED");
    return PRH_ERR_STATE;
  }
  size_t ntiles = 1;
  for (int d = 0; d < 3; d++) ntiles *= (nmesh[d] + kPT - 1) / kPT;
  if (ntiles > (size_t(1) << (32 - kPTShift))) {
    set_error("mesh has too many tiles for 32-bit tile keys");
    return PRH_ERR_ARG;
  }
  const Frame<double> F = make_frame<double>(nmesh, boxsize, boxcenter);
  hipStream_t s = as_stream(stream);
  if (pdt == PRH_F64)
    return tile_paint_dispatch<double>(ws, mesh, F, static_cast<const double*>(pos), w, wdt, n, wrap, s);
  if (pdt == PRH_F32)
    return tile_paint_dispatch<float>(ws, mesh, F, static_cast<const float*>(pos), w, wdt, n, wrap, s);
  set_error("positions dtype must be PRH_F32 or PRH_F64");
  return PRH_ERR_ARG;
}

template <typename R>
static int do_read_cic(const R* mesh, const size_t* nmesh, const R* pos, R* out, size_t n,
                       void* stream) {
  if (!nmesh || (n && (!mesh || !pos || !out))) return PRH_ERR_ARG;
  if (n == 0) return PRH_OK;
  k_read_cic<R><<<ceil_div(n, kCicThreads), kCicThreads, 0, as_stream(stream)>>>(
      mesh, static_cast<int>(nmesh[0]), static_cast<int>(nmesh[1]), static_cast<int>(nmesh[2]),
      pos, out, n);
  PRH_LAUNCHED();
  return PRH_OK;
}

// cell = 2.0 * boxsize / nmesh evaluated in double and stored in R (src/mesh.c:197)
template <typename R>
static void fd_cells(const size_t* nmesh, const R* boxsize, R cell[3]) {
  for (int d = 0; d < 3; d++)
    cell[d] = static_cast<R>(2.0 * static_cast<double>(boxsize[d]) / static_cast<double>(nmesh[d]));
}

template <typename R>
static int do_read_fd(const R* mesh, const size_t* nmesh, const R* boxsize, const R* pos, R* out,
                      size_t n, void* stream) {
  if (!nmesh || !boxsize || (n && (!mesh || !pos || !out))) return PRH_ERR_ARG;
  if (n == 0) return PRH_OK;
  R cell[3];
  fd_cells(nmesh, boxsize, cell);
  k_read_fd<R><<<ceil_div(n, kCicThreads), kCicThreads, 0, as_stream(stream)>>>(
      mesh, static_cast<int>(nmesh[0]), static_cast<int>(nmesh[1]), static_cast<int>(nmesh[2]),
      cell[0], cell[1], cell[2], pos, out, n);
  PRH_LAUNCHED();
  return PRH_OK;
}

// Shifts of n points.  With a workspace holding paint scratch for this mesh,
// the points are binned by 8^3 tile (a counting sort, in chunks of the paint
// scratch size) and read in binned order; otherwise in input order.
template <typename R, typename P, typename O>
static int read_shifts_all(prh_workspace* ws, const R* phi, const ReadArgs<R>& A, const P* pos,
                           size_t n, O* out, uint32_t* misc, hipStream_t s) {
  const Frame<R>& F = A.F;
  const size_t ostride = A.field == PRH_FIELD_VALUE ? 1 : 3;
  // algorithmic: positions in, results out, the potential read once
  const double bytes = static_cast<double>(n) * (3 * sizeof(P) + ostride * sizeof(O)) +
                       static_cast<double>(F.n[0]) * F.n[1] * F.n[2] * sizeof(R);
  const bool binned = (ws->flags & PRH_WS_PAINT) && n >= 65536 && ws->n[0] == size_t(F.n[0]) &&
                      ws->n[1] == size_t(F.n[1]) && ws->n[2] == size_t(F.n[2]);
  if (!binned) {
    timer_before(PRH_TIMER_READ_SHIFTS, s);
    k_read_shifts<R, P, O><<<ceil_div(n, kCicThreads), kCicThreads, 0, s>>>(phi, A, pos, n, out, misc);
    timer_after(PRH_TIMER_READ_SHIFTS, s, bytes);
    PRH_LAUNCHED();
    return PRH_OK;
  }
  PaintScratch& S = ws->paint;
  auto painted = [&](const PaintScratch::Order& o) {
    return o.pos == static_cast<const void*>(pos) && o.n == n && o.pdt == Tag<P>::code;
  };
  const bool last = painted(S.order);
  const uint32_t* porder = last ? S.sorted : painted(S.order_prev) ? S.sorted_prev : nullptr;
  const PaintScratch::Order* po = last ? &S.order : &S.order_prev;
  if (porder && po->copy) {
    // the paint kept these positions in cell order: if they are unchanged (same hash), stream
    // the copy; the hash pass reads the positions once, in order
    const unsigned hgrid = ceil_div(n, static_cast<size_t>(kCicThreads) * kHashItems);
    if (hgrid <= S.nhpart) {
      k_pos_hash<P><<<hgrid, kCicThreads, 0, s>>>(pos, n, S.hpart);
      PRH_LAUNCHED();
      k_hash_final<<<1, 1024, 0, s>>>(S.hpart, hgrid, S.hash + 1);
      PRH_LAUNCHED();
      PRH_HIP(hipMemcpyAsync(ws->red_host, S.hash + 1, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
      PRH_HIP(hipStreamSynchronize(s));
      uint64_t h;
      std::memcpy(&h, ws->red_host, sizeof(uint64_t));
      if (h == po->hash) {
        const P* spos = static_cast<const P*>(last ? S.spos : S.spos_prev);
        const uint32_t* ts = last ? S.tstart : S.tstart_prev;
        timer_before(PRH_TIMER_READ_SHIFTS, s);
        if (po->tiles && ts) {
          TileGrid TG;
          for (int d = 0; d < 3; d++) {
            TG.n[d] = F.n[d];
            TG.t[d] = (F.n[d] + kPT - 1) / kPT;
          }
          const size_t ntiles = static_cast<size_t>(TG.t[0]) * TG.t[1] * TG.t[2];
          const uint32_t per = xcd_per(ntiles);
          if (A.field == PRH_FIELD_DISP_RSD && A.local && A.shifted && !A.wrap)
            k_read_tiles<R, P, O, true><<<per * kXcds, kReadTileThreads, 0, s>>>(
                phi, A, TG, static_cast<uint32_t>(ntiles), per, ts, spos, porder, out, misc);
          else
            k_read_tiles<R, P, O, false><<<per * kXcds, kReadTileThreads, 0, s>>>(
                phi, A, TG, static_cast<uint32_t>(ntiles), per, ts, spos, porder, out, misc);
        } else {
          k_read_sorted<R, P, O><<<ceil_div(n, kCicThreads), kCicThreads, 0, s>>>(phi, A, spos, n, out,
                                                                                 misc, porder);
        }
        timer_after(PRH_TIMER_READ_SHIFTS, s, bytes);
        PRH_LAUNCHED();
        return PRH_OK;
      }
    }
  }
  if (porder) {
    // one of the last two catalogs painted: its cell order (a permutation of [0, n)) serves as
    // the bins
    timer_before(PRH_TIMER_READ_SHIFTS, s);
    k_read_binned<R, P, O><<<ceil_div(n, kCicThreads), kCicThreads, 0, s>>>(phi, A, pos, n, out, misc,
                                                                           porder);
    timer_after(PRH_TIMER_READ_SHIFTS, s, bytes);
    PRH_LAUNCHED();
    return PRH_OK;
  }
  // the tile order goes to the memory of the point-order payload (dead after a paint), so that
  // both paint orders are kept
  uint32_t* order = static_cast<uint32_t*>(S.payload);
  if (n <= S.chunk && S.bins.pos == static_cast<const void*>(pos) && S.bins.n == n && S.bins.pdt == Tag<P>::code) {
    // the same positions were binned for the previous read (e.g. another component's mesh)
    timer_before(PRH_TIMER_READ_SHIFTS, s);
    k_read_binned<R, P, O><<<ceil_div(n, kCicThreads), kCicThreads, 0, s>>>(phi, A, pos, n, out, misc, order);
    timer_after(PRH_TIMER_READ_SHIFTS, s, bytes);
    PRH_LAUNCHED();
    return PRH_OK;
  }
  S.bins = PaintScratch::Bins();
  const int t0 = (F.n[0] + kRT - 1) / kRT, t1 = (F.n[1] + kRT - 1) / kRT, t2 = (F.n[2] + kRT - 1) / kRT;
  const size_t ntile = static_cast<size_t>(t0) * t1 * t2;
  const unsigned nscan = ceil_div(ntile, kScanBlock);
  if (n <= S.chunk) {
    S.bins.pos = pos;
    S.bins.n = n;
    S.bins.pdt = Tag<P>::code;
  }
  for (size_t o = 0; o < n; o += S.chunk) {
    const size_t m = n - o < S.chunk ? n - o : S.chunk;
    const unsigned grid = ceil_div(m, kCicThreads);
    k_bin_key<R, P><<<grid, kCicThreads, 0, s>>>(pos + 3 * o, m, F, A.wrap, t1, t2, S.key, S.rank, S.count);
    PRH_LAUNCHED();
    k_scan_reduce<<<nscan, kCicThreads, 0, s>>>(S.count, ntile, S.partial);
    PRH_LAUNCHED();
    k_scan_partials<<<1, kCicThreads, 0, s>>>(S.partial, static_cast<int>(nscan));
    PRH_LAUNCHED();
    k_scan_final<<<nscan, kCicThreads, 0, s>>>(S.count, ntile, S.partial, S.start);
    PRH_LAUNCHED();
    k_group_scatter<<<grid, kCicThreads, 0, s>>>(S.key, S.rank, m, S.start, order);
    PRH_LAUNCHED();
    timer_before(PRH_TIMER_READ_SHIFTS, s);
    k_read_binned<R, P, O><<<grid, kCicThreads, 0, s>>>(phi, A, pos + 3 * o, m, out + ostride * o,
                                                        misc, order);
    timer_after(PRH_TIMER_READ_SHIFTS, s,
                static_cast<double>(m) * (3 * sizeof(P) + ostride * sizeof(O)) +
                    static_cast<double>(F.n[0]) * F.n[1] * F.n[2] * sizeof(R));
    PRH_LAUNCHED();
  }
  return PRH_OK;
}

// Shifts of the FFT reconstructions (recon.py:255-325 without wrap): the displacement of point i
// is (s0[i], s1[i], s2[i]), read by CIC from the three displacement meshes in the positions dtype.
template <typename R, typename P, typename O>
__global__ __launch_bounds__(kCicThreads) void k_shift_algebra(const P* __restrict__ pos, const P* __restrict__ s0,
                                                               const P* __restrict__ s1, const P* __restrict__ s2,
                                                               size_t n, ReadArgs<R> A, O* __restrict__ out) {
  const size_t i = blockIdx.x * static_cast<size_t>(kCicThreads) + threadIdx.x;
  if (i >= n) return;
  const P x[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
  const P s[3] = {s0[i], s1[i], s2[i]};
  shift_out<R, P, O>(i, x, s, A, A.field, A.shifted, A.local, 0, out);
}

template <typename R>
static int do_shift_algebra(const void* pos, int pdt, const void* s0, const void* s1, const void* s2, size_t n,
                            double f, const R* los, int field, int shifted, void* out, int odt, void* stream) {
  if (n && (!pos || !s0 || !s1 || !s2 || !out)) return PRH_ERR_ARG;
  if (field != PRH_FIELD_DISP && field != PRH_FIELD_RSD && field != PRH_FIELD_DISP_RSD) {
    set_error("unknown field %d", field);
    return PRH_ERR_ARG;
  }
  if ((pdt != PRH_F32 && pdt != PRH_F64) || (odt != PRH_F32 && odt != PRH_F64)) {
    set_error("positions/output dtype must be PRH_F32 or PRH_F64");
    return PRH_ERR_ARG;
  }
  const int q = pdt > Tag<R>::code ? pdt : Tag<R>::code;
  const int expect = (field == PRH_FIELD_RSD && los) ? q : pdt;
  if (odt != expect) {
    set_error("output dtype %d does not follow numpy promotion (expected %d)", odt, expect);
    return PRH_ERR_ARG;
  }
  if (n == 0) return PRH_OK;
  ReadArgs<R> A{};
  A.fgrowth = f;
  for (int d = 0; d < 3; d++) A.los[d] = los ? los[d] : R(0);
  A.local = los ? 0 : 1;
  A.field = field;
  A.wrap = 0;
  A.shifted = shifted;
  const unsigned grid = static_cast<unsigned>(ceil_div(n, static_cast<size_t>(kCicThreads)));
  hipStream_t s = as_stream(stream);
  if (pdt == PRH_F64)
    k_shift_algebra<R, double, double><<<grid, kCicThreads, 0, s>>>(
        static_cast<const double*>(pos), static_cast<const double*>(s0), static_cast<const double*>(s1),
        static_cast<const double*>(s2), n, A, static_cast<double*>(out));
  else if (odt == PRH_F32)
    k_shift_algebra<R, float, float><<<grid, kCicThreads, 0, s>>>(
        static_cast<const float*>(pos), static_cast<const float*>(s0), static_cast<const float*>(s1),
        static_cast<const float*>(s2), n, A, static_cast<float*>(out));
  else
    k_shift_algebra<R, float, double><<<grid, kCicThreads, 0, s>>>(
        static_cast<const float*>(pos), static_cast<const float*>(s0), static_cast<const float*>(s1),
        static_cast<const float*>(s2), n, A, static_cast<double*>(out));
  PRH_LAUNCHED();
  return PRH_OK;
}

template <typename R>
static int do_read_shifts(prh_workspace* ws, const R* phi, const size_t* nmesh, const R* boxsize, const R* boxcenter,
                          const void* pos, int pdt, size_t n, double f, const R* los, int field,
                          int wrap, int shifted, void* out, int odt, int* nonfinite,
                          void* stream) {
  if (!ws || !nmesh || !boxsize || !boxcenter || (n && (!phi || !pos || !out))) return PRH_ERR_ARG;
  if (field != PRH_FIELD_DISP && field != PRH_FIELD_RSD && field != PRH_FIELD_DISP_RSD &&
      field != PRH_FIELD_VALUE) {
    set_error("unknown field %d", field);
    return PRH_ERR_ARG;
  }
  if (field == PRH_FIELD_VALUE && shifted) {
    set_error("PRH_FIELD_VALUE reads mesh values: shifted must be 0");
    return PRH_ERR_ARG;
  }
  if ((pdt != PRH_F32 && pdt != PRH_F64) || (odt != PRH_F32 && odt != PRH_F64)) {
    set_error("positions/output dtype must be PRH_F32 or PRH_F64");
    return PRH_ERR_ARG;
  }
  // expected output dtype under numpy's promotion rules
  const int q = pdt > Tag<R>::code ? pdt : Tag<R>::code;
  int expect = pdt;
  if (field == PRH_FIELD_RSD && los) expect = q;
  if (shifted && wrap) expect = expect > Tag<R>::code ? expect : Tag<R>::code;
  if (odt != expect) {
    set_error("output dtype %d does not follow numpy promotion (expected %d)", odt, expect);
    return PRH_ERR_ARG;
  }
  if (nonfinite) *nonfinite = 0;
  if (n == 0) return PRH_OK;
  uint32_t* misc = ws->misc;
  hipStream_t s = as_stream(stream);
  k_zero_misc<<<1, 64, 0, s>>>(misc);
  PRH_LAUNCHED();
  ReadArgs<R> A;
  A.F = make_frame<R>(nmesh, boxsize, boxcenter);
  fd_cells(nmesh, boxsize, A.cell);
  A.fgrowth = f;
  for (int d = 0; d < 3; d++) A.los[d] = los ? los[d] : R(0);
  A.local = los ? 0 : 1;
  A.field = field;
  A.wrap = wrap;
  A.shifted = shifted;
  int st;
  if (pdt == PRH_F64) {
    st = read_shifts_all<R, double, double>(ws, phi, A, static_cast<const double*>(pos), n,
                                            static_cast<double*>(out), misc, s);
  } else if (odt == PRH_F32) {
    st = read_shifts_all<R, float, float>(ws, phi, A, static_cast<const float*>(pos), n,
                                          static_cast<float*>(out), misc, s);
  } else {
    st = read_shifts_all<R, float, double>(ws, phi, A, static_cast<const float*>(pos), n,
                                           static_cast<double*>(out), misc, s);
  }
  PRH_TRY(st);
  uint32_t* host = ws->flag_host;
  PRH_HIP(hipMemcpyAsync(host, misc, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  PRH_HIP(hipStreamSynchronize(s));
  if (nonfinite && (host[0] & kErrNonFinite)) *nonfinite = 1;
  if (!wrap && (host[0] & kErrRange)) {
    set_error("positions not in box range");
    return PRH_ERR_RANGE;
  }
  return PRH_OK;
}

}  // namespace prh

using namespace prh;

extern "C" {

int prh_assign_cic_f32(prh_workspace* ws, float* mesh, const size_t* nmesh, const float* positions,
                       const float* weights, size_t npositions, void* stream) {
  return do_assign_cic<float>(ws, mesh, nmesh, positions, weights, npositions, stream);
}
int prh_assign_cic_f64(prh_workspace* ws, double* mesh, const size_t* nmesh,
                       const double* positions, const double* weights, size_t npositions,
                       void* stream) {
  return do_assign_cic<double>(ws, mesh, nmesh, positions, weights, npositions, stream);
}

int prh_read_cic_f32(const float* mesh, const size_t* nmesh, const float* positions, float* values,
                     size_t npositions, void* stream) {
  return do_read_cic<float>(mesh, nmesh, positions, values, npositions, stream);
}
int prh_read_cic_f64(const double* mesh, const size_t* nmesh, const double* positions,
                     double* values, size_t npositions, void* stream) {
  return do_read_cic<double>(mesh, nmesh, positions, values, npositions, stream);
}

int prh_read_finite_difference_cic_f32(const float* mesh, const size_t* nmesh,
                                       const float* boxsize, const float* positions,
                                       float* shifts, size_t npositions, void* stream) {
  return do_read_fd<float>(mesh, nmesh, boxsize, positions, shifts, npositions, stream);
}
int prh_read_finite_difference_cic_f64(const double* mesh, const size_t* nmesh,
                                       const double* boxsize, const double* positions,
                                       double* shifts, size_t npositions, void* stream) {
  return do_read_fd<double>(mesh, nmesh, boxsize, positions, shifts, npositions, stream);
}

int prh_paint_f32(prh_workspace* ws, float* mesh, const size_t* nmesh, const float* boxsize,
                  const float* boxcenter, const void* positions, int pos_dtype,
                  const void* weights, int weight_dtype, size_t npositions, int wrap,
                  void* stream) {
  return do_paint<float>(ws, mesh, nmesh, boxsize, boxcenter, positions, pos_dtype, weights,
                         weight_dtype, npositions, wrap, stream);
}
int prh_paint_f64(prh_workspace* ws, double* mesh, const size_t* nmesh, const double* boxsize,
                  const double* boxcenter, const void* positions, int pos_dtype,
                  const void* weights, int weight_dtype, size_t npositions, int wrap,
                  void* stream) {
  return do_paint<double>(ws, mesh, nmesh, boxsize, boxcenter, positions, pos_dtype, weights,
                          weight_dtype, npositions, wrap, stream);
}

int prh_paint_mixed_f64(prh_workspace* ws, double* mesh, const size_t* nmesh, const double* boxsize,
                        const double* boxcenter, const void* positions, int pdtype, const void* weights,
                        int wdtype, size_t npositions, int wrap, void* stream) {
  return do_paint_mixed(ws, mesh, nmesh, boxsize, boxcenter, positions, pdtype, weights, wdtype, npositions,
                        wrap, stream);
}

int prh_shift_algebra_f32(const void* positions, int pos_dtype, const void* s0, const void* s1, const void* s2,
                          size_t npositions, double f, const float* los, int field, int shifted, void* out,
                          int out_dtype, void* stream) {
  return do_shift_algebra<float>(positions, pos_dtype, s0, s1, s2, npositions, f, los, field, shifted, out,
                                 out_dtype, stream);
}
int prh_shift_algebra_f64(const void* positions, int pos_dtype, const void* s0, const void* s1, const void* s2,
                          size_t npositions, double f, const double* los, int field, int shifted, void* out,
                          int out_dtype, void* stream) {
  return do_shift_algebra<double>(positions, pos_dtype, s0, s1, s2, npositions, f, los, field, shifted, out,
                                  out_dtype, stream);
}

int prh_read_shifts_f32(prh_workspace* ws, const float* phi, const size_t* nmesh, const float* boxsize,
                        const float* boxcenter, const void* positions, int pos_dtype,
                        size_t npositions, double f, const float* los, int field, int wrap,
                        int shifted, void* out, int out_dtype, int* nonfinite, void* stream) {
  return do_read_shifts<float>(ws, phi, nmesh, boxsize, boxcenter, positions, pos_dtype, npositions,
                               f, los, field, wrap, shifted, out, out_dtype, nonfinite, stream);
}
int prh_read_shifts_f64(prh_workspace* ws, const double* phi, const size_t* nmesh, const double* boxsize,
                        const double* boxcenter, const void* positions, int pos_dtype,
                        size_t npositions, double f, const double* los, int field, int wrap,
                        int shifted, void* out, int out_dtype, int* nonfinite, void* stream) {
  return do_read_shifts<double>(ws, phi, nmesh, boxsize, boxcenter, positions, pos_dtype, npositions,
                                f, los, field, wrap, shifted, out, out_dtype, nonfinite, stream);
}

}  // extern "C"
