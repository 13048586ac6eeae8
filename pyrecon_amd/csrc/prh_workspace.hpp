// Internal layout of prh_workspace (opaque in the C ABI).
#pragma once

#include "prh_common.hpp"

namespace prh {

constexpr int kMaxLevels = 24;

// One multigrid hierarchy in precision R.  Level 0 borrows the caller's
// iterate and source; it owns only a second iterate (tmp).  Levels >= 1 own
// an iterate v, a source f and a second iterate tmp (Jacobi ping-pong and
// residual storage).  When `full_top` is set, level 0 also owns v and f
// (the FP32 hierarchy of the mixed-precision solve).
struct Hierarchy {
  int nlev = 0;
  size_t n[kMaxLevels][3] = {};
  void* v[kMaxLevels] = {};
  void* f[kMaxLevels] = {};
  void* tmp[kMaxLevels] = {};
};

// CIC paint scratch for one chunk of at most `chunk` points.
// Tiles of the mixed path's paint and read: 2^kPaintTileBits cells per axis,
// with the upper halo (kPaintTile + 1)^3 accumulation cells per tile.
constexpr int kPaintTileBits = 4;
constexpr int kPaintTile = 1 << kPaintTileBits;
constexpr size_t kPaintTileCells = size_t(kPaintTile + 1) * (kPaintTile + 1) * (kPaintTile + 1);
inline size_t paint_tiles(const size_t* n) {
  return ((n[0] + kPaintTile - 1) / kPaintTile) * ((n[1] + kPaintTile - 1) / kPaintTile) *
         ((n[2] + kPaintTile - 1) / kPaintTile);
}

struct PaintScratch {
  size_t chunk = 0;
  size_t ncell = 0;
  uint32_t* count = nullptr;    // ntile      per-tile point counts of the read binning (kept zero)
  uint32_t* start = nullptr;    // ncell + 1  first sorted position of each cell (read binning: tile)
  uint32_t* partial = nullptr;  // scan block partials
  size_t npartial = 0;
  uint32_t* key = nullptr;      // chunk      base cell of each point; radix sort keys
  uint32_t* rank = nullptr;     // chunk      radix sort keys (other buffer); read binning ranks
  uint32_t* vtmp = nullptr;     // chunk      radix sort values between passes
  uint32_t* hist = nullptr;     // nhist      per-tile digit counts of one radix pass
  uint32_t* goff = nullptr;     // nhist + 1  their exclusive scan
  size_t nhist = 0;
  uint32_t* bnd = nullptr;      // ncell / 8192 + 2  first sorted position of each 8192-cell block
  uint32_t* sorted = nullptr;   // chunk      point indices grouped by cell, ascending within a cell
  void* payload = nullptr;      // chunk      (t_x, t_y, t_z, w) in the mesh precision, point order
  void* spay = nullptr;         // chunk      the same payload in cell-grouped order
  uint32_t* misc = nullptr;     // [0] error flags
  // After a paint of one chunk, `sorted` is a permutation of that catalog's
  // indices in cell order.  A later read of the same catalog (same positions
  // pointer, count and dtype) visits its points in that order instead of
  // binning them again.  Any permutation gives the same results (each point is
  // evaluated on its own); the order only sets the cache locality.  The two
  // arrays alternate between paints, so the orders of the last two catalogs
  // painted (data and randoms) are both kept; `sorted` is the one of the last.
  uint32_t* sorted_prev = nullptr;  // chunk      the previous paint's order
  // The FP32 tile paint lays the positions out in its (tile) order anyway
  // (`spos`, 3 values of the positions dtype per point); painting one chunk, it
  // also keeps a 64-bit hash of the positions' bytes in input order.  A read of
  // positions with the same pointer, count, dtype and hash streams that copy
  // instead of gathering the positions at random.
  void* spos = nullptr;         // chunk x 3 doubles
  void* spos_prev = nullptr;    // chunk x 3 doubles
  uint64_t* hpart = nullptr;    // nhpart   per-block partial hashes
  size_t nhpart = 0;
  uint64_t* hash = nullptr;     // [2]      hash of the last paint's positions; of a read's
  double* tbnd = nullptr;       // ntile x 729  boundary cells of the FP32 tile paint (PRH_WS_MIXED)
  // first sorted position of each kPaintTile^3 tile (and the end) of the last two tile
  // paints, for reads that walk the painted positions tile by tile
  uint32_t* tstart = nullptr;       // ntile + 1 (PRH_WS_MIXED)
  uint32_t* tstart_prev = nullptr;  // ntile + 1
  struct Order {
    const void* pos = nullptr;
    size_t n = 0;
    int pdt = 0;
    bool copy = false;   // spos holds the positions in this order
    bool tiles = false;  // ... in tile order, with the tile starts in tstart
    uint64_t hash = 0;
  };
  Order order, order_prev;  // tags of `sorted` and `sorted_prev`
  // The read binning of the last read of positions that no kept paint order covers lives in
  // `payload` until a paint (or a growth of the scratch) overwrites it: a read of the same
  // positions (pointer, count, dtype) in one pass reuses it.  Any permutation of [0, n) gives the
  // same results; the bins only set the locality.
  struct Bins {
    const void* pos = nullptr;
    size_t n = 0;
    int pdt = 0;
  };
  Bins bins;
};

// Tables of the separable Fourier-space smoothing: per axis, the twiddles
// exp(-2 pi i k / n) and the per-call filter gains, in the mesh precision.
struct SmoothState {
  bool have = false;  // false when a dimension is longer than the direct transform allows
  bool direct[3] = {};  // axis transformed by the direct DFT (a prime factor > 7, or too long for the radix passes)
  void* tw[3] = {};
  void* gain[3] = {};
  double* tw64[3] = {};  // exp(-2 pi i k / n), k < n, in double (direct DFT axes)
};

}  // namespace prh

struct prh_workspace {
  size_t n[3] = {0, 0, 0};
  int dtype = 0;
  unsigned flags = 0;
  size_t bytes = 0;
  int device = 0;
  prh::Hierarchy mg;     // workspace precision
  prh::Hierarchy mixed;  // FP32 hierarchy incl. level 0 (PRH_WS_MIXED)
  prh::PaintScratch paint;
  prh::SmoothState smooth;
  double* red = nullptr;       // reduction partials (device), then the results
  uint32_t* misc = nullptr;    // device status flags of the read kernels
  size_t nred = 0;
  double* red_host = nullptr;  // pinned host landing for reductions
  unsigned* flag_host = nullptr;
};

namespace prh {
int ws_alloc(prh_workspace* ws, void** p, size_t bytes);
int smooth_setup(prh_workspace* ws);
size_t smooth_bytes(const size_t* n, int dtype);
constexpr unsigned kReduceBlocks = 1024;

// Opt-in event timing of one kernel class (PRH_TIMER_*): when enabled, the
// launch sites bracket each launch with a pair of HIP events on the launch
// stream; prh_timer_read sums the elapsed times.
bool timer_on(int which);
void timer_before(int which, hipStream_t s);
void timer_after(int which, hipStream_t s, double bytes);
}  // namespace prh
