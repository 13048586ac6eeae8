// Shared definitions of the pyrecon_amd HIP library (gfx950).
//
// Conventions
//   * Meshes are C-order (n0, n1, n2): x slowest, z fastest, as in the
//     reference (src/mesh.c:30-31, src/multigrid.c:48-50).
//   * Every kernel that must agree bit-for-bit with the reference's C is
//     built with -ffp-contract=off and evaluates each expression in the same
//     order and precision as the reference; per-level constants that the
//     reference computes with double intermediates are computed here on the
//     host the same way (LevelGeom::make).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../include/pyrecon_amd.h"

namespace prh {

// Thread-local last-error text, exposed through prh_last_error().
void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
const char* get_error();

#define PRH_HIP(call)                                                              \
  do {                                                                             \
    hipError_t _e = (call);                                                        \
    if (_e != hipSuccess) {                                                        \
      ::prh::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(_e),     \
                       __FILE__, __LINE__);                                        \
      return PRH_ERR_HIP;                                                          \
    }                                                                              \
  } while (0)

#define PRH_TRY(call)             \
  do {                            \
    int _s = (call);              \
    if (_s != PRH_OK) return _s;  \
  } while (0)

#define PRH_LAUNCHED()                                                                     \
  do {                                                                                     \
    hipError_t _e = hipGetLastError();                                                     \
    if (_e != hipSuccess) {                                                                \
      ::prh::set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(_e), __FILE__, \
                       __LINE__);                                                          \
      return PRH_ERR_HIP;                                                                  \
    }                                                                                      \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned ceil_div(size_t a, size_t b) { return static_cast<unsigned>((a + b - 1) / b); }

// Geometry of one multigrid level, with the constants of
// src/multigrid.c:52-59 evaluated exactly as the reference does:
//   cell   = boxsize / n                     (R arithmetic)
//   cell2  = cell * cell                     (R)
//   icell2 = 1. / cell2                      (double, stored as R)
//   offset = (boxcenter - boxsize / 2.) / cell  (double, stored as R)
//   losn   = los / cell                      (R)
template <typename R>
struct LevelGeom {
  int n[3];
  R h2[3], ih2[3], org[3], losc[3];
  R beta;
  int local;

  static LevelGeom make(const size_t* nm, const R* boxsize, const R* boxcenter, R beta_,
                        const R* los) {
    LevelGeom g;
    for (int d = 0; d < 3; d++) {
      g.n[d] = static_cast<int>(nm[d]);
      R cell = boxsize[d] / static_cast<R>(nm[d]);
      g.h2[d] = cell * cell;
      g.ih2[d] = static_cast<R>(1. / static_cast<double>(g.h2[d]));
      g.org[d] = static_cast<R>((static_cast<double>(boxcenter[d]) -
                                 static_cast<double>(boxsize[d]) / 2.) /
                                static_cast<double>(cell));
      g.losc[d] = los ? los[d] / cell : R(0);
    }
    g.beta = beta_;
    g.local = (los == nullptr);
    return g;
  }
};

inline bool coarsenable(const size_t* n) {
  for (int d = 0; d < 3; d++)
    if (!(n[d] > 4 && n[d] % 2 == 0)) return false;
  return true;
}

// Per-mesh-type tag for the C-ABI dispatch.
template <typename R>
struct Tag;
template <>
struct Tag<float> {
  static constexpr int code = PRH_F32;
};
template <>
struct Tag<double> {
  static constexpr int code = PRH_F64;
};

}  // namespace prh
