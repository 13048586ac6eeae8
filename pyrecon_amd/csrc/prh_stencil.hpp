// Device-side evaluation of the local/global line-of-sight operator
//   (A v)_0 = D v_0 - sum_i c_i (v_{+i} + v_{-i})
//             - (g/2) sum_{i<j} p_i p_j (v_{+i+j} + v_{-i-j} - v_{-i+j} - v_{+i-j})
//             - [local] g sum_i p_i (v_{+i} - v_{-i})
// with g = beta / sum_i h_i^2 p_i^2, c_i = 1/h_i^2 + g p_i^2, D = 2 sum_i c_i
// (src/multigrid.c:63-96, 134-157).  Each expression is rounded in the same
// order as the reference so that results match it bit for bit; the library is
// compiled with -ffp-contract=off, so no multiply-add is fused.
#pragma once

#include "prh_common.hpp"

namespace prh {

// The 19 values of a cell neighbourhood: centre, 6 faces, 12 edges.
template <typename R>
struct Hood {
  R c;
  R xp, xm, yp, ym, zp, zm;
  R xpyp, xmym, xmyp, xpym;
  R xpzp, xmzm, xmzp, xpzm;
  R ypzp, ymzm, ymzp, ypzm;
};

template <typename R>
struct Coef {
  R px, py, pz, g, cx, cy, cz;
};

// cell_coef() with g given (the same expressions after g).
template <typename R>
__device__ __forceinline__ Coef<R> cell_coef_g(const LevelGeom<R>& G, int ix, int iy, int iz, R g) {
  Coef<R> k;
  if (G.local) {
    k.px = static_cast<R>(ix) + G.org[0];
    k.py = static_cast<R>(iy) + G.org[1];
    k.pz = static_cast<R>(iz) + G.org[2];
  } else {
    k.px = G.losc[0];
    k.py = G.losc[1];
    k.pz = G.losc[2];
  }
  k.g = g;
  k.cx = G.ih2[0] + k.g * k.px * k.px;
  k.cy = G.ih2[1] + k.g * k.py * k.py;
  k.cz = G.ih2[2] + k.g * k.pz * k.pz;
  return k;
}

template <typename R>
__device__ __forceinline__ Coef<R> cell_coef(const LevelGeom<R>& G, int ix, int iy, int iz) {
  Coef<R> k;
  if (G.local) {
    k.px = static_cast<R>(ix) + G.org[0];
    k.py = static_cast<R>(iy) + G.org[1];
    k.pz = static_cast<R>(iz) + G.org[2];
  } else {
    k.px = G.losc[0];
    k.py = G.losc[1];
    k.pz = G.losc[2];
  }
  k.g = G.beta / (G.h2[0] * k.px * k.px + G.h2[1] * k.py * k.py + G.h2[2] * k.pz * k.pz);
  k.cx = G.ih2[0] + k.g * k.px * k.px;
  k.cy = G.ih2[1] + k.g * k.py * k.py;
  k.cz = G.ih2[2] + k.g * k.pz * k.pz;
  return k;
}

// Off-diagonal groups t0..t3, rounded as in the reference.
template <typename R>
__device__ __forceinline__ void off_terms(const Coef<R>& k, const Hood<R>& h, R& t0, R& t1, R& t2,
                                          R& t3) {
  t0 = k.cx * (h.xp + h.xm);
  t1 = k.cy * (h.yp + h.ym);
  t2 = k.cz * (h.zp + h.zm);
  t3 = k.g / R(2) *
       (k.px * k.py * (h.xpyp + h.xmym - h.xmyp - h.xpym) +
        k.px * k.pz * (h.xpzp + h.xmzm - h.xmzp - h.xpzm) +
        k.py * k.pz * (h.ypzp + h.ymzm - h.ymzp - h.ypzm));
}

template <typename R>
__device__ __forceinline__ R drift_term(const Coef<R>& k, const Hood<R>& h) {
  return k.g * (k.px * (h.xp - h.xm) + k.py * (h.yp - h.ym) + k.pz * (h.zp - h.zm));
}

// One damped-Jacobi update of the cell: (1-w) v + w (f + N v) / D.
template <typename R>
__device__ __forceinline__ R jacobi_update(const Coef<R>& k, const Hood<R>& h, R f, R omega,
                                           R keep, int local) {
  R t0, t1, t2, t3;
  off_terms(k, h, t0, t1, t2, t3);
  R jac = f + t0 + t1 + t2 + t3;
  if (local) jac += drift_term(k, h);
  jac /= R(2) * (k.cx + k.cy + k.cz);
  return keep * h.c + omega * jac;
}

// Residual f - A v of the cell.
template <typename R>
__device__ __forceinline__ R residual_value(const Coef<R>& k, const Hood<R>& h, R f, int local) {
  R t0, t1, t2, t3;
  off_terms(k, h, t0, t1, t2, t3);
  R av = R(2) * (k.cx + k.cy + k.cz) * h.c - (t0 + t1 + t2 + t3);
  if (local) av -= drift_term(k, h);
  return f - av;
}

// ---- split-coefficient forms used by the 2.5D-blocked kernels -------------
// The per-cell coefficients factor into parts that depend on x only (px,
// A = (h0^2 px) px), on y only (py, B), on z only (pz, C) and on (y, z)
// (py pz).  g = beta / ((A + B) + C) and c_i = 1/h_i^2 + (g p_i) p_i then
// round exactly like cell_coef().  For a global line of sight every
// coefficient is a constant of the level, computed once on the host with the
// same expressions (GlobalCoef::make).
template <typename R>
struct GlobalCoef {
  R g, g2, cx, cy, cz, d, pxpy, pxpz, pypz;
  static GlobalCoef make(const LevelGeom<R>& G) {
    GlobalCoef k;
    const R px = G.losc[0], py = G.losc[1], pz = G.losc[2];
    k.g = G.beta / (G.h2[0] * px * px + G.h2[1] * py * py + G.h2[2] * pz * pz);
    k.cx = G.ih2[0] + k.g * px * px;
    k.cy = G.ih2[1] + k.g * py * py;
    k.cz = G.ih2[2] + k.g * pz * pz;
    k.g2 = k.g / R(2);
    k.d = R(2) * (k.cx + k.cy + k.cz);
    k.pxpy = px * py;
    k.pxpz = px * pz;
    k.pypz = py * pz;
    return k;
  }
};

template <typename R>
__device__ __forceinline__ void edge_sums(const Hood<R>& h, R& e1, R& e2, R& e3) {
  e1 = h.xpyp + h.xmym - h.xmyp - h.xpym;
  e2 = h.xpzp + h.xmzm - h.xmzp - h.xpzm;
  e3 = h.ypzp + h.ymzm - h.ymzp - h.ypzm;
}

// Exact forms (same rounding as jacobi_update / residual_value).
template <typename R, bool LOCAL>
__device__ __forceinline__ R jacobi_split(const Hood<R>& h, R f, R px, R py, R pz, R pxpy, R pxpz,
                                          R pypz, R g, R g2, R cx, R cy, R cz, R d, R omega,
                                          R keep) {
  R e1, e2, e3;
  edge_sums(h, e1, e2, e3);
  const R t0 = cx * (h.xp + h.xm);
  const R t1 = cy * (h.yp + h.ym);
  const R t2 = cz * (h.zp + h.zm);
  const R t3 = g2 * (pxpy * e1 + pxpz * e2 + pypz * e3);
  R jac = f + t0 + t1 + t2 + t3;
  if (LOCAL) jac += g * (px * (h.xp - h.xm) + py * (h.yp - h.ym) + pz * (h.zp - h.zm));
  jac /= d;
  return keep * h.c + omega * jac;
}

template <typename R, bool LOCAL>
__device__ __forceinline__ R residual_split(const Hood<R>& h, R f, R px, R py, R pz, R pxpy,
                                            R pxpz, R pypz, R g, R g2, R cx, R cy, R cz, R d) {
  R e1, e2, e3;
  edge_sums(h, e1, e2, e3);
  const R t0 = cx * (h.xp + h.xm);
  const R t1 = cy * (h.yp + h.ym);
  const R t2 = cz * (h.zp + h.zm);
  const R t3 = g2 * (pxpy * e1 + pxpz * e2 + pypz * e3);
  R av = d * h.c - (t0 + t1 + t2 + t3);
  if (LOCAL) av -= g * (px * (h.xp - h.xm) + py * (h.yp - h.ym) + pz * (h.zp - h.zm));
  return f - av;
}

// Fast FP32 forms for the correction solves of the mixed-precision path:
// fused multiply-adds and the hardware reciprocal (1 ulp) instead of
// correctly rounded division.  Not bit-compatible with the reference; the
// FP64 defect-correction loop around them absorbs their rounding.
__device__ __forceinline__ float rcp_fast(float x) { return __builtin_amdgcn_rcpf(x); }

template <bool LOCAL>
__device__ __forceinline__ float jacobi_fast(const Hood<float>& h, float f, float px, float py,
                                             float pz, float pxpy, float pxpz, float pypz, float g,
                                             float g2, float cx, float cy, float cz, float dinv,
                                             float omega, float keep) {
  float e1, e2, e3;
  edge_sums(h, e1, e2, e3);
  float s = fmaf(cx, h.xp + h.xm, f);
  s = fmaf(cy, h.yp + h.ym, s);
  s = fmaf(cz, h.zp + h.zm, s);
  s = fmaf(g2, fmaf(pxpy, e1, fmaf(pxpz, e2, pypz * e3)), s);
  if (LOCAL) s = fmaf(g, fmaf(px, h.xp - h.xm, fmaf(py, h.yp - h.ym, pz * (h.zp - h.zm))), s);
  return fmaf(keep, h.c, omega * (s * dinv));
}

template <bool LOCAL>
__device__ __forceinline__ float residual_fast(const Hood<float>& h, float f, float px, float py,
                                               float pz, float pxpy, float pxpz, float pypz,
                                               float g, float g2, float cx, float cy, float cz,
                                               float d) {
  float e1, e2, e3;
  edge_sums(h, e1, e2, e3);
  float s = fmaf(cx, h.xp + h.xm, f);
  s = fmaf(cy, h.yp + h.ym, s);
  s = fmaf(cz, h.zp + h.zm, s);
  s = fmaf(g2, fmaf(pxpy, e1, fmaf(pxpz, e2, pypz * e3)), s);
  if (LOCAL) s = fmaf(g, fmaf(px, h.xp - h.xm, fmaf(py, h.yp - h.ym, pz * (h.zp - h.zm))), s);
  return fmaf(-d, h.c, s);
}

// Gather the neighbourhood of (ix, iy, iz) from a periodic C-order mesh.
// I: index type (size_t; int for meshes below 2^31 cells, e.g. levels in LDS).
template <typename R, typename I = size_t>
__device__ __forceinline__ Hood<R> gather_hood(const R* __restrict__ v, int n0, int n1, int n2,
                                               int ix, int iy, int iz) {
  const I s0 = static_cast<I>(n1) * n2, s1 = static_cast<I>(n2);
  const I x0 = ix * s0, xp = (ix + 1 == n0 ? 0 : ix + 1) * s0,
          xm = (ix == 0 ? n0 - 1 : ix - 1) * s0;
  const I y0 = iy * s1, yp = (iy + 1 == n1 ? 0 : iy + 1) * s1,
          ym = (iy == 0 ? n1 - 1 : iy - 1) * s1;
  const I z0 = iz, zp = (iz + 1 == n2 ? 0 : iz + 1), zm = (iz == 0 ? n2 - 1 : iz - 1);
  Hood<R> h;
  h.c = v[x0 + y0 + z0];
  h.xp = v[xp + y0 + z0];
  h.xm = v[xm + y0 + z0];
  h.yp = v[x0 + yp + z0];
  h.ym = v[x0 + ym + z0];
  h.zp = v[x0 + y0 + zp];
  h.zm = v[x0 + y0 + zm];
  h.xpyp = v[xp + yp + z0];
  h.xmym = v[xm + ym + z0];
  h.xmyp = v[xm + yp + z0];
  h.xpym = v[xp + ym + z0];
  h.xpzp = v[xp + y0 + zp];
  h.xmzm = v[xm + y0 + zm];
  h.xmzp = v[xm + y0 + zp];
  h.xpzm = v[xp + y0 + zm];
  h.ypzp = v[x0 + yp + zp];
  h.ymzm = v[x0 + ym + zm];
  h.ymzp = v[x0 + ym + zp];
  h.ypzm = v[x0 + yp + zm];
  return h;
}

template <typename R>
__device__ __forceinline__ Hood<R> zero_hood() {
  Hood<R> h;
  h.c = h.xp = h.xm = h.yp = h.ym = h.zp = h.zm = R(0);
  h.xpyp = h.xmym = h.xmyp = h.xpym = R(0);
  h.xpzp = h.xmzm = h.xmzp = h.xpzm = R(0);
  h.ypzp = h.ymzm = h.ymzp = h.ypzm = R(0);
  return h;
}

}  // namespace prh
