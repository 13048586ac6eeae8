/*
 * CPU oracle for the MultiGridReconstruction hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C99 + OpenMP, the arithmetic of the reference's
 * native kernels (pyrecon 1.0.0, src/mesh.c and src/multigrid.c).  It exists to
 * check the HIP library; nothing in the shipped package links or calls it.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * The code layout is this project's own (a per-level geometry record, one
 * operator-evaluation routine shared by the smoother and the residual, a
 * preallocated level hierarchy instead of per-call malloc), but the ORDER OF
 * FLOATING-POINT OPERATIONS inside each expression follows the reference
 * exactly, so that results are bit-identical to it (compile with
 * -ffp-contract=off; the reference's own build has no FMA on x86-64).
 *
 * Compiled twice: with -DORC_SINGLE (float, symbols *_f32) and without
 * (double, symbols *_f64).
 *
 * Parity pinning: every function here is compared bit-for-bit with the
 * reference's own C sources built by oracle/Makefile into oracle/_ref/, and
 * with closed forms (tests/test_oracle_kernels.py).
 */
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#ifdef ORC_SINGLE
typedef float real_t;
#define ORC(name) orc_##name##_f32
#else
typedef double real_t;
#define ORC(name) orc_##name##_f64
#endif

/* ------------------------------------------------------------------------- */
/* Mesh-unit CIC helpers                                                      */
/* ------------------------------------------------------------------------- */

/* Scatter-add of weighted points, in point order (the reference loop is serial,
 * src/mesh.c:32, so each cell receives its contributions in increasing point
 * index; this order fixes the rounding and is what the GPU paint reproduces).
 * Base index: ((int)p) % n; fractional part: p - base   (src/mesh.c:35-47).
 * Corner weight: ((wx * wy) * wz) * w                     (src/mesh.c:53-60). */
int ORC(assign_cic)(real_t *mesh, const size_t *nmesh, const real_t *pos,
                    const real_t *w, size_t npos) {
  const size_t n0 = nmesh[0], n1 = nmesh[1], n2 = nmesh[2];
  const size_t s1 = n2, s0 = n1 * n2;
  for (size_t i = 0; i < npos; i++) {
    const real_t *p = pos + 3 * i;
    size_t b[3];
    b[0] = ((int)p[0]) % n0;
    b[1] = ((int)p[1]) % n1;
    b[2] = ((int)p[2]) % n2;
    real_t fr[3][2];
    for (int d = 0; d < 3; d++) {
      real_t t = p[d] - b[d];
      fr[d][0] = 1 - t;
      fr[d][1] = t;
    }
    const size_t x[2] = {b[0] * s0, ((b[0] + 1) % n0) * s0};
    const size_t y[2] = {b[1] * s1, ((b[1] + 1) % n1) * s1};
    const size_t z[2] = {b[2], (b[2] + 1) % n2};
    const real_t wi = w[i];
    for (int a = 0; a < 2; a++)
      for (int c = 0; c < 2; c++)
        for (int e = 0; e < 2; e++)
          mesh[x[a] + y[c] + z[e]] += fr[0][a] * fr[1][c] * fr[2][e] * wi;
  }
  return 0;
}

/* Interpolation at points (src/mesh.c:269-310): fraction p - (int)p, base
 * ((int)p) % n, terms added in the order (x,y,z) = 000,001,010,011,100,...   */
int ORC(read_cic)(const real_t *mesh, const size_t *nmesh, const real_t *pos,
                  real_t *out, size_t npos) {
  const size_t n0 = nmesh[0], n1 = nmesh[1], n2 = nmesh[2];
  const size_t s1 = n2, s0 = n1 * n2;
  long long np = (long long)npos;
#pragma omp parallel for schedule(static)
  for (long long ii = 0; ii < np; ii++) {
    const real_t *p = pos + 3 * ii;
    real_t t[3];
    size_t b[3];
    for (int d = 0; d < 3; d++) t[d] = p[d] - (int)p[d];
    b[0] = ((int)p[0]) % n0;
    b[1] = ((int)p[1]) % n1;
    b[2] = ((int)p[2]) % n2;
    const size_t x[2] = {b[0] * s0, ((b[0] + 1) % n0) * s0};
    const size_t y[2] = {b[1] * s1, ((b[1] + 1) % n1) * s1};
    const size_t z[2] = {b[2], (b[2] + 1) % n2};
    real_t acc = 0;
    int first = 1;
    for (int a = 0; a < 2; a++)
      for (int c = 0; c < 2; c++)
        for (int e = 0; e < 2; e++) {
          real_t wx = a ? t[0] : 1 - t[0];
          real_t wy = c ? t[1] : 1 - t[1];
          real_t wz = e ? t[2] : 1 - t[2];
          real_t term = mesh[x[a] + y[c] + z[e]] * wx * wy * wz;
          if (first) { acc = term; first = 0; }
          else acc += term;
        }
    out[ii] = acc;
  }
  return 0;
}

/* Central-difference gradient read with CIC weights (src/mesh.c:187-266).
 * For each of the 8 corners (order 000,100,010,001,110,101,011,111 in (x,y,z))
 * the three differences (mesh[c+e_d] - mesh[c-e_d]) are multiplied by the
 * corner weight ((wx*wy)*wz) and accumulated per axis; the result is divided
 * by cell_d = 2*boxsize_d/n_d, computed in double and stored as real_t.       */
int ORC(read_finite_difference_cic)(const real_t *mesh, const size_t *nmesh,
                                    const real_t *boxsize, const real_t *pos,
                                    real_t *out, size_t npos) {
  const size_t n0 = nmesh[0], n1 = nmesh[1], n2 = nmesh[2];
  const size_t s1 = n2, s0 = n1 * n2;
  real_t cell[3];
  for (int d = 0; d < 3; d++) cell[d] = 2.0 * boxsize[d] / nmesh[d];
  /* corner visiting order of the reference */
  static const int corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                   {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
  long long np = (long long)npos;
#pragma omp parallel for schedule(static)
  for (long long ii = 0; ii < np; ii++) {
    const real_t *p = pos + 3 * ii;
    real_t t[3];
    for (int d = 0; d < 3; d++) t[d] = p[d] - (int)p[d];
    size_t b0 = ((int)p[0]) % n0, b1 = ((int)p[1]) % n1, b2 = ((int)p[2]) % n2;
    /* offsets -1, 0, +1, +2 along each axis */
    size_t X[4], Y[4], Z[4];
    for (int k = 0; k < 4; k++) {
      X[k] = ((b0 + n0 - 1 + k) % n0) * s0;
      Y[k] = ((b1 + n1 - 1 + k) % n1) * s1;
      Z[k] = (b2 + n2 - 1 + k) % n2;
    }
    real_t g[3] = {0, 0, 0};
    for (int c = 0; c < 8; c++) {
      const int a = corner[c][0], bb = corner[c][1], e = corner[c][2];
      real_t wx = a ? t[0] : 1 - t[0];
      real_t wy = bb ? t[1] : 1 - t[1];
      real_t wz = e ? t[2] : 1 - t[2];
      real_t wt = wx * wy * wz;
      /* corner cell sits at offset index 1+a etc. in X/Y/Z */
      const size_t cx = X[1 + a], cy = Y[1 + bb], cz = Z[1 + e];
      real_t dx = (mesh[X[2 + a] + cy + cz] - mesh[X[a] + cy + cz]) * wt;
      real_t dy = (mesh[cx + Y[2 + bb] + cz] - mesh[cx + Y[bb] + cz]) * wt;
      real_t dz = (mesh[cx + cy + Z[2 + e]] - mesh[cx + cy + Z[e]]) * wt;
      if (c == 0) { g[0] = dx; g[1] = dy; g[2] = dz; }
      else { g[0] += dx; g[1] += dy; g[2] += dz; }
    }
    for (int d = 0; d < 3; d++) out[3 * ii + d] = g[d] / cell[d];
  }
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Multigrid                                                                  */
/* ------------------------------------------------------------------------- */

/* Geometry of one level (src/multigrid.c:52-59, 115-122). */
typedef struct {
  size_t n[3];
  size_t s0, s1, size;
  real_t h2[3], ih2[3], org[3], losc[3];
  real_t beta;
  int local;
} level_t;

static void level_setup(level_t *L, const size_t *n, const real_t *boxsize,
                        const real_t *boxcenter, real_t beta, const real_t *los) {
  for (int d = 0; d < 3; d++) {
    L->n[d] = n[d];
    real_t h = boxsize[d] / n[d];
    L->h2[d] = h * h;
    L->ih2[d] = 1. / L->h2[d];
    L->org[d] = (boxcenter[d] - boxsize[d] / 2.) / h;
    L->losc[d] = los ? los[d] / h : 0;
  }
  L->s1 = n[2];
  L->s0 = n[1] * n[2];
  L->size = n[0] * n[1] * n[2];
  L->beta = beta;
  L->local = (los == NULL);
}

/* Neighbourhood of one cell for the 19-point operator: faces, the 12 edges,
 * the centre.  Index d in {0,1,2} maps to offset {-1,0,+1}.                  */
typedef struct {
  real_t c;                        /* centre */
  real_t xp, xm, yp, ym, zp, zm;   /* faces */
  real_t xpyp, xmym, xmyp, xpym;   /* x-y edges */
  real_t xpzp, xmzm, xmzp, xpzm;   /* x-z edges */
  real_t ypzp, ymzm, ymzp, ypzm;   /* y-z edges */
} hood_t;

static inline void gather(const level_t *L, const real_t *v, size_t ix, size_t iy,
                          size_t iz, hood_t *h) {
  const size_t x0 = ix * L->s0, xp = ((ix + 1) % L->n[0]) * L->s0,
               xm = ((ix + L->n[0] - 1) % L->n[0]) * L->s0;
  const size_t y0 = iy * L->s1, yp = ((iy + 1) % L->n[1]) * L->s1,
               ym = ((iy + L->n[1] - 1) % L->n[1]) * L->s1;
  const size_t z0 = iz, zp = (iz + 1) % L->n[2], zm = (iz + L->n[2] - 1) % L->n[2];
  h->c = v[x0 + y0 + z0];
  h->xp = v[xp + y0 + z0]; h->xm = v[xm + y0 + z0];
  h->yp = v[x0 + yp + z0]; h->ym = v[x0 + ym + z0];
  h->zp = v[x0 + y0 + zp]; h->zm = v[x0 + y0 + zm];
  h->xpyp = v[xp + yp + z0]; h->xmym = v[xm + ym + z0];
  h->xmyp = v[xm + yp + z0]; h->xpym = v[xp + ym + z0];
  h->xpzp = v[xp + y0 + zp]; h->xmzm = v[xm + y0 + zm];
  h->xmzp = v[xm + y0 + zp]; h->xpzm = v[xp + y0 + zm];
  h->ypzp = v[x0 + yp + zp]; h->ymzm = v[x0 + ym + zm];
  h->ymzp = v[x0 + ym + zp]; h->ypzm = v[x0 + yp + zm];
}

/* Operator coefficients at a cell (src/multigrid.c:63-77). */
typedef struct { real_t px, py, pz, g, cx, cy, cz; } coef_t;

static inline void coefs(const level_t *L, size_t ix, size_t iy, size_t iz, coef_t *k) {
  if (L->local) {
    k->px = ix + L->org[0];
    k->py = iy + L->org[1];
    k->pz = iz + L->org[2];
  } else {
    k->px = L->losc[0];
    k->py = L->losc[1];
    k->pz = L->losc[2];
  }
  k->g = L->beta / (L->h2[0] * k->px * k->px + L->h2[1] * k->py * k->py +
                    L->h2[2] * k->pz * k->pz);
  k->cx = L->ih2[0] + k->g * k->px * k->px;
  k->cy = L->ih2[1] + k->g * k->py * k->py;
  k->cz = L->ih2[2] + k->g * k->pz * k->pz;
}

/* The four groups of off-diagonal terms, each rounded as in the reference. */
static inline void terms(const coef_t *k, const hood_t *h, real_t t[4]) {
  t[0] = k->cx * (h->xp + h->xm);
  t[1] = k->cy * (h->yp + h->ym);
  t[2] = k->cz * (h->zp + h->zm);
  t[3] = k->g / 2 * (k->px * k->py * (h->xpyp + h->xmym - h->xmyp - h->xpym) +
                     k->px * k->pz * (h->xpzp + h->xmzm - h->xmzp - h->xpzm) +
                     k->py * k->pz * (h->ypzp + h->ymzm - h->ymzp - h->ypzm));
}

/* First-derivative term of the local line of sight (src/multigrid.c:91-95). */
static inline real_t drift(const coef_t *k, const hood_t *h) {
  return k->g * (k->px * (h->xp - h->xm) + k->py * (h->yp - h->ym) +
                 k->pz * (h->zp - h->zm));
}

/* damped Jacobi: niter sweeps, each computes every cell from the previous
 * iterate (src/multigrid.c:44-104).  tmp holds the new iterate.            */
static void smooth(const level_t *L, real_t *v, const real_t *f, real_t *tmp,
                   real_t omega, int niter) {
  const real_t keep = 1 - omega;
  for (int it = 0; it < niter; it++) {
    long long nx = (long long)L->n[0];
#pragma omp parallel for schedule(static)
    for (long long ix = 0; ix < nx; ix++) {
      for (size_t iy = 0; iy < L->n[1]; iy++) {
        for (size_t iz = 0; iz < L->n[2]; iz++) {
          hood_t h;
          coef_t k;
          real_t t[4];
          gather(L, v, (size_t)ix, iy, iz, &h);
          coefs(L, (size_t)ix, iy, iz, &k);
          terms(&k, &h, t);
          const size_t ii = ix * L->s0 + iy * L->s1 + iz;
          real_t jac = f[ii] + t[0] + t[1] + t[2] + t[3];
          if (L->local) jac += drift(&k, &h);
          jac /= 2 * (k.cx + k.cy + k.cz);
          tmp[ii] = keep * h.c + omega * jac;
        }
      }
    }
    memcpy(v, tmp, L->size * sizeof(real_t));
  }
}

/* r = f - A v (src/multigrid.c:107-164). */
static void defect(const level_t *L, const real_t *v, const real_t *f, real_t *r) {
  long long nx = (long long)L->n[0];
#pragma omp parallel for schedule(static)
  for (long long ix = 0; ix < nx; ix++) {
    for (size_t iy = 0; iy < L->n[1]; iy++) {
      for (size_t iz = 0; iz < L->n[2]; iz++) {
        hood_t h;
        coef_t k;
        real_t t[4];
        gather(L, v, (size_t)ix, iy, iz, &h);
        coefs(L, (size_t)ix, iy, iz, &k);
        terms(&k, &h, t);
        const size_t ii = ix * L->s0 + iy * L->s1 + iz;
        real_t av = 2 * (k.cx + k.cy + k.cz) * h.c - (t[0] + t[1] + t[2] + t[3]);
        if (L->local) av -= drift(&k, &h);
        r[ii] = f[ii] - av;
      }
    }
  }
}

/* Full-weighting restriction, weights 8/4/2/1 over 64 (src/multigrid.c:212-266).
 * Summation order: centre, 6 faces, 12 edges, 8 corners, as in the reference. */
static void restrict_fw(const size_t *nf, const real_t *fine, real_t *coarse) {
  const size_t s1 = nf[2], s0 = nf[1] * nf[2];
  const size_t c0 = nf[0] / 2, c1 = nf[1] / 2, c2 = nf[2] / 2;
  long long ncx = (long long)c0;
#pragma omp parallel for schedule(static)
  for (long long jx = 0; jx < ncx; jx++) {
    const size_t fx = 2 * jx;
    const size_t x0 = fx * s0, xp = ((fx + 1) % nf[0]) * s0, xm = ((fx + nf[0] - 1) % nf[0]) * s0;
    for (size_t jy = 0; jy < c1; jy++) {
      const size_t fy = 2 * jy;
      const size_t y0 = fy * s1, yp = ((fy + 1) % nf[1]) * s1, ym = ((fy + nf[1] - 1) % nf[1]) * s1;
      for (size_t jz = 0; jz < c2; jz++) {
        const size_t z0 = 2 * jz, zp = (z0 + 1) % nf[2], zm = (z0 + nf[2] - 1) % nf[2];
        const real_t *u = fine;
        real_t faces = u[xp + y0 + z0] + u[xm + y0 + z0] + u[x0 + yp + z0] +
                       u[x0 + ym + z0] + u[x0 + y0 + zp] + u[x0 + y0 + zm];
        real_t edges = u[xp + yp + z0] + u[xm + yp + z0] + u[xp + ym + z0] +
                       u[xm + ym + z0] + u[xp + y0 + zp] + u[xm + y0 + zp] +
                       u[xp + y0 + zm] + u[xm + y0 + zm] + u[x0 + yp + zp] +
                       u[x0 + ym + zp] + u[x0 + yp + zm] + u[x0 + ym + zm];
        real_t s = 8 * u[x0 + y0 + z0] + 4 * faces + 2 * edges;
        s += u[xp + yp + zp]; s += u[xm + yp + zp]; s += u[xp + ym + zp]; s += u[xm + ym + zp];
        s += u[xp + yp + zm]; s += u[xm + yp + zm]; s += u[xp + ym + zm]; s += u[xm + ym + zm];
        coarse[(size_t)jx * c1 * c2 + jy * c2 + jz] = s / 64.0;
      }
    }
  }
}

/* Trilinear prolongation (src/multigrid.c:167-209), either overwriting the
 * fine array (add == 0) or adding to it (add == 1, the V-cycle correction of
 * src/multigrid.c:296, done with the rounded prolonged value).             */
static void prolong_tl(const size_t *nc, const real_t *coarse, real_t *fine, int add) {
  const size_t s1 = nc[2], s0 = nc[1] * nc[2];
  const size_t f1 = 2 * nc[2], f0 = 4 * nc[1] * nc[2];
  long long ncx = (long long)nc[0];
#pragma omp parallel for schedule(static)
  for (long long jx = 0; jx < ncx; jx++) {
    const size_t x0 = jx * s0, xp = ((jx + 1) % nc[0]) * s0;
    const size_t X0 = 2 * jx * f0, X1 = X0 + f0;
    for (size_t jy = 0; jy < nc[1]; jy++) {
      const size_t y0 = jy * s1, yp = ((jy + 1) % nc[1]) * s1;
      const size_t Y0 = 2 * jy * f1, Y1 = Y0 + f1;
      for (size_t jz = 0; jz < nc[2]; jz++) {
        const size_t z0 = jz, zp = (jz + 1) % nc[2];
        const size_t Z0 = 2 * jz, Z1 = Z0 + 1;
        const real_t *u = coarse;
        const real_t a = u[x0 + y0 + z0];
        real_t val[8];
        size_t idx[8];
        idx[0] = X0 + Y0 + Z0; val[0] = a;
        idx[1] = X1 + Y0 + Z0; val[1] = (a + u[xp + y0 + z0]) / 2;
        idx[2] = X0 + Y1 + Z0; val[2] = (a + u[x0 + yp + z0]) / 2;
        idx[3] = X0 + Y0 + Z1; val[3] = (a + u[x0 + y0 + zp]) / 2;
        idx[4] = X1 + Y1 + Z0; val[4] = (a + u[xp + y0 + z0] + u[x0 + yp + z0] + u[xp + yp + z0]) / 4;
        idx[5] = X0 + Y1 + Z1; val[5] = (a + u[x0 + yp + z0] + u[x0 + y0 + zp] + u[x0 + yp + zp]) / 4;
        idx[6] = X1 + Y0 + Z1; val[6] = (a + u[xp + y0 + z0] + u[x0 + y0 + zp] + u[xp + y0 + zp]) / 4;
        idx[7] = X1 + Y1 + Z1;
        val[7] = (a + u[xp + y0 + z0] + u[x0 + yp + z0] + u[x0 + y0 + zp] + u[xp + yp + z0] +
                  u[xp + y0 + zp] + u[x0 + yp + zp] + u[xp + yp + zp]) / 8;
        for (int q = 0; q < 8; q++) {
          if (add) fine[idx[q]] += val[q];
          else fine[idx[q]] = val[q];
        }
      }
    }
  }
}

static int coarsenable(const size_t *n) {
  for (int d = 0; d < 3; d++)
    if (!(n[d] > 4 && n[d] % 2 == 0)) return 0;
  return 1;
}

/* Preallocated hierarchy: level 0 uses caller arrays; levels >= 1 own v, f;
 * every level owns a scratch array (Jacobi iterate / residual). */
typedef struct {
  int nlev;
  level_t L[32];
  real_t *v[32], *f[32], *tmp[32];
} hier_t;

static int hier_build(hier_t *H, const size_t *n, const real_t *boxsize, const real_t *boxcenter,
                      real_t beta, const real_t *los) {
  size_t m[3] = {n[0], n[1], n[2]};
  int l = 0;
  memset(H, 0, sizeof(*H));
  for (;;) {
    level_setup(&H->L[l], m, boxsize, boxcenter, beta, los);
    H->tmp[l] = (real_t *)malloc(H->L[l].size * sizeof(real_t));
    if (!H->tmp[l]) return -1;
    if (l > 0) {
      H->v[l] = (real_t *)calloc(H->L[l].size, sizeof(real_t));
      H->f[l] = (real_t *)malloc(H->L[l].size * sizeof(real_t));
      if (!H->v[l] || !H->f[l]) return -1;
    }
    l++;
    if (!coarsenable(m) || l >= 32) break;
    for (int d = 0; d < 3; d++) m[d] /= 2;
  }
  H->nlev = l;
  return 0;
}

static void hier_free(hier_t *H) {
  for (int l = 0; l < H->nlev; l++) {
    free(H->tmp[l]);
    if (l > 0) { free(H->v[l]); free(H->f[l]); }
  }
}

/* V-cycle at level l on (v, f) (src/multigrid.c:270-300). */
static void vcyc(hier_t *H, int l, real_t *v, const real_t *f, real_t omega, int nsweep) {
  level_t *L = &H->L[l];
  smooth(L, v, f, H->tmp[l], omega, nsweep);
  if (l + 1 < H->nlev) {
    defect(L, v, f, H->tmp[l]);
    restrict_fw(L->n, H->tmp[l], H->f[l + 1]);
    memset(H->v[l + 1], 0, H->L[l + 1].size * sizeof(real_t));
    vcyc(H, l + 1, H->v[l + 1], H->f[l + 1], omega, nsweep);
    prolong_tl(H->L[l + 1].n, H->v[l + 1], v, 1);
  }
  smooth(L, v, f, H->tmp[l], omega, nsweep);
}

/* Full multigrid (src/multigrid.c:303-328): restrict the source to the
 * coarsest level, start from zero there, then at each level prolong the
 * coarser solution (overwriting the iterate) and run ncycle V-cycles.
 * At the finest level the caller's v is used as-is when no coarser level
 * exists.  The source arrays of coarser levels double as V-cycle scratch
 * once their own FMG stage is complete, exactly like the reference's
 * freshly allocated temporaries. */
static void fmg_run(hier_t *H, const real_t *f, real_t *v, real_t omega, int nsweep, int ncycle) {
  const int nl = H->nlev;
  /* FMG sources: keep a private copy per level so V-cycle scratch never aliases them */
  real_t **src = (real_t **)calloc(nl, sizeof(real_t *));
  src[0] = (real_t *)f;
  for (int l = 1; l < nl; l++) {
    src[l] = (real_t *)malloc(H->L[l].size * sizeof(real_t));
    restrict_fw(H->L[l - 1].n, src[l - 1], src[l]);
  }
  real_t **sol = (real_t **)calloc(nl, sizeof(real_t *));
  sol[0] = v;
  for (int l = 1; l < nl; l++) sol[l] = (real_t *)calloc(H->L[l].size, sizeof(real_t));
  for (int l = nl - 1; l >= 0; l--) {
    if (l + 1 < nl) prolong_tl(H->L[l + 1].n, sol[l + 1], sol[l], 0);
    for (int c = 0; c < ncycle; c++) vcyc(H, l, sol[l], src[l], omega, nsweep);
  }
  for (int l = 1; l < nl; l++) { free(src[l]); free(sol[l]); }
  free(src);
  free(sol);
}

/* Exported entry points with the reference's argument meaning. */
void ORC(jacobi)(real_t *v, const real_t *f, const size_t *nmesh, const real_t *boxsize,
                 const real_t *boxcenter, real_t beta, real_t omega, int niter, const real_t *los) {
  level_t L;
  level_setup(&L, nmesh, boxsize, boxcenter, beta, los);
  real_t *tmp = (real_t *)malloc(L.size * sizeof(real_t));
  smooth(&L, v, f, tmp, omega, niter);
  free(tmp);
}

void ORC(residual)(const real_t *v, const real_t *f, real_t *r, const size_t *nmesh,
                   const real_t *boxsize, const real_t *boxcenter, real_t beta, const real_t *los) {
  level_t L;
  level_setup(&L, nmesh, boxsize, boxcenter, beta, los);
  defect(&L, v, f, r);
}

void ORC(prolong)(const real_t *v2h, real_t *v1h, const size_t *nmesh) {
  prolong_tl(nmesh, v2h, v1h, 0);
}

void ORC(reduce)(const real_t *v1h, real_t *v2h, const size_t *nmesh) {
  restrict_fw(nmesh, v1h, v2h);
}

int ORC(vcycle)(real_t *v, const real_t *f, const size_t *nmesh, const real_t *boxsize,
                const real_t *boxcenter, real_t beta, real_t omega, int niter, const real_t *los) {
  hier_t H;
  if (hier_build(&H, nmesh, boxsize, boxcenter, beta, los)) { hier_free(&H); return -1; }
  vcyc(&H, 0, v, f, omega, niter);
  hier_free(&H);
  return 0;
}

int ORC(fmg)(const real_t *f, real_t *v, const size_t *nmesh, const real_t *boxsize,
             const real_t *boxcenter, real_t beta, real_t omega, int niter, int ncycle,
             const real_t *los) {
  hier_t H;
  if (hier_build(&H, nmesh, boxsize, boxcenter, beta, los)) { hier_free(&H); return -1; }
  fmg_run(&H, f, v, omega, niter, ncycle);
  hier_free(&H);
  return 0;
}
