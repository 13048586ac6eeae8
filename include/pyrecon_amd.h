/*
 * pyrecon_amd -- C ABI of the MI355X (gfx950) MultiGridReconstruction library.
 *
 * The reference (pyrecon 1.0.0) binds two native libraries per precision
 * through ctypes: mesh_{float,double}.so (src/mesh.c) and
 * multigrid_{float,double}.so (src/multigrid.c); see pyrecon/mesh.py:126-154,
 * 529-653 and pyrecon/multigrid.py:115-133.  This header declares the
 * replacement: one shared library, libpyrecon_amd.so, whose entry points take
 * DEVICE pointers for every mesh, position, weight and output array and HOST
 * pointers for the small geometry arrays (nmesh, boxsize, boxcenter, los,
 * radius), plus a HIP stream (void*, NULL = default stream).  Every call
 * returns an int status (PRH_OK or a negative PRH_ERR_*); the message of the
 * last failure on the calling thread is returned by prh_last_error().
 *
 * Suffix _f32 / _f64 selects the mesh precision, like the reference's
 * FLOAT32 build switch (src/utils.h:9-19).  All kernel entry points are
 * asynchronous on the given stream unless stated otherwise.
 */
#ifndef PYRECON_AMD_H
#define PYRECON_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
#define PRH_OK 0
#define PRH_ERR_ARG (-1)       /* invalid argument */
#define PRH_ERR_RANGE (-2)     /* positions outside the box (reference: MeshError, mesh.py:549-551) */
#define PRH_ERR_NONFINITE (-3) /* NaN/inf positions or weights (reference: MeshError, mesh.py:552-553) */
#define PRH_ERR_HIP (-4)       /* HIP runtime error */
#define PRH_ERR_NOMEM (-5)     /* device allocation failed */
#define PRH_ERR_FFT (-6)       /* unused (kept for numbering) */
#define PRH_ERR_CAPACITY (-7)  /* request larger than the workspace was built for */
#define PRH_ERR_STATE (-8)     /* call out of order */

/* ---- element type codes (positions, weights, outputs) --------------------- */
#define PRH_F32 4
#define PRH_F64 8

/* ---- shift fields (reference: recon.py:283-302, multigrid.py:137-158) ----- */
#define PRH_FIELD_DISP 0
#define PRH_FIELD_RSD 1
#define PRH_FIELD_DISP_RSD 2
#define PRH_FIELD_VALUE 3 /* CIC value of the mesh (RealMesh.read_cic), out (N,) */

/* ---- runtime plumbing (no reference counterpart: device memory is what
 *      replaces the reference's numpy buffers) ---------------------------- */
const char *prh_last_error(void);
int prh_version(void);
int prh_device_count(int *count);
int prh_set_device(int device);
int prh_get_device_name(char *buffer, size_t size);
int prh_synchronize(void *stream);
int prh_stream_create(void **stream);
int prh_stream_destroy(void *stream);
int prh_malloc(void **ptr, size_t bytes);
int prh_free(void *ptr);
int prh_host_alloc(void **ptr, size_t bytes); /* pinned host memory */
int prh_host_free(void *ptr);
int prh_memcpy_htod(void *dst, const void *src, size_t bytes, void *stream);
int prh_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *stream);
/* Copies between pageable host memory and the device through the library's
 * pinned staging buffers (no reference counterpart: the reference's arrays
 * stay on the host).  Host threads copy one chunk while the transfer of the
 * previous one runs.  htod returns once src has been read (the transfers are
 * queued on stream); dtoh returns once dst holds the data. */
int prh_memcpy_htod_staged(void *dst, const void *src, size_t bytes, void *stream);
int prh_memcpy_dtoh_staged(void *dst, const void *src, size_t bytes, void *stream);
int prh_memcpy_dtod(void *dst, const void *src, size_t bytes, void *stream);
int prh_memset(void *dst, int value, size_t bytes, void *stream);
int prh_mem_info(size_t *free_bytes, size_t *total_bytes);
int prh_event_create(void **event);
int prh_event_destroy(void *event);
int prh_event_record(void *event, void *stream);
int prh_event_elapsed_ms(void *start, void *stop, float *ms); /* synchronizes on stop */
int prh_event_synchronize(void *event);
int prh_stream_wait_event(void *stream, void *event); /* later work on stream waits for event */

/* ---- workspace ------------------------------------------------------------
 * Owns everything a reconstruction of one mesh shape needs besides the
 * caller's meshes: the multigrid hierarchy (replacing the malloc/calloc per
 * level per call of src/multigrid.c:280-297, 314-324), the paint scratch
 * (points are painted in chunks of at most paint_chunk points), the Fourier
 * tables (twiddles, gains), and the FP32 hierarchy of the mixed-precision
 * solve.  Nothing is allocated by the hot calls.  flags: PRH_WS_* below.
 * A workspace may be used by one stream at a time.                          */
typedef struct prh_workspace prh_workspace;
#define PRH_WS_MULTIGRID 1u /* multigrid hierarchy in the workspace precision */
#define PRH_WS_PAINT 2u     /* CIC paint scratch */
#define PRH_WS_FFT 4u       /* Fourier tables (twiddles, gains) of the smoothing and transforms */
#define PRH_WS_MIXED 8u     /* FP32 hierarchy for prh_fmg_mixed (FP64 workspaces) */
#define PRH_WS_ALL 15u
int prh_workspace_create(prh_workspace **ws, const size_t *nmesh, int dtype, size_t paint_chunk,
                         unsigned flags);
int prh_workspace_destroy(prh_workspace *ws);
int prh_workspace_bytes(const prh_workspace *ws, size_t *bytes);
/* Grows the paint scratch to chunks of paint_chunk points (no reference
 * counterpart: the reference paints every catalog in one pass from host
 * memory), keeping the cell and tile orders of the last two paints for their
 * reads.  Synchronizes the device.  PRH_ERR_NOMEM leaves the workspace as it
 * was; a smaller paint_chunk than the current one changes nothing. */
int prh_workspace_reserve_paint(prh_workspace *ws, size_t paint_chunk);
/* bytes a workspace with these arguments would allocate (admission control) */
int prh_workspace_footprint(const size_t *nmesh, int dtype, size_t paint_chunk, unsigned flags,
                            size_t *bytes);

/* ---- reference-shaped kernels (mesh units, device pointers) ---------------
 * Same arguments and meaning as the reference C functions; positions are
 * already in mesh units, (N, 3) C-order, cast to the mesh precision.       */

/* replaces assign_cic, src/mesh.c:26-63.  Deterministic: the result equals
 * the reference's serial accumulation in point order, bit for bit.        */
int prh_assign_cic_f32(prh_workspace *ws, float *mesh, const size_t *nmesh, const float *positions,
                       const float *weights, size_t npositions, void *stream);
int prh_assign_cic_f64(prh_workspace *ws, double *mesh, const size_t *nmesh,
                       const double *positions, const double *weights, size_t npositions,
                       void *stream);
/* replaces read_cic, src/mesh.c:269-310 */
int prh_read_cic_f32(const float *mesh, const size_t *nmesh, const float *positions, float *values,
                     size_t npositions, void *stream);
int prh_read_cic_f64(const double *mesh, const size_t *nmesh, const double *positions,
                     double *values, size_t npositions, void *stream);
/* replaces read_finite_difference_cic, src/mesh.c:187-266 */
int prh_read_finite_difference_cic_f32(const float *mesh, const size_t *nmesh,
                                       const float *boxsize, const float *positions,
                                       float *shifts, size_t npositions, void *stream);
int prh_read_finite_difference_cic_f64(const double *mesh, const size_t *nmesh,
                                       const double *boxsize, const double *positions,
                                       double *shifts, size_t npositions, void *stream);
/* replaces jacobi, src/multigrid.c:44-104 (ws supplies the second iterate) */
int prh_jacobi_f32(prh_workspace *ws, float *v, const float *f, const size_t *nmesh,
                   const float *boxsize, const float *boxcenter, float beta, float damping_factor,
                   int niterations, const float *los, void *stream);
/* The mixed-precision solve's FP32 smoother (no reference counterpart; for
 * tests): niterations fast damped-Jacobi sweeps (hardware reciprocal,
 * explicit FMA; not the reference's bits), two per launch on levels with whole
 * rows of 256 or 512 cells (PRH_TUNE_PAIRED_SWEEPS), from zero if zero_start. */
int prh_jacobi_fast_f32(prh_workspace *ws, float *v, const float *f, const size_t *nmesh,
                        const float *boxsize, const float *boxcenter, float beta, float damping_factor,
                        int niterations, const float *los, int zero_start, void *stream);
int prh_jacobi_f64(prh_workspace *ws, double *v, const double *f, const size_t *nmesh,
                   const double *boxsize, const double *boxcenter, double beta,
                   double damping_factor, int niterations, const double *los, void *stream);
/* replaces residual, src/multigrid.c:107-164 */
int prh_residual_f32(const float *v, const float *f, float *r, const size_t *nmesh,
                     const float *boxsize, const float *boxcenter, float beta, const float *los,
                     void *stream);
int prh_residual_f64(const double *v, const double *f, double *r, const size_t *nmesh,
                     const double *boxsize, const double *boxcenter, double beta,
                     const double *los, void *stream);
/* replaces prolong, src/multigrid.c:167-209 (nmesh = coarse shape) */
int prh_prolong_f32(const float *v2h, float *v1h, const size_t *nmesh, void *stream);
int prh_prolong_f64(const double *v2h, double *v1h, const size_t *nmesh, void *stream);
/* replaces reduce, src/multigrid.c:212-266 (nmesh = fine shape) */
int prh_reduce_f32(const float *v1h, float *v2h, const size_t *nmesh, void *stream);
int prh_reduce_f64(const double *v1h, double *v2h, const size_t *nmesh, void *stream);
/* replaces vcycle, src/multigrid.c:270-300 */
int prh_vcycle_f32(prh_workspace *ws, float *v, const float *f, const size_t *nmesh,
                   const float *boxsize, const float *boxcenter, float beta, float damping_factor,
                   int niterations, const float *los, void *stream);
int prh_vcycle_f64(prh_workspace *ws, double *v, const double *f, const size_t *nmesh,
                   const double *boxsize, const double *boxcenter, double beta,
                   double damping_factor, int niterations, const double *los, void *stream);
/* replaces fmg, src/multigrid.c:303-328: v1h is the initial iterate and
 * receives the solution (the reference's return value is v1h too).       */
int prh_fmg_f32(prh_workspace *ws, const float *f1h, float *v1h, const size_t *nmesh,
                const float *boxsize, const float *boxcenter, float beta,
                float jacobi_damping_factor, int jacobi_niterations, int vcycle_niterations,
                const float *los, void *stream);
int prh_fmg_f64(prh_workspace *ws, const double *f1h, double *v1h, const size_t *nmesh,
                const double *boxsize, const double *boxcenter, double beta,
                double jacobi_damping_factor, int jacobi_niterations, int vcycle_niterations,
                const double *los, void *stream);

/* Allocation-free full multigrid.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * result as prh_fmg_f64 without a workspace: the multigrid hierarchy lives in
 * a file-scope device array sized for meshes of at most 2^18 (64^3) cells
 * (3.0 MB, held by every GPU the library runs on until the process ends);
 * larger meshes return PRH_ERR_CAPACITY.  Calls no allocating routine.  All
 * calls on one GPU share that array: calls that overlap on one GPU race.   */
int prh_fmg_noalloc_f64(const double *f1h, double *v1h, const size_t *nmesh,
                        const double *boxsize, const double *boxcenter, double beta,
                        double jacobi_damping_factor, int jacobi_niterations,
                        int vcycle_niterations, const double *los, void *stream);

/* ---- fused pipeline entry points (what the drop-in Python class calls) ----
 * positions: device (N, 3) array of pos_dtype (PRH_F32/PRH_F64) in
 * Cartesian coordinates; the mesh-unit transform, the wrap or range check and
 * the finite check of pyrecon/mesh.py:529-568 are done on the device with
 * numpy's type promotion.  boxsize/boxcenter are host arrays in the mesh
 * precision (MeshInfo stores them so, mesh.py:447-455).                    */

/* replaces RealMesh.assign_cic (mesh.py:529-568) + assign_cic (mesh.c:26).
 * weights may be NULL (unit weights).  Returns PRH_ERR_RANGE / _NONFINITE
 * (mesh untouched) like the reference's MeshError.  Synchronizes stream.  */
int prh_paint_f32(prh_workspace *ws, float *mesh, const size_t *nmesh, const float *boxsize,
                  const float *boxcenter, const void *positions, int pos_dtype,
                  const void *weights, int weight_dtype, size_t npositions, int wrap,
                  void *stream);
int prh_paint_f64(prh_workspace *ws, double *mesh, const size_t *nmesh, const double *boxsize,
                  const double *boxcenter, const void *positions, int pos_dtype,
                  const void *weights, int weight_dtype, size_t npositions, int wrap,
                  void *stream);

/* the mixed-precision variant of prh_paint_f64 (same arguments and checks;
 * replaces the same assign_cic, mesh.c:26): points are sorted by 8^3-cell
 * tile, each tile sums its FP64 corner weights as 64-bit integers at the scale
 * 2^(62 - e) of its own weight bound 2^e, and tiles are combined in FP64 in a
 * fixed order: the reference's sums to within their own rounding (not its
 * bits), identical bits on repeated calls.  Needs PRH_WS_PAINT | PRH_WS_MIXED. */
int prh_paint_mixed_f64(prh_workspace *ws, double *mesh, const size_t *nmesh,
                        const double *boxsize, const double *boxcenter, const void *positions,
                        int pos_dtype, const void *weights, int weight_dtype, size_t npositions,
                        int wrap, void *stream);

/* replaces RealMesh.smooth_gaussian(method='fft') (mesh.py:655-680): in-place
 * multiplication of the spectrum by exp(-(R k)^2 / 2), done as a 1-D
 * forward transform, gain and inverse transform along each axis in turn.   */
int prh_smooth_gaussian_f32(prh_workspace *ws, float *mesh, const size_t *nmesh,
                            const float *boxsize, const double *radius, void *stream);
int prh_smooth_gaussian_f64(prh_workspace *ws, double *mesh, const size_t *nmesh,
                            const double *boxsize, const double *radius, void *stream);

/* ---- FFT reconstructions (plane_parallel_fft.py, iterative_fft.py,
 *      iterative_fft_particle.py) ---------------------------------------------
 * 3-D c2c transform of an (n0, n1, n2) mesh, as numpy.fft.fftn / ifftn
 * without the 1 / N of the inverse (RealMesh.to_complex and
 * ComplexMesh.to_real, mesh.py:697-705, 827-844, 974-988).  Complex meshes are
 * interleaved (re, im) in the mesh precision.  Forward (inverse = 0): src real
 * (io = PRH_FFT_REAL_IN) or complex -> dst complex.  Inverse: src complex,
 * transformed in place along x and y -> dst complex, or its real part
 * (io = PRH_FFT_REAL_OUT).  dst may equal src for complex -> complex.       */
#define PRH_FFT_COMPLEX 0
#define PRH_FFT_REAL_IN 1
#define PRH_FFT_REAL_OUT 2
int prh_fft3d_f32(prh_workspace *ws, const void *src, void *dst, const size_t *nmesh, int inverse,
                  int io, void *stream);
int prh_fft3d_f64(prh_workspace *ws, const void *src, void *dst, const size_t *nmesh, int inverse,
                  int io, void *stream);

/* Fourier-space operators on a complex spectrum, k_i = 2 pi fftfreq(n_i,
 * L_i / n_i) (ComplexMesh.coords, mesh.py:812-820); dst may equal src; norm
 * multiplies the output (1 / N before an inverse transform), except for
 * PRH_SPEC_INV_K2.  los: double[3] (PLANE_PARALLEL, LOS_SQ), else NULL.     */
#define PRH_SPEC_PLANE_PARALLEL 0 /* i k_a (src / k^2) / (1 + beta mu^2)    plane_parallel_fft.py:45-53 */
#define PRH_SPEC_INV_K2 1         /* src / k^2, 0 at k = 0 (prod_sum, exp=-1) iterative_fft.py:37-39 */
#define PRH_SPEC_LOS_SQ 2         /* (k . los)^2 src                        iterative_fft.py:44 */
#define PRH_SPEC_KA_KB 3          /* k_a k_b src                            iterative_fft.py:58 */
#define PRH_SPEC_GRAD 4           /* i k_a (src / k^2), 0 at k = 0          iterative_fft.py:75-81 */
#define PRH_SPEC_I_KA 5           /* (src i) k_a                            iterative_fft_particle.py:112 */
int prh_spectral_f32(prh_workspace *ws, int op, const void *src, void *dst, const size_t *nmesh,
                     const float *boxsize, const double *los, double beta, double norm, int a, int b,
                     void *stream);
int prh_spectral_f64(prh_workspace *ws, int op, const void *src, void *dst, const size_t *nmesh,
                     const double *boxsize, const double *los, double beta, double norm, int a, int b,
                     void *stream);

/* RSD removal of one IterativeFFTReconstruction step (iterative_fft.py:41-67):
 * a < 0 (global line of sight): delta -= factor * t; else (local line of
 * sight) delta -= factor * x_a x_b t / |x|^2 at the mesh coordinates
 * x_i = offset_i + cellsize_i j_i (RealMesh.coords, mesh.py:522-527).       */
int prh_rsd_update_f32(float *delta, const float *t, const size_t *nmesh, const float *boxsize,
                       const float *boxcenter, double factor, int a, int b, void *stream);
int prh_rsd_update_f64(double *delta, const double *t, const size_t *nmesh, const double *boxsize,
                       const double *boxcenter, double factor, int a, int b, void *stream);

/* Deterministic FP64 sums of one or two meshes (recon.py:227 np.sum calls).
 * sums: host double[2]; synchronizes stream.                               */
int prh_mesh_sums_f32(prh_workspace *ws, const float *a, const float *b, size_t size,
                      double *sums, void *stream);
int prh_mesh_sums_f64(prh_workspace *ws, const double *a, const double *b, size_t size,
                      double *sums, void *stream);

/* replaces the arithmetic of BaseReconstruction.set_density_contrast
 * (recon.py:229-244): delta = data - alpha*randoms; delta /= bias_alpha *
 * randoms where randoms > threshold, 0 elsewhere.  delta may alias data.
 * With randoms == NULL (recon.py:246-247): delta = (data / mean - 1) / bias,
 * where alpha carries mean.  Scalars are in the mesh precision, computed by
 * the caller as numpy does.                                                */
int prh_density_contrast_f32(const float *data, const float *randoms, float *delta, size_t size,
                             float alpha, float threshold, float bias_alpha, float bias,
                             void *stream);
int prh_density_contrast_f64(const double *data, const double *randoms, double *delta,
                             size_t size, double alpha, double threshold, double bias_alpha,
                             double bias, void *stream);

/* replaces the sums of OriginalMultiGridReconstruction.set_density_contrast
 * (multigrid.py:79-83): sums[0] = sum of data over the cells that
 * 0 < randoms < ran_min does not zero, sums[1] = sum of randoms >= ran_min.
 * Deterministic FP64; sums: host double[2]; synchronizes stream.           */
int prh_martin_sums_f32(prh_workspace *ws, const float *data, const float *randoms, size_t size,
                        float ran_min, double *sums, void *stream);
int prh_martin_sums_f64(prh_workspace *ws, const double *data, const double *randoms,
                        size_t size, double ran_min, double *sums, void *stream);

/* multigrid.py:84-95: delta = (data / (alpha * randoms) - 1) / bias where
 * randoms >= ran_min, 0 / bias elsewhere (delta may alias data); sums[0] =
 * FP64 sum of the non-zero delta, sums[1] = their count.  Host double[2];
 * synchronizes stream.                                                     */
int prh_martin_contrast_f32(prh_workspace *ws, const float *data, const float *randoms,
                            float *delta, size_t size, float ran_min, float alpha, float bias,
                            double *sums, void *stream);
int prh_martin_contrast_f64(prh_workspace *ws, const double *data, const double *randoms,
                            double *delta, size_t size, double ran_min, double alpha,
                            double bias, double *sums, void *stream);

/* multigrid.py:95-96: delta[delta != 0] -= mean.                           */
int prh_subtract_nonzero_mean_f32(float *delta, size_t size, float mean, void *stream);
int prh_subtract_nonzero_mean_f64(double *delta, size_t size, double mean, void *stream);

/* replaces BaseMesh.__mul__ / __div__ / __rdiv__ and their in-place forms
 * (pyrecon/mesh.py:156-183) for a Python scalar: out[i] = a[i] * s (op
 * PRH_OP_MUL), a[i] / s (PRH_OP_DIV) or s / a[i] (PRH_OP_RDIV), with s cast to
 * the mesh precision first (numpy's rule for a Python scalar).  out may be a.  */
#define PRH_OP_MUL 0
#define PRH_OP_DIV 1
#define PRH_OP_RDIV 2
int prh_mesh_scalar_op_f32(const float *a, float *out, size_t size, int op, double scalar,
                           void *stream);
int prh_mesh_scalar_op_f64(const double *a, double *out, size_t size, int op, double scalar,
                           void *stream);

/* Grouped full multigrid (no reference counterpart; for serving several
 * reconstructions at once): count (1..8) independent problems of the same
 * nmesh and precision, job j with its own workspace ws[j], source f1h[j],
 * iterate v1h[j], geometry boxsize[3j..3j+2], boxcenter[3j..3j+2], beta[j]
 * and line of sight los[j] (los == NULL or los[j] == NULL: local).  The
 * levels handled by the generic kernels are launched once for all jobs; every
 * job's result is bit-identical to prh_fmg_* (or prh_fmg_mixed) on it alone. */
int prh_fmg_group_f32(prh_workspace *const *ws, int count, const float *const *f1h,
                      float *const *v1h, const size_t *nmesh, const float *boxsize,
                      const float *boxcenter, const float *beta, float jacobi_damping_factor,
                      int jacobi_niterations, int vcycle_niterations, const float *const *los,
                      void *stream);
int prh_fmg_group_f64(prh_workspace *const *ws, int count, const double *const *f1h,
                      double *const *v1h, const size_t *nmesh, const double *boxsize,
                      const double *boxcenter, const double *beta, double jacobi_damping_factor,
                      int jacobi_niterations, int vcycle_niterations, const double *const *los,
                      void *stream);
int prh_fmg_mixed_group(prh_workspace *const *ws, int count, const double *const *f1h,
                        double *const *v1h, const size_t *nmesh, const double *boxsize,
                        const double *boxcenter, const double *beta, double jacobi_damping_factor,
                        int jacobi_niterations, int vcycle_niterations, const double *const *los,
                        int nrefine, int vcycles_per_refine, void *stream);

/* Synthetic catalog on the device (no reference counterpart; load generation
 * for serving and benchmarks).  positions (n, 3) and weights (n,) (may be
 * NULL) of dtype PRH_F32/PRH_F64; point i depends on (seed, i) only.  With
 * nmodes == 0 points are uniform in the box; otherwise uniform points are
 * displaced by psi = sum_m modes[6m+3] sin(k_m . q + modes[6m+4]) k_m with
 * k_m = modes[6m..6m+2] (device array of 6 * nmodes doubles), shifted by
 * f (psi . rhat) rhat along the local line of sight and wrapped into the box.
 * Weights are uniform in [wlo, whi).                                       */
int prh_synth_points(void *positions, void *weights, int dtype, size_t n, uint64_t seed,
                     const double *boxsize, const double *boxcenter, const double *modes,
                     int nmodes, double f, double wlo, double whi, void *stream);

/* replaces read_shifts / read_shifted_positions of MultiGridReconstruction
 * (multigrid.py:137-158, recon.py:304-325, mesh.py:612-653): finite-
 * difference CIC read of phi, the RSD term along the local (los == NULL) or
 * global line of sight, and, if shifted != 0, positions - shifts (wrapped when
 * wrap != 0).  out is (N, 3) of out_dtype, which the caller derives with
 * numpy's promotion rules (see pyrecon_amd/multigrid.py).  Returns
 * PRH_ERR_RANGE if wrap == 0 and a position is outside the box; sets
 * *nonfinite (host int, may be NULL) when positions hold NaN/inf (the
 * reference only warns, mesh.py:637-639).  ws holds the device status
 * flags (any workspace; one stream at a time).  Synchronizes stream.
 * field == PRH_FIELD_DISP with shifted == 0 is RealMesh.
 * read_finite_difference_cic (mesh.py:612-653) of the mesh phi; field ==
 * PRH_FIELD_VALUE (shifted == 0) is RealMesh.read_cic (mesh.py:570-610): out
 * is (N,) in the positions dtype.                                          */
int prh_read_shifts_f32(prh_workspace *ws, const float *phi, const size_t *nmesh, const float *boxsize,
                        const float *boxcenter, const void *positions, int pos_dtype,
                        size_t npositions, double f, const float *los, int field, int wrap,
                        int shifted, void *out, int out_dtype, int *nonfinite, void *stream);
int prh_read_shifts_f64(prh_workspace *ws, const double *phi, const size_t *nmesh, const double *boxsize,
                        const double *boxcenter, const void *positions, int pos_dtype,
                        size_t npositions, double f, const double *los, int field, int wrap,
                        int shifted, void *out, int out_dtype, int *nonfinite, void *stream);
/* The shift algebra of the FFT reconstructions (read_shifts / read_shifted_positions,
 * pyrecon/recon.py:255-325, without wrap), after the three displacement meshes have been read
 * by CIC at the positions (prh_read_shifts_* with PRH_FIELD_VALUE): s0, s1, s2 hold the
 * components in the positions dtype.  field as above (not PRH_FIELD_VALUE); los in the mesh
 * dtype, or NULL for the local line of sight; shifted returns positions - shifts; out (N, 3) in
 * numpy's promotion of the reference's expressions. */
int prh_shift_algebra_f32(const void *positions, int pos_dtype, const void *s0, const void *s1, const void *s2,
                          size_t npositions, double f, const float *los, int field, int shifted, void *out,
                          int out_dtype, void *stream);
int prh_shift_algebra_f64(const void *positions, int pos_dtype, const void *s0, const void *s1, const void *s2,
                          size_t npositions, double f, const double *los, int field, int shifted, void *out,
                          int out_dtype, void *stream);

/* Mixed-precision solve (no reference counterpart; same contract as
 * prh_fmg_f64, which it tracks).  The coarse stages of the full multigrid
 * (levels below the finest) run in FP32; the vcycle_niterations
 * finest-level V-cycles run as FP64 defect corrections: r = f - A v (FP64,
 * stored FP32), e = vcycles_per_refine FP32 V-cycles on r from zero (fast
 * FP32 kernels), v += e (FP64).  Since a V-cycle is affine in (v, f), this
 * reproduces the FP64 iterate up to FP32 rounding of the corrections.
 * nrefine extra corrections may follow.  Requires PRH_WS_MIXED.  If
 * residual_norms is not NULL it receives vcycle_niterations + nrefine + 1
 * FP64 RMS residual norms (host), one before each correction and one at the
 * end, and the call synchronizes.                                          */
int prh_fmg_mixed(prh_workspace *ws, const double *f1h, double *v1h, const size_t *nmesh,
                  const double *boxsize, const double *boxcenter, double beta,
                  double jacobi_damping_factor, int jacobi_niterations, int vcycle_niterations,
                  const double *los, int nrefine, int vcycles_per_refine, double *residual_norms,
                  void *stream);

/* ---- measurement (no reference counterpart) --------------------------------
 * While a timer is enabled, every launch of that kernel class is bracketed by
 * HIP events on its own stream (up to 65536 launches); prh_timer_read
 * synchronizes and returns the summed device time, the number of launches
 * and the algorithmic bytes they moved (e.g. 3 n^3 words per Jacobi sweep).
 * Process-wide; not for concurrent use from several threads.              */
#define PRH_TIMER_JACOBI_FINEST 0 /* damped-Jacobi sweeps on the finest level */
#define PRH_TIMER_PAINT_GATHER 1  /* CIC paint merge-gather kernel */
#define PRH_TIMER_READ_SHIFTS 2   /* fused finite-difference read + shift kernel */
#define PRH_TIMER_PAIR_FINEST 3   /* two fused FP32 sweeps on the finest level (mixed path) */
#define PRH_TIMER_FFT_PASS 4      /* one axis pass of a 3-D transform (prh_fft3d_*) */
#define PRH_TIMER_COUNT 5
int prh_timer_enable(int which, int enable);
int prh_timer_reset(int which);
int prh_timer_read(int which, double *total_ms, long long *launches, double *bytes);
/* Process-wide tuning switches (no reference counterpart); results do not
 * depend on them.  PRH_TUNE_PAIRED_SWEEPS (default 0): 1 makes the mixed-
 * precision solve run two Jacobi sweeps per launch on levels with whole rows
 * of 256 or 512 cells (the same bits; slower on MI355X, where the sweep is
 * nearly VALU-bound).  Returns the previous value, or PRH_ERR_ARG for an
 * unknown option. */
#define PRH_TUNE_PAIRED_SWEEPS 0
int prh_set_tuning(int option, int value);

/* a plain device-to-device copy kernel (HBM traffic calibration for profiles) */
int prh_copy_kernel(void *dst, const void *src, size_t bytes, void *stream);

/* RMS of r = f - A v in FP64 (diagnostics and tests); host result, syncs. */
int prh_residual_norm_f64(prh_workspace *ws, const double *v, const double *f, const size_t *nmesh,
                          const double *boxsize, const double *boxcenter, double beta,
                          const double *los, double *rms, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* PYRECON_AMD_H */
