/*
 * p3dfft.h -- C interface of libp3dfft_amd (MI355X / gfx950).
 *
 * Drop-in for the reference's C header (include/p3dfft.h of P3DFFT 2.7.x) on
 * the ftran_r2c / btran_c2r path: the same entry-point names, argument order
 * and by-reference conventions, and the same inline Cp3dfft_* convenience
 * wrappers.  Link libp3dfft_amd.so (double) or, with SINGLE_PREC defined,
 * libp3dfft_amd_sp.so (float).
 *
 * Each declaration cites the reference entry point it replaces (paths in the
 * reference tree).
 */
#ifndef P3DFFT_AMD_H
#define P3DFFT_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifdef SINGLE_PREC
typedef float p3dfft_real;
#else
typedef double p3dfft_real;
#endif

/* ---- the reference's BIND(C) entry points ---------------------------- */

/* build/setup.F90:83 (p3dfft_setup_c, BIND(C,NAME='p3dfft_setup')).
 * dims: processor grid, must be {1,1} (one GPU = one rank); comm: ignored;
 * nxc/nyc/nzc: cut (pruned) sizes, 1..n; ow: nonzero allows btran to
 * overwrite its input; memsize: 3 ints out.  Axes up to 2^24 points; those
 * above 8192 run in several passes over a library-owned scratch buffer. */
void p3dfft_setup(int *dims, int *nx, int *ny, int *nz, int *comm, int *nxc, int *nyc, int *nzc, int *ow,
                  int *memsize);

/* build/module.F90:214 (p3dfft_get_dims_w).  conf 1: real X-pencil, 2: complex
 * Z-pencil (order (Z,Y,X) under the stride-1 layout), 3: padded in-place size. */
void p3dfft_get_dims(int *start, int *end, int *size, int *conf);

/* build/ftran.F90:469 (p3dfft_ftran_r2c_w).  A: real (nx,ny,nz) host array;
 * B: complex (nxhpc,nyc,nzc) host array as interleaved reals; op: 3 chars, the
 * third selects the Z transform: 't'/'f' FFT, 'c' DCT-I, 's' DST-I, 'n'/'0' none. */
void p3dfft_ftran_r2c(p3dfft_real *A, p3dfft_real *B, unsigned char *op);

/* build/ftran.F90:787-814 (module procedure p3dfft_ftran_r2c_1d): the X r2c
 * stage alone, real (nx, jisize, kjsize) -> complex (nx/2+1, jisize, kjsize),
 * all modes, unnormalised.  Host arrays. */
void p3dfft_ftran_r2c_1d(p3dfft_real *A, p3dfft_real *B);
/* Extension: the same on HBM arrays, asynchronous on `stream`. */
int p3dfft_ftran_r2c_1d_dev(const p3dfft_real *in, p3dfft_real *out, void *stream);
/* build/btran.F90:377 (p3dfft_btran_c2r_w).  op(1:1) selects the Z transform. */
void p3dfft_btran_c2r(p3dfft_real *A, p3dfft_real *B, unsigned char *op);

/* build/ftran.F90:359 (p3dfft_cheby_w): 'ffc' forward, scaling and the
 * Chebyshev derivative recurrence along Z. */
void p3dfft_cheby(p3dfft_real *A, p3dfft_real *B, p3dfft_real *Lz);

/* build/ftran.F90:84 (p3dfft_ftran_r2c_many_w), build/btran.F90:84
 * (p3dfft_btran_c2r_many_w), build/ftran.F90:321 (p3dfft_ftran_cheby_many_w):
 * nv variables, variable j at j*dim_in / j*dim_out elements of the array type. */
void p3dfft_ftran_r2c_many(p3dfft_real *A, int *dim_in, p3dfft_real *B, int *dim_out, int *nv, unsigned char *op);
void p3dfft_btran_c2r_many(p3dfft_real *A, int *dim_in, p3dfft_real *B, int *dim_out, int *nv, unsigned char *op);
void p3dfft_cheby_many(p3dfft_real *A, int *dim_in, p3dfft_real *B, int *dim_out, int *nv, p3dfft_real *Lz);

/* build/module.F90:301 (p3dfft_clean_w) */
void p3dfft_clean(void);

/* build/module.F90:709 / :719 (get_timers_w / set_timers_w): 12 accumulating
 * stage timers in seconds, filled from HIP events. */
void get_timers(double *timers);
void set_timers(void);

/* ---- inline C convenience layer (reference include/p3dfft.h:165-264) --- */
static inline void Cp3dfft_setup(int *dims, int nx, int ny, int nz, int comm, int nxc, int nyc, int nzc,
                                 int overwrite, int *memsize) {
  p3dfft_setup(dims, &nx, &ny, &nz, &comm, &nxc, &nyc, &nzc, &overwrite, memsize);
}
static inline void Cp3dfft_clean(void) { p3dfft_clean(); }
static inline void Cp3dfft_get_dims(int *start, int *end, int *size, int conf) {
  p3dfft_get_dims(start, end, size, &conf);
}
static inline void Cget_timers(double *timers) { get_timers(timers); }
static inline void Cset_timers(void) { set_timers(); }
static inline void Cp3dfft_ftran_r2c(p3dfft_real *A, p3dfft_real *B, unsigned char *op) {
  p3dfft_ftran_r2c(A, B, op);
}
static inline void Cp3dfft_btran_c2r(p3dfft_real *A, p3dfft_real *B, unsigned char *op) {
  p3dfft_btran_c2r(A, B, op);
}
static inline void Cp3dfft_cheby(p3dfft_real *A, p3dfft_real *B, p3dfft_real Lz) { p3dfft_cheby(A, B, &Lz); }
static inline void Cp3dfft_ftran_r2c_many(p3dfft_real *A, int dim_in, p3dfft_real *B, int dim_out, int nv,
                                          unsigned char *op) {
  p3dfft_ftran_r2c_many(A, &dim_in, B, &dim_out, &nv, op);
}
static inline void Cp3dfft_btran_c2r_many(p3dfft_real *A, int dim_in, p3dfft_real *B, int dim_out, int nv,
                                          unsigned char *op) {
  p3dfft_btran_c2r_many(A, &dim_in, B, &dim_out, &nv, op);
}
static inline void Cp3dfft_cheby_many(p3dfft_real *A, int dim_in, p3dfft_real *B, int dim_out, int nv,
                                      p3dfft_real Lz) {
  p3dfft_cheby_many(A, &dim_in, B, &dim_out, &nv, &Lz);
}

/* ---- extensions (no counterpart in the reference) -------------------- */

/* Device-pointer variants of ftran/btran/cheby: in/out are HBM addresses laid
 * out as the host arrays above; the work is queued on `stream` (a hipStream_t,
 * NULL = default stream) and the call returns without synchronising.
 * Return 0, or -1 on error (message via p3dfft_amd_last_error). */
int p3dfft_ftran_r2c_dev(const p3dfft_real *in, p3dfft_real *out, const unsigned char *op, void *stream);
int p3dfft_btran_c2r_dev(p3dfft_real *in, p3dfft_real *out, const unsigned char *op, void *stream);
int p3dfft_cheby_dev(const p3dfft_real *in, p3dfft_real *out, p3dfft_real Lz, void *stream);
/* device-pointer _many variants; dims by value */
int p3dfft_ftran_r2c_many_dev(const p3dfft_real *in, long long dim_in, p3dfft_real *out, long long dim_out, int nv,
                              const unsigned char *op, void *stream);
int p3dfft_btran_c2r_many_dev(p3dfft_real *in, long long dim_in, p3dfft_real *out, long long dim_out, int nv,
                              const unsigned char *op, void *stream);
int p3dfft_cheby_many_dev(const p3dfft_real *in, long long dim_in, p3dfft_real *out, long long dim_out, int nv,
                          p3dfft_real Lz, void *stream);

/* ---- power spectrum (the sample drivers' compute_spectrum,
 * sample/C/driver_spec.c:298-383 and sample/FORTRAN/driver_spec.F90:263-340,
 * as a library routine).  B: the current setup's spectrum in the conf-2
 * layout; E[ik], ik = 0..kmax: the sum of k2*|B|^2 over stored elements with
 * ik = int(sqrt(k2)+1/2), k2 = kx^2+ky^2+kz^2 (kx the stored x index, ky and
 * kz folded to n-j above n/2), accumulated in FP64 in both builds.  Returns 0,
 * or -1 when 4*kmax^2 is below the sum of (size-1)^2 over the three axes. */
int p3dfft_power_spectrum_dev(const p3dfft_real *B, int kmax, double *E, void *stream);
int p3dfft_power_spectrum(const p3dfft_real *B, int kmax, double *E);
/* Allocation-free p3dfft_power_spectrum_dev.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * per-workgroup partial sums go to a static 16 MiB device array that each
 * GPU holds until the process ends; a setup and kmax whose partials need
 * more is refused (-1, message) and p3dfft_power_spectrum_dev serves it.
 * Results are identical to p3dfft_power_spectrum_dev.  Calls that overlap on
 * one GPU (e.g. on two streams) share the array and race. */
int p3dfft_power_spectrum_dev_noalloc(const p3dfft_real *B, int kmax, double *E, void *stream);

/* ---- share mode: one rank of a P1 x P2 grid (no counterpart entry points in
 * the reference; they split build/ftran.F90:489-780 and build/btran.F90:396-679
 * at the MPI_Alltoallv calls of fcomm1/fcomm2/bcomm1/bcomm2).  Buffers are
 * device pointers in complex elements laid out as the reference's send and
 * receive buffers; the caller moves them between ranks. ----------------- */
int p3dfft_setup_share(int *dims, int *ipid, int *jpid, int *nx, int *ny, int *nz, int *nxc, int *nyc, int *nzc,
                       int *ow, int *memsize);
/* per-peer offsets/counts (complex elements): forward row exchange (p1 peers),
 * forward column exchange (p2 peers); backward exchanges swap send and recv */
/* 1 (default): the share-mode stages fuse the pack/unpack maps into the FFT
 * kernels' stores/loads where the lengths have specialised kernels; 0: separate
 * pack/unpack passes.  The buffers are identical either way. */
void p3dfft_share_set_fused(int on);
int p3dfft_share_counts(long long *fsnd_st, long long *fsnd_ct, long long *frcv_st, long long *frcv_ct,
                        long long *ksnd_st, long long *ksnd_ct, long long *krcv_st, long long *krcv_ct);
/* which: 0 fcomm1 pack, 1 fcomm1 unpack, 2 fcomm2 pack, 3 bcomm1 unpack,
 * 4 bcomm2 pack, 5 bcomm2 unpack */
int p3dfft_share_pack(int which, const p3dfft_real *src, p3dfft_real *dst, void *stream);
int p3dfft_share_ftran_x(const p3dfft_real *in, p3dfft_real *send1, void *stream);
int p3dfft_share_ftran_y(const p3dfft_real *recv1, p3dfft_real *send2, void *stream);
int p3dfft_share_ftran_z(p3dfft_real *recv2, p3dfft_real *out, const unsigned char *op, void *stream);
int p3dfft_share_btran_z(p3dfft_real *in, p3dfft_real *send, const unsigned char *op, void *stream);
int p3dfft_share_btran_y(const p3dfft_real *recv, p3dfft_real *send, void *stream);
int p3dfft_share_btran_x(const p3dfft_real *recv, p3dfft_real *out, void *stream);

/* ---- spectral Poisson solves on resident fields (served / batched path;
 * no counterpart in the reference): r2c, multiply by -1/|k|^2 (k=0 -> 0) and
 * 1/(nx ny nz), c2r, in place on `field` (real (nx,ny,nz) in HBM), using
 * `work` (>= p3dfft_poisson_work_bytes) on `stream`.  Independent of
 * p3dfft_setup; any number of solves may be in flight on different streams. */
size_t p3dfft_poisson_work_bytes(int nx, int ny, int nz);
int p3dfft_poisson_solve(int nx, int ny, int nz, p3dfft_real *field, void *work, size_t work_bytes, void *stream);
/* 1 (default): Z forward, multiply and Z inverse fused in one kernel; 0: the
 * composition of the forward transform, a multiply pass and the backward one */
void p3dfft_poisson_set_fused(int on);
/* The multiply step alone, on a spectrum (nx/2+1, ny, nz) from
 * p3dfft_ftran_r2c_dev: times -1/(|k|^2 nx ny nz), k = 0 set to 0. */
int p3dfft_poisson_multiply(p3dfft_real *spec, int nx, int ny, int nz, void *stream);
/* A fixed batch of n solves (shapes: 3n ints nx,ny,nz; fields[i], works[i],
 * work_bytes[i] as for p3dfft_poisson_solve) captured once into a HIP graph,
 * spread over nstreams concurrent branches.  Each launch replays every solve
 * in place on its field, asynchronously on `stream`.  Plans, twiddle tables
 * and kernel attributes are set up before capture, so the graph holds kernel
 * launches and branch joins only. */
int p3dfft_poisson_graph_create(int n, const int *shapes, p3dfft_real *const *fields, void *const *works,
                                const size_t *work_bytes, int nstreams, void **graph);
int p3dfft_poisson_graph_launch(void *graph, void *stream);
int p3dfft_poisson_graph_destroy(void *graph);
/* A fixed batch of n solves (arguments as for p3dfft_poisson_graph_create) run
 * as grouped launches: each pass of every solve whose lengths have
 * specialised kernels is one launch per pass and transform length; the
 * other solves run one by one after them.  Results are bit-identical to
 * p3dfft_poisson_solve of each field.  launch is asynchronous on `stream`. */
int p3dfft_poisson_batch_create(int n, const int *shapes, p3dfft_real *const *fields, void *const *works,
                                const size_t *work_bytes, void **batch);
int p3dfft_poisson_batch_launch(void *batch, void *stream);
int p3dfft_poisson_batch_info(void *batch, int *grouped_launches, int *single_solves);
int p3dfft_poisson_batch_destroy(void *batch);
/* Batched 3D transforms of mixed shapes without a setup: direction 1 is r2c
 * 'fft' of reals[i] (nx,ny,nz) into spectra[i] (nx/2+1,ny,nz), unnormalised
 * (bit-identical to p3dfft_ftran_r2c_dev of that grid); -1 is c2r 'tff' of
 * spectra[i] (overwritten) into reals[i].  One launch per pass and transform
 * length; every length must have a specialised kernel (32, 48, 64, 96, 128,
 * 160, 192, 256, 320, 384, 512, 1024, 2048). */
int p3dfft_amd_batch_create(int direction, int n, const int *shapes, p3dfft_real *const *reals,
                            p3dfft_real *const *spectra, void **batch);
int p3dfft_amd_batch_launch(void *batch, void *stream);
int p3dfft_amd_batch_destroy(void *batch);
/* Capture the library's asynchronous _dev calls issued on `stream` (a created
 * stream, not NULL) between begin and end into a HIP graph; launch replays
 * them.  The stage timers pause during a capture; calls that allocate on
 * first use (p3dfft_power_spectrum_dev) must have run once before it. */
int p3dfft_amd_capture_begin(void *stream);
int p3dfft_amd_capture_end(void *stream, void **graph);
int p3dfft_amd_graph_launch(void *graph, void *stream);
int p3dfft_amd_graph_destroy(void *graph);
/* The continuous-batching loop of served solves: request k solves field
 * order[k] (shapes 3 ints per field, fields, work_bytes per field) and arrives
 * arrivals[k] seconds after the start; admitted once arrived, a stream slot
 * (of `slots`) is free and its work area fits a first-fit arena over the HBM
 * region [arena, arena+arena_bytes); when the oldest waiting request does not
 * fit, up to `lookahead` later ones that do may go first, at most
 * `max_bypass` times per waiting request.  sequential != 0: one request at a
 * time.  Outputs per-request latency (s, from GPU event timestamps), the
 * makespan (s) and the peak work bytes. */
int p3dfft_serve_run(int n, const int *shapes, p3dfft_real *const *fields, const size_t *work_bytes,
                     const int *order, const double *arrivals, int slots, void *arena, size_t arena_bytes,
                     int sequential, int lookahead, int max_bypass, double *latency, double *makespan,
                     size_t *peak);
/* build the plan and twiddle table of a transform length ahead of time */
int p3dfft_amd_prepare_length(int n);
/* i.i.d. N(0,1) values from a counter-based generator keyed by seed */
int p3dfft_amd_fill_normal(p3dfft_real *p, long long n, unsigned long long seed, void *stream);
void *p3dfft_amd_stream_create(void);
int p3dfft_amd_stream_destroy(void *stream);
int p3dfft_amd_stream_synchronize(void *stream);
int p3dfft_amd_event_query(void *event);

/* Select the (Z,Y,X) complex layout for the next p3dfft_setup: the
 * reference's compile-time STRIDE1 option as a run-time switch. */
void p3dfft_amd_set_stride1(int on);

/* Error handling: mode 0 (default) prints and terminates the process on the
 * errors where the reference calls MPI_Abort; mode 1 prints, records the
 * message and returns. */
void p3dfft_amd_set_error_mode(int mode);
const char *p3dfft_amd_last_error(void);
void p3dfft_amd_clear_error(void);

/* Bytes of X/Y intermediate per z-plane chunk (default 0 = one chunk):
 * each chunk's intermediate is re-read while it is still in the last-level cache. */
void p3dfft_amd_set_chunk_bytes(size_t bytes);
/* Stage timers on (default) or off. */
void p3dfft_amd_set_timing(int on);
/* Compile-time specialised kernels for the lengths that have them (default 1);
 * 0 sends every length through the run-time-generic kernels. */
void p3dfft_amd_set_specialised(int on);
/* 1 (default): consecutive passes over a buffer walk their tiles in opposite
 * orders, so each pass starts on data the previous one wrote last (still in
 * the last-level cache); 0: every pass walks forward. Results are identical. */
void p3dfft_amd_set_tile_order(int alternate);
/* 1: the specialised column kernels use 64-bit element offsets even where
 * 32-bit ones fit (default 0; views over 4 GiB always use 64-bit offsets). */
void p3dfft_amd_set_wide_offsets(int on);
/* Tile order of the column passes over one contiguous inner axis (the Z
 * pass): 0 XCD-aware (each XCD a contiguous range of tiles), 1 dispatch
 * order (the tiles running at once cover one contiguous range of each row),
 * 2 (default) chosen by the view's shape.  Results are identical. */
void p3dfft_amd_set_flat_order(int mode);
/* X and Y of each z-plane in one workgroup (one launch and one HBM round trip
 * fewer per direction) for the specialised (nx, ny) pairs 32/64/128 without
 * X/Y cuts: 0 never, 1 wherever they apply, 2 (default) on small grids only.
 * Their time is reported in timer slot 5 (forward) / 12 (backward). */
void p3dfft_amd_set_plane_kernels(int mode);
/* X rows of 512 points: 1 through the wave kernels (one wavefront per row
 * pair; butterfly exchanges by cross-lane permlane/DPP/bpermute operations,
 * no workgroup barrier), 0 (default) through the LDS-tile kernels, which
 * measured faster.  No reference counterpart; results agree to rounding. */
void p3dfft_amd_set_x_wave(int on);
/* DCT-I Z passes ('c' ops) without the fused Chebyshev recurrence: 1 (default)
 * through the direct-load column kernel, which reads the even extension's
 * mirrored rows from the input; 0 through the kernel that builds the
 * extension in LDS first.  No reference counterpart; results agree to
 * rounding. */
void p3dfft_amd_set_dct_direct(int on);
/* Column passes of the lengths whose tile takes more than half the LDS (one
 * workgroup per CU: 2048 points): 1 (default) through a persistent kernel
 * (one workgroup per CU looping over tiles, the next tile's loads issued
 * while the current one computes), 0 one workgroup per tile.
 * No reference counterpart; results agree to rounding. */
void p3dfft_amd_set_col_pipe(int on);
/* 1 (default): a forward without Y/Z cuts whose output does not overlap its
 * input runs X r2c straight into the output and the Y pass in place there
 * (the Z pass then reads lines the Y pass left in the last-level cache);
 * 0: X r2c into the work buffer, Y from there into the output. */
void p3dfft_amd_set_x_into_out(int on);
/* Plan mode for the next p3dfft_setup (counterpart of the reference's
 * FFTW_MEASURE planning, build/fft_spec.F90:101-108).  0 (default): the
 * default plan of every length.  1: the setup times candidate plans of each
 * transform length on its own grid (specialised kernels; generic kernels with
 * 32/64/128 KiB tiles and 64/128/256-byte column segments; other radix
 * orders) and the schedule switches, and keeps the fastest.  Measured plans
 * persist for the process; results stay within each precision's tolerance. */
void p3dfft_amd_set_plan_mode(int mode);
/* the last measurement's choices as text; returns the full length */
int p3dfft_amd_plan_report(char *buf, int len);
/* default plans again (no active setup and no live Poisson graph) */
int p3dfft_amd_reset_plans(void);
/* sizeof(p3dfft_real) of this build: 8 or 4. */
int p3dfft_amd_precision(void);

/* Processor-grid geometry of the current setup (1x1, or the P1 x P2 grid of
 * a share-mode setup): the reference's module arrays proc_dims and
 * proc_id2coords, which build/setup.F90:224-229,551-574 gather from every
 * rank.  proc_dims[(c-1) + 2*((f-1) + 9*p)] for conf c = 1, 2, field
 * f = 1..9 (start x,y,z, end x,y,z, size x,y,z as p3dfft_get_dims) and rank
 * p; id2coords[2p], id2coords[2p+1] = (ipid, jpid) of rank p, with rank
 * p = jpid * P1 + ipid (the reference's Cartesian order).  Arrays of 18P and
 * 2P ints (null skips one).  Returns P, or -1 before setup. */
int p3dfft_amd_proc_dims(int *proc_dims, int *id2coords);
/* get_proc_parts (build/module.F90:888-1054, a Fortran module procedure
 * without a C binding in the reference): the pieces of the box with 1-based
 * base point (base_x, base_y, base_z) and sizes (size_x, size_y, size_z) in
 * configuration conf (1: real X pencils, 2: complex Z pencils) held by each
 * rank, in the reference's order.  parts is proc_parts(P, 7) in Fortran
 * order (parts[(col-1)*P + (q-1)]): column 1 the rank id, 2-4 the base point,
 * 5-7 the sizes, x-y-z order; unused rows -1.  *ierr = 0, 1 (conf not 1 or
 * 2) or -1 (box not inside the grid).  Returns the number of pieces or -1. */
int p3dfft_amd_get_proc_parts(int base_x, int base_y, int base_z, int size_x, int size_y, int size_z, int conf,
                              int *parts, int *ierr);

/* HBM helpers for callers without their own HIP bindings. */
void *p3dfft_amd_malloc(size_t bytes);
int p3dfft_amd_free(void *p);
int p3dfft_amd_memcpy_h2d(void *dst, const void *src, size_t bytes);
int p3dfft_amd_memcpy_d2h(void *dst, const void *src, size_t bytes);
int p3dfft_amd_memcpy_d2d(void *dst, const void *src, size_t bytes);
int p3dfft_amd_memset(void *dst, int value, size_t bytes);
int p3dfft_amd_synchronize(void);
/* streaming copy kernel (bytes a multiple of 16): HBM bandwidth reference */
int p3dfft_amd_stream_copy(const void *in, void *out, size_t bytes, void *stream);
int p3dfft_amd_device_count(void);
void *p3dfft_amd_event_create(void);
int p3dfft_amd_event_record(void *event, void *stream);
float p3dfft_amd_event_elapsed_ms(void *start, void *stop);
int p3dfft_amd_event_destroy(void *event);

#ifdef __cplusplus
}
#endif
#endif
