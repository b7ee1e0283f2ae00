/*
 * p3dfft_oracle.h -- CPU restatement of P3DFFT's forward/backward 3D FFT path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness checker for the
 * HIP implementation in p3dfft_amd/.  Only tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * Everything is FP64.  Complex arrays are interleaved (re, im) doubles, laid
 * out like the Fortran arrays of the reference (first index fastest).
 *
 * What each function restates (paths relative to the reference tree):
 *   orc_map_data_to_proc  build/setup.F90:608-635   (MapDataToProc)
 *   orc_setup             build/setup.F90:107-176, 279-313, 381-398, 580-603
 *   orc_get_dims          build/module.F90:225-273
 *   orc_ftran_r2c         build/ftran.F90:489-780 (1x1 grid; STRIDE1 output per
 *                         build/reorder.F90:359-488)
 *   orc_btran_c2r         build/btran.F90:396-679 (1x1 grid; STRIDE1 input per
 *                         build/reorder.F90:109-245, 495-530)
 *   orc_cheby             build/ftran.F90:383-464
 *   1D kernels            FFTW conventions used by build/fft_exec.F90:
 *                         r2c/c2r/c2c (unnormalised, forward e^{-i}),
 *                         REDFT00 (DCT-I) and RODFT00 (DST-I) applied to the
 *                         real and imaginary parts separately
 *                         (fft_exec.F90:646-701, 866-921).
 *   share-mode maps       build/fcomm1.F90:209-326, fcomm2.F90:250-386,
 *                         bcomm1.F90:250-380, bcomm2.F90:213-318
 *   orc_ranks_fwd_bwd     ftran.F90:489-780 and btran.F90:396-679 on a
 *                         p1 x p2 grid, the ranks as threads, the
 *                         MPI_Alltoallv exchanges as copies between them
 */
#ifndef P3DFFT_ORACLE_H
#define P3DFFT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_plan {
  int nx, ny, nz;          /* global grid */
  int nxc, nyc, nzc;       /* cut (pruned) sizes */
  int ow;                  /* overwrite flag of btran */
  int stride1;             /* complex layout (Z,Y,X) instead of (X,Y,Z) */
  int p1, p2;              /* processor grid dims(1) x dims(2) */
  int ipid, jpid;          /* this rank's grid coordinates */
  /* derived quantities, named as in build/setup.F90 */
  int nxh, nxhp, nxhc, nxhpc, nyh, nzh, nyhc, nzhc, nycph, nzcph;
  int iistart, iiend, iisize;   /* split of nxhpc over p1 */
  int jistart, jiend, jisize;   /* split of ny over p1 */
  int jjstart, jjend, jjsize;   /* split of nyc over p2 */
  int kjstart, kjend, kjsize;   /* split of nz over p2 */
  int padi;
  long long nm;
  int memsize[3];
} orc_plan;

/* ---- decomposition ---- */
void orc_map_data_to_proc(int data, int proc, int *st, int *en, int *sz);
int  orc_setup(orc_plan *p, const int dims[2], int ipid, int jpid,
               int nx, int ny, int nz, int nxc, int nyc, int nzc,
               int ow, int stride1);
void orc_get_dims(const orc_plan *p, int conf, int st[3], int en[3], int sz[3]);

/* ---- 1D building blocks (interleaved complex) ---- */
void orc_fft(double *data, int n, int sign);            /* in place, fast */
void orc_dft_naive(const double *in, double *out, int n, int sign);
void orc_r2c(const double *in, double *out, int n);     /* n real -> n/2+1 */
void orc_c2r(const double *in, double *out, int n);     /* n/2+1 -> n real */
void orc_dct1(double *data, int n, int cstride);        /* complex column, REDFT00 on re & im */
void orc_dst1(double *data, int n, int cstride);        /* RODFT00 on re & im */
void orc_dct1_naive(const double *in, double *out, int n);   /* real data */
void orc_dst1_naive(const double *in, double *out, int n);

/* ---- 3D path on the 1x1 grid ---- */
/* in: real (nx,ny,nz); out: complex (nxhpc,nyc,nzc) or (nzc,nyc,nxhpc).
 * Returns 0, or -1 on an unknown op character (reference: MPI_Abort). */
int orc_ftran_r2c(const orc_plan *p, const double *in, double *out, const char *op);
/* in: complex spectrum (overwritten when ow && nz==nzc, as the reference does);
 * out: real (nx,ny,nz). */
int orc_btran_c2r(const orc_plan *p, double *in, double *out, const char *op);
int orc_cheby(const orc_plan *p, const double *in, double *out, double Lz);

/* ---- share mode: one rank of a p1 x p2 grid (element = complex) ---- */
/* forward row transpose: pack buf2 (nxhp, jisize, kjsize) into the send
 * buffer (sum over peers of iisz(i)*jisize*kjsize). */
void orc_fcomm1_pack(const orc_plan *p, const double *buf2, double *send);
/* unpack the row receive buffer into (iisize, ny, kjsize). */
void orc_fcomm1_unpack(const orc_plan *p, const double *recv, double *dest);
/* forward column transpose: pack (iisize, ny, kjsize) into the send buffer
 * (sum over peers of iisize*jjsz(i)*kjsize), dropping pruned Y modes. */
void orc_fcomm2_pack(const orc_plan *p, const double *src, double *send);
/* the column receive buffer is the output (iisize, jjsize, nz) itself. */
/* backward column transpose: unpack the receive buffer into (iisize, ny, kjsize)
 * with the pruned Y modes zero-filled. */
void orc_bcomm1_unpack(const orc_plan *p, const double *recv, double *dest);
/* backward row transpose: pack (iisize, ny, kjsize) into the send buffer. */
void orc_bcomm2_pack(const orc_plan *p, const double *src, double *send);
/* unpack the row receive buffer into (nxhp, jisize, kjsize), zeroing x>nxhpc. */
void orc_bcomm2_unpack(const orc_plan *p, const double *recv, double *dest);
/* element offsets (complex units) of each peer's block: row (p1 entries) and
 * column (p2 entries) send/recv displacements and counts. */
void orc_share_counts(const orc_plan *p, long long *fsnd_st, long long *fsnd_ct,
                      long long *frcv_st, long long *frcv_ct,
                      long long *ksnd_st, long long *ksnd_ct,
                      long long *krcv_st, long long *krcv_ct);

/* ---- the distributed CPU path, p1 x p2 ranks as OpenMP threads ---- */
/* ops 'fft' / 'tff' without cuts; in: global real (nx,ny,nz); spec
 * (optional): global spectrum (nxhp,ny,nz) after the last forward; out
 * (optional): global real after the last backward; t_fwd / t_bwd: reps wall
 * times (s); slots (optional): p1*p2*reps*12 values, rank r's 12 timer
 * slots of repetition it at ((r*reps)+it)*12 (s, the reference's timers(1..12)
 * meaning).  Returns 0 or -1. */
int orc_ranks_fwd_bwd(int nx, int ny, int nz, int p1, int p2, const double *in, double *spec, double *out,
                      int reps, double *t_fwd, double *t_bwd, double *slots);

/* number of OpenMP threads the oracle uses (0 = leave default) */
void orc_set_threads(int n);
int  orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
