/*
 * p3dfft_oracle.c -- FP64 CPU restatement of P3DFFT's ftran_r2c / btran_c2r
 * path (1x1 processor grid) plus its decomposition arithmetic and the
 * pack/unpack index maps of the row/column transposes.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker for the HIP library in
 * p3dfft_amd/, never part of it.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.
 *
 * The reference delegates its 1D transforms to FFTW (absent here).  The FFT
 * below is an independent FP64 mixed-radix decimation-in-time FFT whose
 * twiddles come from exact argument reduction in long double; it follows the
 * FFTW conventions the reference relies on (unnormalised, forward e^{-2 pi i/n},
 * r2c keeps k=0..n/2, c2r drops Im of k=0 and k=n/2, REDFT00/RODFT00).
 * Citations are file:line in the reference tree.
 */
#include "p3dfft_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* threads                                                             */
/* ------------------------------------------------------------------ */
static int g_threads = 0;
void orc_set_threads(int n) {
  g_threads = n;
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#endif
}
int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* decomposition: build/setup.F90:608-635 (MapDataToProc)              */
/* remainder goes to the LAST ranks                                    */
/* ------------------------------------------------------------------ */
void orc_map_data_to_proc(int data, int proc, int *st, int *en, int *sz) {
  int size = data / proc;
  int nu = data - size * proc;
  int nl = proc - nu;
  int i;
  st[0] = 1;
  sz[0] = size;
  en[0] = size;
  for (i = 1; i <= nl - 1; i++) {
    st[i] = st[i - 1] + size;
    sz[i] = size;
    en[i] = en[i - 1] + size;
  }
  size = size + 1;
  for (i = nl; i <= proc - 1; i++) {
    st[i] = en[i - 1] + 1;
    sz[i] = size;
    en[i] = en[i - 1] + size;
  }
  en[proc - 1] = data;
  sz[proc - 1] = data - st[proc - 1] + 1;
}

/* build/setup.F90:107-176 (sizes), 279-313 (local extents), 381-398 (padi,
 * nm), 580-603 (memsize).  Returns 0 or -1 on invalid input (the reference
 * aborts there, setup.F90:125-128, 181-184). */
int orc_setup(orc_plan *p, const int dims[2], int ipid, int jpid, int nx, int ny,
              int nz, int nxc, int nyc, int nzc, int ow, int stride1) {
  int *st, *en, *sz, P;
  long long padd, pad1;
  if (nx <= 0 || ny <= 0 || nz <= 0) return -1;
  if (dims[0] <= 0 || dims[1] <= 0) return -1;
  if (ipid < 0 || ipid >= dims[0] || jpid < 0 || jpid >= dims[1]) return -1;
  memset(p, 0, sizeof(*p));
  p->nx = nx; p->ny = ny; p->nz = nz;
  p->nxc = nxc > 0 ? nxc : nx;
  p->nyc = nyc > 0 ? nyc : ny;
  p->nzc = nzc > 0 ? nzc : nz;
  p->ow = ow; p->stride1 = stride1;
  p->p1 = dims[0]; p->p2 = dims[1]; p->ipid = ipid; p->jpid = jpid;
  p->nxh = nx / 2; p->nxhp = p->nxh + 1;
  p->nxhc = p->nxc / 2; p->nxhpc = p->nxhc + 1;
  p->nyh = ny / 2; p->nzh = nz / 2;
  p->nyhc = p->nyc / 2; p->nzhc = p->nzc / 2;
  p->nycph = (p->nyc + 1) / 2; p->nzcph = (p->nzc + 1) / 2;

  P = p->p1 > p->p2 ? p->p1 : p->p2;
  st = (int *)malloc(sizeof(int) * P);
  en = (int *)malloc(sizeof(int) * P);
  sz = (int *)malloc(sizeof(int) * P);
  orc_map_data_to_proc(p->nxhpc, p->p1, st, en, sz);
  p->iistart = st[ipid]; p->iiend = en[ipid]; p->iisize = sz[ipid];
  orc_map_data_to_proc(ny, p->p1, st, en, sz);
  p->jistart = st[ipid]; p->jiend = en[ipid]; p->jisize = sz[ipid];
  orc_map_data_to_proc(p->nyc, p->p2, st, en, sz);
  p->jjstart = st[jpid]; p->jjend = en[jpid]; p->jjsize = sz[jpid];
  orc_map_data_to_proc(nz, p->p2, st, en, sz);
  p->kjstart = st[jpid]; p->kjend = en[jpid]; p->kjsize = sz[jpid];
  free(st); free(en); free(sz);

  /* setup.F90:382-393 */
  {
    long long a = (long long)p->iisize * p->jjsize * nz;
    long long b = (long long)p->iisize * ny * p->kjsize;
    long long m = a > b ? a : b;
    long long d = (long long)p->nxhp * p->jisize;
    padd = m - (long long)p->nxhp * p->jisize * p->kjsize;
    if (padd <= 0) p->padi = 0;
    else p->padi = (int)((padd % d == 0) ? padd / d : padd / d + 1);
    p->nm = (long long)p->nxhp * p->jisize * (p->kjsize + p->padi);
  }
  /* setup.F90:580-603 */
  {
    long long a = (long long)nz * p->jjsize * p->iisize;
    long long b = (long long)ny * p->kjsize * p->iisize;
    long long d = (long long)nx * p->jisize;
    long long padi2;
    pad1 = 2 * (a > b ? a : b) - (long long)nx * p->jisize * p->kjsize;
    if (pad1 <= 0) pad1 = 0;
    padi2 = (pad1 % d != 0) ? pad1 / d + 1 : pad1 / d;
    p->memsize[0] = nx;
    p->memsize[1] = p->jisize;
    p->memsize[2] = p->kjsize + (int)padi2;
  }
  return 0;
}

/* build/module.F90:225-273 */
void orc_get_dims(const orc_plan *p, int conf, int s[3], int e[3], int z[3]) {
  if (conf == 1) {
    s[0] = 1; e[0] = p->nx; z[0] = p->nx;
    s[1] = p->jistart; e[1] = p->jiend; z[1] = p->jisize;
    s[2] = p->kjstart; e[2] = p->kjend; z[2] = p->kjsize;
  } else if (conf == 2) {
    if (p->stride1) {
      s[2] = p->iistart; e[2] = p->iiend; z[2] = p->iisize;
      s[1] = p->jjstart; e[1] = p->jjend; z[1] = p->jjsize;
      s[0] = 1; e[0] = p->nzc; z[0] = p->nzc;
    } else {
      s[0] = p->iistart; e[0] = p->iiend; z[0] = p->iisize;
      s[1] = p->jjstart; e[1] = p->jjend; z[1] = p->jjsize;
      s[2] = 1; e[2] = p->nzc; z[2] = p->nzc;
    }
  } else if (conf == 3) {
    int i;
    for (i = 0; i < 3; i++) { s[i] = 0; e[i] = p->memsize[i]; z[i] = p->memsize[i]; }
  }
}

/* ------------------------------------------------------------------ */
/* twiddles: exact quadrant reduction, long double evaluation          */
/* returns exp(-2 pi i e / N)                                          */
/* ------------------------------------------------------------------ */
static void twid(long long e, long long N, double *c, double *s) {
  long long q, r;
  long double a, cs, sn, t;
  e %= N;
  if (e < 0) e += N;
  /* 4e = q*N + r, 0 <= r < N: angle = 2 pi e/N = (pi/2)(q + r/N) */
  q = (4 * e) / N;
  r = (4 * e) - q * N;
  a = (long double)r / (long double)N * 1.5707963267948966192313216916397514L;
  if (r == 0) { cs = 1.0L; sn = 0.0L; }
  else if (2 * r == N) { cs = sqrtl(0.5L); sn = cs; }
  else { cs = cosl(a); sn = sinl(a); }
  /* rotate by q quarter turns: (cs + i sn) * i^q */
  switch (q & 3) {
    case 0: break;
    case 1: t = cs; cs = -sn; sn = t; break;
    case 2: cs = -cs; sn = -sn; break;
    case 3: t = cs; cs = sn; sn = -t; break;
  }
  *c = (double)cs;
  *s = (double)(-sn); /* e^{-i angle} */
}

/* ------------------------------------------------------------------ */
/* 1D complex FFT: recursive mixed-radix DIT over a cached length plan */
/* ------------------------------------------------------------------ */
typedef struct fplan {
  int n;
  int nf;
  int fac[64];
  double *w; /* exp(-2 pi i k/n), k=0..n-1, interleaved */
} fplan;

static void fplan_init(fplan *fp, int n) {
  int m = n, f, k;
  fp->n = n;
  fp->nf = 0;
  /* prefer radix 4, then 2, 3, 5, then any prime */
  while (m % 4 == 0) { fp->fac[fp->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { fp->fac[fp->nf++] = 2; m /= 2; }
  for (f = 3; (long long)f * f <= m; f += 2)
    while (m % f == 0) { fp->fac[fp->nf++] = f; m /= f; }
  if (m > 1) fp->fac[fp->nf++] = m;
  fp->w = (double *)malloc(sizeof(double) * 2 * (n > 0 ? n : 1));
  for (k = 0; k < n; k++) twid(k, n, &fp->w[2 * k], &fp->w[2 * k + 1]);
}

/* x: input with stride xs (complex units); y: contiguous output of length n. */
static void dit(const double *x, long long xs, double *y, int n, const int *fac,
                const double *W, int N, int sign, double *tmp) {
  int p, m, k, r, q;
  long long step;
  if (n == 1) { y[0] = x[0]; y[1] = x[1]; return; }
  p = fac[0];
  m = n / p;
  for (r = 0; r < p; r++) dit(x + 2 * r * xs, xs * p, y + 2 * (long long)r * m, m, fac + 1, W, N, sign, tmp);
  step = N / n;
  for (k = 0; k < m; k++) {
    for (r = 0; r < p; r++) {
      long long e = ((long long)r * k % n) * step;
      double wr = W[2 * e], wi = sign < 0 ? W[2 * e + 1] : -W[2 * e + 1];
      double ar = y[2 * ((long long)r * m + k)], ai = y[2 * ((long long)r * m + k) + 1];
      tmp[2 * r] = ar * wr - ai * wi;
      tmp[2 * r + 1] = ar * wi + ai * wr;
    }
    if (p == 2) {
      y[2 * k] = tmp[0] + tmp[2];
      y[2 * k + 1] = tmp[1] + tmp[3];
      y[2 * (m + k)] = tmp[0] - tmp[2];
      y[2 * (m + k) + 1] = tmp[1] - tmp[3];
    } else if (p == 4) {
      /* X_q = sum_r t_r (-i sign')^{rq}; forward uses -i */
      double s = sign < 0 ? 1.0 : -1.0;
      double a0r = tmp[0] + tmp[4], a0i = tmp[1] + tmp[5];
      double a1r = tmp[0] - tmp[4], a1i = tmp[1] - tmp[5];
      double b0r = tmp[2] + tmp[6], b0i = tmp[3] + tmp[7];
      double b1r = tmp[2] - tmp[6], b1i = tmp[3] - tmp[7];
      /* (-i*s) * b1 = (s*b1i, -s*b1r) */
      y[2 * k] = a0r + b0r;
      y[2 * k + 1] = a0i + b0i;
      y[2 * (m + k)] = a1r + s * b1i;
      y[2 * (m + k) + 1] = a1i - s * b1r;
      y[2 * (2 * m + k)] = a0r - b0r;
      y[2 * (2 * m + k) + 1] = a0i - b0i;
      y[2 * (3 * m + k)] = a1r - s * b1i;
      y[2 * (3 * m + k) + 1] = a1i + s * b1r;
    } else {
      long long pstep = N / p;
      for (q = 0; q < p; q++) {
        double sr = 0.0, si = 0.0;
        for (r = 0; r < p; r++) {
          long long e = ((long long)r * q % p) * pstep;
          double wr = W[2 * e], wi = sign < 0 ? W[2 * e + 1] : -W[2 * e + 1];
          sr += tmp[2 * r] * wr - tmp[2 * r + 1] * wi;
          si += tmp[2 * r] * wi + tmp[2 * r + 1] * wr;
        }
        y[2 * ((long long)q * m + k)] = sr;
        y[2 * ((long long)q * m + k) + 1] = si;
      }
    }
  }
}

/* small cache of plans keyed by length; plans are never freed (test lib) */
#define NPLAN 64
static fplan g_plans[NPLAN];
static int g_nplans = 0;

static const fplan *get_plan(int n) {
  const fplan *res = NULL;
  int i;
#ifdef _OPENMP
#pragma omp critical(orc_plan_cache)
#endif
  {
    for (i = 0; i < g_nplans; i++)
      if (g_plans[i].n == n) { res = &g_plans[i]; break; }
    if (!res) {
      if (g_nplans == NPLAN) {
        /* recycle the oldest slot (only used by tests with many sizes) */
        free(g_plans[0].w);
        memmove(&g_plans[0], &g_plans[1], sizeof(fplan) * (NPLAN - 1));
        g_nplans--;
      }
      fplan_init(&g_plans[g_nplans], n);
      res = &g_plans[g_nplans];
      g_nplans++;
    }
  }
  return res;
}

/* out-of-place transform of a strided complex line into contiguous y */
static void fft_line(const fplan *fp, const double *x, long long xs, double *y, int sign,
                     double *tmp) {
  dit(x, xs, y, fp->n, fp->fac, fp->w, fp->n, sign, tmp);
}

static int max_factor(const fplan *fp) {
  int i, m = 2;
  for (i = 0; i < fp->nf; i++) if (fp->fac[i] > m) m = fp->fac[i];
  return m;
}

void orc_fft(double *data, int n, int sign) {
  const fplan *fp = get_plan(n);
  double *y = (double *)malloc(sizeof(double) * 2 * n);
  double *tmp = (double *)malloc(sizeof(double) * 2 * max_factor(fp));
  fft_line(fp, data, 1, y, sign, tmp);
  memcpy(data, y, sizeof(double) * 2 * n);
  free(y);
  free(tmp);
}

void orc_dft_naive(const double *in, double *out, int n, int sign) {
  int k, j;
  for (k = 0; k < n; k++) {
    long double sr = 0, si = 0;
    for (j = 0; j < n; j++) {
      double c, s;
      twid((long long)j * k, n, &c, &s);
      if (sign > 0) s = -s;
      sr += (long double)in[2 * j] * c - (long double)in[2 * j + 1] * s;
      si += (long double)in[2 * j] * s + (long double)in[2 * j + 1] * c;
    }
    out[2 * k] = (double)sr;
    out[2 * k + 1] = (double)si;
  }
}

void orc_r2c(const double *in, double *out, int n) {
  double *c = (double *)malloc(sizeof(double) * 2 * n);
  int j;
  for (j = 0; j < n; j++) { c[2 * j] = in[j]; c[2 * j + 1] = 0.0; }
  orc_fft(c, n, -1);
  memcpy(out, c, sizeof(double) * 2 * (n / 2 + 1));
  free(c);
}

/* FFTW c2r semantics: Hermitian extension; imaginary parts of k=0 and (n even)
 * k=n/2 do not contribute. */
static void c2r_line(const fplan *fp, const double *in, long long is, double *out, long long os,
                     double *c, double *y, double *tmp) {
  int n = fp->n, k, h = n / 2;
  c[0] = in[0]; c[1] = 0.0;
  for (k = 1; k <= h; k++) {
    c[2 * k] = in[2 * k * is];
    c[2 * k + 1] = in[2 * k * is + 1];
  }
  if (n % 2 == 0) c[2 * h + 1] = 0.0;
  for (k = 1; k < n - h; k++) { /* k' = n-k > h */
    c[2 * (n - k)] = c[2 * k];
    c[2 * (n - k) + 1] = -c[2 * k + 1];
  }
  fft_line(fp, c, 1, y, +1, tmp);
  for (k = 0; k < n; k++) out[k * os] = y[2 * k];
}

void orc_c2r(const double *in, double *out, int n) {
  const fplan *fp = get_plan(n);
  double *c = (double *)malloc(sizeof(double) * 2 * n);
  double *y = (double *)malloc(sizeof(double) * 2 * n);
  double *tmp = (double *)malloc(sizeof(double) * 2 * max_factor(fp));
  c2r_line(fp, in, 1, out, 1, c, y, tmp);
  free(c); free(y); free(tmp);
}

/* REDFT00 / RODFT00 by definition (real data) */
void orc_dct1_naive(const double *in, double *out, int n) {
  int k, j;
  for (k = 0; k < n; k++) {
    long double s = (long double)in[0] + ((k & 1) ? -1.0L : 1.0L) * in[n - 1];
    for (j = 1; j <= n - 2; j++) {
      double c, sn;
      twid((long long)j * k, 2LL * (n - 1), &c, &sn);
      s += 2.0L * in[j] * c;
    }
    out[k] = (double)s;
  }
}

void orc_dst1_naive(const double *in, double *out, int n) {
  int k, j;
  for (k = 0; k < n; k++) {
    long double s = 0;
    for (j = 0; j < n; j++) {
      double c, sn;
      twid((long long)(j + 1) * (k + 1), 2LL * (n + 1), &c, &sn);
      s += 2.0L * in[j] * (-sn); /* sin(pi (j+1)(k+1)/(n+1)) */
    }
    out[k] = (double)s;
  }
}

/* DCT-I of a complex column (stride cstride complex units) via the DFT of the
 * even extension of length 2(n-1): the DFT is real-linear, so the real and
 * imaginary parts are transformed independently as the reference does with two
 * r2r executes (fft_exec.F90:646-701). */
static void dct1_col(double *d, int n, long long cs, double *e, double *y, double *tmp) {
  int m = 2 * (n - 1), j;
  const fplan *fp;
  if (n == 1) { return; } /* REDFT00 undefined for n=1; leave data as is */
  fp = get_plan(m);
  for (j = 0; j < n; j++) { e[2 * j] = d[2 * j * cs]; e[2 * j + 1] = d[2 * j * cs + 1]; }
  for (j = 1; j < n - 1; j++) { e[2 * (m - j)] = e[2 * j]; e[2 * (m - j) + 1] = e[2 * j + 1]; }
  fft_line(fp, e, 1, y, -1, tmp);
  for (j = 0; j < n; j++) { d[2 * j * cs] = y[2 * j]; d[2 * j * cs + 1] = y[2 * j + 1]; }
}

/* DST-I via the odd extension of length 2(n+1): Y_k = i * O_{k+1}. */
static void dst1_col(double *d, int n, long long cs, double *e, double *y, double *tmp) {
  int m = 2 * (n + 1), j;
  const fplan *fp = get_plan(m);
  e[0] = 0; e[1] = 0;
  e[2 * (n + 1)] = 0; e[2 * (n + 1) + 1] = 0;
  for (j = 1; j <= n; j++) {
    double re = d[2 * (j - 1) * cs], im = d[2 * (j - 1) * cs + 1];
    e[2 * j] = re; e[2 * j + 1] = im;
    e[2 * (m - j)] = -re; e[2 * (m - j) + 1] = -im;
  }
  fft_line(fp, e, 1, y, -1, tmp);
  for (j = 0; j < n; j++) {
    double orr = y[2 * (j + 1)], oi = y[2 * (j + 1) + 1];
    d[2 * j * cs] = -oi;     /* i * (orr + i oi) = -oi + i orr */
    d[2 * j * cs + 1] = orr;
  }
}

void orc_dct1(double *data, int n, int cstride) {
  int m = 2 * (n - 1) > 2 ? 2 * (n - 1) : 2;
  double *e = (double *)malloc(sizeof(double) * 2 * m);
  double *y = (double *)malloc(sizeof(double) * 2 * m);
  double *tmp = (double *)malloc(sizeof(double) * 2 * m);
  dct1_col(data, n, cstride, e, y, tmp);
  free(e); free(y); free(tmp);
}

void orc_dst1(double *data, int n, int cstride) {
  int m = 2 * (n + 1);
  double *e = (double *)malloc(sizeof(double) * 2 * m);
  double *y = (double *)malloc(sizeof(double) * 2 * m);
  double *tmp = (double *)malloc(sizeof(double) * 2 * m);
  dst1_col(data, n, cstride, e, y, tmp);
  free(e); free(y); free(tmp);
}

/* ------------------------------------------------------------------ */
/* batched line transforms over 3D arrays                              */
/* ------------------------------------------------------------------ */

/* r2c along x for all rows: in (nx, nrows) real -> out (nxhp, nrows)
 * (build/fft_exec.F90:495-540, plan init_plan.F90:105-106) */
static void batch_r2c(const double *in, double *out, int nx, long long nrows) {
  const fplan *fp = get_plan(nx);
  int nxhp = nx / 2 + 1;
  int mf = max_factor(fp);
#ifdef _OPENMP
#pragma omp parallel
#endif
  {
    double *c = (double *)malloc(sizeof(double) * 2 * nx);
    double *y = (double *)malloc(sizeof(double) * 2 * nx);
    double *tmp = (double *)malloc(sizeof(double) * 2 * mf);
    long long r;
    int j;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (r = 0; r < nrows; r++) {
      for (j = 0; j < nx; j++) { c[2 * j] = in[r * nx + j]; c[2 * j + 1] = 0.0; }
      fft_line(fp, c, 1, y, -1, tmp);
      memcpy(out + 2 * r * nxhp, y, sizeof(double) * 2 * nxhp);
    }
    free(c); free(y); free(tmp);
  }
}

/* c2r along x: in (nxhp, nrows) complex -> out (nx, nrows) real
 * (build/fft_exec.F90:298-342) */
static void batch_c2r(const double *in, double *out, int nx, long long nrows) {
  const fplan *fp = get_plan(nx);
  int nxhp = nx / 2 + 1;
  int mf = max_factor(fp);
#ifdef _OPENMP
#pragma omp parallel
#endif
  {
    double *c = (double *)malloc(sizeof(double) * 2 * nx);
    double *y = (double *)malloc(sizeof(double) * 2 * nx);
    double *tmp = (double *)malloc(sizeof(double) * 2 * mf);
    long long r;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (r = 0; r < nrows; r++) c2r_line(fp, in + 2 * r * nxhp, 1, out + r * nx, 1, c, y, tmp);
    free(c); free(y); free(tmp);
  }
}

/* complex transform along the middle axis of A(n1, n, n3) (complex units),
 * in place.  kind: 'f' c2c forward, 'b' c2c backward, 'c' DCT-I, 's' DST-I. */
static void batch_mid(double *A, long long n1, int n, long long n3, char kind) {
  int m = n;
  const fplan *fp;
  int mf;
  if (kind == 'c') m = 2 * (n - 1);
  if (kind == 's') m = 2 * (n + 1);
  if (m < 1) m = 1;
  fp = get_plan(m);
  mf = max_factor(fp);
  if (mf < m) mf = m;
#ifdef _OPENMP
#pragma omp parallel
#endif
  {
    double *c = (double *)malloc(sizeof(double) * 2 * (m + 2));
    double *y = (double *)malloc(sizeof(double) * 2 * (m + 2));
    double *tmp = (double *)malloc(sizeof(double) * 2 * (mf + 2));
    long long l, total = n1 * n3;
    int j;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (l = 0; l < total; l++) {
      long long i1 = l % n1, i3 = l / n1;
      double *base = A + 2 * (i1 + i3 * n1 * (long long)n);
      if (kind == 'f' || kind == 'b') {
        fft_line(fp, base, n1, y, kind == 'f' ? -1 : +1, tmp);
        for (j = 0; j < n; j++) { base[2 * j * n1] = y[2 * j]; base[2 * j * n1 + 1] = y[2 * j + 1]; }
      } else if (kind == 'c') {
        dct1_col(base, n, n1, c, y, tmp);
      } else if (kind == 's') {
        dst1_col(base, n, n1, c, y, tmp);
      }
    }
    free(c); free(y); free(tmp);
  }
}

/* ------------------------------------------------------------------ */
/* 3D forward (build/ftran.F90:489-780), 1x1 grid                      */
/* ------------------------------------------------------------------ */
static int z_kind(char c, int fwd, char *kind) {
  if (c == 't' || c == 'f') { *kind = fwd ? 'f' : 'b'; return 0; }
  if (c == 'c') { *kind = 'c'; return 0; }
  if (c == 's') { *kind = 's'; return 0; }
  if (c == 'n' || c == '0') { *kind = 0; return 0; }
  return -1; /* unknown: the reference prints and calls MPI_Abort */
}

int orc_ftran_r2c(const orc_plan *p, const double *in, double *out, const char *op) {
  int nx = p->nx, ny = p->ny, nz = p->nz;
  int nxhp = p->nxhp, nxhpc = p->nxhpc, nyc = p->nyc, nzc = p->nzc;
  int nycph = p->nycph, nzcph = p->nzcph;
  int dny = ny - nyc, dnz = nz - nzc;
  char kind;
  double *buf2, *buf, *zb;
  long long x, y, z;
  if (p->p1 != 1 || p->p2 != 1) return -2;
  if (z_kind(op[2], 1, &kind)) return -1;

  /* X r2c: IN (nx, ny, nz) -> buf2 (nxhp, ny, nz)   ftran.F90:525-534 */
  buf2 = (double *)malloc(sizeof(double) * 2 * (size_t)nxhp * ny * nz);
  batch_r2c(in, buf2, nx, (long long)ny * nz);
  /* seg_copy_x keeps x = 1..nxhpc               ftran.F90:554, module.F90:689-706 */
  buf = (double *)malloc(sizeof(double) * 2 * (size_t)nxhpc * ny * nz);
  for (z = 0; z < nz; z++)
    for (y = 0; y < ny; y++)
      memcpy(buf + 2 * (nxhpc * (y + ny * z)), buf2 + 2 * (nxhp * (y + ny * z)),
             sizeof(double) * 2 * nxhpc);
  free(buf2);
  /* Y c2c forward per z-plane                   ftran.F90:576-584 */
  batch_mid(buf, nxhpc, ny, nz, 'f');
  /* seg_copy_y: keep y <= nycph, then y+dny     ftran.F90:736-738 / 703-704 */
  zb = (double *)malloc(sizeof(double) * 2 * (size_t)nxhpc * nyc * nz);
  for (z = 0; z < nz; z++)
    for (y = 0; y < nyc; y++) {
      long long sy = y < nycph ? y : y + dny;
      memcpy(zb + 2 * (nxhpc * (y + (long long)nyc * z)), buf + 2 * (nxhpc * (sy + (long long)ny * z)),
             sizeof(double) * 2 * nxhpc);
    }
  free(buf);
  /* Z op on (nxhpc, nyc, nz)                    ftran.F90:740-763 */
  if (kind) batch_mid(zb, (long long)nxhpc * nyc, nz, 1, kind);
  /* seg_copy_z: keep z <= nzcph, then z+dnz (identity when dnz == 0) and
   * store in the (X,Y,Z) layout, or (Z,Y,X) under STRIDE1
   * (reorder.F90:359-488 uses the same split points). */
  for (z = 0; z < nzc; z++) {
    long long sz = z < nzcph ? z : z + dnz;
    for (y = 0; y < nyc; y++)
      for (x = 0; x < nxhpc; x++) {
        const double *s = zb + 2 * (x + nxhpc * (y + (long long)nyc * sz));
        double *d = p->stride1 ? out + 2 * (z + nzc * (y + (long long)nyc * x))
                               : out + 2 * (x + nxhpc * (y + (long long)nyc * z));
        d[0] = s[0]; d[1] = s[1];
      }
  }
  free(zb);
  return 0;
}

/* ------------------------------------------------------------------ */
/* 3D backward (build/btran.F90:396-679), 1x1 grid                     */
/* ------------------------------------------------------------------ */
int orc_btran_c2r(const orc_plan *p, double *in, double *out, const char *op) {
  int nx = p->nx, ny = p->ny, nz = p->nz;
  int nxhp = p->nxhp, nxhpc = p->nxhpc, nyc = p->nyc, nzc = p->nzc;
  int nycph = p->nycph, nzcph = p->nzcph;
  int dny = ny - nyc, dnz = nz - nzc;
  char kind;
  double *zb, *buf, *buf1;
  long long x, y, z;
  if (p->p1 != 1 || p->p2 != 1) return -2;
  if (z_kind(op[0], 0, &kind)) return -1;

  zb = (double *)malloc(sizeof(double) * 2 * (size_t)nxhpc * nyc * nz);
  if (!p->stride1 && p->ow && nz == nzc) {
    /* Z op in place on the caller's array, then ar_copy   btran.F90:525-550 */
    if (kind) batch_mid(in, (long long)nxhpc * nyc, nz, 1, kind);
    memcpy(zb, in, sizeof(double) * 2 * (size_t)nxhpc * nyc * nz);
  } else {
    /* seg_copy_z + seg_zero_z, then Z op           btran.F90:556-582
     * non-STRIDE1: keep z < nzcph and z >= nz-nzcph (from z-dnz);
     * STRIDE1 (reorder.F90:208-245): keep z < nzcph and z >= nzcph+dnz. */
    for (z = 0; z < nz; z++) {
      long long sz;
      int keep;
      if (z < nzcph) { keep = 1; sz = z; }
      else if (!p->stride1 && z >= nz - nzcph) { keep = 1; sz = z - dnz; }
      else if (p->stride1 && z >= nzcph + dnz) { keep = 1; sz = z - dnz; }
      else { keep = 0; sz = 0; }
      for (y = 0; y < nyc; y++)
        for (x = 0; x < nxhpc; x++) {
          double *d = zb + 2 * (x + nxhpc * (y + (long long)nyc * z));
          if (keep) {
            const double *s = p->stride1 ? in + 2 * (sz + nzc * (y + (long long)nyc * x))
                                         : in + 2 * (x + nxhpc * (y + (long long)nyc * sz));
            d[0] = s[0]; d[1] = s[1];
          } else { d[0] = 0; d[1] = 0; }
        }
    }
    if (kind) batch_mid(zb, (long long)nxhpc * nyc, nz, 1, kind);
  }
  /* seg_copy_y + seg_zero_y (btran.F90:584-587; STRIDE1 reorder.F90:208-245):
   * (nxhpc, nyc, nz) -> (nxhpc, ny, nz) */
  buf = (double *)malloc(sizeof(double) * 2 * (size_t)nxhpc * ny * nz);
  for (z = 0; z < nz; z++)
    for (y = 0; y < ny; y++) {
      long long sy;
      int keep;
      if (y < nycph) { keep = 1; sy = y; }
      else if (!p->stride1 && y >= ny - nycph) { keep = 1; sy = y - dny; }
      else if (p->stride1 && y >= nycph + dny) { keep = 1; sy = y - dny; }
      else { keep = 0; sy = 0; }
      double *d = buf + 2 * (nxhpc * (y + (long long)ny * z));
      if (keep) memcpy(d, zb + 2 * (nxhpc * (sy + (long long)nyc * z)), sizeof(double) * 2 * nxhpc);
      else memset(d, 0, sizeof(double) * 2 * nxhpc);
    }
  free(zb);
  /* Y c2c backward                                  btran.F90:615-624 */
  batch_mid(buf, nxhpc, ny, nz, 'b');
  /* seg_copy_x + seg_zero_x (x > nxhpc)              btran.F90:658-659 */
  buf1 = (double *)malloc(sizeof(double) * 2 * (size_t)nxhp * ny * nz);
  for (z = 0; z < nz; z++)
    for (y = 0; y < ny; y++) {
      double *d = buf1 + 2 * (nxhp * (y + (long long)ny * z));
      memcpy(d, buf + 2 * (nxhpc * (y + (long long)ny * z)), sizeof(double) * 2 * nxhpc);
      memset(d + 2 * nxhpc, 0, sizeof(double) * 2 * (nxhp - nxhpc));
    }
  free(buf);
  /* X c2r                                           btran.F90:661-670 */
  batch_c2r(buf1, out, nx, (long long)ny * nz);
  free(buf1);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Chebyshev derivative (build/ftran.F90:383-464)                      */
/* ------------------------------------------------------------------ */
int orc_cheby(const orc_plan *p, const double *in, double *out, double Lz) {
  int nzc = p->nzc, rc;
  long long ncol = (long long)p->nxhpc * p->nyc, c, n, total;
  double scale = 1.0 / ((double)p->nx * (double)p->ny * (double)(nzc - 1));
  double L = 4.0 / Lz;
  rc = orc_ftran_r2c(p, in, out, "ffc");
  if (rc) return rc;
  total = ncol * nzc;
  for (n = 0; n < 2 * total; n++) out[n] *= scale;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (c = 0; c < ncol; c++) {
    /* element k (1-based) of column c */
#define EL(k) (p->stride1 ? out + 2 * ((long long)(k) - 1 + (long long)nzc * c) \
                          : out + 2 * (c + ncol * ((long long)(k) - 1)))
    double oldr, oldi, newr, newi, tr, ti;
    int k;
    double *e;
    e = EL(nzc - 1); oldr = e[0]; oldi = e[1];
    { double *t = EL(nzc);
      e[0] = L * (double)(nzc - 1) * t[0] * 0.5;
      e[1] = L * (double)(nzc - 1) * t[1] * 0.5;
      t[0] = 0.0; t[1] = 0.0; }
    for (k = nzc - 2; k >= 1; k--) {
      double *ek = EL(k), *ek2 = EL(k + 2);
      newr = ek[0]; newi = ek[1];
      ek[0] = L * (double)k * oldr + ek2[0];
      ek[1] = L * (double)k * oldi + ek2[1];
      tr = newr; ti = newi;
      newr = oldr; newi = oldi;
      oldr = tr; oldi = ti;
    }
    e = EL(1); e[0] *= 0.5; e[1] *= 0.5;
#undef EL
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* share mode: pack/unpack maps of the transposes (non-STRIDE1,        */
/* non-USE_EVEN).  Offsets in complex elements.                        */
/* ------------------------------------------------------------------ */
typedef struct splits {
  int *iist, *iien, *iisz, *jist, *jien, *jisz, *jjst, *jjen, *jjsz, *kjst, *kjen, *kjsz;
} splits;

static void splits_make(const orc_plan *p, splits *s) {
  s->iist = malloc(sizeof(int) * p->p1); s->iien = malloc(sizeof(int) * p->p1); s->iisz = malloc(sizeof(int) * p->p1);
  s->jist = malloc(sizeof(int) * p->p1); s->jien = malloc(sizeof(int) * p->p1); s->jisz = malloc(sizeof(int) * p->p1);
  s->jjst = malloc(sizeof(int) * p->p2); s->jjen = malloc(sizeof(int) * p->p2); s->jjsz = malloc(sizeof(int) * p->p2);
  s->kjst = malloc(sizeof(int) * p->p2); s->kjen = malloc(sizeof(int) * p->p2); s->kjsz = malloc(sizeof(int) * p->p2);
  orc_map_data_to_proc(p->nxhpc, p->p1, s->iist, s->iien, s->iisz);
  orc_map_data_to_proc(p->ny, p->p1, s->jist, s->jien, s->jisz);
  orc_map_data_to_proc(p->nyc, p->p2, s->jjst, s->jjen, s->jjsz);
  orc_map_data_to_proc(p->nz, p->p2, s->kjst, s->kjen, s->kjsz);
}
static void splits_free(splits *s) {
  free(s->iist); free(s->iien); free(s->iisz); free(s->jist); free(s->jien); free(s->jisz);
  free(s->jjst); free(s->jjen); free(s->jjsz); free(s->kjst); free(s->kjen); free(s->kjsz);
}

/* setup.F90:480-518, divided by 2*p3dfft_type to count complex elements */
void orc_share_counts(const orc_plan *p, long long *fsnd_st, long long *fsnd_ct,
                      long long *frcv_st, long long *frcv_ct, long long *ksnd_st,
                      long long *ksnd_ct, long long *krcv_st, long long *krcv_ct) {
  splits s;
  int i;
  splits_make(p, &s);
  for (i = 0; i < p->p1; i++) {
    fsnd_st[i] = (long long)(s.iist[i] - 1) * p->jisize * p->kjsize;
    fsnd_ct[i] = (long long)s.iisz[i] * p->jisize * p->kjsize;
    frcv_st[i] = (long long)(s.jist[i] - 1) * p->iisize * p->kjsize;
    frcv_ct[i] = (long long)s.jisz[i] * p->iisize * p->kjsize;
  }
  for (i = 0; i < p->p2; i++) {
    ksnd_st[i] = (long long)(s.jjst[i] - 1) * p->iisize * p->kjsize;
    ksnd_ct[i] = (long long)p->iisize * p->kjsize * s.jjsz[i];
    krcv_st[i] = (long long)(s.kjst[i] - 1) * p->iisize * p->jjsize;
    krcv_ct[i] = (long long)p->iisize * p->jjsize * s.kjsz[i];
  }
  splits_free(&s);
}

#define CP(dst, di, src, si) do { (dst)[2*(di)] = (src)[2*(si)]; (dst)[2*(di)+1] = (src)[2*(si)+1]; } while (0)

/* fcomm1.F90:239-253 */
void orc_fcomm1_pack(const orc_plan *p, const double *buf2, double *send) {
  splits s;
  int i, x, y, z;
  splits_make(p, &s);
  for (i = 0; i < p->p1; i++) {
    long long pos = (long long)(s.iist[i] - 1) * p->jisize * p->kjsize;
    for (z = 1; z <= p->kjsize; z++)
      for (y = 1; y <= p->jisize; y++)
        for (x = s.iist[i]; x <= s.iien[i]; x++) {
          long long si = (x - 1) + (long long)p->nxhp * ((y - 1) + (long long)p->jisize * (z - 1));
          CP(send, pos, buf2, si);
          pos++;
        }
  }
  splits_free(&s);
}

/* fcomm1.F90:284-320 (non-STRIDE1) */
void orc_fcomm1_unpack(const orc_plan *p, const double *recv, double *dest) {
  splits s;
  int i, x, y, z;
  splits_make(p, &s);
  for (i = 0; i < p->p1; i++)
    for (z = 1; z <= p->kjsize; z++) {
      long long pos = (long long)(s.jist[i] - 1) * p->iisize * p->kjsize +
                      (long long)(z - 1) * p->iisize * s.jisz[i];
      for (y = s.jist[i]; y <= s.jien[i]; y++)
        for (x = 1; x <= p->iisize; x++) {
          long long di = (x - 1) + (long long)p->iisize * ((y - 1) + (long long)p->ny * (z - 1));
          CP(dest, di, recv, pos);
          pos++;
        }
    }
  splits_free(&s);
}

/* fcomm2.F90:321-386 (pack_fcomm2) */
void orc_fcomm2_pack(const orc_plan *p, const double *src, double *send) {
  splits s;
  int i, x, y, z, dny = p->ny - p->nyc, nyhc = p->nyhc;
  splits_make(p, &s);
  for (i = 0; i < p->p2; i++) {
    long long pos0 = (long long)(s.jjst[i] - 1) * p->iisize * p->kjsize;
    for (z = 1; z <= p->kjsize; z++) {
      long long pos = pos0 + (long long)(z - 1) * s.jjsz[i] * p->iisize;
#define PK(yy) for (x = 1; x <= p->iisize; x++) { \
        long long si = (x - 1) + (long long)p->iisize * (((yy) - 1) + (long long)p->ny * (z - 1)); \
        CP(send, pos, src, si); pos++; }
      if (s.jjen[i] <= nyhc) {
        for (y = s.jjst[i]; y <= s.jjen[i]; y++) PK(y)
      } else if (s.jjst[i] >= nyhc + 1) {
        for (y = s.jjst[i] + dny; y <= s.jjen[i] + dny; y++) PK(y)
      } else {
        for (y = s.jjst[i]; y <= nyhc; y++) PK(y)
        for (y = p->ny - nyhc + 1; y <= s.jjen[i] + dny; y++) PK(y)
      }
#undef PK
    }
  }
  splits_free(&s);
}

/* bcomm1.F90:309-380 (unpack_bcomm1); the centre zero-fill runs after each
 * peer's unpack in the reference, so its final effect is a zero-fill of
 * y = nyhc+1..ny-nyhc after all unpacks. */
void orc_bcomm1_unpack(const orc_plan *p, const double *recv, double *dest) {
  splits s;
  int i, x, y, z, dny = p->ny - p->nyc, nyhc = p->nyhc;
  splits_make(p, &s);
  for (i = 0; i < p->p2; i++) {
    long long pos0 = (long long)(s.jjst[i] - 1) * p->iisize * p->kjsize;
    for (z = 1; z <= p->kjsize; z++) {
      long long pos = pos0 + (long long)(z - 1) * s.jjsz[i] * p->iisize;
#define UP(yy) for (x = 1; x <= p->iisize; x++) { \
        long long di = (x - 1) + (long long)p->iisize * (((yy) - 1) + (long long)p->ny * (z - 1)); \
        CP(dest, di, recv, pos); pos++; }
      if (s.jjen[i] <= nyhc) {
        for (y = s.jjst[i]; y <= s.jjen[i]; y++) UP(y)
      } else if (s.jjst[i] >= nyhc + 1) {
        for (y = s.jjst[i] + dny; y <= s.jjen[i] + dny; y++) UP(y)
      } else {
        for (y = s.jjst[i]; y <= nyhc; y++) UP(y)
        for (y = p->ny - nyhc + 1; y <= s.jjen[i] + dny; y++) UP(y)
      }
#undef UP
    }
    for (z = 1; z <= p->kjsize; z++)
      for (y = nyhc + 1; y <= p->ny - nyhc; y++)
        for (x = 1; x <= p->iisize; x++) {
          long long di = (x - 1) + (long long)p->iisize * ((y - 1) + (long long)p->ny * (z - 1));
          dest[2 * di] = 0.0; dest[2 * di + 1] = 0.0;
        }
  }
  splits_free(&s);
}

/* bcomm2.F90:233-273 (pack, non-STRIDE1) */
void orc_bcomm2_pack(const orc_plan *p, const double *src, double *send) {
  splits s;
  int i, x, y, z;
  splits_make(p, &s);
  for (i = 0; i < p->p1; i++)
    for (z = 1; z <= p->kjsize; z++) {
      long long pos = (long long)(s.jist[i] - 1) * p->iisize * p->kjsize +
                      (long long)(z - 1) * p->iisize * s.jisz[i];
      for (y = s.jist[i]; y <= s.jien[i]; y++)
        for (x = 1; x <= p->iisize; x++) {
          long long si = (x - 1) + (long long)p->iisize * ((y - 1) + (long long)p->ny * (z - 1));
          CP(send, pos, src, si);
          pos++;
        }
    }
  splits_free(&s);
}

/* bcomm2.F90:290-313 (unpack) */
void orc_bcomm2_unpack(const orc_plan *p, const double *recv, double *dest) {
  splits s;
  int i, x, y, z;
  splits_make(p, &s);
  for (i = 0; i < p->p1; i++) {
    long long pos0 = (long long)(s.iist[i] - 1) * p->jisize * p->kjsize;
    for (z = 1; z <= p->kjsize; z++) {
      long long pos = pos0;
      for (y = 1; y <= p->jisize; y++)
        for (x = s.iist[i]; x <= s.iien[i]; x++) {
          long long di = (x - 1) + (long long)p->nxhp * ((y - 1) + (long long)p->jisize * (z - 1));
          CP(dest, di, recv, pos);
          pos++;
        }
      pos0 += (long long)s.iisz[i] * p->jisize;
    }
  }
  for (z = 1; z <= p->kjsize; z++)
    for (y = 1; y <= p->jisize; y++)
      for (x = p->nxhpc + 1; x <= p->nxhp; x++) {
        long long di = (x - 1) + (long long)p->nxhp * ((y - 1) + (long long)p->jisize * (z - 1));
        dest[2 * di] = 0.0; dest[2 * di + 1] = 0.0;
      }
  splits_free(&s);
}

/* ------------------------------------------------------------------ */
/* The reference's distributed CPU path on a p1 x p2 grid, with the ranks
 * as OpenMP threads of one process (each rank single-threaded, as one MPI
 * rank is): per rank X r2c of its X pencil (ftran.F90:525-551), fcomm1
 * pack, the row exchange (the MPI_Alltoallv of fcomm1.F90:255-262 as
 * copies between the ranks' buffers), fcomm1 unpack, Y c2c, fcomm2 pack,
 * the column exchange (fcomm2.F90:313), Z c2c on the receive buffer (the Z
 * pencil); backward: Z c2c, the bcomm1 exchange (bcomm1.F90:295-296) and
 * unpack, Y c2c, bcomm2 pack, exchange and unpack, X c2r
 * (btran.F90:396-679).  Ops 'fft' / 'tff', no cuts.  in: global real
 * (nx, ny, nz); spec (optional): the global spectrum (nxhp, ny, nz)
 * assembled from the Z pencils after the last forward; out: global real
 * after the last backward.  t_fwd / t_bwd (reps each): wall seconds of each
 * forward / backward, from barrier to barrier.  Returns 0, or -1 on invalid
 * input.                                                               */
/* ------------------------------------------------------------------ */
static void ranks_line(const fplan *fp, double *base, long long stride, int sign, double *y, double *tmp) {
  int j;
  fft_line(fp, base, stride, y, sign, tmp);
  for (j = 0; j < fp->n; j++) { base[2 * j * stride] = y[2 * j]; base[2 * j * stride + 1] = y[2 * j + 1]; }
}

/* wall clock of the rank threads (0 without OpenMP: no stage times) */
static double ranks_now(void) {
#ifdef _OPENMP
  return omp_get_wtime();
#else
  return 0.0;
#endif
}

int orc_ranks_fwd_bwd(int nx, int ny, int nz, int p1, int p2, const double *in, double *spec, double *out,
                      int reps, double *t_fwd, double *t_bwd, double *slots) {
  int P = p1 * p2, ok = 1;
  orc_plan *pl;
  double **xr, **b2, **s1, **r1, **yp, **s2, **zp;
  const fplan *fx, *fy, *fz;
  int mf;
  if (nx <= 0 || ny <= 0 || nz <= 0 || p1 <= 0 || p2 <= 0 || reps <= 0) return -1;
  /* every rank needs a non-empty pencil of each split (nx/2+1 and ny over
   * p1, ny and nz over p2); orc_setup divides by those sizes */
  if (p1 > nx / 2 + 1 || p1 > ny || p2 > ny || p2 > nz) return -1;
  pl = (orc_plan *)calloc(P, sizeof(orc_plan));
  {
    int dims[2] = {p1, p2}, r;
    for (r = 0; r < P; r++)
      if (orc_setup(&pl[r], dims, r % p1, r / p1, nx, ny, nz, nx, ny, nz, 1, 0)) ok = 0;
    for (r = 0; r < P; r++)
      if (pl[r].iisize <= 0 || pl[r].jisize <= 0 || pl[r].jjsize <= 0 || pl[r].kjsize <= 0) ok = 0;
  }
  if (!ok) { free(pl); return -1; }
  fx = get_plan(nx); fy = get_plan(ny); fz = get_plan(nz);
  mf = max_factor(fx);
  if (max_factor(fy) > mf) mf = max_factor(fy);
  if (max_factor(fz) > mf) mf = max_factor(fz);
  xr = (double **)calloc(P, sizeof(double *)); b2 = (double **)calloc(P, sizeof(double *));
  s1 = (double **)calloc(P, sizeof(double *)); r1 = (double **)calloc(P, sizeof(double *));
  yp = (double **)calloc(P, sizeof(double *)); s2 = (double **)calloc(P, sizeof(double *));
  zp = (double **)calloc(P, sizeof(double *));
#ifdef _OPENMP
#pragma omp parallel num_threads(P)
#endif
  {
#ifdef _OPENMP
    const int r = omp_get_thread_num();
#else
    const int r = 0;
#endif
    const orc_plan *p = &pl[r];
    const int ip = p->ipid, jp = p->jpid;
    const long long nrows = (long long)p->jisize * p->kjsize;
    const long long nb2 = (long long)p->nxhp * nrows;
    const long long nyp = (long long)p->iisize * ny * p->kjsize;
    const long long nzp = (long long)p->iisize * p->jjsize * nz;
    long long *fs_st = malloc(sizeof(long long) * p1), *fs_ct = malloc(sizeof(long long) * p1);
    long long *fr_st = malloc(sizeof(long long) * p1), *fr_ct = malloc(sizeof(long long) * p1);
    long long *ks_st = malloc(sizeof(long long) * p2), *ks_ct = malloc(sizeof(long long) * p2);
    long long *kr_st = malloc(sizeof(long long) * p2), *kr_ct = malloc(sizeof(long long) * p2);
    double *c = malloc(sizeof(double) * 2 * (nx + ny + nz));
    double *y = malloc(sizeof(double) * 2 * (nx + ny + nz));
    double *tmp = malloc(sizeof(double) * 2 * (mf + 2));
    long long l, q;
    int it, i, j;
    orc_share_counts(p, fs_st, fs_ct, fr_st, fr_ct, ks_st, ks_ct, kr_st, kr_ct);
    xr[r] = malloc(sizeof(double) * nx * nrows);
    b2[r] = malloc(sizeof(double) * 2 * nb2);
    s1[r] = malloc(sizeof(double) * 2 * (long long)p->nxhpc * nrows);
    r1[r] = malloc(sizeof(double) * 2 * nyp);
    yp[r] = malloc(sizeof(double) * 2 * nyp);
    s2[r] = malloc(sizeof(double) * 2 * nyp);
    zp[r] = malloc(sizeof(double) * 2 * nzp);
    /* this rank's X pencil of the global field */
    for (l = 0; l < nrows; l++) {
      const long long yy = p->jistart - 1 + l % p->jisize, zz = p->kjstart - 1 + l / p->jisize;
      memcpy(xr[r] + l * nx, in + (yy + zz * (long long)ny) * nx, sizeof(double) * nx);
    }
#ifdef _OPENMP
#pragma omp barrier
#endif
    for (it = 0; it < reps; it++) {
      double t0 = 0, ts;
      /* this rank's 12 timer slots of this repetition, as the reference
       * fills timers(1..12) (module.F90:106): 1-4 the fcomm1 / fcomm2 /
       * bcomm1 / bcomm2 exchanges, 5 X r2c, 6 fcomm1 pack and unpack, 7 Y
       * forward, 8 Z forward, 9 Z backward and the bcomm1 unpack, 10 Y
       * backward, 11 bcomm2 pack and unpack, 12 X c2r (ftran.F90:528-532,
       * 579-584, 600-680; btran.F90:519-592, 608-624, 639-642; fcomm1.F90:
       * 236-322; bcomm2.F90:228-315); the fcomm2 pack is untimed there
       * (fcomm2.F90:265-271) and here */
      double *sl = slots ? slots + ((long long)r * reps + it) * 12 : NULL;
      double sink[12];
      if (!sl) sl = sink;
      for (i = 0; i < 12; i++) sl[i] = 0.0;
#ifdef _OPENMP
#pragma omp barrier
      if (r == 0) t0 = omp_get_wtime();
#endif
      /* X r2c */
      ts = ranks_now();
      for (l = 0; l < nrows; l++) {
        for (j = 0; j < nx; j++) { c[2 * j] = xr[r][l * nx + j]; c[2 * j + 1] = 0.0; }
        fft_line(fx, c, 1, y, -1, tmp);
        memcpy(b2[r] + 2 * l * p->nxhp, y, sizeof(double) * 2 * p->nxhp);
      }
      sl[4] += ranks_now() - ts; ts = ranks_now();
      orc_fcomm1_pack(p, b2[r], s1[r]);
      sl[5] += ranks_now() - ts; ts = ranks_now();
#ifdef _OPENMP
#pragma omp barrier
#endif
      /* row exchange: the block from row peer i (same jpid) */
      for (i = 0; i < p1; i++) {
        const orc_plan *pp = &pl[jp * p1 + i];
        const long long off = (long long)(p->iistart - 1) * pp->jisize * pp->kjsize;
        memcpy(r1[r] + 2 * fr_st[i], s1[jp * p1 + i] + 2 * off, sizeof(double) * 2 * fr_ct[i]);
      }
      sl[0] += ranks_now() - ts; ts = ranks_now();
      orc_fcomm1_unpack(p, r1[r], yp[r]);
      sl[5] += ranks_now() - ts; ts = ranks_now();
      /* Y c2c */
      for (q = 0; q < (long long)p->iisize * p->kjsize; q++) {
        const long long xx = q % p->iisize, zz = q / p->iisize;
        ranks_line(fy, yp[r] + 2 * (xx + zz * (long long)p->iisize * ny), p->iisize, -1, y, tmp);
      }
      sl[6] += ranks_now() - ts;
      orc_fcomm2_pack(p, yp[r], s2[r]);
      ts = ranks_now();
#ifdef _OPENMP
#pragma omp barrier
#endif
      /* column exchange: the block from column peer j (same ipid) lands at
       * its z range of the Z pencil */
      for (j = 0; j < p2; j++) {
        const orc_plan *pp = &pl[j * p1 + ip];
        const long long off = (long long)(p->jjstart - 1) * pp->iisize * pp->kjsize;
        memcpy(zp[r] + 2 * kr_st[j], s2[j * p1 + ip] + 2 * off, sizeof(double) * 2 * kr_ct[j]);
      }
      sl[1] += ranks_now() - ts; ts = ranks_now();
      /* Z c2c */
      for (q = 0; q < (long long)p->iisize * p->jjsize; q++)
        ranks_line(fz, zp[r] + 2 * q, (long long)p->iisize * p->jjsize, -1, y, tmp);
      sl[7] += ranks_now() - ts;
#ifdef _OPENMP
#pragma omp barrier
      if (r == 0) { t_fwd[it] = omp_get_wtime() - t0; t0 = omp_get_wtime(); }
#endif
      if (spec && it == reps - 1) {
        for (q = 0; q < (long long)p->jjsize * nz; q++) {
          const long long yy = p->jjstart - 1 + q % p->jjsize, zz = q / p->jjsize;
          memcpy(spec + 2 * (p->iistart - 1 + p->nxhp * (yy + (long long)ny * zz)), zp[r] + 2 * q * p->iisize,
                 sizeof(double) * 2 * p->iisize);
        }
#ifdef _OPENMP
#pragma omp barrier
        if (r == 0) t0 = omp_get_wtime();
#endif
      }
      /* backward: Z c2c */
      ts = ranks_now();
      for (q = 0; q < (long long)p->iisize * p->jjsize; q++)
        ranks_line(fz, zp[r] + 2 * q, (long long)p->iisize * p->jjsize, +1, y, tmp);
      sl[8] += ranks_now() - ts; ts = ranks_now();
#ifdef _OPENMP
#pragma omp barrier
#endif
      /* bcomm1: from column peer j its slice z in this rank's z range of
       * its Z pencil, into this rank's receive buffer at the fcomm2 send
       * displacement of j */
      for (j = 0; j < p2; j++) {
        const orc_plan *pp = &pl[j * p1 + ip];
        const long long off = (long long)(p->kjstart - 1) * pp->iisize * pp->jjsize;
        memcpy(s2[r] + 2 * ks_st[j], zp[j * p1 + ip] + 2 * off, sizeof(double) * 2 * ks_ct[j]);
      }
      sl[2] += ranks_now() - ts; ts = ranks_now();
      orc_bcomm1_unpack(p, s2[r], yp[r]);
      sl[8] += ranks_now() - ts; ts = ranks_now();
      /* Y c2c */
      for (q = 0; q < (long long)p->iisize * p->kjsize; q++) {
        const long long xx = q % p->iisize, zz = q / p->iisize;
        ranks_line(fy, yp[r] + 2 * (xx + zz * (long long)p->iisize * ny), p->iisize, +1, y, tmp);
      }
      sl[9] += ranks_now() - ts; ts = ranks_now();
      orc_bcomm2_pack(p, yp[r], r1[r]);
      sl[10] += ranks_now() - ts; ts = ranks_now();
#ifdef _OPENMP
#pragma omp barrier
#endif
      /* bcomm2: from row peer i its block for this rank, into the receive
       * buffer at the fcomm1 send displacement of i */
      for (i = 0; i < p1; i++) {
        const orc_plan *pp = &pl[jp * p1 + i];
        const long long off = (long long)(p->jistart - 1) * pp->iisize * pp->kjsize;
        memcpy(s1[r] + 2 * fs_st[i], r1[jp * p1 + i] + 2 * off, sizeof(double) * 2 * fs_ct[i]);
      }
      sl[3] += ranks_now() - ts; ts = ranks_now();
      orc_bcomm2_unpack(p, s1[r], b2[r]);
      sl[10] += ranks_now() - ts; ts = ranks_now();
      /* X c2r */
      for (l = 0; l < nrows; l++) c2r_line(fx, b2[r] + 2 * l * p->nxhp, 1, xr[r] + l * nx, 1, c, y, tmp);
      sl[11] += ranks_now() - ts;
#ifdef _OPENMP
#pragma omp barrier
      if (r == 0) t_bwd[it] = omp_get_wtime() - t0;
#endif
      /* the next forward transforms the input again */
      if (it < reps - 1 || out) {
        if (out && it == reps - 1)
          for (l = 0; l < nrows; l++) {
            const long long yy = p->jistart - 1 + l % p->jisize, zz = p->kjstart - 1 + l / p->jisize;
            memcpy(out + (yy + zz * (long long)ny) * nx, xr[r] + l * nx, sizeof(double) * nx);
          }
        for (l = 0; l < nrows; l++) {
          const long long yy = p->jistart - 1 + l % p->jisize, zz = p->kjstart - 1 + l / p->jisize;
          memcpy(xr[r] + l * nx, in + (yy + zz * (long long)ny) * nx, sizeof(double) * nx);
        }
      }
    }
#ifdef _OPENMP
#pragma omp barrier
#endif
    free(xr[r]); free(b2[r]); free(s1[r]); free(r1[r]); free(yp[r]); free(s2[r]); free(zp[r]);
    free(fs_st); free(fs_ct); free(fr_st); free(fr_ct); free(ks_st); free(ks_ct); free(kr_st); free(kr_ct);
    free(c); free(y); free(tmp);
  }
  free(xr); free(b2); free(s1); free(r1); free(yp); free(s2); free(zp); free(pl);
  return 0;
}
