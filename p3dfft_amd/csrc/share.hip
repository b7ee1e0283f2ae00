// share.hip -- pack/unpack kernels of the row and column transposes for one
// rank of a P1 x P2 processor grid ("share mode").
//
// Each kernel is a gather: one thread per output element computes where the
// reference's pack or unpack loop would have taken that element from, so the
// buffers are bit-identical to the reference's send and receive buffers
// (non-STRIDE1, MPI_Alltoallv layout):
//   k_pack_fcomm1    build/fcomm1.F90:238-253   buf2 (nxhp,jisize,kjsize) -> row send buffer
//   k_unpack_fcomm1  build/fcomm1.F90:284-320   row receive buffer -> (iisize,ny,kjsize)
//   k_pack_fcomm2    build/fcomm2.F90:321-386   (iisize,ny,kjsize) -> column send buffer
//   k_unpack_bcomm1  build/bcomm1.F90:309-380   column receive buffer -> (iisize,ny,kjsize)
//   k_pack_bcomm2    build/bcomm2.F90:233-273   (iisize,ny,kjsize) -> row send buffer
//   k_unpack_bcomm2  build/bcomm2.F90:290-313   row receive buffer -> (nxhp,jisize,kjsize)
// Block starts and sizes per peer (build/setup.F90:480-518, divided by the
// complex element size) come from the MapDataToProc splits.  All indices are
// 0-based; the splits arrive as 0-based starts.
#include "engine.hpp"
#include "share.hpp"

namespace p3amd {

struct SplitTab {
  const int* st;  // 0-based start of each peer's range
  const int* sz;  // size of each peer's range
  int np;
};

// last peer whose block start (st[i] * unit) is <= t; empty peers come first
// (MapDataToProc gives the extra elements to the last ranks), so the last
// match is the non-empty owner.
__device__ __forceinline__ int owner(const int* st, int np, long long t, long long unit) {
  int lo = 0, hi = np - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((long long)st[mid] * unit <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <class T>
__global__ void k_pack_fcomm1(const cpx<T>* __restrict__ buf2, cpx<T>* __restrict__ send, SplitTab ii,
                              int nxhp, int jisize, int kjsize, long long total) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= total) return;
  const long long unit = (long long)jisize * kjsize;
  const int i = owner(ii.st, ii.np, t, unit);
  const long long off = t - (long long)ii.st[i] * unit;
  const int xl = (int)(off % ii.sz[i]);
  const long long r = off / ii.sz[i];
  const int y = (int)(r % jisize);
  const long long z = r / jisize;
  send[t] = buf2[(ii.st[i] + xl) + (long long)nxhp * (y + (long long)jisize * z)];
}

template <class T>
__global__ void k_unpack_fcomm1(const cpx<T>* __restrict__ recv, cpx<T>* __restrict__ dest, SplitTab ji,
                                int iisize, int ny, int kjsize, long long total) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int x = (int)(t % iisize);
  const long long r = t / iisize;
  const int y = (int)(r % ny);
  const long long z = r / ny;
  const int i = owner(ji.st, ji.np, y, 1);
  dest[t] = recv[(long long)ji.st[i] * iisize * kjsize + z * iisize * ji.sz[i] + (long long)(y - ji.st[i]) * iisize + x];
}

// source row (0-based, in the (iisize, ny, kjsize) array) of pruned-Y position
// yh (0-based) in column peer j's block, or -1 where the reference's loops
// leave the position unwritten (the spanning block of an odd nyc)
__device__ __forceinline__ int fcomm2_src_row(int yh, int jst, int jsz, int nyhc, int ny, int dny) {
  const int jen = jst + jsz - 1;  // 0-based inclusive
  if (jen < nyhc) return yh;                 // clearly in the first half
  if (jst >= nyhc) return yh + dny;          // clearly in the second half
  if (yh < nyhc) return yh;                  // spanning: first part
  const int y = ny - nyhc + (yh - nyhc);     // spanning: second part
  return y <= jen + dny ? y : -1;
}

template <class T>
__global__ void k_pack_fcomm2(const cpx<T>* __restrict__ src, cpx<T>* __restrict__ send, SplitTab jj, int iisize,
                              int ny, int kjsize, int nyhc, int dny, long long total) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= total) return;
  const long long unit = (long long)iisize * kjsize;
  const int j = owner(jj.st, jj.np, t, unit);
  const long long off = t - (long long)jj.st[j] * unit;
  const int x = (int)(off % iisize);
  const long long r = off / iisize;
  const int yl = (int)(r % jj.sz[j]);
  const long long z = r / jj.sz[j];
  const int y = fcomm2_src_row(jj.st[j] + yl, jj.st[j], jj.sz[j], nyhc, ny, dny);
  if (y >= 0) send[t] = src[x + (long long)iisize * (y + (long long)ny * z)];
}

template <class T>
__global__ void k_unpack_bcomm1(const cpx<T>* __restrict__ recv, cpx<T>* __restrict__ dest, SplitTab jj,
                                int iisize, int ny, int kjsize, int nyhc, int nyc, int dny, long long total) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int x = (int)(t % iisize);
  const long long r = t / iisize;
  const int y = (int)(r % ny);
  const long long z = r / ny;
  cpx<T> v = mk<T>((T)0, (T)0);
  if (y < nyhc || y >= ny - nyhc) {
    // invert the pack map: which block position lands in row y
    int yh, j;
    if (y < nyhc) {
      yh = y;
      j = owner(jj.st, jj.np, yh, 1);
    } else {
      yh = y - dny;
      j = owner(jj.st, jj.np, yh, 1);
      if (jj.st[j] < nyhc) {  // spanning block: rows follow ny-nyhc+k
        yh = nyhc + (y - (ny - nyhc));
        j = owner(jj.st, jj.np, yh, 1);
      }
    }
    if (yh >= 0 && yh < nyc && fcomm2_src_row(yh, jj.st[j], jj.sz[j], nyhc, ny, dny) == y)
      v = recv[(long long)jj.st[j] * iisize * kjsize + z * (long long)jj.sz[j] * iisize +
               (long long)(yh - jj.st[j]) * iisize + x];
  }
  dest[t] = v;
}

template <class T>
__global__ void k_pack_bcomm2(const cpx<T>* __restrict__ src, cpx<T>* __restrict__ send, SplitTab ji, int iisize,
                              int ny, int kjsize, long long total) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= total) return;
  const long long unit = (long long)iisize * kjsize;
  const int i = owner(ji.st, ji.np, t, unit);
  const long long off = t - (long long)ji.st[i] * unit;
  const int x = (int)(off % iisize);
  const long long r = off / iisize;
  const int yl = (int)(r % ji.sz[i]);
  const long long z = r / ji.sz[i];
  send[t] = src[x + (long long)iisize * ((ji.st[i] + yl) + (long long)ny * z)];
}

template <class T>
__global__ void k_unpack_bcomm2(const cpx<T>* __restrict__ recv, cpx<T>* __restrict__ dest, SplitTab ii, int nxhp,
                                int nxhpc, int jisize, int kjsize, long long total) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int x = (int)(t % nxhp);
  const long long r = t / nxhp;
  const int y = (int)(r % jisize);
  const long long z = r / jisize;
  cpx<T> v = mk<T>((T)0, (T)0);
  if (x < nxhpc) {
    const int i = owner(ii.st, ii.np, x, 1);
    v = recv[(long long)ii.st[i] * jisize * kjsize + z * (long long)ii.sz[i] * jisize + (long long)y * ii.sz[i] +
             (x - ii.st[i])];
  }
  dest[t] = v;
}

// ---------------------------------------------------------------------------
// Table-addressed forms of the six maps: one thread per element of the
// contiguous side, the other side addressed through the per-row tables
// (no peer search, no 64-bit division).  Used when the element counts fit in
// 32 bits; the search forms above serve larger ranks.
// ---------------------------------------------------------------------------

// n / d for 32-bit n by multiply-high and shift (the round-up reciprocal
// method of Granlund and Montgomery, 1994): m = floor(2^32 (2^s - d) / d) + 1
// with s = ceil(log2 d), then n / d = (umulhi(n, m) + n) >> s.
struct Div32 {
  unsigned m, s, d;
};

static Div32 make_div32(unsigned d) {
  unsigned s = 0;
  while ((1ull << s) < d) s++;
  Div32 r;
  r.s = s;
  r.d = d;
  r.m = (unsigned)(((((unsigned long long)1 << s) - d) << 32) / d + 1);
  return r;
}

__device__ __forceinline__ unsigned div32(unsigned n, Div32 q) {
  return (unsigned)(((unsigned long long)__umulhi(n, q.m) + n) >> q.s);
}

template <class T>
__global__ void k_pack_fcomm1_t(const cpx<T>* __restrict__ buf2, cpx<T>* __restrict__ send, const long long* xo,
                                const int* xs, Div32 dx, int nxhp, unsigned total) {
  const unsigned q = blockIdx.x * 256u + threadIdx.x;
  if (q >= total) return;
  const unsigned r = div32(q, dx), x = q - r * dx.d;
  send[xo[x] + (long long)r * xs[x] + x] = buf2[x + (long long)nxhp * r];
}

template <class T>
__global__ void k_unpack_bcomm2_t(const cpx<T>* __restrict__ recv, cpx<T>* __restrict__ dest, const long long* xo,
                                  const int* xs, Div32 dx, int nxhpc, unsigned total) {
  const unsigned q = blockIdx.x * 256u + threadIdx.x;
  if (q >= total) return;
  const unsigned r = div32(q, dx), x = q - r * dx.d;  // dx.d = nxhp
  dest[q] = (int)x < nxhpc ? recv[xo[x] + (long long)r * xs[x] + x] : mk<T>((T)0, (T)0);
}

// (x, y, z) of element q of an (iisize, rows, kjsize) array
__device__ __forceinline__ void split3(unsigned q, Div32 di, Div32 dr, unsigned& x, unsigned& y, unsigned& z) {
  const unsigned r = div32(q, di);
  x = q - r * di.d;
  z = div32(r, dr);
  y = r - z * dr.d;
}

template <class T>
__global__ void k_unpack_fcomm1_t(const cpx<T>* __restrict__ recv, cpx<T>* __restrict__ dest, const long long* ya,
                                  const long long* yb, Div32 di, Div32 dy, unsigned total) {
  const unsigned q = blockIdx.x * 256u + threadIdx.x;
  if (q >= total) return;
  unsigned x, y, z;
  split3(q, di, dy, x, y, z);
  dest[q] = recv[ya[y] + (long long)z * yb[y] + x];
}

template <class T>
__global__ void k_pack_bcomm2_t(const cpx<T>* __restrict__ src, cpx<T>* __restrict__ send, const long long* ya,
                                const long long* yb, Div32 di, Div32 dy, unsigned total) {
  const unsigned q = blockIdx.x * 256u + threadIdx.x;
  if (q >= total) return;
  unsigned x, y, z;
  split3(q, di, dy, x, y, z);
  send[ya[y] + (long long)z * yb[y] + x] = src[q];
}

template <class T>
__global__ void k_pack_fcomm2_t(const cpx<T>* __restrict__ src, cpx<T>* __restrict__ send, const long long* ha,
                                const long long* hb, const int* hy, Div32 di, Div32 dh, int ny, unsigned total) {
  const unsigned q = blockIdx.x * 256u + threadIdx.x;
  if (q >= total) return;
  unsigned x, yh, z;
  split3(q, di, dh, x, yh, z);  // q over (iisize, nyc, kjsize)
  const int y = hy[yh];
  if (y >= 0) send[ha[yh] + (long long)z * hb[yh] + x] = src[x + (long long)di.d * (y + (long long)ny * z)];
}

template <class T>
__global__ void k_unpack_bcomm1_t(const cpx<T>* __restrict__ recv, cpx<T>* __restrict__ dest, const long long* ha,
                                  const long long* hb, const int* yh_of, Div32 di, Div32 dy, unsigned total) {
  const unsigned q = blockIdx.x * 256u + threadIdx.x;
  if (q >= total) return;
  unsigned x, y, z;
  split3(q, di, dy, x, y, z);
  const int yh = yh_of[y];
  dest[q] = yh >= 0 ? recv[ha[yh] + (long long)z * hb[yh] + x] : mk<T>((T)0, (T)0);
}

// host restatements of owner() and fcomm2_src_row() for the tables
static int owner_h(const std::vector<int>& st, long long t) {
  int lo = 0, hi = (int)st.size() - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((long long)st[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

static int fcomm2_src_row_h(int yh, int jst, int jsz, int nyhc, int ny, int dny) {
  const int jen = jst + jsz - 1;
  if (jen < nyhc) return yh;
  if (jst >= nyhc) return yh + dny;
  if (yh < nyhc) return yh;
  const int y = ny - nyhc + (yh - nyhc);
  return y <= jen + dny ? y : -1;
}

template <class V>
static V* upload_vec(const std::vector<V>& v) {
  V* d = nullptr;
  check_hip(hipMalloc(&d, std::max<size_t>(v.size(), 1) * sizeof(V)), "hipMalloc(share table)");
  check_hip(hipMemcpy(d, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice), "hipMemcpy(share table)");
  return d;
}

void share_build_tables(ShareGeom& s, const std::vector<int>& iist, const std::vector<int>& iisz,
                        const std::vector<int>& jist, const std::vector<int>& jisz, const std::vector<int>& jjst,
                        const std::vector<int>& jjsz) {
  share_free_tables(s);
  const long long unit_x = (long long)s.jisize * s.kjsize, unit_y = (long long)s.iisize * s.kjsize;
  std::vector<long long> xo(s.nxhpc);
  std::vector<int> xs(s.nxhpc);
  for (int x = 0; x < s.nxhpc; x++) {
    const int i = owner_h(iist, x);
    xo[x] = (long long)iist[i] * unit_x - iist[i];
    xs[x] = iisz[i];
  }
  std::vector<long long> ya(s.ny), yb(s.ny);
  for (int y = 0; y < s.ny; y++) {
    const int i = owner_h(jist, y);
    ya[y] = (long long)jist[i] * unit_y + (long long)(y - jist[i]) * s.iisize;
    yb[y] = (long long)s.iisize * jisz[i];
  }
  const int nyhc = s.nyc / 2, dny = s.ny - s.nyc;
  std::vector<long long> ha(s.nyc), hb(s.nyc);
  std::vector<int> hy(s.nyc), yh_of(s.ny, -1);
  for (int yh = 0; yh < s.nyc; yh++) {
    const int j = owner_h(jjst, yh);
    ha[yh] = (long long)jjst[j] * unit_y + (long long)(yh - jjst[j]) * s.iisize;
    hb[yh] = (long long)jjsz[j] * s.iisize;
    hy[yh] = fcomm2_src_row_h(yh, jjst[j], jjsz[j], nyhc, s.ny, dny);
  }
  // row y of the unpacked Y pencil is filled from the position yh whose
  // source row is y (the inverse of the pack map); rows no position maps to
  // (the pruned middle, the unwritten end of a spanning block) stay zero
  for (int y = 0; y < s.ny; y++) {
    if (!(y < nyhc || y >= s.ny - nyhc)) continue;
    int yh, j;
    if (y < nyhc) {
      yh = y;
      j = owner_h(jjst, yh);
    } else {
      yh = y - dny;
      j = owner_h(jjst, yh);
      if (jjst[j] < nyhc) {
        yh = nyhc + (y - (s.ny - nyhc));
        j = owner_h(jjst, yh);
      }
    }
    if (yh >= 0 && yh < s.nyc && fcomm2_src_row_h(yh, jjst[j], jjsz[j], nyhc, s.ny, dny) == y) yh_of[y] = yh;
  }
  s.d_xo = upload_vec(xo); s.d_xs = upload_vec(xs);
  s.d_ya = upload_vec(ya); s.d_yb = upload_vec(yb);
  s.d_ha = upload_vec(ha); s.d_hb = upload_vec(hb); s.d_hy = upload_vec(hy);
  s.d_yh = upload_vec(yh_of);
  if ((long long)s.iisize * s.ny * s.kjsize < (1LL << 31)) {
    std::vector<int2> yt(s.ny), ysf(s.ny, make_int2(-1, 0)), ylb(s.ny, make_int2(-1, 0));
    for (int y = 0; y < s.ny; y++) yt[y] = make_int2((int)ya[y], (int)yb[y]);
    for (int yh = 0; yh < s.nyc; yh++)
      if (hy[yh] >= 0) ysf[hy[yh]] = make_int2((int)ha[yh], (int)hb[yh]);
    for (int y = 0; y < s.ny; y++)
      if (yh_of[y] >= 0) ylb[y] = make_int2((int)ha[yh_of[y]], (int)hb[yh_of[y]]);
    s.d_yt = upload_vec(yt); s.d_ysf = upload_vec(ysf); s.d_ylb = upload_vec(ylb);
  }
  if ((long long)s.nxhpc * unit_x < (1LL << 31)) {
    std::vector<int2> xt(s.nxhpc);
    for (int x = 0; x < s.nxhpc; x++) xt[x] = make_int2((int)(xo[x] + x), xs[x]);
    s.d_xt = upload_vec(xt);
  }
}

void share_free_tables(ShareGeom& s) {
  for (void** p : {(void**)&s.d_xo, (void**)&s.d_xs, (void**)&s.d_ya, (void**)&s.d_yb, (void**)&s.d_ha,
                   (void**)&s.d_hb, (void**)&s.d_hy, (void**)&s.d_yh, (void**)&s.d_xt, (void**)&s.d_yt,
                   (void**)&s.d_ysf, (void**)&s.d_ylb}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
}

// the table forms need every element index below 2^32
static bool fits32(long long total) { return total < (1LL << 32) - 256; }

static unsigned nblocks(long long total) { return (unsigned)((total + 255) / 256); }

void share_pack_fcomm1(const ShareGeom& s, const cplx_t* buf2, cplx_t* send, hipStream_t st) {
  const long long total = (long long)s.nxhpc * s.jisize * s.kjsize;
  if (total <= 0) return;
  if (s.d_xo && fits32(total)) {
    k_pack_fcomm1_t<real_t><<<nblocks(total), 256, 0, st>>>(buf2, send, s.d_xo, s.d_xs, make_div32(s.nxhpc), s.nxhp,
                                                           (unsigned)total);
    check_hip(hipGetLastError(), "k_pack_fcomm1_t");
    return;
  }
  k_pack_fcomm1<real_t><<<nblocks(total), 256, 0, st>>>(buf2, send, SplitTab{s.d_iist, s.d_iisz, s.p1}, s.nxhp,
                                                       s.jisize, s.kjsize, total);
  check_hip(hipGetLastError(), "k_pack_fcomm1");
}

void share_unpack_fcomm1(const ShareGeom& s, const cplx_t* recv, cplx_t* dest, hipStream_t st) {
  const long long total = (long long)s.iisize * s.ny * s.kjsize;
  if (total <= 0) return;
  if (s.d_ya && fits32(total)) {
    k_unpack_fcomm1_t<real_t><<<nblocks(total), 256, 0, st>>>(recv, dest, s.d_ya, s.d_yb, make_div32(s.iisize),
                                                             make_div32(s.ny), (unsigned)total);
    check_hip(hipGetLastError(), "k_unpack_fcomm1_t");
    return;
  }
  k_unpack_fcomm1<real_t><<<nblocks(total), 256, 0, st>>>(recv, dest, SplitTab{s.d_jist, s.d_jisz, s.p1},
                                                         s.iisize, s.ny, s.kjsize, total);
  check_hip(hipGetLastError(), "k_unpack_fcomm1");
}

void share_pack_fcomm2(const ShareGeom& s, const cplx_t* src, cplx_t* send, hipStream_t st) {
  const long long total = (long long)s.iisize * s.nyc * s.kjsize;
  if (total <= 0) return;
  if (s.d_ha && fits32(total) && fits32((long long)s.iisize * s.ny * s.kjsize)) {
    k_pack_fcomm2_t<real_t><<<nblocks(total), 256, 0, st>>>(src, send, s.d_ha, s.d_hb, s.d_hy, make_div32(s.iisize),
                                                           make_div32(s.nyc), s.ny, (unsigned)total);
    check_hip(hipGetLastError(), "k_pack_fcomm2_t");
    return;
  }
  k_pack_fcomm2<real_t><<<nblocks(total), 256, 0, st>>>(src, send, SplitTab{s.d_jjst, s.d_jjsz, s.p2}, s.iisize,
                                                       s.ny, s.kjsize, s.nyc / 2, s.ny - s.nyc, total);
  check_hip(hipGetLastError(), "k_pack_fcomm2");
}

void share_unpack_bcomm1(const ShareGeom& s, const cplx_t* recv, cplx_t* dest, hipStream_t st) {
  const long long total = (long long)s.iisize * s.ny * s.kjsize;
  if (total <= 0) return;
  if (s.d_ha && fits32(total)) {
    k_unpack_bcomm1_t<real_t><<<nblocks(total), 256, 0, st>>>(recv, dest, s.d_ha, s.d_hb, s.d_yh,
                                                             make_div32(s.iisize), make_div32(s.ny),
                                                             (unsigned)total);
    check_hip(hipGetLastError(), "k_unpack_bcomm1_t");
    return;
  }
  k_unpack_bcomm1<real_t><<<nblocks(total), 256, 0, st>>>(recv, dest, SplitTab{s.d_jjst, s.d_jjsz, s.p2},
                                                         s.iisize, s.ny, s.kjsize, s.nyc / 2, s.nyc,
                                                         s.ny - s.nyc, total);
  check_hip(hipGetLastError(), "k_unpack_bcomm1");
}

void share_pack_bcomm2(const ShareGeom& s, const cplx_t* src, cplx_t* send, hipStream_t st) {
  const long long total = (long long)s.iisize * s.ny * s.kjsize;
  if (total <= 0) return;
  if (s.d_ya && fits32(total)) {
    k_pack_bcomm2_t<real_t><<<nblocks(total), 256, 0, st>>>(src, send, s.d_ya, s.d_yb, make_div32(s.iisize),
                                                           make_div32(s.ny), (unsigned)total);
    check_hip(hipGetLastError(), "k_pack_bcomm2_t");
    return;
  }
  k_pack_bcomm2<real_t><<<nblocks(total), 256, 0, st>>>(src, send, SplitTab{s.d_jist, s.d_jisz, s.p1}, s.iisize,
                                                       s.ny, s.kjsize, total);
  check_hip(hipGetLastError(), "k_pack_bcomm2");
}

void share_unpack_bcomm2(const ShareGeom& s, const cplx_t* recv, cplx_t* dest, hipStream_t st) {
  const long long total = (long long)s.nxhp * s.jisize * s.kjsize;
  if (total <= 0) return;
  if (s.d_xo && fits32(total)) {
    k_unpack_bcomm2_t<real_t><<<nblocks(total), 256, 0, st>>>(recv, dest, s.d_xo, s.d_xs, make_div32(s.nxhp),
                                                             s.nxhpc, (unsigned)total);
    check_hip(hipGetLastError(), "k_unpack_bcomm2_t");
    return;
  }
  k_unpack_bcomm2<real_t><<<nblocks(total), 256, 0, st>>>(recv, dest, SplitTab{s.d_iist, s.d_iisz, s.p1}, s.nxhp,
                                                         s.nxhpc, s.jisize, s.kjsize, total);
  check_hip(hipGetLastError(), "k_unpack_bcomm2");
}

}  // namespace p3amd
