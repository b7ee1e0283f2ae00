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

static unsigned nblocks(long long total) { return (unsigned)((total + 255) / 256); }

void share_pack_fcomm1(const ShareGeom& s, const cplx_t* buf2, cplx_t* send, hipStream_t st) {
  const long long total = (long long)s.nxhpc * s.jisize * s.kjsize;
  if (total <= 0) return;
  k_pack_fcomm1<real_t><<<nblocks(total), 256, 0, st>>>(buf2, send, SplitTab{s.d_iist, s.d_iisz, s.p1}, s.nxhp,
                                                       s.jisize, s.kjsize, total);
  check_hip(hipGetLastError(), "k_pack_fcomm1");
}

void share_unpack_fcomm1(const ShareGeom& s, const cplx_t* recv, cplx_t* dest, hipStream_t st) {
  const long long total = (long long)s.iisize * s.ny * s.kjsize;
  if (total <= 0) return;
  k_unpack_fcomm1<real_t><<<nblocks(total), 256, 0, st>>>(recv, dest, SplitTab{s.d_jist, s.d_jisz, s.p1},
                                                         s.iisize, s.ny, s.kjsize, total);
  check_hip(hipGetLastError(), "k_unpack_fcomm1");
}

void share_pack_fcomm2(const ShareGeom& s, const cplx_t* src, cplx_t* send, hipStream_t st) {
  const long long total = (long long)s.iisize * s.nyc * s.kjsize;
  if (total <= 0) return;
  k_pack_fcomm2<real_t><<<nblocks(total), 256, 0, st>>>(src, send, SplitTab{s.d_jjst, s.d_jjsz, s.p2}, s.iisize,
                                                       s.ny, s.kjsize, s.nyc / 2, s.ny - s.nyc, total);
  check_hip(hipGetLastError(), "k_pack_fcomm2");
}

void share_unpack_bcomm1(const ShareGeom& s, const cplx_t* recv, cplx_t* dest, hipStream_t st) {
  const long long total = (long long)s.iisize * s.ny * s.kjsize;
  if (total <= 0) return;
  k_unpack_bcomm1<real_t><<<nblocks(total), 256, 0, st>>>(recv, dest, SplitTab{s.d_jjst, s.d_jjsz, s.p2},
                                                         s.iisize, s.ny, s.kjsize, s.nyc / 2, s.nyc,
                                                         s.ny - s.nyc, total);
  check_hip(hipGetLastError(), "k_unpack_bcomm1");
}

void share_pack_bcomm2(const ShareGeom& s, const cplx_t* src, cplx_t* send, hipStream_t st) {
  const long long total = (long long)s.iisize * s.ny * s.kjsize;
  if (total <= 0) return;
  k_pack_bcomm2<real_t><<<nblocks(total), 256, 0, st>>>(src, send, SplitTab{s.d_jist, s.d_jisz, s.p1}, s.iisize,
                                                       s.ny, s.kjsize, total);
  check_hip(hipGetLastError(), "k_pack_bcomm2");
}

void share_unpack_bcomm2(const ShareGeom& s, const cplx_t* recv, cplx_t* dest, hipStream_t st) {
  const long long total = (long long)s.nxhp * s.jisize * s.kjsize;
  if (total <= 0) return;
  k_unpack_bcomm2<real_t><<<nblocks(total), 256, 0, st>>>(recv, dest, SplitTab{s.d_iist, s.d_iisz, s.p1}, s.nxhp,
                                                         s.nxhpc, s.jisize, s.kjsize, total);
  check_hip(hipGetLastError(), "k_unpack_bcomm2");
}

}  // namespace p3amd
