// share.hpp -- one rank of a P1 x P2 processor grid ("share mode").
//
// The rank keeps its real pencil shapes and runs the reference's per-rank
// stages; the exchanges between them (MPI_Alltoallv in the reference) are left
// to the caller, who receives the exact send buffers the reference would pass
// to MPI and hands back the receive buffers.
#pragma once
#include <vector>

#include "engine.hpp"

namespace p3amd {

struct ShareGeom {
  int p1 = 1, p2 = 1, ipid = 0, jpid = 0;
  int nx = 0, ny = 0, nz = 0, nxc = 0, nyc = 0, nzc = 0;
  int nxhp = 0, nxhpc = 0;
  int iisize = 0, jisize = 0, jjsize = 0, kjsize = 0;
  // device copies of the 0-based split starts and sizes
  int* d_iist = nullptr;  // nxhpc over p1
  int* d_iisz = nullptr;
  int* d_jist = nullptr;  // ny over p1
  int* d_jisz = nullptr;
  int* d_jjst = nullptr;  // nyc over p2
  int* d_jjsz = nullptr;
  // Per-row address tables of the six maps (built once by share_build_tables).
  // An element of a peer block is  table offset + outer index * table stride
  // + inner index, so the kernels stream contiguous runs with no search:
  long long* d_xo = nullptr;   // [nxhpc] fcomm1 send / bcomm2 recv: block start - x start
  int* d_xs = nullptr;         // [nxhpc] run length (the owning row peer's x count)
  long long* d_ya = nullptr;   // [ny] fcomm1 recv / bcomm2 send: offset of row y in its block
  long long* d_yb = nullptr;   // [ny] z stride of that block
  long long* d_ha = nullptr;   // [nyc] fcomm2 send / bcomm1 recv: offset of position yh
  long long* d_hb = nullptr;   // [nyc] z stride of that block
  int* d_hy = nullptr;         // [nyc] source row of position yh, -1 where left unwritten
  int* d_yh = nullptr;         // [ny] position yh that fills row y, -1 for a zero row
  // [nxhpc] {xo + x, xs} as 32-bit pairs for the X kernels that write the
  // fcomm1 send buffer / read the bcomm2 receive buffer directly (null when
  // those buffers exceed 2^31 elements)
  int2* d_xt = nullptr;
  // [ny] row tables of the Y column kernel that reads / writes the exchange
  // buffers directly: {ya, yb} (fcomm1 recv, bcomm2 send), the fcomm2 send
  // position of FFT row p (-1: not sent) and the bcomm1 recv position that
  // fills row y (-1: zero row); null when the buffers exceed 2^31 elements
  int2* d_yt = nullptr;
  int2* d_ysf = nullptr;
  int2* d_ylb = nullptr;
};

// Build (and upload) the address tables from the 0-based splits.
void share_build_tables(ShareGeom& s, const std::vector<int>& iist, const std::vector<int>& iisz,
                        const std::vector<int>& jist, const std::vector<int>& jisz, const std::vector<int>& jjst,
                        const std::vector<int>& jjsz);
void share_free_tables(ShareGeom& s);

void share_pack_fcomm1(const ShareGeom& s, const cplx_t* buf2, cplx_t* send, hipStream_t st);
void share_unpack_fcomm1(const ShareGeom& s, const cplx_t* recv, cplx_t* dest, hipStream_t st);
void share_pack_fcomm2(const ShareGeom& s, const cplx_t* src, cplx_t* send, hipStream_t st);
void share_unpack_bcomm1(const ShareGeom& s, const cplx_t* recv, cplx_t* dest, hipStream_t st);
void share_pack_bcomm2(const ShareGeom& s, const cplx_t* src, cplx_t* send, hipStream_t st);
void share_unpack_bcomm2(const ShareGeom& s, const cplx_t* recv, cplx_t* dest, hipStream_t st);

}  // namespace p3amd
