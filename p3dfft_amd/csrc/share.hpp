// share.hpp -- one rank of a P1 x P2 processor grid ("share mode").
//
// The rank keeps its real pencil shapes and runs the reference's per-rank
// stages; the exchanges between them (MPI_Alltoallv in the reference) are left
// to the caller, who receives the exact send buffers the reference would pass
// to MPI and hands back the receive buffers.
#pragma once
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
};

void share_pack_fcomm1(const ShareGeom& s, const cplx_t* buf2, cplx_t* send, hipStream_t st);
void share_unpack_fcomm1(const ShareGeom& s, const cplx_t* recv, cplx_t* dest, hipStream_t st);
void share_pack_fcomm2(const ShareGeom& s, const cplx_t* src, cplx_t* send, hipStream_t st);
void share_unpack_bcomm1(const ShareGeom& s, const cplx_t* recv, cplx_t* dest, hipStream_t st);
void share_pack_bcomm2(const ShareGeom& s, const cplx_t* src, cplx_t* send, hipStream_t st);
void share_unpack_bcomm2(const ShareGeom& s, const cplx_t* recv, cplx_t* dest, hipStream_t st);

}  // namespace p3amd
