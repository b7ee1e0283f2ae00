// api_share.cpp -- C-ABI of share mode: one rank of a P1 x P2 grid at its real
// pencil shapes, with the reference's per-rank stages separated at the points
// where it calls MPI_Alltoallv (build/ftran.F90:489-780 and build/btran.F90:
// 396-679 with iproc > 1 and jproc > 1, non-STRIDE1).  Each stage takes the
// receive buffer of the previous exchange and produces the send buffer of the
// next, laid out exactly as the reference's send buffers (MPI_Alltoallv with
// the counts and displacements of build/setup.F90:480-518).  A multi-GPU
// runtime moves the buffers between ranks; on one GPU the tests exchange them.
#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "engine.hpp"
#include "p3dfft_amd_internal.hpp"
#include "share.hpp"

using namespace p3amd;

namespace p3amd {

static void share_free() {
  for (int** p : {&g.sg.d_iist, &g.sg.d_iisz, &g.sg.d_jist, &g.sg.d_jisz, &g.sg.d_jjst, &g.sg.d_jjsz}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  share_free_tables(g.sg);
  if (g.w3) (void)hipFree(g.w3);
  g.w3 = nullptr;
}

static std::vector<int> zero_based(const std::vector<int>& st1) {
  std::vector<int> v(st1.size());
  for (size_t i = 0; i < v.size(); i++) v[i] = st1[i] - 1;
  return v;
}

static int* upload_starts(const std::vector<int>& st1) {
  std::vector<int> v(st1.size());
  for (size_t i = 0; i < v.size(); i++) v[i] = st1[i] - 1;  // 1-based -> 0-based
  int* d = nullptr;
  check_hip(hipMalloc(&d, std::max<size_t>(v.size(), 1) * sizeof(int)), "hipMalloc(split)");
  check_hip(hipMemcpy(d, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy(split)");
  return d;
}

static int* upload(const std::vector<int>& v) {
  int* d = nullptr;
  check_hip(hipMalloc(&d, std::max<size_t>(v.size(), 1) * sizeof(int)), "hipMalloc(split)");
  check_hip(hipMemcpy(d, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy(split)");
  return d;
}

void share_release() { share_free(); }

// pack/unpack fused into the FFT kernels' loads and stores where the lengths
// have specialised kernels (default); off: separate pack/unpack passes
static bool g_share_fused = true;

static bool require_share() {
  if (!require_setup()) return false;
  if (!g.share) return fatal(" p3dfft_amd: the share-mode stages need p3dfft_setup_share");
  return true;
}

}  // namespace p3amd

#define GUARD_BEGIN try {
#define GUARD_END                                  \
  }                                                \
  catch (const Error& e) {                         \
    fatal(std::string("p3dfft_amd: ") + e.msg);    \
    return -1;                                     \
  }

extern "C" {

// One rank (ipid, jpid) of a dims[0] x dims[1] grid: the arguments of
// p3dfft_setup (build/setup.F90:107) plus the rank's grid coordinates, which
// the reference takes from MPI_Cart_coords (setup.F90:205-223).
int p3dfft_setup_share(int* dims, int* ipid, int* jpid, int* nx, int* ny, int* nz, int* nxc, int* nyc, int* nzc,
                       int* ow, int* memsize) {
  GUARD_BEGIN
  if (*nx <= 0 || *ny <= 0 || *nz <= 0) {
    fatal(" Invalid dimensions :" + std::to_string(*nx) + " " + std::to_string(*ny) + " " + std::to_string(*nz));
    return -1;
  }
  if (g.set) {
    fatal(" P3DFFT Setup error: the problem is already initialized. ");
    return -1;
  }
  if (dims[0] <= 0 || dims[1] <= 0 || *ipid < 0 || *ipid >= dims[0] || *jpid < 0 || *jpid >= dims[1]) {
    fatal(" Invalid processor geometry: " + std::to_string(dims[0]) + " " + std::to_string(dims[1]));
    return -1;
  }
  if (*nxc < 1 || *nxc > *nx || *nyc < 1 || *nyc > *ny || *nzc < 1 || *nzc > *nz) {
    fatal(" Invalid cut dimensions");
    return -1;
  }
  if (g.stride1_next) {
    fatal(" p3dfft_amd: share mode supports the (X,Y,Z) layout only");
    return -1;
  }
  for (int i = 0; i < 12; i++) g.timers[i] = 0.0;
  g.nx = *nx; g.ny = *ny; g.nz = *nz;
  g.nxc = *nxc; g.nyc = *nyc; g.nzc = *nzc;
  g.ow = (*ow != 0);
  g.stride1 = false;
  compute_geometry(g, dims, *ipid, *jpid);
  const int lim = max_len();
  if (g.nx > lim || g.ny > lim || g.nz > lim) {
    fatal(" p3dfft_amd: grid dimension exceeds the supported transform length " + std::to_string(lim));
    return -1;
  }
  ShareGeom& s = g.sg;
  s.p1 = g.p1; s.p2 = g.p2; s.ipid = g.ipid; s.jpid = g.jpid;
  s.nx = g.nx; s.ny = g.ny; s.nz = g.nz; s.nxc = g.nxc; s.nyc = g.nyc; s.nzc = g.nzc;
  s.nxhp = g.nxhp; s.nxhpc = g.nxhpc;
  s.iisize = g.iisize; s.jisize = g.jisize; s.jjsize = g.jjsize; s.kjsize = g.kjsize;
  s.d_iist = upload_starts(g.iist); s.d_iisz = upload(g.iisz);
  s.d_jist = upload_starts(g.jist); s.d_jisz = upload(g.jisz);
  s.d_jjst = upload_starts(g.jjst); s.d_jjsz = upload(g.jjsz);
  share_build_tables(s, zero_based(g.iist), g.iisz, zero_based(g.jist), g.jisz, zero_based(g.jjst), g.jjsz);
  // work: w1 = buf2 (nxhp, jisize, kjsize); w2 = Y pencil (iisize, ny, kjsize);
  // w3 = Z pencil (iisize, jjsize, nz)
  check_hip(hipMalloc(&g.w1, std::max<size_t>(1, (size_t)g.nxhp * g.jisize * g.kjsize) * sizeof(cplx_t)), "hipMalloc");
  check_hip(hipMalloc(&g.w2, std::max<size_t>(1, (size_t)g.iisize * g.ny * g.kjsize) * sizeof(cplx_t)), "hipMalloc");
  check_hip(hipMalloc(&g.w3, std::max<size_t>(1, (size_t)g.iisize * g.jjsize * g.nz) * sizeof(cplx_t)), "hipMalloc");
  get_len(g.nx);
  get_len(g.ny);
  get_len(g.nz);
  g.share = true;
  g.set = true;
  memsize[0] = g.memsize[0];
  memsize[1] = g.memsize[1];
  memsize[2] = g.memsize[2];
  return 0;
  GUARD_END
}

// Element offsets and counts (complex units) of each peer's block in the four
// exchange buffers: forward row (p1 entries), forward column (p2 entries); the
// backward exchanges use the same blocks with send and receive swapped
// (build/setup.F90:480-518).
// 1 (default): share-mode stages fuse the pack/unpack maps into the FFT
// kernels where they can; 0: separate pack/unpack passes (same buffers).
void p3dfft_share_set_fused(int on) { g_share_fused = (on != 0); }

int p3dfft_share_counts(long long* fsnd_st, long long* fsnd_ct, long long* frcv_st, long long* frcv_ct,
                        long long* ksnd_st, long long* ksnd_ct, long long* krcv_st, long long* krcv_ct) {
  if (!require_share()) return -1;
  for (int i = 0; i < g.p1; i++) {
    fsnd_st[i] = (long long)(g.iist[i] - 1) * g.jisize * g.kjsize;
    fsnd_ct[i] = (long long)g.iisz[i] * g.jisize * g.kjsize;
    frcv_st[i] = (long long)(g.jist[i] - 1) * g.iisize * g.kjsize;
    frcv_ct[i] = (long long)g.jisz[i] * g.iisize * g.kjsize;
  }
  for (int i = 0; i < g.p2; i++) {
    ksnd_st[i] = (long long)(g.jjst[i] - 1) * g.iisize * g.kjsize;
    ksnd_ct[i] = (long long)g.iisize * g.kjsize * g.jjsz[i];
    krcv_st[i] = (long long)(g.kjst[i] - 1) * g.iisize * g.jjsize;
    krcv_ct[i] = (long long)g.iisize * g.jjsize * g.kjsz[i];
  }
  return 0;
}

// Pure pack / unpack (data movement only), for callers that run their own
// transforms and for bit-exactness checks.  which: 0 fcomm1 pack, 1 fcomm1
// unpack, 2 fcomm2 pack, 3 bcomm1 unpack, 4 bcomm2 pack, 5 bcomm2 unpack.
int p3dfft_share_pack(int which, const p3amd::real_t* src, p3amd::real_t* dst, void* stream) {
  if (!require_share()) return -1;
  GUARD_BEGIN
  const cplx_t* a = (const cplx_t*)src;
  cplx_t* b = (cplx_t*)dst;
  hipStream_t st = (hipStream_t)stream;
  switch (which) {
    case 0: share_pack_fcomm1(g.sg, a, b, st); break;
    case 1: share_unpack_fcomm1(g.sg, a, b, st); break;
    case 2: share_pack_fcomm2(g.sg, a, b, st); break;
    case 3: share_unpack_bcomm1(g.sg, a, b, st); break;
    case 4: share_pack_bcomm2(g.sg, a, b, st); break;
    case 5: share_unpack_bcomm2(g.sg, a, b, st); break;
    default: fatal(" p3dfft_amd: unknown pack selector"); return -1;
  }
  return 0;
  GUARD_END
}

// forward stage 1: X r2c of the local pencil (nx, jisize, kjsize) and the
// fcomm1 pack (build/ftran.F90:525-544, fcomm1.F90:238-253) -> row send buffer
int p3dfft_share_ftran_x(const p3amd::real_t* in, p3amd::real_t* send1, void* stream) {
  if (!require_share()) return -1;
  GUARD_BEGIN
  hipStream_t st = (hipStream_t)stream;
  tstart(g, 5, st);
  // the X kernel writes the send buffer through the fcomm1 map when the
  // length has a specialised kernel; otherwise X into buf2, then the pack
  const bool fused = g.sg.d_xt && g_share_fused &&
                     launch_x_r2c_tab(in, (cplx_t*)send1, (long long)g.jisize * g.kjsize, g.nx, g.nxhpc,
                                      g.sg.d_xt, st);
  tstop(g, 5, st);
  if (fused) return 0;
  tstart(g, 5, st);
  launch_x_r2c(in, g.w1, (long long)g.jisize * g.kjsize, g.nx, g.nxhp, g.nx, g.nxhpc, st);
  tstop(g, 5, st);
  tstart(g, 6, st);
  share_pack_fcomm1(g.sg, g.w1, (cplx_t*)send1, st);
  tstop(g, 6, st);
  return 0;
  GUARD_END
}

// forward stage 2: fcomm1 unpack, Y c2c per z-plane, fcomm2 pack
// (fcomm1.F90:284-320, ftran.F90:576-584, fcomm2.F90:321-386)
int p3dfft_share_ftran_y(const p3amd::real_t* recv1, p3amd::real_t* send2, void* stream) {
  if (!require_share()) return -1;
  GUARD_BEGIN
  hipStream_t st = (hipStream_t)stream;
  if (g_share_fused && g.sg.d_yt) {
    // the Y kernel reads the fcomm1 receive buffer and writes the fcomm2
    // send buffer through the maps (row tables)
    ColSpec y{};
    y.in = (const cplx_t*)recv1; y.out = (cplx_t*)send2;
    y.n_inner = g.iisize; y.n_outer = g.kjsize; y.n = g.ny; y.kind = COL_C2C; y.fwd = true;
    tstart(g, 7, st);
    const bool done = launch_col_tab(y, g.sg.d_yt, g.sg.d_ysf, st);
    tstop(g, 7, st);
    if (done) return 0;
  }
  tstart(g, 6, st);
  share_unpack_fcomm1(g.sg, (const cplx_t*)recv1, g.w2, st);
  tstop(g, 6, st);
  ColSpec y{};
  y.in = g.w2; y.out = g.w2;
  y.si_in = 1; y.sl_in = g.iisize; y.so_in = (long long)g.iisize * g.ny;
  y.si_out = 1; y.sl_out = g.iisize; y.so_out = (long long)g.iisize * g.ny;
  y.n_inner = g.iisize; y.n_outer = g.kjsize; y.n = g.ny; y.kind = COL_C2C; y.fwd = true;
  y.nld = g.ny; y.ld_lo = g.ny; y.ld_hi = g.ny; y.ld_shift = 0;
  y.nst = g.ny; y.st_lo = g.ny; y.st_shift = 0;
  tstart(g, 7, st);
  launch_col(y, st);
  tstop(g, 7, st);
  tstart(g, 8, st);
  share_pack_fcomm2(g.sg, g.w2, (cplx_t*)send2, st);
  tstop(g, 8, st);
  return 0;
  GUARD_END
}

static int zop_kind(char c, bool* none) {
  *none = false;
  if (c == 't' || c == 'f') return COL_C2C;
  if (c == 'c') return COL_DCT1;
  if (c == 's') return COL_DST1;
  if (c == 'n' || c == '0') { *none = true; return COL_COPY; }
  return -1;
}

// forward stage 3: Z op on the received Z pencil recv2 (iisize, jjsize, nz)
// into out (iisize, jjsize, nzc), keeping z < nzcph and the top modes
// (ftran.F90:595-689).  recv2 == out is allowed when nzc == nz.
int p3dfft_share_ftran_z(p3amd::real_t* recv2, p3amd::real_t* out, const unsigned char* op, void* stream) {
  if (!require_share()) return -1;
  GUARD_BEGIN
  bool none;
  const int kind = zop_kind((char)op[2], &none);
  if (kind < 0) { fatal(std::string("Unknown transform type: ") + (char)op[2]); return -1; }
  if (!zop_len_ok(g.nz, kind)) return -1;
  hipStream_t st = (hipStream_t)stream;
  const long long nxy = (long long)g.iisize * g.jjsize;
  tstart(g, 8, st);
  if (!(none && g.nz == g.nzc && recv2 == out)) {
    ColSpec z{};
    z.in = (const cplx_t*)recv2; z.out = (cplx_t*)out;
    z.n_inner = (int)nxy; z.n_outer = 1;
    z.si_in = 1; z.sl_in = nxy; z.si_out = 1; z.sl_out = nxy;
    z.n = g.nz; z.kind = kind; z.fwd = true;
    z.nld = g.nz; z.ld_lo = g.nz; z.ld_hi = g.nz; z.ld_shift = 0;
    z.nst = g.nzc; z.st_lo = g.nzcph; z.st_shift = g.nz - g.nzc;
    launch_col(z, st);
  }
  tstop(g, 8, st);
  return 0;
  GUARD_END
}

// backward stage 1: Z op on the rank's spectrum in (iisize, jjsize, nzc) into
// the bcomm1 send buffer send (iisize, jjsize, nz) with the middle z modes
// zero-filled (btran.F90:517-594; bcomm1 sends this array as is, by z blocks).
// send == in is the reference's overwrite path (nz == nzc).
int p3dfft_share_btran_z(p3amd::real_t* in, p3amd::real_t* send, const unsigned char* op, void* stream) {
  if (!require_share()) return -1;
  GUARD_BEGIN
  bool none;
  const int kind = zop_kind((char)op[0], &none);
  if (kind < 0) { fatal(std::string(" Unknown transform type: ") + (char)op[0]); return -1; }
  if (!zop_len_ok(g.nz, kind)) return -1;
  hipStream_t st = (hipStream_t)stream;
  const long long nxy = (long long)g.iisize * g.jjsize;
  const int dnz = g.nz - g.nzc;
  tstart(g, 9, st);
  if (!(none && dnz == 0 && in == send)) {
    ColSpec z{};
    z.in = (const cplx_t*)in; z.out = (cplx_t*)send;
    z.n_inner = (int)nxy; z.n_outer = 1;
    z.si_in = 1; z.sl_in = nxy; z.si_out = 1; z.sl_out = nxy;
    z.n = g.nz; z.kind = kind; z.fwd = false;
    z.nld = g.nz; z.ld_lo = g.nzcph; z.ld_hi = g.nz - g.nzcph; z.ld_shift = dnz;
    z.nst = g.nz; z.st_lo = g.nz; z.st_shift = 0;
    launch_col(z, st);
  }
  tstop(g, 9, st);
  return 0;
  GUARD_END
}

// backward stage 2: bcomm1 unpack (zero-filling the pruned Y modes), Y c2c
// backward, bcomm2 pack (bcomm1.F90:309-380, btran.F90:615-624, bcomm2.F90:233-273)
int p3dfft_share_btran_y(const p3amd::real_t* recv, p3amd::real_t* send, void* stream) {
  if (!require_share()) return -1;
  GUARD_BEGIN
  hipStream_t st = (hipStream_t)stream;
  if (g_share_fused && g.sg.d_yt) {
    // the Y kernel reads the bcomm1 receive buffer (zero rows included) and
    // writes the bcomm2 send buffer through the maps (row tables)
    ColSpec y{};
    y.in = (const cplx_t*)recv; y.out = (cplx_t*)send;
    y.n_inner = g.iisize; y.n_outer = g.kjsize; y.n = g.ny; y.kind = COL_C2C; y.fwd = false;
    tstart(g, 10, st);
    const bool done = launch_col_tab(y, g.sg.d_ylb, g.sg.d_yt, st);
    tstop(g, 10, st);
    if (done) return 0;
  }
  tstart(g, 9, st);
  share_unpack_bcomm1(g.sg, (const cplx_t*)recv, g.w2, st);
  tstop(g, 9, st);
  ColSpec y{};
  y.in = g.w2; y.out = g.w2;
  y.si_in = 1; y.sl_in = g.iisize; y.so_in = (long long)g.iisize * g.ny;
  y.si_out = 1; y.sl_out = g.iisize; y.so_out = (long long)g.iisize * g.ny;
  y.n_inner = g.iisize; y.n_outer = g.kjsize; y.n = g.ny; y.kind = COL_C2C; y.fwd = false;
  y.nld = g.ny; y.ld_lo = g.ny; y.ld_hi = g.ny; y.ld_shift = 0;
  y.nst = g.ny; y.st_lo = g.ny; y.st_shift = 0;
  tstart(g, 10, st);
  launch_col(y, st);
  tstop(g, 10, st);
  tstart(g, 11, st);
  share_pack_bcomm2(g.sg, g.w2, (cplx_t*)send, st);
  tstop(g, 11, st);
  return 0;
  GUARD_END
}

// backward stage 3: bcomm2 unpack (zeroing x >= nxhpc) and X c2r into the
// rank's real pencil (bcomm2.F90:290-313, btran.F90:646-656)
int p3dfft_share_btran_x(const p3amd::real_t* recv, p3amd::real_t* out, void* stream) {
  if (!require_share()) return -1;
  GUARD_BEGIN
  hipStream_t st = (hipStream_t)stream;
  tstart(g, 12, st);
  const bool fused = g.sg.d_xt && g_share_fused &&
                     launch_x_c2r_tab((const cplx_t*)recv, out, (long long)g.jisize * g.kjsize, g.nx, g.nxhpc,
                                      g.sg.d_xt, st);
  tstop(g, 12, st);
  if (fused) return 0;
  tstart(g, 11, st);
  share_unpack_bcomm2(g.sg, (const cplx_t*)recv, g.w1, st);
  tstop(g, 11, st);
  tstart(g, 12, st);
  launch_x_c2r(g.w1, out, (long long)g.jisize * g.kjsize, g.nxhp, g.nx, g.nx, g.nxhp, st);
  tstop(g, 12, st);
  return 0;
  GUARD_END
}

}  // extern "C"
