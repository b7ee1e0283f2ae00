// api.cpp -- the C-ABI of libp3dfft_amd: P3DFFT's BIND(C) entry points for the
// ftran_r2c / btran_c2r path, re-implemented over the HIP engine.
//
// State is module-global and there is one active setup at a time, as in the
// reference (build/module.F90:101-176, build/setup.F90:130-135).  The single
// GPU is rank 0 of a 1x1 processor grid: both transposes of the reference are
// local (build/ftran.F90:539,595 and build/btran.F90:430,646 take the iproc=1 /
// jproc=1 branches), so they are folded into the load/store indexing of the
// neighbouring FFT kernels.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "engine.hpp"
#include "p3dfft_amd_internal.hpp"

using namespace p3amd;

namespace p3amd {

State g;

std::atomic<int> g_live_captures{0};

static int g_err_mode = 0;  // 0: print and terminate (reference: MPI_Abort); 1: print and return
static std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

// Reference behaviour for fatal errors is "print, then MPI_Abort".
bool fatal(const std::string& msg) {
  fprintf(stderr, "%s\n", msg.c_str());
  fflush(stderr);
  g_last_error = msg;
  if (g_err_mode == 0) std::exit(1);
  return false;
}

// Messages that the reference prints before returning (no abort).
void notice(const std::string& msg) {
  fprintf(stderr, "%s\n", msg.c_str());
  fflush(stderr);
  g_last_error = msg;
}

// ---------------------------------------------------------------------------
// timers: the reference's 12 accumulating wall-clock slots
// (build/module.F90:106, 709-744), filled from HIP events recorded on the
// stream around each stage and resolved lazily.
// ---------------------------------------------------------------------------
static hipEvent_t ev_get() {
  hipEvent_t e;
  if (!g.evpool.empty()) {
    e = g.evpool.back();
    g.evpool.pop_back();
    return e;
  }
  check_hip(hipEventCreate(&e), "hipEventCreate");
  return e;
}

void timers_resolve() {
  for (auto& p : g.pend) {
    float ms = 0.f;
    if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess)
      g.timers[p.slot - 1] += ms * 1e-3;
    g.evpool.push_back(p.a);
    g.evpool.push_back(p.b);
  }
  g.pend.clear();
}

void tstart(State& S, int slot, hipStream_t s) {
  if (!S.timing) return;
  hipEvent_t a = ev_get();
  check_hip(hipEventRecord(a, s), "hipEventRecord");
  S.open[slot - 1] = a;
}

void tstop(State& S, int slot, hipStream_t s) {
  if (!S.timing) return;
  hipEvent_t b = ev_get();
  check_hip(hipEventRecord(b, s), "hipEventRecord");
  S.pend.push_back(Pending{slot, S.open[slot - 1], b});
  if (S.pend.size() > 8192) timers_resolve();
}

// ---------------------------------------------------------------------------
// decomposition (build/setup.F90:608-635, remainder to the last ranks)
// ---------------------------------------------------------------------------
void map_data_to_proc(int data, int proc, int* st, int* en, int* sz) {
  int size = data / proc;
  const int nu = data - size * proc;
  const int nl = proc - nu;
  st[0] = 1;
  sz[0] = size;
  en[0] = size;
  for (int i = 1; i <= nl - 1; i++) {
    st[i] = st[i - 1] + size;
    sz[i] = size;
    en[i] = en[i - 1] + size;
  }
  size = size + 1;
  for (int i = nl; i <= proc - 1; i++) {
    st[i] = en[i - 1] + 1;
    sz[i] = size;
    en[i] = en[i - 1] + size;
  }
  en[proc - 1] = data;
  sz[proc - 1] = data - st[proc - 1] + 1;
}

// Sizes and local extents (build/setup.F90:148-176, 279-313, 580-603).
bool compute_geometry(State& s, const int dims[2], int ipid, int jpid) {
  s.p1 = dims[0];
  s.p2 = dims[1];
  s.ipid = ipid;
  s.jpid = jpid;
  s.nxh = s.nx / 2;
  s.nxhp = s.nxh + 1;
  s.nxhc = s.nxc / 2;
  s.nxhpc = s.nxhc + 1;
  s.nyh = s.ny / 2;
  s.nzh = s.nz / 2;
  s.nyhc = s.nyc / 2;
  s.nzhc = s.nzc / 2;
  s.nycph = (s.nyc + 1) / 2;
  s.nzcph = (s.nzc + 1) / 2;
  const int P = s.p1 > s.p2 ? s.p1 : s.p2;
  std::vector<int> st(P), en(P), sz(P);
  map_data_to_proc(s.nxhpc, s.p1, st.data(), en.data(), sz.data());
  s.iist.assign(st.begin(), st.begin() + s.p1);
  s.iisz.assign(sz.begin(), sz.begin() + s.p1);
  s.iistart = st[ipid]; s.iiend = en[ipid]; s.iisize = sz[ipid];
  map_data_to_proc(s.ny, s.p1, st.data(), en.data(), sz.data());
  s.jist.assign(st.begin(), st.begin() + s.p1);
  s.jisz.assign(sz.begin(), sz.begin() + s.p1);
  s.jistart = st[ipid]; s.jiend = en[ipid]; s.jisize = sz[ipid];
  map_data_to_proc(s.nyc, s.p2, st.data(), en.data(), sz.data());
  s.jjst.assign(st.begin(), st.begin() + s.p2);
  s.jjsz.assign(sz.begin(), sz.begin() + s.p2);
  s.jjstart = st[jpid]; s.jjend = en[jpid]; s.jjsize = sz[jpid];
  map_data_to_proc(s.nz, s.p2, st.data(), en.data(), sz.data());
  s.kjst.assign(st.begin(), st.begin() + s.p2);
  s.kjsz.assign(sz.begin(), sz.begin() + s.p2);
  s.kjstart = st[jpid]; s.kjend = en[jpid]; s.kjsize = sz[jpid];
  long long a = (long long)s.nz * s.jjsize * s.iisize;
  long long b = (long long)s.ny * s.kjsize * s.iisize;
  long long d = (long long)s.nx * s.jisize;
  long long pad1 = 2 * (a > b ? a : b) - (long long)s.nx * s.jisize * s.kjsize;
  if (pad1 <= 0) pad1 = 0;
  long long padi = (pad1 % d != 0) ? pad1 / d + 1 : pad1 / d;
  s.memsize[0] = s.nx;
  s.memsize[1] = s.jisize;
  s.memsize[2] = s.kjsize + (int)padi;
  return true;
}

size_t real_elems() { return (size_t)g.nx * g.jisize * g.kjsize; }
size_t spec_elems() { return (size_t)g.iisize * g.jjsize * g.nzc; }

void ensure_staging() {
  const size_t bytes = std::max(real_elems() * sizeof(real_t), spec_elems() * sizeof(cplx_t));
  if (g.stage_bytes >= bytes) return;
  if (g.stage_a) (void)hipFree(g.stage_a);
  if (g.stage_b) (void)hipFree(g.stage_b);
  g.stage_a = g.stage_b = nullptr;
  g.stage_bytes = 0;
  check_hip(hipMalloc(&g.stage_a, bytes), "hipMalloc(staging)");
  check_hip(hipMalloc(&g.stage_b, bytes), "hipMalloc(staging)");
  g.stage_bytes = bytes;
}

// The Z op's transform length fits the engine: nz for 'f'/'t', the even
// extension 2(nz-1) for the DCT-I and the odd one 2(nz+1) for the DST-I.
bool zop_len_ok(int nz, int kind) {
  const long long m = kind == COL_DCT1 ? 2LL * (nz - 1) : (kind == COL_DST1 ? 2LL * (nz + 1) : (long long)nz);
  if (m <= max_len_any()) return true;
  return fatal(std::string(" p3dfft_amd: the ") + (kind == COL_DCT1 ? "DCT-I" : "DST-I") +
               " along z of nz = " + std::to_string(nz) + " needs transform length " + std::to_string(m) +
               ", above the supported " + std::to_string(max_len_any()));
}

static int zkind(char c, bool* none) {
  *none = false;
  if (c == 't' || c == 'f') return COL_C2C;
  if (c == 'c') return COL_DCT1;
  if (c == 's') return COL_DST1;
  if (c == 'n' || c == '0') { *none = true; return COL_COPY; }
  return -1;
}

// z-planes per X/Y chunk: the chunk intermediate is sized to stay resident in
// the last-level cache; small problems run as one chunk
int chunk_planes(State& S) {
  const size_t plane = (size_t)S.nxhpc * S.ny * sizeof(cplx_t);
  if (S.chunk_bytes == 0 || plane * (size_t)S.nz <= 2 * S.chunk_bytes) return S.nz;
  size_t zc = S.chunk_bytes / plane;
  if (zc < 1) zc = 1;
  return (int)std::min<size_t>(zc, (size_t)S.nz);
}

// forward without Y/Z cuts: X r2c into the output, Y in place (see
// p3dfft_amd_set_x_into_out)
static bool g_x_into_out = true;

// ---------------------------------------------------------------------------
// forward: X r2c -> Y c2c (per z-plane) -> Z op, build/ftran.F90:489-780
// ---------------------------------------------------------------------------
bool forward_dev(State& S, const real_t* in, cplx_t* out, const unsigned char* op, hipStream_t s, double cheby_L,
                 bool* cheby_done) {
  if (cheby_done) *cheby_done = false;
  bool none;
  const int kind = zkind((char)op[2], &none);
  if (kind < 0) return fatal(std::string("Unknown transform type: ") + (char)op[2]);
  if (!zop_len_ok(S.nz, kind)) return false;
  const int nx = S.nx, ny = S.ny, nz = S.nz, nxhpc = S.nxhpc, nyc = S.nyc, nzc = S.nzc;
  const int dny = ny - nyc, dnz = nz - nzc;
  const long long nxy = (long long)nxhpc * nyc;
  // X r2c then Y c2c, z-plane chunk by chunk: the (nxhpc, ny, zc) intermediate
  // of a chunk is re-read by the Y kernel right after the X kernel wrote it,
  // so it is served from the last-level cache instead of HBM.
  // seg_copy_x and seg_copy_y (keep y < nycph, then y + dny) are folded into
  // the stores.
  const bool direct = (dnz == 0 && !S.stride1);
  cplx_t* ydst = direct ? out : S.w2;
  // a complex output plane is larger than a real input plane: when the output
  // overlaps the input (in-place use) a chunk would overwrite input planes of
  // later chunks, so such calls run as one chunk
  const char* ib = (const char*)in;
  const char* ob = (const char*)ydst;
  const bool alias = ob < ib + (size_t)nx * ny * nz * sizeof(real_t) &&
                     ib < ob + (size_t)nxhpc * nyc * nz * sizeof(cplx_t);
  const int zc = alias ? nz : chunk_planes(S);
  // unpruned in Y, no overlap: X r2c straight into the output and Y in place
  // there, so the Y pass reads (and leaves in the last-level cache) the lines
  // the Z pass reads next
  const bool xin = g_x_into_out && direct && dny == 0 && !alias && zc == nz;
  // small grids: X and Y of each z-plane in one workgroup (one launch and one
  // HBM round trip fewer); the fused pass is timed in slot 5 (X)
  bool plane = false;
  if (direct && dny == 0 && !alias && zc == nz && nxhpc == nx / 2 + 1 && nx % 2 == 0 && ny % 2 == 0) {
    tstart(S, 5, s);
    plane = launch_xy_plane(true, in, out, nx, ny, nz, s);
    tstop(S, 5, s);
  }
  if (xin && !plane) {
    tstart(S, 5, s);
    launch_x_r2c(in, out, (long long)ny * nz, nx, nxhpc, nx, nxhpc, s);
    tstop(S, 5, s);
    ColSpec y{};
    y.in = out; y.out = out;
    y.si_in = 1; y.sl_in = nxhpc; y.so_in = nxy;
    y.si_out = 1; y.sl_out = nxhpc; y.so_out = nxy;
    y.n_inner = nxhpc; y.n_outer = nz; y.n = ny; y.kind = COL_C2C; y.fwd = true;
    y.nld = ny; y.ld_lo = ny; y.ld_hi = ny; y.ld_shift = 0;
    y.nst = ny; y.st_lo = ny; y.st_shift = 0;
    y.rev = tile_order_alternate();
    tstart(S, 7, s);
    launch_col(y, s);
    tstop(S, 7, s);
  }
  for (int z0 = 0; z0 < (xin || plane ? 0 : nz); z0 += zc) {
    const int nzk = std::min(zc, nz - z0);
    tstart(S, 5, s);
    launch_x_r2c(in + (size_t)z0 * ny * nx, S.w1, (long long)ny * nzk, nx, nxhpc, nx, nxhpc, s);
    tstop(S, 5, s);
    ColSpec y{};
    y.in = S.w1; y.out = ydst + (size_t)z0 * nxy;
    y.si_in = 1; y.sl_in = nxhpc; y.so_in = (long long)nxhpc * ny;
    y.si_out = 1; y.sl_out = nxhpc; y.so_out = nxy;
    y.n_inner = nxhpc; y.n_outer = nzk; y.n = ny; y.kind = COL_C2C; y.fwd = true;
    y.nld = ny; y.ld_lo = ny; y.ld_hi = ny; y.ld_shift = 0;
    y.nst = nyc; y.st_lo = S.nycph; y.st_shift = dny;
    y.rev = tile_order_alternate();  // starts where the X pass finished
    tstart(S, 7, s);
    launch_col(y, s);
    tstop(S, 7, s);
  }
  // Z op; seg_copy_z (keep z < nzcph, then z + dnz) and the STRIDE1 layout
  // folded into the store
  tstart(S, 8, s);
  if (!(direct && none)) {
    ColSpec z{};
    z.in = ydst; z.out = out;
    if (S.stride1) {
      z.n_inner = nxhpc; z.n_outer = nyc;
      z.si_in = 1; z.sl_in = nxy; z.so_in = nxhpc;
      z.si_out = (long long)nzc * nyc; z.sl_out = 1; z.so_out = nzc;
    } else {
      z.n_inner = (int)nxy; z.n_outer = 1;
      z.si_in = 1; z.sl_in = nxy; z.so_in = 0;
      z.si_out = 1; z.sl_out = nxy; z.so_out = 0;
    }
    z.n = nz; z.kind = kind; z.fwd = true;
    z.nld = nz; z.ld_lo = nz; z.ld_hi = nz; z.ld_shift = 0;
    z.nst = nzc; z.st_lo = S.nzcph; z.st_shift = dnz;
    if (cheby_done && kind == COL_DCT1) {
      z.cheby = true;
      z.cheby_scale = 1.0 / ((double)S.nx * (double)S.ny * (double)(S.nzc - 1));
      z.cheby_L = cheby_L;
    }
    const bool done = launch_col(z, s);
    if (cheby_done) *cheby_done = done && z.cheby;
  }
  tstop(S, 8, s);
  return true;
}

// ---------------------------------------------------------------------------
// backward: Z op -> Y c2c -> X c2r, build/btran.F90:396-679
// ---------------------------------------------------------------------------
bool backward_dev(State& S, cplx_t* in, real_t* out, const unsigned char* op, hipStream_t s) {
  bool none;
  const int kind = zkind((char)op[0], &none);
  if (kind < 0) return fatal(std::string(" Unknown transform type: ") + (char)op[0]);
  if (!zop_len_ok(S.nz, kind)) return false;
  const int nx = S.nx, ny = S.ny, nz = S.nz, nxhpc = S.nxhpc, nyc = S.nyc, nzc = S.nzc;
  const int dny = ny - nyc, dnz = nz - nzc;
  const long long nxy = (long long)nxhpc * nyc;
  cplx_t* zdst;
  tstart(S, 9, s);
  if (!S.stride1 && S.ow && nz == nzc) {
    // Z op in place on the caller's array (btran.F90:525-550): the input
    // spectrum is overwritten, as in the reference
    if (!none) {
      ColSpec z{};
      z.in = in; z.out = in;
      z.n_inner = (int)nxy; z.n_outer = 1;
      z.si_in = 1; z.sl_in = nxy; z.si_out = 1; z.sl_out = nxy;
      z.n = nz; z.kind = kind; z.fwd = false;
      z.nld = nz; z.ld_lo = nz; z.ld_hi = nz; z.ld_shift = 0;
      z.nst = nz; z.st_lo = nz; z.st_shift = 0;
      z.rev = tile_order_alternate();  // starts where a forward Z pass finished
      launch_col(z, s);
    }
    zdst = in;
  } else {
    // seg_copy_z + seg_zero_z folded into the load (btran.F90:556-558;
    // STRIDE1: reorder.F90:208-245 keeps z >= nzcph+dnz)
    ColSpec z{};
    z.in = in; z.out = S.w2;
    if (S.stride1) {
      z.n_inner = nxhpc; z.n_outer = nyc;
      z.si_in = (long long)nzc * nyc; z.sl_in = 1; z.so_in = nzc;
      z.si_out = 1; z.sl_out = nxy; z.so_out = nxhpc;
    } else {
      z.n_inner = (int)nxy; z.n_outer = 1;
      z.si_in = 1; z.sl_in = nxy; z.so_in = 0;
      z.si_out = 1; z.sl_out = nxy; z.so_out = 0;
    }
    z.n = nz; z.kind = kind; z.fwd = false;
    z.nld = nz; z.ld_lo = S.nzcph; z.ld_hi = S.stride1 ? S.nzcph + dnz : nz - S.nzcph; z.ld_shift = dnz;
    z.nst = nz; z.st_lo = nz; z.st_shift = 0;
    launch_col(z, s);
    zdst = S.w2;
  }
  tstop(S, 9, s);
  // Y c2c backward then X c2r, chunk by chunk over z-planes (the chunk's
  // intermediate stays in the last-level cache); seg_copy_y + seg_zero_y and
  // seg_copy_x + seg_zero_x (x > nxhpc) are folded into the loads
  const int zc = chunk_planes(S);
  // unpruned Y: in place on the Z result (a work buffer, or the caller's
  // spectrum that overwrite mode already destroys) unless the output array
  // overlaps it; the in-place column pass measured 0.205 ms against 0.236 ms
  // into a separate buffer at 512^3 FP32
  bool yip = dny == 0 && !S.stride1;
  if (yip) {
    const char* zb = (const char*)zdst;
    const char* ob = (const char*)out;
    yip = !(ob < zb + (size_t)nxy * nz * sizeof(cplx_t) && zb < ob + (size_t)nx * ny * nz * sizeof(real_t));
  }
  // small grids: Y and X c2r of each z-plane in one workgroup, timed in slot
  // 12 (X)
  if (yip && zc == nz && nxhpc == nx / 2 + 1 && nx % 2 == 0 && ny % 2 == 0) {
    tstart(S, 12, s);
    const bool plane = launch_xy_plane(false, zdst, out, nx, ny, nz, s);
    tstop(S, 12, s);
    if (plane) return true;
  }
  for (int z0 = 0; z0 < nz; z0 += zc) {
    const int nzk = std::min(zc, nz - z0);
    cplx_t* ybuf = yip ? zdst + (size_t)z0 * nxy : S.w1;
    ColSpec y{};
    y.in = zdst + (size_t)z0 * nxy; y.out = ybuf;
    y.si_in = 1; y.sl_in = nxhpc; y.so_in = nxy;
    y.si_out = 1; y.sl_out = nxhpc; y.so_out = (long long)nxhpc * ny;
    y.n_inner = nxhpc; y.n_outer = nzk; y.n = ny; y.kind = COL_C2C; y.fwd = false;
    y.nld = ny; y.ld_lo = S.nycph; y.ld_hi = S.stride1 ? S.nycph + dny : ny - S.nycph; y.ld_shift = dny;
    y.nst = ny; y.st_lo = ny; y.st_shift = 0;
    tstart(S, 10, s);
    launch_col(y, s);
    tstop(S, 10, s);
    tstart(S, 12, s);
    launch_x_c2r(ybuf, out + (size_t)z0 * ny * nx, (long long)ny * nzk, nxhpc, nx, nx, nxhpc, s,
                 tile_order_alternate());  // starts where the Y pass finished
    tstop(S, 12, s);
  }
  return true;
}

bool cheby_dev(State& S, const real_t* in, cplx_t* out, double Lz, hipStream_t s) {
  static const unsigned char ffc[4] = {'f', 'f', 'c', 0};
  if (S.nzc < 2) return fatal("p3dfft_cheby: nzc must be at least 2");
  bool fused = false;
  if (!forward_dev(S, in, out, ffc, s, 4.0 / Lz, &fused)) return false;
  if (fused) return true;  // the recurrence ran in the DCT-I pass
  const double scale = 1.0 / ((double)S.nx * (double)S.ny * (double)(S.nzc - 1));
  const long long ncol = (long long)S.nxhpc * S.nyc;
  if (S.stride1) launch_cheby(out, ncol, S.nzc, S.nzc, 1, scale, 4.0 / Lz, s);
  else launch_cheby(out, ncol, S.nzc, 1, ncol, scale, 4.0 / Lz, s);
  return true;
}

// ---------------------------------------------------------------------------
// multivariable transforms (build/ftran.F90:104-317, build/btran.F90:104-304):
// nv variables, variable j of the input at j*dim_in, of the output at
// j*dim_out (elements of the array's type).  Each stage covers all variables
// in one launch where the layouts allow it.
// ---------------------------------------------------------------------------
bool ensure_many(State& S, int nv) {
  if (nv <= S.nv_alloc) return true;
  if (g_live_captures > 0) {
    // a captured graph replays launches on the current work buffers
    notice(" p3dfft_amd: a _many call with more variables than any earlier one needs larger work buffers; "
           "destroy the captured graphs first (p3dfft_amd_graph_destroy)");
    return false;
  }
  if (S.w1) (void)hipFree(S.w1);
  if (S.w2) (void)hipFree(S.w2);
  S.w1 = S.w2 = nullptr;
  S.nv_alloc = 0;
  check_hip(hipMalloc(&S.w1, (size_t)S.nxhpc * S.ny * S.nz * nv * sizeof(cplx_t)), "hipMalloc(work)");
  check_hip(hipMalloc(&S.w2, (size_t)S.nxhpc * S.nyc * S.nz * nv * sizeof(cplx_t)), "hipMalloc(work)");
  S.nv_alloc = nv;
  return true;
}

bool forward_many_dev(State& S, const real_t* in, long long dim_in, cplx_t* out, long long dim_out, int nv,
                      const unsigned char* op, hipStream_t s) {
  bool none;
  const int kind = zkind((char)op[2], &none);
  if (kind < 0) return fatal(std::string("Unknown transform type: ") + (char)op[2]);
  if (!zop_len_ok(S.nz, kind)) return false;
  if (nv <= 0) return true;
  if (!ensure_many(S, nv)) return false;
  const int nx = S.nx, ny = S.ny, nz = S.nz, nxhpc = S.nxhpc, nyc = S.nyc, nzc = S.nzc;
  const long long nxy = (long long)nxhpc * nyc;
  const long long nreal = (long long)nx * ny * nz;
  const long long nspec = nxy * nzc;
  // no cuts, packed variables, no overlap: X r2c of every variable straight
  // into the output, then Y and Z in place there (as the single-variable path)
  const char* ib = (const char*)in;
  const char* ob = (const char*)out;
  const bool alias = ob < ib + (size_t)nreal * nv * sizeof(real_t) && ib < ob + (size_t)nspec * nv * sizeof(cplx_t);
  if (g_x_into_out && !S.stride1 && ny == nyc && nz == nzc && dim_in == nreal && dim_out == nspec && !alias) {
    tstart(S, 5, s);
    launch_x_r2c(in, out, (long long)ny * nz * nv, nx, nxhpc, nx, nxhpc, s);
    tstop(S, 5, s);
    ColSpec y{};
    y.in = out; y.out = out;
    y.si_in = 1; y.sl_in = nxhpc; y.so_in = nxy;
    y.si_out = 1; y.sl_out = nxhpc; y.so_out = nxy;
    y.n_inner = nxhpc; y.n_outer = nz * nv; y.n = ny; y.kind = COL_C2C; y.fwd = true;
    y.nld = ny; y.ld_lo = ny; y.ld_hi = ny; y.ld_shift = 0;
    y.nst = ny; y.st_lo = ny; y.st_shift = 0;
    y.rev = tile_order_alternate();
    tstart(S, 7, s);
    launch_col(y, s);
    tstop(S, 7, s);
    tstart(S, 8, s);
    if (!none) {
      ColSpec z{};
      z.in = out; z.out = out;
      z.n_inner = (int)nxy; z.n_outer = nv;
      z.si_in = 1; z.sl_in = nxy; z.so_in = nspec;
      z.si_out = 1; z.sl_out = nxy; z.so_out = nspec;
      z.n = nz; z.kind = kind; z.fwd = true;
      z.nld = nz; z.ld_lo = nz; z.ld_hi = nz; z.ld_shift = 0;
      z.nst = nz; z.st_lo = nz; z.st_shift = 0;
      launch_col(z, s);
    }
    tstop(S, 8, s);
    return true;
  }
  // X r2c of every variable into w1 [nv][nz][ny][nxhpc]
  tstart(S, 5, s);
  if (dim_in == nreal) {
    launch_x_r2c(in, S.w1, (long long)ny * nz * nv, nx, nxhpc, nx, nxhpc, s);
  } else {
    for (int j = 0; j < nv; j++)
      launch_x_r2c(in + (size_t)j * dim_in, S.w1 + (size_t)j * nxhpc * ny * nz, (long long)ny * nz, nx, nxhpc,
                   nx, nxhpc, s);
  }
  tstop(S, 5, s);
  // Y c2c over all nz*nv planes into w2 [nv][nz][nyc][nxhpc]
  ColSpec y{};
  y.in = S.w1; y.out = S.w2;
  y.si_in = 1; y.sl_in = nxhpc; y.so_in = (long long)nxhpc * ny;
  y.si_out = 1; y.sl_out = nxhpc; y.so_out = nxy;
  y.n_inner = nxhpc; y.n_outer = nz * nv; y.n = ny; y.kind = COL_C2C; y.fwd = true;
  y.nld = ny; y.ld_lo = ny; y.ld_hi = ny; y.ld_shift = 0;
  y.nst = nyc; y.st_lo = S.nycph; y.st_shift = ny - nyc;
  tstart(S, 7, s);
  launch_col(y, s);
  tstop(S, 7, s);
  // Z op from w2 into every output variable
  tstart(S, 8, s);
  ColSpec z{};
  z.n = nz; z.kind = kind; z.fwd = true;
  z.nld = nz; z.ld_lo = nz; z.ld_hi = nz; z.ld_shift = 0;
  z.nst = nzc; z.st_lo = S.nzcph; z.st_shift = nz - nzc;
  if (!S.stride1) {
    z.in = S.w2; z.out = out;
    z.n_inner = (int)nxy; z.n_outer = nv;
    z.si_in = 1; z.sl_in = nxy; z.so_in = nxy * nz;
    z.si_out = 1; z.sl_out = nxy; z.so_out = dim_out;
    launch_col(z, s);
  } else {
    for (int j = 0; j < nv; j++) {
      z.in = S.w2 + (size_t)j * nxy * nz; z.out = out + (size_t)j * dim_out;
      z.n_inner = nxhpc; z.n_outer = nyc;
      z.si_in = 1; z.sl_in = nxy; z.so_in = nxhpc;
      z.si_out = (long long)nzc * nyc; z.sl_out = 1; z.so_out = nzc;
      launch_col(z, s);
    }
  }
  tstop(S, 8, s);
  return true;
}

bool backward_many_dev(State& S, cplx_t* in, long long dim_in, real_t* out, long long dim_out, int nv,
                       const unsigned char* op, hipStream_t s) {
  bool none;
  const int kind = zkind((char)op[0], &none);
  if (kind < 0) return fatal(std::string(" Unknown transform type: ") + (char)op[0]);
  if (!zop_len_ok(S.nz, kind)) return false;
  if (nv <= 0) return true;
  if (!ensure_many(S, nv)) return false;
  const int nx = S.nx, ny = S.ny, nz = S.nz, nxhpc = S.nxhpc, nyc = S.nyc, nzc = S.nzc;
  const int dny = ny - nyc, dnz = nz - nzc;
  const long long nxy = (long long)nxhpc * nyc;
  // Z op: in place on the caller's arrays (overwrite, nz == nzc, XYZ layout)
  // or from the caller's arrays into w2 [nv][nz][nyc][nxhpc]
  const bool inplace = !S.stride1 && S.ow && nz == nzc;
  tstart(S, 9, s);
  ColSpec z{};
  z.n = nz; z.kind = kind; z.fwd = false;
  z.nst = nz; z.st_lo = nz; z.st_shift = 0;
  if (inplace) {
    if (!none) {
      z.in = in; z.out = in;
      z.n_inner = (int)nxy; z.n_outer = nv;
      z.si_in = 1; z.sl_in = nxy; z.so_in = dim_in;
      z.si_out = 1; z.sl_out = nxy; z.so_out = dim_in;
      z.nld = nz; z.ld_lo = nz; z.ld_hi = nz; z.ld_shift = 0;
      z.rev = tile_order_alternate();
      launch_col(z, s);
    }
  } else {
    z.nld = nz; z.ld_lo = S.nzcph; z.ld_hi = S.stride1 ? S.nzcph + dnz : nz - S.nzcph; z.ld_shift = dnz;
    if (!S.stride1) {
      z.in = in; z.out = S.w2;
      z.n_inner = (int)nxy; z.n_outer = nv;
      z.si_in = 1; z.sl_in = nxy; z.so_in = dim_in;
      z.si_out = 1; z.sl_out = nxy; z.so_out = nxy * nz;
      launch_col(z, s);
    } else {
      for (int j = 0; j < nv; j++) {
        z.in = in + (size_t)j * dim_in; z.out = S.w2 + (size_t)j * nxy * nz;
        z.n_inner = nxhpc; z.n_outer = nyc;
        z.si_in = (long long)nzc * nyc; z.sl_in = 1; z.so_in = nzc;
        z.si_out = 1; z.sl_out = nxy; z.so_out = nxhpc;
        launch_col(z, s);
      }
    }
  }
  tstop(S, 9, s);
  // Y c2c backward into w1 [nv][nz][ny][nxhpc]
  ColSpec y{};
  y.out = S.w1;
  y.si_in = 1; y.sl_in = nxhpc; y.so_in = nxy;
  y.si_out = 1; y.sl_out = nxhpc; y.so_out = (long long)nxhpc * ny;
  y.n_inner = nxhpc; y.n = ny; y.kind = COL_C2C; y.fwd = false;
  y.nld = ny; y.ld_lo = S.nycph; y.ld_hi = S.stride1 ? S.nycph + dny : ny - S.nycph; y.ld_shift = dny;
  y.nst = ny; y.st_lo = ny; y.st_shift = 0;
  tstart(S, 10, s);
  if (!inplace || dim_in == nxy * nz) {
    y.in = inplace ? in : S.w2;
    y.n_outer = nz * nv;
    launch_col(y, s);
  } else {
    for (int j = 0; j < nv; j++) {
      y.in = in + (size_t)j * dim_in;
      y.out = S.w1 + (size_t)j * nxhpc * ny * nz;
      y.n_outer = nz;
      launch_col(y, s);
    }
  }
  tstop(S, 10, s);
  // X c2r of every variable
  tstart(S, 12, s);
  if (dim_out == (long long)nx * ny * nz) {
    launch_x_c2r(S.w1, out, (long long)ny * nz * nv, nxhpc, nx, nx, nxhpc, s, tile_order_alternate());
  } else {
    for (int j = 0; j < nv; j++)
      launch_x_c2r(S.w1 + (size_t)j * nxhpc * ny * nz, out + (size_t)j * dim_out, (long long)ny * nz, nxhpc, nx,
                   nx, nxhpc, s);
  }
  tstop(S, 12, s);
  return true;
}

bool cheby_many_dev(State& S, const real_t* in, long long dim_in, cplx_t* out, long long dim_out, int nv, double Lz,
                    hipStream_t s) {
  static const unsigned char ffc[4] = {'f', 'f', 'c', 0};
  if (S.nzc < 2) return fatal("p3dfft_cheby: nzc must be at least 2");
  if (!forward_many_dev(S, in, dim_in, out, dim_out, nv, ffc, s)) return false;
  const double scale = 1.0 / ((double)S.nx * (double)S.ny * (double)(S.nzc - 1));
  const long long ncol = (long long)S.nxhpc * S.nyc;
  for (int j = 0; j < nv; j++) {
    cplx_t* o = out + (size_t)j * dim_out;
    if (S.stride1) launch_cheby(o, ncol, S.nzc, S.nzc, 1, scale, 4.0 / Lz, s);
    else launch_cheby(o, ncol, S.nzc, 1, ncol, scale, 4.0 / Lz, s);
  }
  return true;
}

// The reference prints these and continues (build/ftran.F90:123-129).
static bool check_many_dims(long long dim_in, long long dim_out, bool fwd) {
  const long long nr = (long long)g.nx * g.jisize * g.kjsize;
  const long long nc = (long long)g.nzc * g.jjsize * g.iisize;
  const long long need_in = fwd ? nr : nc, need_out = fwd ? nc : nr;
  if (dim_in < need_in) {
    notice(std::string(fwd ? " ftran" : " btran") + " error: input array dimensions are too low: " +
           std::to_string(dim_in) + " while expecting " + std::to_string(need_in));
    return false;
  }
  if (dim_out < need_out) {
    notice(std::string(fwd ? " ftran" : " btran") + " error: output array dimensions are too low: " +
           std::to_string(dim_out) + " while expecting " + std::to_string(need_out));
    return false;
  }
  return true;
}

bool require_setup() {
  if (!g.set) {
    notice(" P3DFFT error: call setup before other routines");
    return false;
  }
  return true;
}

// the whole-transform entry points run the 1x1 grid only
static bool require_full() {
  if (!require_setup()) return false;
  if (g.share) return fatal(" p3dfft_amd: a share-mode setup runs the p3dfft_share_* stages only");
  return true;
}

// ---------------------------------------------------------------------------
// plan-time measurement (the counterpart of the reference's FFTW_MEASURE
// planning, build/fft_spec.F90:101-108): at setup, time candidate plans of
// every transform length on the setup's own grid and keep the fastest
// ---------------------------------------------------------------------------
static int g_plan_mode = 0;  // 0: default plans; 1: measure at setup
static std::string g_plan_report;

}  // namespace p3amd
extern "C" int p3dfft_amd_fill_normal(p3amd::real_t* p, long long n, unsigned long long seed, void* stream);
namespace p3amd {

// ms per forward+backward pair on scratch arrays (one untimed pair first)
static double time_pair(State& S, const real_t* rin, cplx_t* cspec, real_t* rout, hipStream_t s) {
  static const unsigned char fop[4] = {'f', 'f', 't', 0}, bop[4] = {'t', 'f', 'f', 0};
  // the stage timers are off while candidates run and the events are
  // released on every exit, also when a candidate throws
  struct Guard {
    State& S;
    bool tm;
    hipEvent_t a = nullptr, b = nullptr;
    ~Guard() {
      if (a) (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
      S.timing = tm;
    }
  } gd{S, S.timing};
  S.timing = false;
  check_hip(hipEventCreate(&gd.a), "hipEventCreate");
  check_hip(hipEventCreate(&gd.b), "hipEventCreate");
  hipEvent_t a = gd.a, b = gd.b;
  auto run = [&](int reps) {
    check_hip(hipEventRecord(a, s), "hipEventRecord");
    for (int i = 0; i < reps; i++) {
      forward_dev(S, rin, cspec, fop, s);
      backward_dev(S, cspec, rout, bop, s);
    }
    check_hip(hipEventRecord(b, s), "hipEventRecord");
    check_hip(hipEventSynchronize(b), "hipEventSynchronize");
    float ms = 0;
    check_hip(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
    return (double)ms / reps;
  };
  const double t1 = run(1);
  const int reps = std::max(2, std::min(50, (int)(5.0 / std::max(t1, 1e-3))));
  return run(reps);
}

static void measure_plans(State& S) {
  const size_t nr = (size_t)S.nx * S.ny * S.nz;
  const size_t nc = (size_t)S.nxhpc * S.nyc * S.nzc;
  real_t *rin = nullptr, *rout = nullptr;
  cplx_t* cspec = nullptr;
  hipStream_t s = S.stream;
  check_hip(hipMalloc(&rin, nr * sizeof(real_t)), "hipMalloc(measure)");
  check_hip(hipMalloc(&rout, nr * sizeof(real_t)), "hipMalloc(measure)");
  check_hip(hipMalloc(&cspec, nc * sizeof(cplx_t)), "hipMalloc(measure)");
  (void)p3dfft_amd_fill_normal(rin, (long long)nr, 12345, (void*)s);
  std::vector<int> lens;
  // lengths above max_len() keep their multi-pass plans
  for (int n : {S.nx, S.ny, S.nz})
    if (n <= max_len()) lens.push_back(n);
  std::sort(lens.begin(), lens.end());
  lens.erase(std::unique(lens.begin(), lens.end()), lens.end());
  std::string rep;
  for (int n : lens) {
    struct Cand {
      std::vector<int> f;
      bool ct;
      int tile, seg;
    };
    std::vector<Cand> cands;
    const std::vector<int> f0 = default_factors(n);
    if (len_has_ct(n)) cands.push_back({f0, true, 64 * 1024, 128});
    for (int tile : {32 * 1024, 64 * 1024, 128 * 1024})
      for (int seg : {64, 128, 256}) cands.push_back({f0, false, tile, seg});
    std::vector<int> f1(f0.rbegin(), f0.rend());
    if (f1 != f0) cands.push_back({f1, false, 64 * 1024, 128});
    std::vector<int> f2 = f0;
    std::sort(f2.begin(), f2.end());
    if (f2 != f0 && f2 != f1) cands.push_back({f2, false, 64 * 1024, 128});
    double best = 1e30;
    size_t bi = 0;
    for (size_t i = 0; i < cands.size(); i++) {
      set_len_plan(n, cands[i].f, cands[i].ct, cands[i].tile, cands[i].seg);
      try {
        const double t = time_pair(S, rin, cspec, rout, s);
        if (t < best) { best = t; bi = i; }
      } catch (const Error&) {
        // a candidate whose tiles do not fit is refused before any launch
        check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
      }
    }
    set_len_plan(n, cands[bi].f, cands[bi].ct, cands[bi].tile, cands[bi].seg);
    rep += "n=" + std::to_string(n) + ": " + len_describe(n) + " (" + std::to_string(cands.size()) +
           " candidates); ";
  }
  // schedule switches: tile order of consecutive passes, X r2c into the
  // output, tile order of the Z pass (XCD-aware or dispatch order)
  double best = 1e30;
  int balt = 1, bxin = 1, bflat = 0;
  for (int alt = 1; alt >= 0; alt--)
    for (int xin = 1; xin >= 0; xin--)
      for (int fo = 0; fo <= 1; fo++) {
        set_tile_order(alt != 0);
        g_x_into_out = xin != 0;
        set_flat_order(fo);
        const double t = time_pair(S, rin, cspec, rout, s);
        if (t < best) { best = t; balt = alt; bxin = xin; bflat = fo; }
      }
  set_tile_order(balt != 0);
  g_x_into_out = bxin != 0;
  set_flat_order(bflat);
  rep += std::string("tile order ") + (balt ? "alternating" : "fixed") + "; X r2c " +
         (bxin ? "into the output" : "into the work buffer") + "; Z tiles " +
         (bflat ? "in dispatch order" : "XCD-aware") + "; fwd+bwd " + std::to_string(best) + " ms";
  g_plan_report = rep;
  check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
  (void)hipFree(rin);
  (void)hipFree(rout);
  (void)hipFree(cspec);
}

void release_state() {
  timers_resolve();
  share_release();
  g.share = false;
  for (auto e : g.evpool) (void)hipEventDestroy(e);
  g.evpool.clear();
  if (g.w1) (void)hipFree(g.w1);
  if (g.w2) (void)hipFree(g.w2);
  if (g.spec_partial) (void)hipFree(g.spec_partial);
  g.spec_partial = nullptr;
  g.spec_partial_bytes = 0;
  if (g.stage_a) (void)hipFree(g.stage_a);
  if (g.stage_b) (void)hipFree(g.stage_b);
  g.w1 = g.w2 = nullptr;
  g.nv_alloc = 0;
  g.stage_a = g.stage_b = nullptr;
  g.stage_bytes = 0;
  free_long_scratch();
  g.set = false;
}

// The power spectrum's view of the current setup's spectrum: dims d (first
// fastest) and fold lengths f (0 = the axis is not folded), after the kmax
// check of the drivers (driver_spec.F90:272-275).
struct SpecDims {
  int d[3], f[3];
};
static bool spectrum_dims(int kmax, SpecDims* s) {
  if (g.stride1) {
    s->d[0] = g.nzc; s->d[1] = g.nyc; s->d[2] = g.nxhpc; s->f[0] = g.nz; s->f[1] = g.ny; s->f[2] = 0;
  } else {
    s->d[0] = g.nxhpc; s->d[1] = g.nyc; s->d[2] = g.nzc; s->f[0] = 0; s->f[1] = g.ny; s->f[2] = g.nz;
  }
  double need = 0;
  for (int i = 0; i < 3; i++) need += (double)(s->d[i] - 1) * (s->d[i] - 1);
  if (kmax < 0 || 4.0 * kmax * kmax < need) {
    set_error("Error: kmax is less than it should be (" + std::to_string(kmax) + ")");
    return false;
  }
  return true;
}

// Static device scratch of the allocation-free entry points: each GPU the
// library runs on holds its own copy of this array until the process ends.
// Host code must not use the array's name as a device address: its address
// on each device comes from hipGetSymbolAddress, taken once per device with
// that device current (at p3dfft_setup, or at the first call on a device)
// and kept in g_scratch_dev.
constexpr size_t kScratchBytes = 16u << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_dev[kMaxDevices] = {};

// the scratch array's address on the current device (null when it cannot be
// had, with the error recorded if `record`)
double* scratch_ptr(bool record = true) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    if (record) set_error("p3dfft_amd: no usable current device for the scratch array");
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      if (record) set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_dev[dev] = (double*)p;
  }
  return g_scratch_dev[dev];
}

}  // namespace p3amd

// ===========================================================================
// exported entry points
// ===========================================================================
#define GUARD_BEGIN try {
#define GUARD_END                                  \
  }                                                \
  catch (const Error& e) {                         \
    fatal(std::string("p3dfft_amd: ") + e.msg);    \
  }

extern "C" {

// build/setup.F90:83-104 (BIND(C) p3dfft_setup) and 107-605
void p3dfft_setup(int* dims, int* nx, int* ny, int* nz, int* comm, int* nxc, int* nyc, int* nzc, int* ow,
                  int* memsize) {
  (void)comm;  // the single GPU is rank 0 of a one-rank communicator
  GUARD_BEGIN
  if (*nx <= 0 || *ny <= 0 || *nz <= 0) {
    fatal(" Invalid dimensions :" + std::to_string(*nx) + " " + std::to_string(*ny) + " " + std::to_string(*nz));
    return;
  }
  if (g.set) {
    fatal(" P3DFFT Setup error: the problem is already initialized. \n"
          " Currently multiple setups not supported.\n"
          " Quit the library using p3dfft_clean before initializing another setup");
    return;
  }
  if (dims[0] <= 0 || dims[1] <= 0 || dims[0] * dims[1] != 1) {
    fatal(" Invalid processor geometry: " + std::to_string(dims[0]) + " " + std::to_string(dims[1]) +
          " for 1 tasks");
    return;
  }
  if (*nxc < 1 || *nxc > *nx || *nyc < 1 || *nyc > *ny || *nzc < 1 || *nzc > *nz) {
    fatal(" Invalid cut dimensions :" + std::to_string(*nxc) + " " + std::to_string(*nyc) + " " +
          std::to_string(*nzc));
    return;
  }
  for (int i = 0; i < 12; i++) g.timers[i] = 0.0;
  g.nx = *nx; g.ny = *ny; g.nz = *nz;
  g.nxc = *nxc; g.nyc = *nyc; g.nzc = *nzc;
  g.ow = (*ow != 0);
  g.stride1 = g.stride1_next;
  compute_geometry(g, dims, 0, 0);
  if (g.stride1) printf(" Using stride-1 layout\n");
  const int lim = max_len_any();
  // the Z length of a DCT-I / DST-I op is checked when that op is requested
  if (g.nx > lim || g.ny > lim || g.nz > lim) {
    fatal(" p3dfft_amd: grid dimension exceeds the supported transform length " + std::to_string(lim));
    return;
  }
  check_hip(hipMalloc(&g.w1, (size_t)g.nxhpc * g.ny * g.nz * sizeof(cplx_t)), "hipMalloc(work)");
  check_hip(hipMalloc(&g.w2, (size_t)g.nxhpc * g.nyc * g.nz * sizeof(cplx_t)), "hipMalloc(work)");
  g.nv_alloc = 1;
  // the scratch array's address on this device, for the allocation-free
  // entry points (a failure is reported by those entry points)
  (void)scratch_ptr(false);
  // build the 1D plans (twiddles are generated on the GPU here, once)
  plan_len(g.nx);
  plan_len(g.ny);
  plan_len(g.nz);
  if (g_plan_mode == 1) measure_plans(g);
  g.set = true;
  memsize[0] = g.memsize[0];
  memsize[1] = g.memsize[1];
  memsize[2] = g.memsize[2];
  GUARD_END
}

// build/module.F90:214-273
void p3dfft_get_dims(int* start, int* end, int* size, int* conf) {
  if (!require_setup()) return;
  if (*conf == 1) {
    start[0] = 1; end[0] = g.nx; size[0] = g.nx;
    start[1] = g.jistart; end[1] = g.jiend; size[1] = g.jisize;
    start[2] = g.kjstart; end[2] = g.kjend; size[2] = g.kjsize;
  } else if (*conf == 2) {
    if (g.stride1) {
      start[2] = g.iistart; end[2] = g.iiend; size[2] = g.iisize;
      start[1] = g.jjstart; end[1] = g.jjend; size[1] = g.jjsize;
      start[0] = 1; end[0] = g.nzc; size[0] = g.nzc;
    } else {
      start[0] = g.iistart; end[0] = g.iiend; size[0] = g.iisize;
      start[1] = g.jjstart; end[1] = g.jjend; size[1] = g.jjsize;
      start[2] = 1; end[2] = g.nzc; size[2] = g.nzc;
    }
  } else if (*conf == 3) {
    for (int i = 0; i < 3; i++) { start[i] = 0; end[i] = g.memsize[i]; size[i] = g.memsize[i]; }
  }
}

// build/ftran.F90:469-485 (host arrays; copied to and from HBM inside the call)
// The sample drivers' power spectrum (sample/C/driver_spec.c:298-383,
// sample/FORTRAN/driver_spec.F90:263-340) of the current setup's spectrum B
// (conf-2 layout, HBM): E[ik] = sum of k2*|B|^2 over the stored elements with
// ik = int(sqrt(k2) + 1/2), k2 = kx^2 + ky^2 + kz^2, kx the stored x index,
// ky and kz folded (n - j above n/2).  E: kmax+1 doubles in HBM.  The kmax
// check is driver_spec.F90:272-275.  Asynchronous on `stream`; the first call
// for a larger kmax allocates its per-plane scratch.
int p3dfft_power_spectrum_dev(const real_t* B, int kmax, double* E, void* stream) {
  if (!require_full()) return -1;
  GUARD_BEGIN
  SpecDims s;
  if (!spectrum_dims(kmax, &s)) return -1;
  const int d0 = s.d[0], d1 = s.d[1], d2 = s.d[2], f0 = s.f[0], f1 = s.f[1], f2 = s.f[2];
  const size_t pb = spectrum_partial_bytes(d2, kmax);
  if (pb > g.spec_partial_bytes) {
    if (g.spec_partial) (void)hipFree(g.spec_partial);
    g.spec_partial = nullptr;
    g.spec_partial_bytes = 0;
    check_hip(hipMalloc(&g.spec_partial, pb), "hipMalloc(spectrum scratch)");
    g.spec_partial_bytes = pb;
  }
  launch_spectrum((const cplx_t*)B, d0, d1, d2, f0, f1, f2, kmax, g.spec_partial, E, (hipStream_t)stream);
  return 0;
  GUARD_END
  return -1;
}

// Allocation-free variant of p3dfft_power_spectrum_dev (no counterpart in the
// reference; for callers that cannot allocate device memory: code under
// stream capture, serving loops, harnesses that forbid it).  The partial
// shell sums go to the library's static per-GPU scratch array instead of an
// allocation; a setup whose partials need more than its 16 MiB is refused.
// The sums do not depend on the array's previous contents.  Calls that
// overlap on one GPU share the array and race.
int p3dfft_power_spectrum_dev_noalloc(const real_t* B, int kmax, double* E, void* stream) {
  if (!require_full()) return -1;
  GUARD_BEGIN
  SpecDims s;
  if (!spectrum_dims(kmax, &s)) return -1;
  const size_t pb = spectrum_partial_bytes(s.d[2], kmax);
  if (pb > kScratchBytes) {
    set_error("p3dfft_power_spectrum_dev_noalloc: the partial sums need " + std::to_string(pb) +
              " bytes, above the " + std::to_string(kScratchBytes) + "-byte scratch array; use "
              "p3dfft_power_spectrum_dev");
    return -1;
  }
  double* partial = scratch_ptr();
  if (!partial) return -1;
  launch_spectrum((const cplx_t*)B, s.d[0], s.d[1], s.d[2], s.f[0], s.f[1], s.f[2], kmax, partial, E,
                  (hipStream_t)stream);
  return 0;
  GUARD_END
  return -1;
}

// The same on a host spectrum, E returned to the host.
int p3dfft_power_spectrum(const real_t* B, int kmax, double* E) {
  if (!require_full()) return -1;
  GUARD_BEGIN
  ensure_staging();
  double* dE = nullptr;
  check_hip(hipMalloc(&dE, (size_t)(kmax + 1) * sizeof(double)), "hipMalloc(spectrum)");
  check_hip(hipMemcpy(g.stage_b, B, spec_elems() * sizeof(cplx_t), hipMemcpyHostToDevice), "hipMemcpy H2D");
  const int rc = p3dfft_power_spectrum_dev((const real_t*)g.stage_b, kmax, dE, g.stream);
  if (rc == 0)
    check_hip(hipMemcpy(E, dE, (size_t)(kmax + 1) * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy D2H");
  (void)hipFree(dE);
  return rc;
  GUARD_END
  return -1;
}

// staging for nv variables of the host-array _many entry points
static void ensure_staging_many(size_t in_bytes, size_t out_bytes) {
  const size_t bytes = std::max(in_bytes, out_bytes);
  if (g.stage_bytes >= bytes) return;
  if (g.stage_a) (void)hipFree(g.stage_a);
  if (g.stage_b) (void)hipFree(g.stage_b);
  g.stage_a = g.stage_b = nullptr;
  g.stage_bytes = 0;
  check_hip(hipMalloc(&g.stage_a, bytes), "hipMalloc(staging)");
  check_hip(hipMalloc(&g.stage_b, bytes), "hipMalloc(staging)");
  g.stage_bytes = bytes;
}

void p3dfft_ftran_r2c(real_t* A, real_t* B, unsigned char* op) {
  if (!require_full()) return;
  GUARD_BEGIN
  ensure_staging();
  check_hip(hipMemcpy(g.stage_a, A, real_elems() * sizeof(real_t), hipMemcpyHostToDevice), "hipMemcpy H2D");
  if (!forward_dev(g, (const real_t*)g.stage_a, (cplx_t*)g.stage_b, op, g.stream)) return;
  check_hip(hipMemcpy(B, g.stage_b, spec_elems() * sizeof(cplx_t), hipMemcpyDeviceToHost), "hipMemcpy D2H");
  GUARD_END
}

// build/ftran.F90:787-814: the X r2c stage alone, real (nx, jisize, kjsize)
// to complex (nxhp, jisize, kjsize) (all nx/2+1 modes, no cut), unnormalised
int p3dfft_ftran_r2c_1d_dev(const real_t* in, real_t* out, void* stream) {
  if (!require_full()) return -1;
  GUARD_BEGIN
  launch_x_r2c(in, (cplx_t*)out, (long long)g.jisize * g.kjsize, g.nx, g.nxhp, g.nx, g.nxhp, (hipStream_t)stream);
  return 0;
  GUARD_END
  return -1;
}

void p3dfft_ftran_r2c_1d(real_t* A, real_t* B) {
  if (!require_full()) return;
  GUARD_BEGIN
  // the staging buffers grow once to the (nxhp, jisize, kjsize) output
  const size_t nout = (size_t)g.nxhp * g.jisize * g.kjsize;
  ensure_staging_many(real_elems() * sizeof(real_t), nout * sizeof(cplx_t));
  check_hip(hipMemcpy(g.stage_a, A, real_elems() * sizeof(real_t), hipMemcpyHostToDevice), "hipMemcpy H2D");
  launch_x_r2c((const real_t*)g.stage_a, (cplx_t*)g.stage_b, (long long)g.jisize * g.kjsize, g.nx, g.nxhp, g.nx,
               g.nxhp, g.stream);
  check_hip(hipMemcpy(B, g.stage_b, nout * sizeof(cplx_t), hipMemcpyDeviceToHost), "hipMemcpy D2H");
  GUARD_END
}

// build/btran.F90:377-392
void p3dfft_btran_c2r(real_t* A, real_t* B, unsigned char* op) {
  if (!require_full()) return;
  GUARD_BEGIN
  ensure_staging();
  check_hip(hipMemcpy(g.stage_a, A, spec_elems() * sizeof(cplx_t), hipMemcpyHostToDevice), "hipMemcpy H2D");
  if (!backward_dev(g, (cplx_t*)g.stage_a, (real_t*)g.stage_b, op, g.stream)) return;
  check_hip(hipMemcpy(B, g.stage_b, real_elems() * sizeof(real_t), hipMemcpyDeviceToHost), "hipMemcpy D2H");
  GUARD_END
}

// build/ftran.F90:359-379
void p3dfft_cheby(real_t* A, real_t* B, real_t* Lz) {
  if (!require_full()) return;
  GUARD_BEGIN
  ensure_staging();
  check_hip(hipMemcpy(g.stage_a, A, real_elems() * sizeof(real_t), hipMemcpyHostToDevice), "hipMemcpy H2D");
  if (!cheby_dev(g, (const real_t*)g.stage_a, (cplx_t*)g.stage_b, (double)*Lz, g.stream)) return;
  check_hip(hipMemcpy(B, g.stage_b, spec_elems() * sizeof(cplx_t), hipMemcpyDeviceToHost), "hipMemcpy D2H");
  GUARD_END
}


// build/ftran.F90:84-100 (BIND(C) p3dfft_ftran_r2c_many) and 104-317
void p3dfft_ftran_r2c_many(real_t* A, int* dim_in, real_t* B, int* dim_out, int* nv, unsigned char* op) {
  if (!require_full()) return;
  GUARD_BEGIN
  // the reference prints and continues; continuing here would index past the
  // staging buffers, so the call returns
  if (!check_many_dims(*dim_in, *dim_out, true)) return;
  const size_t ib = (size_t)*dim_in * *nv * sizeof(real_t), ob = (size_t)*dim_out * *nv * sizeof(cplx_t);
  ensure_staging_many(ib, ob);
  check_hip(hipMemcpy(g.stage_a, A, ib, hipMemcpyHostToDevice), "hipMemcpy H2D");
  if (!forward_many_dev(g, (const real_t*)g.stage_a, *dim_in, (cplx_t*)g.stage_b, *dim_out, *nv, op, g.stream)) return;
  check_hip(hipMemcpy(B, g.stage_b, ob, hipMemcpyDeviceToHost), "hipMemcpy D2H");
  GUARD_END
}

// build/btran.F90:84-100 and 104-304
void p3dfft_btran_c2r_many(real_t* A, int* dim_in, real_t* B, int* dim_out, int* nv, unsigned char* op) {
  if (!require_full()) return;
  GUARD_BEGIN
  if (!check_many_dims(*dim_in, *dim_out, false)) return;
  const size_t ib = (size_t)*dim_in * *nv * sizeof(cplx_t), ob = (size_t)*dim_out * *nv * sizeof(real_t);
  ensure_staging_many(ib, ob);
  check_hip(hipMemcpy(g.stage_a, A, ib, hipMemcpyHostToDevice), "hipMemcpy H2D");
  if (!backward_many_dev(g, (cplx_t*)g.stage_a, *dim_in, (real_t*)g.stage_b, *dim_out, *nv, op, g.stream)) return;
  check_hip(hipMemcpy(B, g.stage_b, ob, hipMemcpyDeviceToHost), "hipMemcpy D2H");
  GUARD_END
}

// build/ftran.F90:321-354 (the reference re-runs the forward per variable
// inside p3dfft_cheby; the result is the same)
void p3dfft_cheby_many(real_t* A, int* dim_in, real_t* B, int* dim_out, int* nv, real_t* Lz) {
  if (!require_full()) return;
  GUARD_BEGIN
  if (!check_many_dims(*dim_in, *dim_out, true)) return;
  const size_t ib = (size_t)*dim_in * *nv * sizeof(real_t), ob = (size_t)*dim_out * *nv * sizeof(cplx_t);
  ensure_staging_many(ib, ob);
  check_hip(hipMemcpy(g.stage_a, A, ib, hipMemcpyHostToDevice), "hipMemcpy H2D");
  if (!cheby_many_dev(g, (const real_t*)g.stage_a, *dim_in, (cplx_t*)g.stage_b, *dim_out, *nv, (double)*Lz, g.stream))
    return;
  check_hip(hipMemcpy(B, g.stage_b, ob, hipMemcpyDeviceToHost), "hipMemcpy D2H");
  GUARD_END
}

// build/module.F90:301-424
void p3dfft_clean(void) {
  if (g_live_captures > 0) {
    // the graphs would replay launches on freed work buffers
    notice(" p3dfft_clean: destroy the captured graphs first (p3dfft_amd_graph_destroy)");
    return;
  }
  GUARD_BEGIN
  release_state();
  GUARD_END
}

// build/module.F90:709-744
void get_timers(double* t) {
  GUARD_BEGIN
  timers_resolve();
  GUARD_END
  for (int i = 0; i < 12; i++) t[i] = g.timers[i];
}

void set_timers(void) {
  GUARD_BEGIN
  timers_resolve();
  GUARD_END
  for (int i = 0; i < 12; i++) g.timers[i] = 0.0;
}

// ---------------------------------------------------------------------------
// extensions (no counterpart in the reference)
// ---------------------------------------------------------------------------

// Device-pointer variants: in/out are HBM addresses, the work is queued on
// `stream` (a hipStream_t, or null for the default stream) and the call
// returns without synchronising.
int p3dfft_ftran_r2c_dev(const real_t* in, real_t* out, const unsigned char* op, void* stream) {
  if (!require_full()) return -1;
  GUARD_BEGIN
  return forward_dev(g, in, (cplx_t*)out, op, (hipStream_t)stream) ? 0 : -1;
  GUARD_END
  return -1;
}

int p3dfft_btran_c2r_dev(real_t* in, real_t* out, const unsigned char* op, void* stream) {
  if (!require_full()) return -1;
  GUARD_BEGIN
  return backward_dev(g, (cplx_t*)in, out, op, (hipStream_t)stream) ? 0 : -1;
  GUARD_END
  return -1;
}

int p3dfft_cheby_dev(const real_t* in, real_t* out, real_t Lz, void* stream) {
  if (!require_full()) return -1;
  GUARD_BEGIN
  return cheby_dev(g, in, (cplx_t*)out, (double)Lz, (hipStream_t)stream) ? 0 : -1;
  GUARD_END
  return -1;
}

int p3dfft_ftran_r2c_many_dev(const real_t* in, long long dim_in, real_t* out, long long dim_out, int nv,
                              const unsigned char* op, void* stream) {
  if (!require_full()) return -1;
  GUARD_BEGIN
  if (!check_many_dims(dim_in, dim_out, true)) return -1;
  return forward_many_dev(g, in, dim_in, (cplx_t*)out, dim_out, nv, op, (hipStream_t)stream) ? 0 : -1;
  GUARD_END
  return -1;
}

int p3dfft_btran_c2r_many_dev(real_t* in, long long dim_in, real_t* out, long long dim_out, int nv,
                              const unsigned char* op, void* stream) {
  if (!require_full()) return -1;
  GUARD_BEGIN
  if (!check_many_dims(dim_in, dim_out, false)) return -1;
  return backward_many_dev(g, (cplx_t*)in, dim_in, out, dim_out, nv, op, (hipStream_t)stream) ? 0 : -1;
  GUARD_END
  return -1;
}

int p3dfft_cheby_many_dev(const real_t* in, long long dim_in, real_t* out, long long dim_out, int nv, real_t Lz,
                          void* stream) {
  if (!require_full()) return -1;
  GUARD_BEGIN
  if (!check_many_dims(dim_in, dim_out, true)) return -1;
  return cheby_many_dev(g, in, dim_in, (cplx_t*)out, dim_out, nv, (double)Lz, (hipStream_t)stream) ? 0 : -1;
  GUARD_END
  return -1;
}

// Select the (Z,Y,X) complex layout for the next setup (the reference's
// compile-time STRIDE1, build/module.F90:245-265).
void p3dfft_amd_set_stride1(int on) { g.stride1_next = (on != 0); }

// 0: fatal errors print and terminate the process (the reference's MPI_Abort);
// 1: they print, record the message and return.
void p3dfft_amd_set_error_mode(int mode) { g_err_mode = mode; }
const char* p3dfft_amd_last_error(void) { return g_last_error.c_str(); }
void p3dfft_amd_clear_error(void) { g_last_error.clear(); }

// Bytes of the X/Y intermediate processed per chunk (0 = no chunking).
void p3dfft_amd_set_chunk_bytes(size_t bytes) { g.chunk_bytes = bytes; }

// Stage timers on/off (default on).
void p3dfft_amd_set_timing(int on) {
  if (!on) timers_resolve();
  g.timing = (on != 0);
}

int p3dfft_amd_precision(void) { return (int)sizeof(real_t); }

// Use the compile-time specialised kernels for the lengths that have them
// (default 1); 0 runs every length through the run-time-generic kernels.
void p3dfft_amd_set_specialised(int on) { set_specialised(on != 0); }
void p3dfft_amd_set_tile_order(int alternate) { set_tile_order(alternate != 0); }
void p3dfft_amd_set_wide_offsets(int on) { set_wide_offsets(on != 0); }
void p3dfft_amd_set_flat_order(int mode) { set_flat_order(mode); }
void p3dfft_amd_set_plane_kernels(int mode) { set_plane_kernels(mode); }
void p3dfft_amd_set_x_wave(int on) { set_x_wave(on != 0); }
void p3dfft_amd_set_dct_direct(int on) { set_dct_direct(on != 0); }
void p3dfft_amd_set_col_pipe(int on) { set_col_pipe(on); }
void p3dfft_amd_set_x_into_out(int on) { g_x_into_out = (on != 0); }

// 0 (default): the default plan of every length; 1: from the next setup on,
// p3dfft_setup measures candidate plans (radix schedule, specialised or
// generic kernels, tile shape) of each length and the schedule switches on
// its grid and keeps the fastest (plans persist for the process).
void p3dfft_amd_set_plan_mode(int mode) { g_plan_mode = mode == 1 ? 1 : 0; }

// The choices of the last measured setup as text; returns its length (the
// text is truncated to len - 1 characters).
int p3dfft_amd_plan_report(char* buf, int len) {
  const std::string& r = g_plan_report;
  if (buf && len > 0) {
    const int n = std::min((int)r.size(), len - 1);
    std::memcpy(buf, r.data(), (size_t)n);
    buf[n] = 0;
  }
  return (int)r.size();
}

// Back to the default plans of every length and the default schedule
// switches (no setup may be active).
int p3dfft_amd_reset_plans(void) {
  if (g.set) {
    set_error("p3dfft_amd_reset_plans: call p3dfft_clean first");
    return -1;
  }
  if (g_live_graphs > 0) {
    set_error("p3dfft_amd_reset_plans: destroy the Poisson graphs first");
    return -1;
  }
  try {
    free_lens();
    free_long();
    set_tile_order(true);
    set_flat_order(2);
    g_x_into_out = true;
    g_plan_report.clear();
    return 0;
  } catch (const Error& e) {
    set_error(e.msg);
    return -1;
  }
}

// Stream capture of the library's asynchronous (_dev) calls into a HIP
// graph: every transform issues kernel launches only (kernel attributes are
// set when a length is planned and the work buffers exist from setup), so a
// captured ftran/btran sequence replays with one launch.  The stage timers
// are off between begin and end (their events cannot be read back from a
// graph); calls that allocate on first use (power spectrum scratch, staging
// of host arrays) must have been made once before the capture.
struct CapturedGraph {
  hipGraphExec_t exec = nullptr;
};
static bool g_cap_timing = true;

int p3dfft_amd_capture_begin(void* stream) {
  if (!stream) {
    set_error("p3dfft_amd_capture_begin: capture needs a created stream, not the null stream");
    return -1;
  }
  const hipError_t e = hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    set_error(std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
    return -1;
  }
  g_cap_timing = g.timing;
  g.timing = false;
  return 0;
}

int p3dfft_amd_capture_end(void* stream, void** graph) {
  g.timing = g_cap_timing;
  if (graph) *graph = nullptr;
  hipGraph_t gr = nullptr;
  hipError_t e = hipStreamEndCapture((hipStream_t)stream, &gr);
  if (e != hipSuccess || !gr) {
    set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    return -1;
  }
  hipGraphExec_t ex = nullptr;
  e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
  (void)hipGraphDestroy(gr);
  if (e != hipSuccess) {
    set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return -1;
  }
  CapturedGraph* c = new CapturedGraph;
  c->exec = ex;
  g_live_graphs++;
  g_live_captures++;
  if (graph) *graph = c;
  return 0;
}

int p3dfft_amd_graph_launch(void* graph, void* stream) {
  if (!graph) {
    set_error("p3dfft_amd_graph_launch: null graph");
    return -1;
  }
  const hipError_t e = hipGraphLaunch(((CapturedGraph*)graph)->exec, (hipStream_t)stream);
  if (e != hipSuccess) {
    set_error(std::string("hipGraphLaunch: ") + hipGetErrorString(e));
    return -1;
  }
  return 0;
}

int p3dfft_amd_graph_destroy(void* graph) {
  if (!graph) return 0;
  CapturedGraph* c = (CapturedGraph*)graph;
  if (c->exec) (void)hipGraphExecDestroy(c->exec);
  delete c;
  g_live_graphs--;
  g_live_captures--;
  return 0;
}

// Device memory helpers for callers without their own HIP runtime bindings.
void* p3dfft_amd_malloc(size_t bytes) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    set_error("hipMalloc failed");
    return nullptr;
  }
  return p;
}
int p3dfft_amd_free(void* p) { return hipFree(p) == hipSuccess ? 0 : -1; }
int p3dfft_amd_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
}
int p3dfft_amd_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}
int p3dfft_amd_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice) == hipSuccess ? 0 : -1;
}
int p3dfft_amd_memset(void* dst, int value, size_t bytes) {
  return hipMemset(dst, value, bytes) == hipSuccess ? 0 : -1;
}
int p3dfft_amd_synchronize(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : -1; }
int p3dfft_amd_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

// Streaming device copy of `bytes` (a multiple of 16) with the library's copy
// kernel: the achievable-HBM-bandwidth reference for roofline figures.
int p3dfft_amd_stream_copy(const void* in, void* out, size_t bytes, void* stream) {
  if (bytes % 16) return -1;
  try {
    launch_copy16(in, out, (long long)(bytes / 16), (hipStream_t)stream);
  } catch (const Error& e) {
    set_error(e.msg);
    return -1;
  }
  return 0;
}

// Events on the null stream, for callers timing their own regions.
void* p3dfft_amd_event_create(void) {
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return (void*)e;
}
int p3dfft_amd_event_record(void* e, void* stream) {
  return hipEventRecord((hipEvent_t)e, (hipStream_t)stream) == hipSuccess ? 0 : -1;
}
float p3dfft_amd_event_elapsed_ms(void* a, void* b) {
  float ms = -1.f;
  if (hipEventSynchronize((hipEvent_t)b) != hipSuccess) return -1.f;
  if (hipEventElapsedTime(&ms, (hipEvent_t)a, (hipEvent_t)b) != hipSuccess) return -1.f;
  return ms;
}
int p3dfft_amd_event_destroy(void* e) { return hipEventDestroy((hipEvent_t)e) == hipSuccess ? 0 : -1; }

}  // extern "C"
