// engine.hpp -- host-side interface of the gfx950 FFT engine (kernels live in
// engine.hip).  One build of the library instantiates one real type:
// double by default, float when P3DFFT_SINGLE is defined (the reference's
// SINGLE_PREC switch, build/module.F90:91-99).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "fft_device.hpp"

namespace p3amd {

#ifdef P3DFFT_SINGLE
typedef float real_t;
#else
typedef double real_t;
#endif
typedef cpx<real_t> cplx_t;
typedef TwT<real_t>::type tw_t;

// Library error: carries a message in the reference's wording.
struct Error {
  std::string msg;
  int code;
};

void check_hip(hipError_t e, const char* what);

// A 1D length with its stage schedule and device twiddle table.
struct Len1D {
  LenDesc d;
  tw_t* tw = nullptr;  // device table: stage twiddles then (if needed) W_n^e
  tw_t* twf = nullptr; // W_n^e, e < n, for the wave kernels of n = 512
  bool has_generic = false;
  // plan choices (plan-time measurement may change them): specialised
  // kernels allowed; data bytes per tile and bytes of adjacent columns per
  // row segment of the run-time-generic kernels
  bool ct_ok = true;
  int tile_bytes = 64 * 1024;
  int col_seg = 128;
};

// Length cache (one per library instance).
const Len1D& get_len(int n);
void free_lens();
// the default radix schedule of a length
std::vector<int> default_factors(int n);
// replace the plan of length n (radix schedule and tile choices); the device
// is synchronised before the old twiddle table is freed
void set_len_plan(int n, const std::vector<int>& factors, bool ct_ok, int tile_bytes, int col_seg);
// a specialised kernel set exists for length n with its default schedule
bool len_has_ct(int n);
// "radix 8x8x8, specialised" / "radix 5x5x5x8, generic, tile 64 KiB, segment 128 B"
std::string len_describe(int n);

// Largest transform length the engine runs in one workgroup.
int max_len();
// Largest transform length accepted at all: longer ones (and DCT-I / DST-I
// extensions above max_len()) run in several passes over a scratch buffer
// (engine_long.hip): four-step, or Bluestein where the length does not split.
int max_len_any();
// build the plan of a length of either kind (twiddles, sub-lengths, filters)
void plan_len(int n);
// free the multi-pass plans and their scratch (no captured launch may use them)
void free_long();
// free only the multi-pass scratch (grown again on the next long transform)
void free_long_scratch();

// Row transforms along the contiguous axis (two rows per complex FFT).
//   r2c: in real rows (stride si_row reals), out complex rows (stride so_row),
//        keeping modes k < nxo.
//   c2r: in complex rows with nxi valid modes (the rest read as zero),
//        out real rows.
void launch_x_r2c(const real_t* in, cplx_t* out, long long nrows, long long si_row, long long so_row,
                  int nx, int nxo, hipStream_t st, bool rev = false);
void launch_x_c2r(const cplx_t* in, real_t* out, long long nrows, long long si_row, long long so_row,
                  int nx, int nxi, hipStream_t st, bool rev = false);
// The same with table-addressed complex rows: mode k of row r at element
// xt[k].x + r * xt[k].y (k < nxo / nxi).  False when nx has no specialised
// kernel (the caller then goes through a contiguous buffer).
bool launch_x_r2c_tab(const real_t* in, cplx_t* out, long long nrows, int nx, int nxo, const int2* xt,
                      hipStream_t st);
bool launch_x_c2r_tab(const cplx_t* in, real_t* out, long long nrows, int nx, int nxi, const int2* xt,
                      hipStream_t st);

// X and Y transforms of whole z-planes in one launch (one workgroup per
// plane, the plane kept in LDS), for small grids of the (nx, ny) pairs that
// have such kernels: forward real (nx, ny, nz) -> half spectrum (nx/2+1, ny,
// nz), backward the reverse (Im of the DC and Nyquist modes ignored).  False
// when the grid does not qualify (the caller runs the separate passes).
bool launch_xy_plane(bool fwd, const void* in, void* out, int nx, int ny, int nz, hipStream_t st);
// 0: never; 1: wherever they apply; 2 (default): on small grids only.
void set_plane_kernels(int mode);
// X rows of 512 points through the wave kernels (one wavefront per row pair,
// butterfly exchanges by cross-lane permlane/DPP/bpermute operations) or,
// by default, the LDS-tile kernels
void set_x_wave(bool on);

// Column transform along a strided axis of a 3D view [outer][len][inner].
enum ColKind { COL_C2C = 0, COL_DCT1 = 1, COL_DST1 = 2, COL_COPY = 3 };
struct ColSpec {
  const cplx_t* in;
  cplx_t* out;
  long long si_in, sl_in, so_in;     // input strides (complex elements): inner, length, outer
  long long si_out, sl_out, so_out;  // output strides
  int n_inner, n_outer;
  int n;          // logical transform length (rows of the column)
  int kind;       // ColKind
  bool fwd;       // forward (e^{-i}) or backward
  int nld;        // rows filled from the input (<= n)
  int ld_lo, ld_hi, ld_shift;  // LDS row e <- input row e (e<ld_lo), e-ld_shift (e>=ld_hi), else 0
  int nst;        // rows written to the output
  int st_lo, st_shift;         // output row e <- LDS row e (e<st_lo) or e+st_shift
  bool rev;       // walk the tiles from the last one (see set_tile_order)
  // DCT-I only: apply the Chebyshev derivative recurrence (scale, L) of
  // launch_cheby to the stored rows in the same pass where the kernel allows
  bool cheby = false;
  double cheby_scale = 0, cheby_L = 0;
};
// Returns true when a requested Chebyshev recurrence was applied (false: the
// caller runs launch_cheby on the output).
bool launch_col(const ColSpec& c, hipStream_t st);
// the multi-pass form of launch_col for an FFT length m above max_len()
void launch_col_long(const ColSpec& c, int m, hipStream_t st);
// the multi-pass X r2c (r2c) or c2r of rows of length nx above max_len();
// nk = nxo (r2c) or nxi (c2r)
void launch_x_long(bool r2c, const void* in, void* out, long long nrows, long long si_row, long long so_row, int nx,
                   int nk, hipStream_t st);
// C2C column transform of n_inner x n_outer columns of length n whose rows
// are table-addressed: row e of column (o, i) at element ltab[e].x + o *
// ltab[e].y + i of c.in (zero when ltab[e].x < 0), FFT row p stored at
// stab[p] likewise in c.out (skipped when < 0).  False when n has no
// specialised kernel.
bool launch_col_tab(const ColSpec& c, const int2* ltab, const int2* stab, hipStream_t st);

// Consecutive passes over one buffer walk their tiles in opposite orders, so
// a pass starts on what the previous pass wrote last (still in the
// last-level cache).  Off: every pass walks forward.
void set_tile_order(bool alternate);
// Specialised column kernels address with 64-bit offsets even where 32-bit
// ones fit (views larger than 4 GiB always use 64-bit ones).
void set_wide_offsets(bool on);
bool tile_order_alternate();
// Tiles of a column pass over one contiguous inner axis (the Z pass): 0 in
// the XCD-aware order, 1 in dispatch order, 2 (default) chosen by shape.
void set_flat_order(int mode);
int flat_order();

// Fused Z-direction spectral solve on a spectrum (nxh, ny, nz) in place:
// forward FFT along z, multiply by -scale/|k|^2 (k=0 -> 0), inverse FFT.
// Returns false when nz has no specialised kernel (the caller composes it).
bool launch_zsolve(cplx_t* spec, int nxh, int ny, int nz, double scale, hipStream_t st);

// Grouped batches over n fields (shapes 3n ints; field i real, works[i] its
// (nx/2+1, ny, nz) complex spectrum) whose lengths all have specialised
// kernels: one launch per pass and transform length.  mode 0: Poisson solves
// in place on the fields (spectrum as work area); 1: forward r2c 'fft' of the
// fields into the spectra; 2: backward c2r 'tff' of the spectra (destroyed)
// into the fields.  Indices of the other fields are returned in `singles`.
struct GroupBatch;
GroupBatch* group_batch_create(int mode, int n, const int* shapes, real_t* const* fields, void* const* works,
                               std::vector<int>* singles);
void group_batch_launch(GroupBatch* B, hipStream_t st);
int group_batch_launches(const GroupBatch* B);
void group_batch_free(GroupBatch* B);

// k2-weighted shell sums of a (d0, d1, d2) half spectrum (a fastest; fold
// lengths f0, f1, f2, 0 = the axis is not folded) into E[0..kmax], through
// `partial` (spectrum_partial_bytes)
size_t spectrum_partial_bytes(int d2, int kmax);
void launch_spectrum(const cplx_t* B, int d0, int d1, int d2, int f0, int f1, int f2, int kmax, double* partial,
                     double* E, hipStream_t st);

// Chebyshev post-processing of a 'ffc' spectrum (build/ftran.F90:408-451):
// scale, then the derivative recurrence along z.  Element (col c, z k) at
// out[c*cs + k*ks].
void launch_cheby(cplx_t* out, long long ncol, int nzc, long long cs, long long ks, double scale,
                  double L, hipStream_t st);

// use the compile-time specialised kernels where a length has them (default on)
void set_specialised(bool on);
// DCT-I column passes without the fused Chebyshev recurrence through the
// direct-load kernel reading the mirrored rows (default on) or through the
// LDS load pass that builds the extension
void set_dct_direct(bool on);
void set_col_pipe(int on);


// out[i] = in[i] over n16 16-byte vectors
void launch_copy16(const void* in, void* out, long long n16, hipStream_t st);

}  // namespace p3amd
