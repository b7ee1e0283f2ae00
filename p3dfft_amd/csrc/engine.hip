// engine.hip -- gfx950 kernels of the 3D FFT path and their launchers.
//
//  k_x_r2c / k_x_c2r : transforms along the contiguous X axis.  Two real rows
//                      are packed into one complex row (x_a + i x_b) and run
//                      through one complex FFT; the two half spectra are
//                      separated (r2c) or merged (c2r) in LDS.  Replaces the
//                      FFTW r2c/c2r plans of build/fft_exec.F90:495-540,298-342.
//  k_col             : transform along a strided axis (Y or Z) of a 3D view.
//                      A workgroup owns TX adjacent inner columns of one outer
//                      index; loads are coalesced along the inner axis (or along
//                      the transform axis for the STRIDE1 layout), the pruning
//                      index maps and zero fills of build/module.F90:502-706 are
//                      applied in the load/store indexing, and DCT-I / DST-I are
//                      run as complex FFTs of the even / odd extension built in
//                      LDS (build/fft_exec.F90:646-701, 866-921).
//  k_gen_tw          : FP64 twiddle generation (sincospi), stored as double2 or
//                      as hi+lo float pairs.
//  k_cheby           : Chebyshev derivative recurrence (build/ftran.F90:408-451)
//                      in FP64, as two interleaved suffix sums over row chunks.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "engine.hpp"
#include "engine_ct.hpp"

namespace p3amd {

void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error{std::string("HIP error in ") + what + ": " + hipGetErrorString(e), -3};
}

typedef TwT<real_t>::type Tw;

// Workgroups are dispatched round-robin over the 8 XCDs.  Remap the flat
// workgroup id so that each XCD owns a contiguous range of tiles: adjacent
// tiles (which share partial cache lines at their edges) then meet in the same
// L2.  The map is a bijection on [0, nwg).
__device__ __forceinline__ int xcd_tile(int bid, int nwg) {
  const int q = nwg >> 3, rm = nwg & 7;
  const int x = bid & 7, slot = bid >> 3;
  return x * q + (x < rm ? x : rm) + slot;
}

// tile of this workgroup: XCD-aware (xcd != 0) or in dispatch order,
// optionally walking from the last tile
__device__ __forceinline__ int tile_of(int rev, int xcd = 1) {
  const int t = xcd ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  return rev ? (int)gridDim.x - 1 - t : t;
}

// ---------------------------------------------------------------------------
// Stockham stages over an LDS tile.  COL layout: element e of sequence b at
// lds[e*B + b] (B sequences side by side, sequence index fastest).  ROW layout:
// lds[b*n + e].
// ---------------------------------------------------------------------------
template <class T, int R, bool FWD, bool COL, int NT, int EPT>
__device__ __forceinline__ void stage_r(cpx<T>* lds, const LenDesc& L, int s, int B, FDiv dB,
                                        const typename TwT<T>::type* tw) {
  constexpr int IT = EPT / R;
  static_assert(IT >= 1, "radix larger than the per-thread element budget");
  const int n = L.n, nR = n / R, Ns = L.ns[s];
  const int total = B * nR;
  cpx<T> v[IT][R];
#pragma unroll
  for (int m = 0; m < IT; m++) {
    const int it = threadIdx.x + m * NT;
    if (it < total) {
      int b, j;
      if (COL) { j = fdiv(it, dB); b = it - j * B; }
      else { b = fdiv(it, L.dnr[s]); j = it - b * nR; }
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int e = j + r * nR;
        v[m][r] = lds[COL ? e * B + b : b * n + e];
      }
      if (s > 0) {
        const int k = j - fdiv(j, L.dns[s]) * Ns;
        const typename TwT<T>::type* w = tw + L.twoff[s] + k * (R - 1);
#pragma unroll
        for (int r = 1; r < R; r++) v[m][r] = twmul<FWD>(v[m][r], w[r - 1]);
      }
      dft<R, FWD>(v[m]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < IT; m++) {
    const int it = threadIdx.x + m * NT;
    if (it < total) {
      int b, j;
      if (COL) { j = fdiv(it, dB); b = it - j * B; }
      else { b = fdiv(it, L.dnr[s]); j = it - b * nR; }
      const int k = j - fdiv(j, L.dns[s]) * Ns;
      const int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int e = base + r * Ns;
        lds[COL ? e * B + b : b * n + e] = v[m][r];
      }
    }
  }
  __syncthreads();
}

// Any radix (used for prime factors above 13): one output per work item,
// y = sum_r x_r W_{Ns R}^{r (k + q Ns)}, twiddles from the full table W_n^e.
template <class T, bool FWD, bool COL, int NT, int EPT>
__device__ __forceinline__ void stage_g(cpx<T>* lds, const LenDesc& L, int s, int B, FDiv dB,
                                        const typename TwT<T>::type* tw) {
  const int n = L.n, R = L.radix[s], nR = n / R, Ns = L.ns[s], M = Ns * R, step = n / M;
  const int total = B * n;
  const typename TwT<T>::type* wf = tw + L.fulloff;
  cpx<T> v[EPT];
#pragma unroll
  for (int m = 0; m < EPT; m++) {
    const int it = threadIdx.x + m * NT;
    if (it < total) {
      int b, jq;
      if (COL) { jq = fdiv(it, dB); b = it - jq * B; }
      else { b = fdiv(it, L.dn); jq = it - b * n; }
      const int q = fdiv(jq, L.dnr[s]);
      const int j = jq - q * nR;
      const int k = j - fdiv(j, L.dns[s]) * Ns;
      const int inc = k + q * Ns;
      int e = 0;
      cpx<T> acc = mk<T>(0, 0);
      for (int r = 0; r < R; r++) {
        const int ee = j + r * nR;
        acc = add(acc, twmul<FWD>(lds[COL ? ee * B + b : b * n + ee], wf[e * step]));
        e += inc;
        if (e >= M) e -= M;
      }
      v[m] = acc;
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < EPT; m++) {
    const int it = threadIdx.x + m * NT;
    if (it < total) {
      int b, jq;
      if (COL) { jq = fdiv(it, dB); b = it - jq * B; }
      else { b = fdiv(it, L.dn); jq = it - b * n; }
      const int q = fdiv(jq, L.dnr[s]);
      const int j = jq - q * nR;
      const int k = j - fdiv(j, L.dns[s]) * Ns;
      const int e = (j - k) * R + k + q * Ns;
      lds[COL ? e * B + b : b * n + e] = v[m];
    }
  }
  __syncthreads();
}

template <class T, bool FWD, bool COL, int NT, int EPT>
__device__ void run_fft(cpx<T>* lds, const LenDesc& L, int B, FDiv dB, const typename TwT<T>::type* tw) {
  for (int s = 0; s < L.nst; s++) {
    switch (L.radix[s]) {
      case 2: stage_r<T, 2, FWD, COL, NT, EPT>(lds, L, s, B, dB, tw); break;
      case 3: stage_r<T, 3, FWD, COL, NT, EPT>(lds, L, s, B, dB, tw); break;
      case 4: stage_r<T, 4, FWD, COL, NT, EPT>(lds, L, s, B, dB, tw); break;
      case 5: stage_r<T, 5, FWD, COL, NT, EPT>(lds, L, s, B, dB, tw); break;
      case 7: stage_r<T, 7, FWD, COL, NT, EPT>(lds, L, s, B, dB, tw); break;
      case 8: stage_r<T, 8, FWD, COL, NT, EPT>(lds, L, s, B, dB, tw); break;
      case 11: stage_r<T, 11, FWD, COL, NT, EPT>(lds, L, s, B, dB, tw); break;
      case 13: stage_r<T, 13, FWD, COL, NT, EPT>(lds, L, s, B, dB, tw); break;
      case 16: stage_r<T, 16, FWD, COL, NT, EPT>(lds, L, s, B, dB, tw); break;
      default: stage_g<T, FWD, COL, NT, EPT>(lds, L, s, B, dB, tw); break;
    }
  }
}

// copy the twiddle table into LDS behind the data tile
template <class T, int NT>
__device__ __forceinline__ const typename TwT<T>::type* stage_twiddles(unsigned char* smem, size_t off,
                                                                      const typename TwT<T>::type* tw,
                                                                      int count) {
  typedef typename TwT<T>::type TwL;
  TwL* tl = (TwL*)(smem + off);
  for (int i = threadIdx.x; i < count; i += NT) tl[i] = tw[i];
  return tl;
}

// ---------------------------------------------------------------------------
// X kernels
// ---------------------------------------------------------------------------
struct XArgs {
  const void* in;
  void* out;
  long long nrows, si_row, so_row;
  int nxo;     // r2c: kept modes; c2r: valid input modes
  int B;       // row pairs per workgroup
  FDiv dnxo;
  LenDesc L;
  const void* tw;
  int tw_lds;  // twiddles staged in LDS
  int tw_off;  // byte offset of the staged twiddles
  int rev;     // walk the tiles from the last one
  // table-addressed rows (share mode): mode k of row r at element
  // xt[k].x + r * xt[k].y of `out` (r2c) / `in` (c2r), for k < nxo
  const int2* xt;
};

extern __shared__ __align__(16) unsigned char g_smem[];

template <class T, int NT, int EPT>
__global__ __launch_bounds__(NT) void k_x_r2c(XArgs a) {
  typedef typename TwT<T>::type TwL;
  cpx<T>* lds = (cpx<T>*)g_smem;
  const int n = a.L.n, B = a.B;
  const TwL* tw = (const TwL*)a.tw;
  if (a.tw_lds) tw = stage_twiddles<T, NT>(g_smem, a.tw_off, tw, a.L.twcount);
  const long long pair0 = (long long)tile_of(a.rev) * B;
  const T* in = (const T*)a.in;
  for (int q = threadIdx.x; q < B * n; q += NT) {
    const int b = fdiv(q, a.L.dn), j = q - b * n;
    const long long ra = 2 * (pair0 + b), rb = ra + 1;
    const T xa = ra < a.nrows ? in[ra * a.si_row + j] : (T)0;
    const T xb = rb < a.nrows ? in[rb * a.si_row + j] : (T)0;
    lds[b * n + j] = mk<T>(xa, xb);
  }
  __syncthreads();
  run_fft<T, true, false, NT, EPT>(lds, a.L, B, FDiv{0, 1}, tw);
  cpx<T>* out = (cpx<T>*)a.out;
  const T h = (T)0.5;
  for (int q = threadIdx.x; q < B * a.nxo; q += NT) {
    const int b = fdiv(q, a.dnxo), k = q - b * a.nxo;
    const cpx<T> zk = lds[b * n + k], zn = lds[b * n + (k == 0 ? 0 : n - k)];
    const long long ra = 2 * (pair0 + b), rb = ra + 1;
    if (ra < a.nrows) out[ra * a.so_row + k] = mk<T>((zk.x + zn.x) * h, (zk.y - zn.y) * h);
    if (rb < a.nrows) out[rb * a.so_row + k] = mk<T>((zk.y + zn.y) * h, (zn.x - zk.x) * h);
  }
}

template <class T, int NT, int EPT>
__global__ __launch_bounds__(NT) void k_x_c2r(XArgs a) {
  typedef typename TwT<T>::type TwL;
  cpx<T>* lds = (cpx<T>*)g_smem;
  const int n = a.L.n, B = a.B, h = n / 2;
  const TwL* tw = (const TwL*)a.tw;
  if (a.tw_lds) tw = stage_twiddles<T, NT>(g_smem, a.tw_off, tw, a.L.twcount);
  const long long pair0 = (long long)tile_of(a.rev) * B;
  const cpx<T>* in = (const cpx<T>*)a.in;
  for (int q = threadIdx.x; q < B * n; q += NT) {
    const int b = fdiv(q, a.L.dn), k = q - b * n;
    const long long ra = 2 * (pair0 + b), rb = ra + 1;
    const int kk = k <= h ? k : n - k;
    cpx<T> A = mk<T>(0, 0), Bv = mk<T>(0, 0);
    if (kk < a.nxo) {
      if (ra < a.nrows) A = in[ra * a.si_row + kk];
      if (rb < a.nrows) Bv = in[rb * a.si_row + kk];
    }
    // FFTW c2r: the imaginary parts of k=0 and k=n/2 (n even) do not contribute
    if (kk == 0 || 2 * kk == n) { A.y = 0; Bv.y = 0; }
    lds[b * n + k] = (k == kk) ? mk<T>(A.x - Bv.y, A.y + Bv.x) : mk<T>(A.x + Bv.y, Bv.x - A.y);
  }
  __syncthreads();
  run_fft<T, false, false, NT, EPT>(lds, a.L, B, FDiv{0, 1}, tw);
  T* out = (T*)a.out;
  for (int q = threadIdx.x; q < B * n; q += NT) {
    const int b = fdiv(q, a.L.dn), j = q - b * n;
    const long long ra = 2 * (pair0 + b), rb = ra + 1;
    const cpx<T> z = lds[b * n + j];
    if (ra < a.nrows) out[ra * a.so_row + j] = z.x;
    if (rb < a.nrows) out[rb * a.so_row + j] = z.y;
  }
}

// ---------------------------------------------------------------------------
// column kernel
// ---------------------------------------------------------------------------
struct ColArgs {
  const void* in;
  void* out;
  long long si_in, sl_in, so_in, si_out, sl_out, so_out;
  int n_inner, n_outer, tiles;
  int TX;
  FDiv dTX;
  int kind;
  int nld, ld_lo, ld_hi, ld_shift;
  int nst, st_lo, st_shift;
  int ld_lenfast, st_lenfast;
  FDiv dnld, dnst;
  int nrows_lds;  // rows of the LDS tile
  LenDesc L;
  const void* tw;
  int tw_lds, tw_off;
  int rev;  // walk the tiles from the last one
  // XCD-aware tile order (tiles that share cache lines at their edges run in
  // one XCD's L2); 0: dispatch order, for views whose tiles never share a
  // line (one outer index: every tile starts at a multiple of TX elements
  // of one contiguous inner axis), where the concurrent tiles then cover one
  // contiguous range of each row
  int xcd;
  const int2* ltab;  // table-addressed rows (see ColIO)
  const int2* stab;
  int tab_off;       // byte offset of the LDS copy of the row tables
  // Chebyshev recurrence fused into a DCT-I column pass (k_col_ct): applied
  // to the output rows in LDS before they are stored; cheby_off: byte offset
  // of its per-wave sums behind the tile
  int cheby, cheby_off;
  double cheby_scale, cheby_L;
  // > 0: the direct-load kernel reads a DCT-I even extension of this many
  // input rows (row e <- row e, e < mirror; row 2(mirror-1) - e above)
  int mirror;
};

template <class T, bool FWD, int NT, int EPT>
__global__ __launch_bounds__(NT) void k_col(ColArgs a) {
  typedef typename TwT<T>::type TwL;
  cpx<T>* lds = (cpx<T>*)g_smem;
  const int TX = a.TX;
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int o = t / a.tiles;
  const int ti = t - o * a.tiles;
  const long long i0 = (long long)ti * TX;
  const int nb = (int)((long long)a.n_inner - i0 < TX ? (long long)a.n_inner - i0 : TX);
  const cpx<T>* in = (const cpx<T>*)a.in + i0 * a.si_in + (long long)o * a.so_in;
  cpx<T>* out = (cpx<T>*)a.out + i0 * a.si_out + (long long)o * a.so_out;
  const TwL* tw = (const TwL*)a.tw;
  if (a.tw_lds) tw = stage_twiddles<T, NT>(g_smem, a.tw_off, tw, a.L.twcount);
  const int off = (a.kind == COL_DST1) ? 1 : 0;
  for (int q = threadIdx.x; q < TX * a.nld; q += NT) {
    int b, e;
    if (a.ld_lenfast) { b = fdiv(q, a.dnld); e = q - b * a.nld; }
    else { e = fdiv(q, a.dTX); b = q - e * TX; }
    cpx<T> v = mk<T>(0, 0);
    if (b < nb) {
      const int src = e < a.ld_lo ? e : (e >= a.ld_hi ? e - a.ld_shift : -1);
      if (src >= 0) v = in[(long long)b * a.si_in + (long long)src * a.sl_in];
    }
    lds[(e + off) * TX + b] = v;
  }
  __syncthreads();
  if (a.kind == COL_DCT1) {
    // even extension of length m = 2(nld-1): row e <- row m-e
    const int m = a.L.n;
    for (int q = threadIdx.x; q < TX * (m - a.nld); q += NT) {
      const int r = fdiv(q, a.dTX), b = q - r * TX, e = a.nld + r;
      lds[e * TX + b] = lds[(m - e) * TX + b];
    }
    __syncthreads();
  } else if (a.kind == COL_DST1) {
    // odd extension of length m = 2(nld+1): rows 0 and nld+1 zero, row e <- -row m-e
    const int m = a.L.n;
    for (int q = threadIdx.x; q < TX * (m - a.nld - 1); q += NT) {
      const int r = fdiv(q, a.dTX), b = q - r * TX, e = a.nld + 1 + r;
      if (e == a.nld + 1) lds[e * TX + b] = mk<T>(0, 0);
      else { const cpx<T> u = lds[(m - e) * TX + b]; lds[e * TX + b] = mk<T>(-u.x, -u.y); }
    }
    for (int b = threadIdx.x; b < TX; b += NT) lds[b] = mk<T>(0, 0);
    __syncthreads();
  }
  if (a.kind != COL_COPY) run_fft<T, FWD, true, NT, EPT>(lds, a.L, TX, a.dTX, tw);
  for (int q = threadIdx.x; q < TX * a.nst; q += NT) {
    int b, e;
    if (a.st_lenfast) { b = fdiv(q, a.dnst); e = q - b * a.nst; }
    else { e = fdiv(q, a.dTX); b = q - e * TX; }
    if (b < nb) {
      const int p = e < a.st_lo ? e : e + a.st_shift;
      cpx<T> v;
      if (a.kind == COL_DST1) { const cpx<T> u = lds[(p + 1) * TX + b]; v = mk<T>(-u.y, u.x); }
      else v = lds[p * TX + b];
      out[(long long)b * a.si_out + (long long)e * a.sl_out] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// compile-time specialised kernels (same arguments and semantics as above).
// TWL: the twiddle table is staged in LDS (decided per length at compile
// time, so twiddle reads are LDS reads).  Addresses are a wave-uniform 64-bit
// tile base plus 32-bit byte offsets (the host checks that they fit).
// ---------------------------------------------------------------------------
template <class T, int NT, bool TWL>
__device__ __forceinline__ const typename TwT<T>::type* ct_twiddles(const void* tw, int off, int count) {
  typedef typename TwT<T>::type TwL;
  if constexpr (TWL) return stage_twiddles<T, NT>(g_smem, off, (const TwL*)tw, count);
  else return (const TwL*)tw;
}

// X r2c of the row pairs pair0 .. pair0+B-1 (nr <= 2B rows exist) through
// the LDS tile `lds`
template <class T, int B, int NT, class RL, bool XT>
__device__ __forceinline__ void x_r2c_ct_tile(cpx<T>* lds, const typename TwT<T>::type* tw, const XArgs& a,
                                              long long pin, long long pair0, int nr) {
  constexpr int N = RL_N<RL>::value;
  constexpr int BP = B + 1;
  const char* in = (const char*)a.in + 2 * pin * a.si_row * (long long)sizeof(T);
  const unsigned si = (unsigned)(a.si_row * sizeof(T));
  {
    // issue every load of the tile before the LDS writes
    constexpr int LI = (B * N + NT - 1) / NT;
    T xa[LI], xb[LI];
#pragma unroll
    for (int i = 0; i < LI; i++) {
      const int q = threadIdx.x + i * NT;
      const int b = q / N, j = q % N;
      const bool ok = ((B * N) % NT == 0) || q < B * N;
      // out-of-range items load the tile's first element (always valid) and
      // are zeroed, so no load sits under a branch
      const bool oka = ok && 2 * b < nr, okb = ok && 2 * b + 1 < nr;
      const unsigned o = (unsigned)(2 * b) * si + (unsigned)j * (unsigned)sizeof(T);
      const T va = *(const T*)(in + (oka ? o : 0u));
      const T vb = *(const T*)(in + (okb ? o + si : 0u));
      xa[i] = oka ? va : (T)0;
      xb[i] = okb ? vb : (T)0;
    }
#pragma unroll
    for (int i = 0; i < LI; i++) {
      const int q = threadIdx.x + i * NT;
      if (((B * N) % NT == 0) || q < B * N) lds[(q % N) * BP + q / N] = mk<T>(xa[i], xb[i]);
    }
  }
  __syncthreads();
  run_fft_ct<T, true, B, BP, NT>(lds, tw, RL{});
  char* out = (char*)a.out + 2 * pair0 * a.so_row * (long long)sizeof(cpx<T>);
  const unsigned so = (unsigned)(a.so_row * sizeof(cpx<T>));
  const T h = (T)0.5;
  const bool full = nr == 2 * B;  // uniform: no per-store row guard
  // separation of mode k of pair b and its stores (FULL: no row guard)
  auto put = [&](int b, int k, auto fl) {
    const cpx<T> zk = lds[k * BP + b], zn = lds[(k == 0 ? 0 : N - k) * BP + b];
    const unsigned o = (unsigned)(2 * b) * so + (unsigned)k * (unsigned)sizeof(cpx<T>);
    const cpx<T> ya = mk<T>((zk.x + zn.x) * h, (zk.y - zn.y) * h), yb = mk<T>((zk.y + zn.y) * h, (zn.x - zk.x) * h);
    if (decltype(fl)::value || 2 * b < nr) *(cpx<T>*)(out + o) = ya;
    if (decltype(fl)::value || 2 * b + 1 < nr) *(cpx<T>*)(out + o + so) = yb;
  };
  constexpr int H = N / 2;
  if constexpr (XT) {
    // the table entries of all this thread's modes first, then the stores:
    // a store placed after its entry's load waits for every memory access
    // issued before it, which otherwise made each item a round trip
    constexpr int LS = (B * (H + 1) + NT - 1) / NT;  // nxo <= N/2 + 1
    int2 tks[LS];
#pragma unroll
    for (int i = 0; i < LS; i++) {
      const int q = threadIdx.x + i * NT;
      const bool ok = q < B * a.nxo;
      const int b = ok ? fdiv(q, a.dnxo) : 0;
      tks[i] = a.xt[ok ? q - b * a.nxo : 0];
    }
#pragma unroll
    for (int i = 0; i < LS; i++) {
      const int q = threadIdx.x + i * NT;
      if (q < B * a.nxo) {
        const int b = fdiv(q, a.dnxo), k = q - b * a.nxo;
        const cpx<T> zk = lds[k * BP + b], zn = lds[(k == 0 ? 0 : N - k) * BP + b];
        const cpx<T> ya = mk<T>((zk.x + zn.x) * h, (zk.y - zn.y) * h), yb = mk<T>((zk.y + zn.y) * h, (zn.x - zk.x) * h);
        const long long r = 2 * (pair0 + b);
        cpx<T>* oc = (cpx<T>*)a.out;
        if (2 * b < nr) oc[tks[i].x + r * tks[i].y] = ya;
        if (2 * b + 1 < nr) oc[tks[i].x + (r + 1) * tks[i].y] = yb;
      }
    }
  } else if ((N % 2) == 0 && a.nxo == H + 1) {
    // every mode stored (no X cut): modes k < N/2 as item b*(N/2) + k
    // (compile-time divisor), the Nyquist mode of pair b by thread b
    // a compile-time trip count and the full-tile test outside the loop, so
    // that the LDS reads of all items can be issued ahead of their stores
    constexpr int LS = (B * H + NT - 1) / NT;
    auto put_all = [&](auto fl) {
#pragma unroll
      for (int i = 0; i < LS; i++) {
        const int q = threadIdx.x + i * NT;
        if ((B * H) % NT == 0 || q < B * H) put(q / H, q % H, fl);
      }
      if (threadIdx.x < B) put(threadIdx.x, H, fl);
    };
    if (full) put_all(std::true_type{});
    else put_all(std::false_type{});
  } else {
    for (int q = threadIdx.x; q < B * a.nxo; q += NT) {
      const int b = fdiv(q, a.dnxo);
      put(b, q - b * a.nxo, std::false_type{});
    }
  }
}

template <class T, int B, int NT, class RL, int W, bool TWL, bool XT>
__global__ __launch_bounds__(NT, W) void k_x_r2c_ct(XArgs a) {
  cpx<T>* lds = (cpx<T>*)g_smem;
  const auto* tw = ct_twiddles<T, NT, TWL>(a.tw, a.tw_off, a.L.twcount);
  const long long pair0 = (long long)tile_of(a.rev) * B;
  // rows of this tile (2B unless it is the last one)
  const int nr = (int)(a.nrows - 2 * pair0 < 2 * B ? a.nrows - 2 * pair0 : 2 * B);
  x_r2c_ct_tile<T, B, NT, RL, XT>(lds, tw, a, pair0, pair0, nr);
}

// X c2r of the row pairs pair0 .. pair0+B-1 (nr <= 2B rows exist)
template <class T, int B, int NT, class RL, bool XT>
__device__ __forceinline__ void x_c2r_ct_tile(cpx<T>* lds, const typename TwT<T>::type* tw, const XArgs& a,
                                              long long pair0, long long pout, int nr) {
  constexpr int N = RL_N<RL>::value;
  constexpr int BP = B + 1;
  constexpr int H = N / 2;
  const char* in = (const char*)a.in + 2 * pair0 * a.si_row * (long long)sizeof(cpx<T>);
  const unsigned si = (unsigned)(a.si_row * sizeof(cpx<T>));
  if (!XT && (N % 2) == 0 && a.nxo == H + 1) {
    // every mode stored (no X cut): the modes k < N/2 of pair b as item
    // b*(N/2) + k (compile-time divisor: shifts for the power-of-two
    // lengths, whole iterations for every thread), the Nyquist mode of pair b
    // by thread b; values and positions as in the general loop below
    constexpr int LH = (B * H + NT - 1) / NT;
    cpx<T> va[LH], vb[LH];
#pragma unroll
    for (int i = 0; i < LH; i++) {
      const int q = threadIdx.x + i * NT;
      const int b = q / H, k = q % H;
      const bool ok = ((B * H) % NT == 0) || q < B * H;
      const bool oka = ok && 2 * b < nr, okb = ok && 2 * b + 1 < nr;
      const unsigned o = (unsigned)(2 * b) * si + (unsigned)k * (unsigned)sizeof(cpx<T>);
      const cpx<T> ua = *(const cpx<T>*)(in + (oka ? o : 0u));
      const cpx<T> ub = *(const cpx<T>*)(in + (okb ? o + si : 0u));
      va[i] = oka ? ua : mk<T>(0, 0);
      vb[i] = okb ? ub : mk<T>(0, 0);
    }
    cpx<T> na = mk<T>(0, 0), nb = mk<T>(0, 0);
    if (threadIdx.x < B) {
      const int b = threadIdx.x;
      const unsigned o = (unsigned)(2 * b) * si + (unsigned)H * (unsigned)sizeof(cpx<T>);
      if (2 * b < nr) na = *(const cpx<T>*)(in + o);
      if (2 * b + 1 < nr) nb = *(const cpx<T>*)(in + o + si);
    }
#pragma unroll
    for (int i = 0; i < LH; i++) {
      const int q = threadIdx.x + i * NT;
      if (((B * H) % NT == 0) || q < B * H) {
        const int b = q / H, k = q % H;
        cpx<T> A = va[i], Bv = vb[i];
        if (k == 0) { A.y = 0; Bv.y = 0; }
        lds[k * BP + b] = mk<T>(A.x - Bv.y, A.y + Bv.x);
        lds[(k == 0 ? 0 : N - k) * BP + b] = mk<T>(A.x + Bv.y, Bv.x - A.y);
      }
    }
    // Nyquist: Im of A and B ignored, Z[N/2] = Re A + i Re B
    if (threadIdx.x < B) lds[H * BP + threadIdx.x] = mk<T>(na.x, nb.x);
  } else {
    // each input element k <= N/2 is loaded once and gives Z[k] and Z[N-k]
    // (Z = A + iB of the two Hermitian half spectra)
    constexpr int K = H + 1;
    constexpr int LI = (B * K + NT - 1) / NT;
    cpx<T> va[LI], vb[LI];
    // table-addressed modes: every table entry of this thread first, then the
    // data loads (a load placed after its entry's load waits for every access
    // issued before it, which otherwise made each item a round trip)
    int2 tks[XT ? LI : 1];
    if constexpr (XT) {
#pragma unroll
      for (int i = 0; i < LI; i++) {
        const int q = threadIdx.x + i * NT;
        const int k = q % K;
        const bool ok = (((B * K) % NT == 0) || q < B * K) && k < a.nxo;
        tks[i] = a.xt[ok ? k : 0];
      }
    }
#pragma unroll
    for (int i = 0; i < LI; i++) {
      const int q = threadIdx.x + i * NT;
      const int b = q / K, k = q % K;
      const bool ok = (((B * K) % NT == 0) || q < B * K) && k < a.nxo;
      // out-of-range items load the tile's first element (always valid) and
      // are zeroed, so no load sits under a branch
      const bool oka = ok && 2 * b < nr, okb = ok && 2 * b + 1 < nr;
      if constexpr (XT) {
        const cpx<T>* ic = (const cpx<T>*)a.in;
        const int2 tk = tks[i];
        const long long r = 2 * (pair0 + b);
        const cpx<T> ua = ic[oka ? tk.x + r * tk.y : 0];
        const cpx<T> ub = ic[okb ? tk.x + (r + 1) * tk.y : 0];
        va[i] = oka ? ua : mk<T>(0, 0);
        vb[i] = okb ? ub : mk<T>(0, 0);
      } else {
        const unsigned o = (unsigned)(2 * b) * si + (unsigned)k * (unsigned)sizeof(cpx<T>);
        const cpx<T> ua = *(const cpx<T>*)(in + (oka ? o : 0u));
        const cpx<T> ub = *(const cpx<T>*)(in + (okb ? o + si : 0u));
        va[i] = oka ? ua : mk<T>(0, 0);
        vb[i] = okb ? ub : mk<T>(0, 0);
      }
    }
#pragma unroll
    for (int i = 0; i < LI; i++) {
      const int q = threadIdx.x + i * NT;
      if (((B * K) % NT == 0) || q < B * K) {
        const int b = q / K, k = q % K;
        cpx<T> A = va[i], Bv = vb[i];
        if (k == 0 || 2 * k == N) { A.y = 0; Bv.y = 0; }
        // at k = 0 and k = N/2 both forms are the same value: the second
        // write repeats the first (row 0 stands in for row N)
        lds[k * BP + b] = mk<T>(A.x - Bv.y, A.y + Bv.x);
        lds[(k == 0 ? 0 : N - k) * BP + b] = mk<T>(A.x + Bv.y, Bv.x - A.y);
      }
    }
  }
  __syncthreads();
  run_fft_ct<T, false, B, BP, NT>(lds, tw, RL{});
  char* out = (char*)a.out + 2 * pout * a.so_row * (long long)sizeof(T);
  const unsigned so = (unsigned)(a.so_row * sizeof(T));
  // a compile-time trip count and the full-tile test outside the loop, so
  // that the LDS reads of all items can be issued ahead of their stores
  constexpr int LS = (B * N + NT - 1) / NT;
  auto store_all = [&](auto full) {
#pragma unroll
    for (int i = 0; i < LS; i++) {
      const int q = threadIdx.x + i * NT;
      if ((B * N) % NT == 0 || q < B * N) {
        const int b = q / N, j = q % N;
        const cpx<T> z = lds[j * BP + b];
        const unsigned o = (unsigned)(2 * b) * so + (unsigned)j * (unsigned)sizeof(T);
        if (decltype(full)::value || 2 * b < nr) *(T*)(out + o) = z.x;
        if (decltype(full)::value || 2 * b + 1 < nr) *(T*)(out + o + so) = z.y;
      }
    }
  };
  if (nr == 2 * B) store_all(std::true_type{});  // uniform: no per-store row guard
  else store_all(std::false_type{});
}

template <class T, int B, int NT, class RL, int W, bool TWL, bool XT>
__global__ __launch_bounds__(NT, W) void k_x_c2r_ct(XArgs a) {
  cpx<T>* lds = (cpx<T>*)g_smem;
  const auto* tw = ct_twiddles<T, NT, TWL>(a.tw, a.tw_off, a.L.twcount);
  const long long pair0 = (long long)tile_of(a.rev) * B;
  const int nr = (int)(a.nrows - 2 * pair0 < 2 * B ? a.nrows - 2 * pair0 : 2 * B);
  x_c2r_ct_tile<T, B, NT, RL, XT>(lds, tw, a, pair0, pair0, nr);
}

// ---------------------------------------------------------------------------
// Wave kernels for X rows of 512 points: one wavefront per row pair, no LDS
// tile and no workgroup barrier.  Lane l holds the 8 values l + 64m of the
// complex row z = x_a + i x_b (m = 0..7, registers).  The length-512 FFT is
// 512 = 8 x 64 with the 64-point part over the lane index, 64 = 8 x 8:
//   radix-8 over the registers (stride 64), twiddle W_512^(l k1);
//   exchange lane bits 3-5 <-> register bits 0-2 (permlane32_swap,
//   permlane16_swap, DPP row_ror:8), radix-8 over the registers, twiddle
//   W_64^(l0 k2);
//   exchange lane bits 0-2 <-> register bits 0-2 (DPP row_ror:4/12 with bank
//   masks, DPP quad_perm), radix-8 over the registers.
// Lane s + 8q then holds Z[q + 8s + 64r] in register r; ds_bpermute gives
// every lane the contiguous run k = l + 64r to store, and the two-for-one
// separation's partner Z[N-k].  Every exchange is a cross-lane VALU (or
// bpermute) operation of one wave.
// ---------------------------------------------------------------------------
// swap lane bit LB with the register pair (lo, hi): a lane whose bit LB is 0
// keeps lo and receives its partner's lo as hi; a lane whose bit is 1 keeps
// hi and receives its partner's hi as lo (partner = lane ^ (1 << LB))
template <int LB>
__device__ __forceinline__ void lswap(unsigned& lo, unsigned& hi) {
  if constexpr (LB == 5) {
    const auto r = __builtin_amdgcn_permlane32_swap(lo, hi, false, false);
    lo = r[0];
    hi = r[1];
  } else if constexpr (LB == 4) {
    const auto r = __builtin_amdgcn_permlane16_swap(lo, hi, false, false);
    lo = r[0];
    hi = r[1];
  } else if constexpr (LB == 3 || LB == 2) {
    // rows of 16 lanes: banks of 4 lanes; bit 3 clear = banks 0,1, bit 2
    // clear = banks 0,2; partner by rotation inside the row (row_ror:n gives
    // lane i the value of lane i - n mod 16)
    constexpr int up = LB == 3 ? 0x128 : 0x12C;    // row_ror:8 / row_ror:12 (lane i <- i + 2^LB)
    constexpr int dn = LB == 3 ? 0x128 : 0x124;    // row_ror:8 / row_ror:4 (lane i <- i - 2^LB)
    constexpr int b0 = LB == 3 ? 0x3 : 0x5, b1 = LB == 3 ? 0xC : 0xA;
    const unsigned t = hi;
    hi = (unsigned)__builtin_amdgcn_update_dpp((int)hi, (int)lo, up, 0xF, b0, false);
    lo = (unsigned)__builtin_amdgcn_update_dpp((int)lo, (int)t, dn, 0xF, b1, false);
  } else {
    // inside a quad: quad_perm [1,0,3,2] (bit 0) / [2,3,0,1] (bit 1)
    constexpr int qp = LB == 0 ? 0xB1 : 0x4E;
    const bool b = (threadIdx.x >> LB) & 1;
    const unsigned send = b ? lo : hi;
    const unsigned recv = (unsigned)__builtin_amdgcn_mov_dpp((int)send, qp, 0xF, 0xF, false);
    if (b) lo = recv;
    else hi = recv;
  }
}

template <int LB>
__device__ __forceinline__ void lswap(float& lo, float& hi) {
  unsigned a = __float_as_uint(lo), b = __float_as_uint(hi);
  lswap<LB>(a, b);
  lo = __uint_as_float(a);
  hi = __uint_as_float(b);
}

template <int LB>
__device__ __forceinline__ void lswap(double& lo, double& hi) {
  const unsigned long long a = (unsigned long long)__double_as_longlong(lo);
  const unsigned long long b = (unsigned long long)__double_as_longlong(hi);
  unsigned al = (unsigned)a, ah = (unsigned)(a >> 32), bl = (unsigned)b, bh = (unsigned)(b >> 32);
  lswap<LB>(al, bl);
  lswap<LB>(ah, bh);
  lo = __longlong_as_double((long long)(((unsigned long long)ah << 32) | al));
  hi = __longlong_as_double((long long)(((unsigned long long)bh << 32) | bl));
}

// transpose lane bit LB with register bit RB of an 8-register array
template <int LB, int RB, class T>
__device__ __forceinline__ void lswap_regs(cpx<T>* z) {
#pragma unroll
  for (int r = 0; r < 8; r++)
    if (!(r & (1 << RB))) {
      lswap<LB>(z[r].x, z[r | (1 << RB)].x);
      lswap<LB>(z[r].y, z[r | (1 << RB)].y);
    }
}

template <class T>
__device__ __forceinline__ T bperm(int lane, T v);
template <>
__device__ __forceinline__ float bperm<float>(int lane, float v) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(lane << 2, __float_as_int(v)));
}
template <>
__device__ __forceinline__ double bperm<double>(int lane, double v) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_ds_bpermute(lane << 2, (int)u);
  const int hi = __builtin_amdgcn_ds_bpermute(lane << 2, (int)(u >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// the 512-point FFT of one wave's row (see above); twf[e] = W_512^e
template <bool FWD, class T>
__device__ __forceinline__ void wave_fft512(cpx<T>* z, const typename TwT<T>::type* twf) {
  const int l = threadIdx.x & 63;
  dft8<FWD>(z);
#pragma unroll
  for (int k = 1; k < 8; k++) z[k] = twmul<FWD>(z[k], twf[l * k]);
  lswap_regs<3, 0>(z);
  lswap_regs<4, 1>(z);
  lswap_regs<5, 2>(z);
  dft8<FWD>(z);
  const int l0 = l & 7;
#pragma unroll
  for (int k = 1; k < 8; k++) z[k] = twmul<FWD>(z[k], twf[8 * l0 * k]);
  lswap_regs<0, 0>(z);
  lswap_regs<1, 1>(z);
  lswap_regs<2, 2>(z);
  dft8<FWD>(z);
}

// X r2c of 512-point rows, one wave per row pair at a time; each workgroup
// (four waves) stages the twiddle table in LDS once and walks row pairs at a
// stride of the grid, loading the next pair before transforming the current
template <class T>
__global__ __launch_bounds__(256) void k_x_r2c_w512(XArgs a, const typename TwT<T>::type* twg) {
  typedef typename TwT<T>::type TwL;
  TwL* twf = (TwL*)g_smem;
  for (int i = threadIdx.x; i < 512; i += 256) twf[i] = twg[i];
  __syncthreads();
  const long long npairs = (a.nrows + 1) / 2;
  const int l = threadIdx.x & 63;
  const int tl = (l >> 3) + 8 * (l & 7);
  const int lp = (64 - l) & 63;
  const int tp = (lp >> 3) + 8 * (lp & 7);
  const long long stride = (long long)gridDim.x * 4;
  long long idx = (long long)tile_of(0) * 4 + (threadIdx.x >> 6);
  // rows of pair index i (walked from the last pair when a.rev)
  auto rows = [&](long long i, const T*& ia, const T*& ib, long long& r0, bool& hb) {
    const long long pair = a.rev ? npairs - 1 - i : i;
    r0 = 2 * pair;
    hb = r0 + 1 < a.nrows;
    ia = (const T*)a.in + r0 * a.si_row;
    ib = hb ? ia + a.si_row : ia;
  };
  if (idx >= npairs) return;
  const T *ia, *ib;
  long long r0;
  bool hb;
  rows(idx, ia, ib, r0, hb);
  T xa[8], xb[8];
#pragma unroll
  for (int m = 0; m < 8; m++) { xa[m] = ia[l + 64 * m]; xb[m] = ib[l + 64 * m]; }
  for (; idx < npairs; idx += stride) {
    cpx<T> z[8];
#pragma unroll
    for (int m = 0; m < 8; m++) z[m] = mk<T>(xa[m], hb ? xb[m] : (T)0);
    const long long cr0 = r0;
    const bool chb = hb;
    if (idx + stride < npairs) {
      rows(idx + stride, ia, ib, r0, hb);
#pragma unroll
      for (int m = 0; m < 8; m++) { xa[m] = ia[l + 64 * m]; xb[m] = ib[l + 64 * m]; }
    }
    wave_fft512<true>(z, twf);
    // lane s + 8q holds Z[q + 8s + 64r].  Lane l stores the modes k = l + 64r
    // (a contiguous 64-mode run per store): Z[k] from lane (l >> 3) + 8(l & 7)
    // register r, Z[512 - k] from the lane of natural index (64 - l) & 63
    // register 7 - r; lane 0's partners Z[64(8 - r)] are its own registers
    // (8 - r) & 7
    cpx<T> zk[5], zn[5];
#pragma unroll
    for (int r = 0; r < 5; r++) {
      zk[r] = mk<T>(bperm<T>(tl, z[r].x), bperm<T>(tl, z[r].y));
      zn[r] = mk<T>(bperm<T>(tp, z[7 - r].x), bperm<T>(tp, z[7 - r].y));
    }
    if (l == 0) {
#pragma unroll
      for (int r = 0; r < 5; r++) zn[r] = z[(8 - r) & 7];
    }
    cpx<T>* oa = (cpx<T>*)a.out + cr0 * a.so_row;
    cpx<T>* ob = oa + a.so_row;
    const T h = (T)0.5;
#pragma unroll
    for (int r = 0; r < 5; r++) {
      const int k = l + 64 * r;
      if (k < a.nxo) {
        const cpx<T> u = zk[r], v = zn[r];
        oa[k] = mk<T>((u.x + v.x) * h, (u.y - v.y) * h);
        if (chb) ob[k] = mk<T>((u.y + v.y) * h, (v.x - u.x) * h);
      }
    }
  }
}

// X c2r of 512-point rows, one wave per row pair at a time (workgroups as
// in k_x_r2c_w512): Z = A + iB from the stored modes k <= 256 (Im of the DC
// and Nyquist modes ignored, modes at and above nxi zero), the upper half
// conj(A[512-k]) + i conj(B[512-k]) taken from the lane holding 512 - k
template <class T>
__global__ __launch_bounds__(256) void k_x_c2r_w512(XArgs a, const typename TwT<T>::type* twg) {
  typedef typename TwT<T>::type TwL;
  TwL* twf = (TwL*)g_smem;
  for (int i = threadIdx.x; i < 512; i += 256) twf[i] = twg[i];
  __syncthreads();
  const long long npairs = (a.nrows + 1) / 2;
  const int l = threadIdx.x & 63;
  const int tl = (l >> 3) + 8 * (l & 7);
  const int src = (64 - l) & 63;
  const long long stride = (long long)gridDim.x * 4;
  long long idx = (long long)tile_of(0) * 4 + (threadIdx.x >> 6);
  auto rows = [&](long long i, const cpx<T>*& ia, const cpx<T>*& ib, long long& r0, bool& hb) {
    const long long pair = a.rev ? npairs - 1 - i : i;
    r0 = 2 * pair;
    hb = r0 + 1 < a.nrows;
    ia = (const cpx<T>*)a.in + r0 * a.si_row;
    ib = hb ? ia + a.si_row : ia;
  };
  if (idx >= npairs) return;
  const cpx<T> *ia, *ib;
  long long r0;
  bool hb;
  rows(idx, ia, ib, r0, hb);
  // modes k = l + 64m (m < 5, k <= 256); out-of-range lanes load element 0
  // (always valid) and zero it
  cpx<T> va[5], vb[5];
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const int k = l + 64 * m;
    const bool ok = k <= 256 && k < a.nxo;
    va[m] = ia[ok ? k : 0];
    vb[m] = ib[ok ? k : 0];
  }
  for (; idx < npairs; idx += stride) {
    cpx<T> z[8], w[4];
#pragma unroll
    for (int m = 0; m < 5; m++) {
      const int k = l + 64 * m;
      const bool ok = k <= 256 && k < a.nxo;
      cpx<T> A = va[m], B = vb[m];
      if (!ok) A = mk<T>(0, 0);
      if (!ok || !hb) B = mk<T>(0, 0);
      if (k == 0 || k == 256) { A.y = 0; B.y = 0; }
      if (m < 4) {
        z[m] = mk<T>(A.x - B.y, A.y + B.x);
        w[m] = mk<T>(A.x + B.y, B.x - A.y);  // conj(A) + i conj(B)
      } else {
        z[4] = mk<T>(A.x - B.y, A.y + B.x);  // lane 0: k = 256
      }
    }
    const long long cr0 = r0;
    const bool chb = hb;
    if (idx + stride < npairs) {
      rows(idx + stride, ia, ib, r0, hb);
#pragma unroll
      for (int m = 0; m < 5; m++) {
        const int k = l + 64 * m;
        const bool ok = k <= 256 && k < a.nxo;
        va[m] = ia[ok ? k : 0];
        vb[m] = ib[ok ? k : 0];
      }
    }
    // k = l + 64m > 256 (m >= 4, l > 0): from lane 64 - l, register 7 - m;
    // lane 0: k = 64m, from its own register 8 - m (m = 5..7)
    cpx<T> up[4];
#pragma unroll
    for (int m = 4; m < 8; m++) up[m - 4] = mk<T>(bperm<T>(src, w[7 - m].x), bperm<T>(src, w[7 - m].y));
    if (l == 0) {
      up[0] = z[4];
#pragma unroll
      for (int m = 5; m < 8; m++) up[m - 4] = w[8 - m];
    }
#pragma unroll
    for (int m = 4; m < 8; m++) z[m] = up[m - 4];
    wave_fft512<false>(z, twf);
    // lane s + 8q holds x[q + 8s + 64r]; lane l stores n = l + 64r, taken from
    // lane (l >> 3) + 8(l & 7) register r (a contiguous run per store)
    T* oa = (T*)a.out + cr0 * a.so_row;
    T* ob = oa + a.so_row;
#pragma unroll
    for (int r = 0; r < 8; r++) {
      const T ua = bperm<T>(tl, z[r].x), ub = bperm<T>(tl, z[r].y);
      oa[l + 64 * r] = ua;
      if (chb) ob[l + 64 * r] = ub;
    }
  }
}

// ---------------------------------------------------------------------------
// Plane kernels for small grids: the X and Y transforms of one z-plane in one
// workgroup, the plane's half spectrum kept in LDS between them.  On small
// grids every pass is one dependent launch, and a dependent launch costs
// more than the plane's data movement, so fusing X and Y saves one launch and
// one HBM round trip per direction.  NX, NY are compile-time lengths with
// specialised radix schedules RLX, RLY (the twiddle tables of those lengths).
// LDS: the X tile (NY/2 row pairs of length NX, element e of pair b at
// lx[e*(NY/2+1) + b]), then the Y tile (NX/2+1 columns of length NY, element
// y of column k at ly[y*(NX/2+1) + k]), then both twiddle tables.
// ---------------------------------------------------------------------------
struct PlaneArgs {
  const void* in;
  void* out;
  const void* twx;  // twiddles of NX, of NY (the same table when NX == NY)
  const void* twy;
  int ntwx, ntwy;   // entries
};

template <class T, class RLX, class RLY>
struct PlaneGeom {
  static constexpr int NX = RL_N<RLX>::value, NY = RL_N<RLY>::value;
  static constexpr int BX = NY / 2, BPX = BX + 1;  // row pairs, padded
  static constexpr int H = NX / 2 + 1;             // columns of the half spectrum
  static constexpr long long X_BYTES = (long long)NX * BPX * sizeof(cpx<T>);
  static constexpr long long Y_BYTES = (long long)NY * H * sizeof(cpx<T>);
  static constexpr long long DATA = X_BYTES + Y_BYTES;
};

template <class T, class RLX, class RLY, int NT>
__device__ __forceinline__ void plane_twiddles(const PlaneArgs& a, const typename TwT<T>::type** tx,
                                               const typename TwT<T>::type** ty) {
  typedef PlaneGeom<T, RLX, RLY> G;
  typedef typename TwT<T>::type TwL;
  TwL* t = (TwL*)(g_smem + G::DATA);
  for (int i = threadIdx.x; i < a.ntwx; i += NT) t[i] = ((const TwL*)a.twx)[i];
  *tx = t;
  if (a.twy == a.twx) {
    *ty = t;
  } else {
    for (int i = threadIdx.x; i < a.ntwy; i += NT) t[a.ntwx + i] = ((const TwL*)a.twy)[i];
    *ty = t + a.ntwx;
  }
}

// forward: X r2c (two-for-one, as k_x_r2c_ct) and Y forward of plane
// blockIdx.x: real rows of NX (row stride NX) -> half-spectrum rows of NX/2+1
template <class T, class RLX, class RLY, int NT>
__global__ __launch_bounds__(NT) void k_xy_plane_fwd(PlaneArgs a) {
  typedef PlaneGeom<T, RLX, RLY> G;
  constexpr int NX = G::NX, NY = G::NY, BX = G::BX, BPX = G::BPX, H = G::H;
  cpx<T>* lx = (cpx<T>*)g_smem;
  cpx<T>* ly = (cpx<T>*)(g_smem + G::X_BYTES);
  const typename TwT<T>::type *tx, *ty;
  plane_twiddles<T, RLX, RLY, NT>(a, &tx, &ty);
  const T* in = (const T*)a.in + (long long)blockIdx.x * NX * NY;
  {
    // compile-time trip counts: every load of the plane issued before the
    // LDS writes (items past the plane read element 0 and are not written)
    constexpr int LI = (BX * NX + NT - 1) / NT;
    T xa[LI], xb[LI];
#pragma unroll
    for (int i = 0; i < LI; i++) {
      const int q = threadIdx.x + i * NT;
      const bool ok = (BX * NX) % NT == 0 || q < BX * NX;
      const int b = q / NX, j = q % NX;
      xa[i] = in[ok ? 2 * b * NX + j : 0];
      xb[i] = in[ok ? (2 * b + 1) * NX + j : 0];
    }
#pragma unroll
    for (int i = 0; i < LI; i++) {
      const int q = threadIdx.x + i * NT;
      if ((BX * NX) % NT == 0 || q < BX * NX) lx[(q % NX) * BPX + q / NX] = mk<T>(xa[i], xb[i]);
    }
  }
  __syncthreads();
  run_fft_ct<T, true, BX, BPX, NT>(lx, tx, RLX{});
  const T h = (T)0.5;
  for (int q = threadIdx.x; q < BX * H; q += NT) {
    const int b = q / H, k = q % H;
    const cpx<T> zk = lx[k * BPX + b], zn = lx[(k == 0 ? 0 : NX - k) * BPX + b];
    ly[(2 * b) * H + k] = mk<T>((zk.x + zn.x) * h, (zk.y - zn.y) * h);
    ly[(2 * b + 1) * H + k] = mk<T>((zk.y + zn.y) * h, (zn.x - zk.x) * h);
  }
  __syncthreads();
  run_fft_ct<T, true, H, H, NT>(ly, ty, RLY{});
  cpx<T>* out = (cpx<T>*)a.out + (long long)blockIdx.x * H * NY;
  constexpr int LQ = (NY * H + NT - 1) / NT;
  cpx<T> v[LQ];
#pragma unroll
  for (int i = 0; i < LQ; i++) {
    const int q = threadIdx.x + i * NT;
    v[i] = ly[(NY * H) % NT == 0 || q < NY * H ? q : 0];
  }
#pragma unroll
  for (int i = 0; i < LQ; i++) {
    const int q = threadIdx.x + i * NT;
    if ((NY * H) % NT == 0 || q < NY * H) out[q] = v[i];
  }
}

// backward: Y backward and X c2r (as k_x_c2r_ct: Im of the DC and Nyquist
// modes ignored) of plane blockIdx.x
template <class T, class RLX, class RLY, int NT>
__global__ __launch_bounds__(NT) void k_yx_plane_bwd(PlaneArgs a) {
  typedef PlaneGeom<T, RLX, RLY> G;
  constexpr int NX = G::NX, NY = G::NY, BX = G::BX, BPX = G::BPX, H = G::H;
  cpx<T>* lx = (cpx<T>*)g_smem;
  cpx<T>* ly = (cpx<T>*)(g_smem + G::X_BYTES);
  const typename TwT<T>::type *tx, *ty;
  plane_twiddles<T, RLX, RLY, NT>(a, &tx, &ty);
  const cpx<T>* in = (const cpx<T>*)a.in + (long long)blockIdx.x * H * NY;
  {
    // every load of the plane issued before the LDS writes
    constexpr int LQ = (NY * H + NT - 1) / NT;
    cpx<T> v[LQ];
#pragma unroll
    for (int i = 0; i < LQ; i++) {
      const int q = threadIdx.x + i * NT;
      v[i] = in[(NY * H) % NT == 0 || q < NY * H ? q : 0];
    }
#pragma unroll
    for (int i = 0; i < LQ; i++) {
      const int q = threadIdx.x + i * NT;
      if ((NY * H) % NT == 0 || q < NY * H) ly[q] = v[i];
    }
  }
  __syncthreads();
  run_fft_ct<T, false, H, H, NT>(ly, ty, RLY{});
  for (int q = threadIdx.x; q < BX * H; q += NT) {
    const int b = q / H, k = q % H;
    cpx<T> A = ly[(2 * b) * H + k], Bv = ly[(2 * b + 1) * H + k];
    if (k == 0 || 2 * k == NX) { A.y = 0; Bv.y = 0; }
    // at k = 0 and k = NX/2 both forms are the same value
    lx[k * BPX + b] = mk<T>(A.x - Bv.y, A.y + Bv.x);
    lx[(k == 0 ? 0 : NX - k) * BPX + b] = mk<T>(A.x + Bv.y, Bv.x - A.y);
  }
  __syncthreads();
  run_fft_ct<T, false, BX, BPX, NT>(lx, tx, RLX{});
  T* out = (T*)a.out + (long long)blockIdx.x * NX * NY;
  constexpr int LI = (BX * NX + NT - 1) / NT;
  cpx<T> z[LI];
#pragma unroll
  for (int i = 0; i < LI; i++) {
    const int q = threadIdx.x + i * NT;
    const bool ok = (BX * NX) % NT == 0 || q < BX * NX;
    z[i] = lx[ok ? (q % NX) * BPX + q / NX : 0];
  }
#pragma unroll
  for (int i = 0; i < LI; i++) {
    const int q = threadIdx.x + i * NT;
    if ((BX * NX) % NT == 0 || q < BX * NX) {
      const int b = q / NX, j = q % NX;
      out[2 * b * NX + j] = z[i].x;
      out[(2 * b + 1) * NX + j] = z[i].y;
    }
  }
}

// Tile of B consecutive columns c0 .. c0+B-1 of the flattened (outer, inner)
// column index c = o * n_inner + i, element (o, i, row e) at
// o*so + i*si + e*sl.  The tile base is column c0's address; lane offsets are
// relative to it and non-negative (so >= n_inner * si).
template <class T, int B>
__device__ __forceinline__ ColIO col_io(const ColArgs& a, long long c0, int nb) {
  ColIO io;
  const unsigned n_in = (unsigned)a.n_inner;
  const unsigned o0 = (unsigned)c0 / n_in, i0 = (unsigned)c0 - o0 * n_in;
  const unsigned c = (unsigned)c0 + threadIdx.x % B;
  const unsigned o = c / n_in, i = c - o * n_in;
  constexpr long long E = sizeof(cpx<T>);
  io.in = (const char*)a.in + ((long long)o0 * a.so_in + (long long)i0 * a.si_in) * E;
  io.out = (char*)a.out + ((long long)o0 * a.so_out + (long long)i0 * a.si_out) * E;
  io.lane_in = (unsigned long long)((((long long)o - o0) * a.so_in + ((long long)i - i0) * a.si_in) * E);
  io.lane_out = (unsigned long long)((((long long)o - o0) * a.so_out + ((long long)i - i0) * a.si_out) * E);
  io.sl_in = (unsigned long long)(a.sl_in * E);
  io.sl_out = (unsigned long long)(a.sl_out * E);
  io.nb = nb;
  io.ld_lo = a.ld_lo; io.ld_hi = a.ld_hi; io.ld_shift = a.ld_shift;
  io.nst = a.nst; io.st_lo = a.st_lo; io.st_shift = a.st_shift;
  io.mirror = a.mirror;
  io.lane_o = o; io.lane_i = i;
  io.ltab = a.ltab; io.stab = a.stab;
  if (a.ltab) {  // table-addressed rows index the whole arrays
    io.in = (const char*)a.in;
    io.out = (char*)a.out;
  }
  return io;
}

// C2C column transform with direct global <-> register first/last stages,
// over tiles of TX consecutive columns of the flattened (outer, inner) index
// (no tile is partial except the last one).  PR: pruning maps.
template <class T, bool FWD, int TX, int NT, class RL, int W, bool TWL, bool PR, class Off>
__global__ __launch_bounds__(NT, W) void k_colio_ct(ColArgs a) {
  cpx<T>* lds = (cpx<T>*)g_smem;
  // the twiddle staging is completed by the barrier at the end of the first stage
  const auto* tw = ct_twiddles<T, NT, TWL>(a.tw, a.tw_off, a.L.twcount);
  const long long ncol = (long long)a.n_inner * a.n_outer;
  const long long c0 = (long long)tile_of(a.rev, a.xcd) * TX;
  const int nb = (int)(ncol - c0 < TX ? ncol - c0 : TX);
  const ColIO io = col_io<T, TX>(a, c0, nb);
  if constexpr (std::is_same<Off, TabRows>::value)
    run_fft_io_tab<T, FWD, TX, NT>(lds, (int2*)(g_smem + a.tab_off), tw, io, a.ltab, a.stab, RL{});
  else
    run_fft_io<T, FWD, TX, NT, PR, Off>(lds, tw, io, RL{});
}

// Persistent form of k_colio_ct (unpruned rows, 32-bit offsets or the
// table-addressed rows) for tiles that take more than half the LDS, i.e. one
// workgroup per CU: gridDim.x workgroups, the 8 XCDs (workgroup w on XCD
// w mod 8) each owning a contiguous range of tiles that its workgroups
// share out round robin (a.xcd; else workgroup w takes tiles w, w + G, ...),
// and each workgroup keeping its next tile's loads in flight while the
// current tile computes.  The row tables (TabRows) are the
// same for every tile and stay in LDS, load table then store table, at
// tab_off.
template <class T, bool FWD, int TX, int NT, class RL, int W, bool TWL, class Off>
__global__ __launch_bounds__(NT, W) void k_colio_pipe(ColArgs a) {
  constexpr int N = RL_N<RL>::value;
  constexpr bool TB = std::is_same<Off, TabRows>::value;
  cpx<T>* lds = (cpx<T>*)g_smem;
  // the twiddle staging is completed by the barrier at the end of the first stage
  const auto* tw = ct_twiddles<T, NT, TWL>(a.tw, a.tw_off, a.L.twcount);
  const long long ncol = (long long)a.n_inner * a.n_outer;
  const int ntiles = (int)((ncol + TX - 1) / TX);
  const int G = (int)gridDim.x, x = (int)(blockIdx.x & 7), slot = (int)(blockIdx.x >> 3);
  const int nslot = G / 8 + (x < G % 8 ? 1 : 0);
  const int q = ntiles / 8, rm = ntiles % 8;
  const int lo = x * q + (x < rm ? x : rm), cnt = q + (x < rm ? 1 : 0);
  int2* tl = (int2*)(g_smem + a.tab_off);
  if constexpr (TB) {
    for (int i = threadIdx.x; i < N; i += NT) {
      tl[i] = a.ltab[i];
      tl[N + i] = a.stab[i];
    }
    __syncthreads();
  }
  auto make_io = [&](int t) -> ColIO {
    const long long c0 = (long long)(a.rev ? ntiles - 1 - t : t) * TX;
    const int nb = (int)(ncol - c0 < TX ? ncol - c0 : TX);
    ColIO io = col_io<T, TX>(a, c0, nb);
    if constexpr (TB) {
      io.ltab = tl;
      io.stab = tl + N;
    }
    return io;
  };
  if (a.xcd) run_fft_io_pipe<T, FWD, TX, NT, false, Off>(lds, tw, lo + slot, lo + cnt, nslot, make_io, RL{});
  // dispatch order: the tiles running at once form one contiguous range
  else run_fft_io_pipe<T, FWD, TX, NT, false, Off>(lds, tw, (int)blockIdx.x, ntiles, G, make_io, RL{});
}

// Chebyshev derivative recurrence (build/ftran.F90:408-451) on the output
// rows of an LDS column tile: output row e of column b sits in LDS row
// p(e) = e (e < st_lo) or e + st_shift.  As in k_cheby, two interleaved
// suffix sums: thread (chunk q, column b), b fastest, owns a chunk of at
// most RM rows of its column, held in registers; chunk sums by parity, then
// the sums of the later chunks (inside a wave by shuffles at a stride of TX
// lanes, across waves through LDS at `wt`), then each chunk walked from its
// top with that carry.  Columns fastest keep the lanes of one LDS access on
// different banks; every LDS access of a thread is issued as a batch.
constexpr int kChebyRM = 8;
// Each thread stores its output rows straight to `out` (column b of the tile
// at out + b*si_out, row e at + e*sl_out; columns b >= nb are not stored).
template <class T, int TX, int NT>
__device__ __forceinline__ void cheby_tile(cpx<T>* lds, double* wt, const ColArgs& a, cpx<T>* out, int nb) {
  static_assert(64 % TX == 0 && NT % 64 == 0, "whole column groups per wave");
  constexpr int P = NT / TX, CW = 64 / TX, NW = NT / 64, RM = kChebyRM;
  const int N = a.nst;
  const int b = threadIdx.x % TX, q = threadIdx.x / TX;
  const int c = q % CW, w = q / CW;  // chunk within the wave, wave
  const int R = (N + P - 1) / P;     // <= RM (checked by the host)
  const int lo = q * R, hi = min(N, lo + R);
  const double scale = a.cheby_scale, L = a.cheby_L;
  cpx<T> t[RM];
  int addr[RM];
#pragma unroll
  for (int i = 0; i < RM; i++) {
    const int e = lo + i;
    addr[i] = (e < a.st_lo ? e : e + a.st_shift) * TX + b;
    t[i] = (e < hi) ? lds[addr[i]] : mk<T>(0, 0);
  }
  // terms L j scale t_j (the top one halved), summed by parity
  double2 x[RM];
  double v[4] = {0, 0, 0, 0};  // even re, even im, odd re, odd im
#pragma unroll
  for (int i = RM - 1; i >= 0; i--) {
    const int j = lo + i;
    double f = (j < hi) ? L * (double)j : 0.0;
    if (j == N - 1) f *= 0.5;
    x[i] = make_double2(f * ((double)t[i].x * scale), f * ((double)t[i].y * scale));
    if (j & 1) { v[2] += x[i].x; v[3] += x[i].y; }
    else { v[0] += x[i].x; v[1] += x[i].y; }
  }
  // inclusive suffix over the chunks of this wave (lanes TX apart)
  double s[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    s[k] = v[k];
#pragma unroll
    for (int d = 1; d < CW; d *= 2) {
      const double u = __shfl_down(s[k], d * TX, 64);
      if (c + d < CW) s[k] += u;
    }
  }
  if (c == 0)
#pragma unroll
    for (int k = 0; k < 4; k++) wt[(w * TX + b) * 4 + k] = s[k];
  __syncthreads();
  double cr[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const double nx = __shfl_down(s[k], TX, 64);
    cr[k] = c + 1 < CW ? nx : 0.0;
  }
#pragma unroll
  for (int u = 1; u < NW; u++)
    if (u > w)
#pragma unroll
      for (int k = 0; k < 4; k++) cr[k] += wt[(u * TX + b) * 4 + k];
  double2 ev = make_double2(cr[0], cr[1]), od = make_double2(cr[2], cr[3]);
  cpx<T>* oc = out + (long long)b * a.si_out;
#pragma unroll
  for (int i = RM - 1; i >= 0; i--) {
    const int r = lo + i;
    if (r < hi) {
      double2 d = (r & 1) ? ev : od;
      if (r == 0) { d.x *= 0.5; d.y *= 0.5; }
      if (b < nb) oc[(long long)r * a.sl_out] = mk<T>((T)d.x, (T)d.y);
      if (r & 1) { od.x += x[i].x; od.y += x[i].y; }
      else { ev.x += x[i].x; ev.y += x[i].y; }
    }
  }
}

// Column transforms through an LDS load/store pass: DCT-I / DST-I
// extensions, length-fastest ("STRIDE1") views, copies, and views whose
// offsets do not fit the 32-bit form.  Loads are issued LU at a time per
// thread before their LDS stores (several loads in flight per thread).
template <class T, bool FWD, int TX, int NT, class RL, int W, bool TWL>
__global__ __launch_bounds__(NT, W) void k_col_ct(ColArgs a) {
  cpx<T>* lds = (cpx<T>*)g_smem;
  const auto* tw = ct_twiddles<T, NT, TWL>(a.tw, a.tw_off, a.L.twcount);
  const int t = tile_of(a.rev, a.xcd);
  const int o = t / a.tiles;
  const int ti = t - o * a.tiles;
  const long long i0 = (long long)ti * TX;
  const int nb = (int)((long long)a.n_inner - i0 < TX ? (long long)a.n_inner - i0 : TX);
  const cpx<T>* in = (const cpx<T>*)a.in + i0 * a.si_in + (long long)o * a.so_in;
  cpx<T>* out = (cpx<T>*)a.out + i0 * a.si_out + (long long)o * a.so_out;
  const int off = (a.kind == COL_DST1) ? 1 : 0;
  constexpr int LU = 8;
  const int nq = a.mirror ? 0 : TX * a.nld;
  if (a.mirror) {
    // DCT-I: the first stage reads the even extension's mirrored rows from
    // the input; the other stages, the recurrence and the store use the tile
    run_fft_in_io<T, TX, NT, unsigned>(lds, tw, col_io<T, TX>(a, (long long)o * a.n_inner + i0, nb), RL{});
  }
  for (int q0 = threadIdx.x; q0 < nq; q0 += LU * NT) {
    cpx<T> v[LU];
    int dst[LU];
#pragma unroll
    for (int u = 0; u < LU; u++) {
      const int q = q0 + u * NT;
      int b, e;
      if (a.ld_lenfast) { b = fdiv(q, a.dnld); e = q - b * a.nld; }
      else { e = q / TX; b = q % TX; }
      const int src = e < a.ld_lo ? e : (e >= a.ld_hi ? e - a.ld_shift : -1);
      const bool ok = q < nq && b < nb && src >= 0;
      // items outside the tile load its first element (always valid) and
      // are zeroed, so no load sits under a branch
      const cpx<T> t = in[ok ? (long long)b * a.si_in + (long long)src * a.sl_in : 0];
      v[u] = ok ? t : mk<T>(0, 0);
      dst[u] = q < nq ? (e + off) * TX + b : -1;
    }
#pragma unroll
    for (int u = 0; u < LU; u++)
      if (dst[u] >= 0) lds[dst[u]] = v[u];
  }
  if (!a.mirror) __syncthreads();
  if (a.kind == COL_DCT1 && !a.mirror) {
    const int m = a.L.n;
    for (int q = threadIdx.x; q < TX * (m - a.nld); q += NT) {
      const int r = q / TX, b = q % TX, e = a.nld + r;
      lds[e * TX + b] = lds[(m - e) * TX + b];
    }
    __syncthreads();
  } else if (a.kind == COL_DST1) {
    const int m = a.L.n;
    for (int q = threadIdx.x; q < TX * (m - a.nld - 1); q += NT) {
      const int r = q / TX, b = q % TX, e = a.nld + 1 + r;
      if (e == a.nld + 1) lds[e * TX + b] = mk<T>(0, 0);
      else { const cpx<T> u = lds[(m - e) * TX + b]; lds[e * TX + b] = mk<T>(-u.x, -u.y); }
    }
    for (int b = threadIdx.x; b < TX; b += NT) lds[b] = mk<T>(0, 0);
    __syncthreads();
  }
  if (a.kind != COL_COPY && !a.mirror) run_fft_ct<T, FWD, TX, TX, NT>(lds, tw, RL{});
  if constexpr (64 % TX == 0 && NT % 64 == 0) {
    if (a.cheby) {
      // the recurrence stores the output rows itself
      cheby_tile<T, TX, NT>(lds, (double*)(g_smem + a.cheby_off), a, out, nb);
      return;
    }
  }
  for (int q = threadIdx.x; q < TX * a.nst; q += NT) {
    int b, e;
    if (a.st_lenfast) { b = fdiv(q, a.dnst); e = q - b * a.nst; }
    else { e = q / TX; b = q % TX; }
    if (b < nb) {
      const int p = e < a.st_lo ? e : e + a.st_shift;
      cpx<T> v;
      if (a.kind == COL_DST1) { const cpx<T> u = lds[(p + 1) * TX + b]; v = mk<T>(-u.y, u.x); }
      else v = lds[p * TX + b];
      out[(long long)b * a.si_out + (long long)e * a.sl_out] = v;
    }
  }
}

// fused Z-solve: the Z column of a spectrum (nxh, ny, nz), forward FFT,
// -scale/|k|^2, inverse FFT, in place
template <class T, int TX, int NT, class RL, int W, bool TWL>
__global__ __launch_bounds__(NT, W) void k_zsolve_ct(ColArgs a, double scale, int nxh, int ny) {
  cpx<T>* lds = (cpx<T>*)g_smem;
  const auto* tw = ct_twiddles<T, NT, TWL>(a.tw, a.tw_off, a.L.twcount);
  const long long ncol = a.n_inner;
  const long long c0 = (long long)tile_of(a.rev, a.xcd) * TX;
  const int nb = (int)(ncol - c0 < TX ? ncol - c0 : TX);
  const ColIO io = col_io<T, TX>(a, c0, nb);
  SolveCtx sc;
  sc.nxh = nxh; sc.ny = ny; sc.nz = a.L.n; sc.scale = scale;
  run_fft_solve<T, TX, NT>(lds, tw, io, c0, sc, RL{});
}

// ---------------------------------------------------------------------------
// Grouped launches: one pass of many arrays of one transform length (the
// fields of a batch of solves of mixed shapes) in one launch.  Each
// workgroup finds its array in a tile -> array table and runs the same tile
// body as the single-array kernel, so results are bit-identical to solving
// the arrays one by one.
// ---------------------------------------------------------------------------
struct GrpDesc {
  const void* in;
  void* out;
  long long nrows;   // X: rows; columns: outer count
  long long sl, so;  // columns: row and outer strides (elements)
  int n_inner;       // columns: inner count
  int nxh, ny;       // Z solve: spectrum geometry
  double scale;      // Z solve: 1 / (nx ny nz)
};
struct GrpArgs {
  const GrpDesc* d;
  const int* tile_req;    // workgroup -> array
  const int* tile_first;  // array -> its first workgroup
};

template <class T, bool R2C, int B, int NT, class RL, int W, bool TWL>
__global__ __launch_bounds__(NT, W) void k_x_grp(XArgs a, GrpArgs g) {
  cpx<T>* lds = (cpx<T>*)g_smem;
  const auto* tw = ct_twiddles<T, NT, TWL>(a.tw, a.tw_off, a.L.twcount);
  const int r = g.tile_req[blockIdx.x];
  const GrpDesc d = g.d[r];
  a.in = d.in;
  a.out = d.out;
  a.nrows = d.nrows;
  const long long pair0 = (long long)((int)blockIdx.x - g.tile_first[r]) * B;
  const int nr = (int)(a.nrows - 2 * pair0 < 2 * B ? a.nrows - 2 * pair0 : 2 * B);
  if constexpr (R2C) x_r2c_ct_tile<T, B, NT, RL, false>(lds, tw, a, pair0, pair0, nr);
  else x_c2r_ct_tile<T, B, NT, RL, false>(lds, tw, a, pair0, pair0, nr);
}

template <class T, bool FWD, int TX, int NT, class RL, int W, bool TWL>
__global__ __launch_bounds__(NT, W) void k_col_grp(ColArgs a, GrpArgs g) {
  cpx<T>* lds = (cpx<T>*)g_smem;
  const auto* tw = ct_twiddles<T, NT, TWL>(a.tw, a.tw_off, a.L.twcount);
  const int r = g.tile_req[blockIdx.x];
  const GrpDesc d = g.d[r];
  a.in = d.in; a.out = d.out;
  a.n_inner = d.n_inner; a.n_outer = (int)d.nrows;
  a.si_in = a.si_out = 1;
  a.sl_in = a.sl_out = d.sl;
  a.so_in = a.so_out = d.so;
  const long long ncol = (long long)a.n_inner * a.n_outer;
  const long long c0 = (long long)((int)blockIdx.x - g.tile_first[r]) * TX;
  const int nb = (int)(ncol - c0 < TX ? ncol - c0 : TX);
  const ColIO io = col_io<T, TX>(a, c0, nb);
  run_fft_io<T, FWD, TX, NT, false, unsigned>(lds, tw, io, RL{});
}

template <class T, int TX, int NT, class RL, int W, bool TWL>
__global__ __launch_bounds__(NT, W) void k_zs_grp(ColArgs a, GrpArgs g) {
  cpx<T>* lds = (cpx<T>*)g_smem;
  const auto* tw = ct_twiddles<T, NT, TWL>(a.tw, a.tw_off, a.L.twcount);
  const int r = g.tile_req[blockIdx.x];
  const GrpDesc d = g.d[r];
  a.in = d.in; a.out = d.out;
  a.n_inner = d.n_inner; a.n_outer = 1;
  a.si_in = a.si_out = 1;
  a.sl_in = a.sl_out = d.sl;
  a.so_in = a.so_out = 0;
  const long long ncol = a.n_inner;
  const long long c0 = (long long)((int)blockIdx.x - g.tile_first[r]) * TX;
  const int nb = (int)(ncol - c0 < TX ? ncol - c0 : TX);
  const ColIO io = col_io<T, TX>(a, c0, nb);
  SolveCtx sc;
  sc.nxh = d.nxh; sc.ny = d.ny; sc.nz = a.L.n; sc.scale = d.scale;
  run_fft_solve<T, TX, NT>(lds, tw, io, c0, sc, RL{});
}

// ---------------------------------------------------------------------------
// k2-weighted shell sums of a stored half spectrum (the drivers' power
// spectrum, sample/C/driver_spec.c:298-383): element (a, b, c), a fastest,
// adds k2*|B|^2 to shell int(sqrt(k2) + 1/2), k2 = ka^2 + kb^2 + kc^2, where an
// axis index j folds to n - j above n/2 on axes with fold length n > 0.
// One workgroup per index c; one histogram per wave in LDS (FP64), each
// thread summing runs of equal shells along a before its LDS add, so the
// sums do not depend on scheduling; the wave histograms and then the
// workgroups' are added in a fixed order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int fold_k(int j, int n) { return (n > 0 && j > n / 2) ? n - j : j; }

// exact shell of an integer k2: the r with r(r-1) < k2 <= r(r+1), which is
// int(sqrt(k2) + 1/2) (no integer k2 lies on a half-integer radius)
__device__ __forceinline__ int shell_of(int k2) {
  int r = (int)(sqrtf((float)k2) + 0.5f);
  if (r * (r + 1) < k2) r++;
  else if (r > 0 && r * (r - 1) >= k2) r--;
  return r;
}

// one workgroup per chunk `ch` of plane c (blockIdx.x = c * nchunk + ch)
template <class T>
__global__ __launch_bounds__(256) void k_spectrum(const cpx<T>* B, int d0, int d1, int f0, int f1, int f2,
                                                  int kmax, int nchunk, double* partial) {
  constexpr int R = 8;  // consecutive elements per thread per step
  double* hist = (double*)g_smem;
  const int nw = blockDim.x / 64, w = threadIdx.x / 64, nb = kmax + 1;
  for (int i = threadIdx.x; i < nw * nb; i += blockDim.x) hist[i] = 0.0;
  __syncthreads();
  double* hw = hist + w * nb;
  const int c = blockIdx.x / nchunk, ch = blockIdx.x - c * nchunk;
  const int kc = fold_k(c, f2);
  const int kc2 = kc * kc;
  const cpx<T>* P = B + (long long)c * d0 * d1;
  const unsigned n = (unsigned)d0 * (unsigned)d1;
  // chunk bounds, multiples of one workgroup step
  const unsigned step = blockDim.x * R;
  const unsigned per = ((n + nchunk - 1) / nchunk + step - 1) / step * step;
  const unsigned lo = ch * per, hi = min(n, lo + per);
  for (unsigned i0 = lo + threadIdx.x * R; i0 < hi; i0 += step) {
    // every load of the step first (items past the chunk load its last
    // element and are skipped below), each element's shell and weighted
    // energy computed outside any branch, so that the loads are issued back
    // to back rather than one per branch with a wait after each
    cpx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = P[i0 + r < hi ? i0 + r : hi - 1];
    unsigned b = i0 / (unsigned)d0, a = i0 - b * (unsigned)d0;
    int kb = fold_k((int)b, f1);
    int ik[R];
    double e[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int ka = fold_k((int)a, f0);
      const int k2 = ka * ka + kb * kb + kc2;
      ik[r] = shell_of(k2);
      e[r] = (double)k2 * ((double)v[r].x * (double)v[r].x + (double)v[r].y * (double)v[r].y);
      if (++a == (unsigned)d0) {
        a = 0;
        b++;
        kb = fold_k((int)b, f1);
      }
    }
    int cur = -1;
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (i0 + r < hi) {
        if (ik[r] != cur) {
          if (cur >= 0 && cur <= kmax) atomicAdd(hw + cur, acc);
          cur = ik[r];
          acc = e[r];
        } else {
          acc += e[r];
        }
      }
    }
    if (cur >= 0 && cur <= kmax) atomicAdd(hw + cur, acc);
  }
  __syncthreads();
  // bin-major partials: bin k of every workgroup contiguous
  for (int k = threadIdx.x; k < nb; k += blockDim.x) {
    double sum = 0.0;
    for (int q = 0; q < nw; q++) sum += hist[q * nb + k];
    partial[(long long)k * gridDim.x + blockIdx.x] = sum;
  }
}

// one workgroup of 256 per bin: strided partial sums, then a fixed tree
__global__ __launch_bounds__(256) void k_spectrum_sum(const double* partial, int nparts, double* E) {
  __shared__ double red[256];
  const double* row = partial + (long long)blockIdx.x * nparts;
  double sum = 0.0;
  for (int p = threadIdx.x; p < nparts; p += 256) sum += row[p];
  red[threadIdx.x] = sum;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) E[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------
template <class TwL>
__global__ void k_gen_tw(const int2* idx, int count, TwL* out);

template <>
__global__ void k_gen_tw<double2>(const int2* idx, int count, double2* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int2 em = idx[i];
  double s, c;
  sincospi(2.0 * (double)em.x / (double)em.y, &s, &c);
  out[i] = make_double2(c, -s);
}

template <>
__global__ void k_gen_tw<float4>(const int2* idx, int count, float4* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int2 em = idx[i];
  double s, c;
  sincospi(2.0 * (double)em.x / (double)em.y, &s, &c);
  s = -s;
  const float ch = (float)c, sh = (float)s;
  out[i] = make_float4(ch, sh, (float)(c - (double)ch), (float)(s - (double)sh));
}

// Chebyshev derivative recurrence along z (build/ftran.F90:408-451):
// with a_{k+1} = scale * row k (0-based rows), d_{N-1} = L (N-1) a_N / 2 and
// d_k = L k a_{k+1} + d_{k+2}, row r receives d_{r+1} = the sum of
// L j scale t_j over the rows j > r of the other parity (the top term
// halved), row 0 is halved again and row N-1 is zero: two interleaved suffix
// sums.  A workgroup owns TX adjacent columns; P threads per column each
// own a chunk of rows: pass 1 sums the chunk's terms by parity into LDS, pass
// 2 walks the chunk from its top with the later chunks' sums as the carry.
// Sums run in FP64 in both builds; within a chunk they are added in the
// reference's order (top down).
template <class T, int TX, int P>
__global__ __launch_bounds__(TX * P) void k_cheby(cpx<T>* o, long long ncol, int N, long long cs, long long ks,
                                                 double scale, double L) {
  __shared__ double2 tot[P][2][TX];
  const int b = threadIdx.x % TX, q = threadIdx.x / TX;
  const long long c = (long long)blockIdx.x * TX + b;
  const bool valid = c < ncol;
  cpx<T>* p = o + (valid ? c : 0) * cs;
  const int R = (N + P - 1) / P;
  const int lo = q * R, hi = min(N, lo + R);
  auto term = [&](int j, cpx<T> t) {
    double f = L * (double)j;
    if (j == N - 1) f *= 0.5;
    return make_double2(f * ((double)t.x * scale), f * ((double)t.y * scale));
  };
  // even-row and odd-row sums (named registers, no indexed private array)
  double2 ev = make_double2(0, 0), od = make_double2(0, 0);
  if (valid)
    for (int j = hi - 1; j >= lo; j--) {
      const double2 w = term(j, p[(long long)j * ks]);
      if (j & 1) { od.x += w.x; od.y += w.y; }
      else { ev.x += w.x; ev.y += w.y; }
    }
  tot[q][0][b] = ev;
  tot[q][1][b] = od;
  __syncthreads();
  // carry: the sums of the later chunks, by row parity
  ev = od = make_double2(0, 0);
  for (int u = P - 1; u > q; u--) {
    ev.x += tot[u][0][b].x; ev.y += tot[u][0][b].y;
    od.x += tot[u][1][b].x; od.y += tot[u][1][b].y;
  }
  if (!valid) return;
  for (int r = hi - 1; r >= lo; r--) {
    cpx<T>* e = p + (long long)r * ks;
    const cpx<T> t = *e;
    double2 d = (r & 1) ? ev : od;
    if (r == 0) { d.x *= 0.5; d.y *= 0.5; }
    *e = mk<T>((T)d.x, (T)d.y);
    const double2 w = term(r, t);
    if (r & 1) { od.x += w.x; od.y += w.y; }
    else { ev.x += w.x; ev.y += w.y; }
  }
}

// plain streaming copy (16-byte vectors), used to measure achievable HBM
// bandwidth and to calibrate traffic counters: one vector per thread, the
// load with the nontemporal hint.  Copy shapes measured on MI355X (GB/s over
// 0.54 / 1.08 / 2.15 GB): this one 6345 / 6427 / 6462; without the hint
// 6090 / 6187 / 6161; 4 vectors per thread, loads before stores,
// 5749 / 5825 / 5734 (6136 / 6203 / 6014 with both hints); 8 vectors per
// thread 5422 / 5454 / 5393; hipMemcpy device to device 4.89 TB/s.
typedef unsigned int copy_vec __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy16(const copy_vec* __restrict__ in, copy_vec* __restrict__ out,
                                                long long n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = __builtin_nontemporal_load(in + i);
}

// ---------------------------------------------------------------------------
// host side: length plans
// ---------------------------------------------------------------------------
static FDiv make_fdiv(int d) {
  FDiv f;
  f.d = d;
  f.m = ((1ULL << 32) + (uint64_t)d - 1) / (uint64_t)d;
  return f;
}

// stage radices: radix 8 as far as possible (two 4s instead of an 8 and a 2),
// then 3, 5, 7, 11, 13, then any remaining prime (generic stage)
static std::vector<int> factorize(int n) {
  // 2048 with radix 16: three stages (two LDS exchanges instead of three),
  // the share-mode length-2048 stages 4-10% faster; radix 16 for 256, 384
  // and 1024 measured no faster (kernel trace of 384^3 FP32: 505 against
  // 503 us per fwd+bwd; FP64 384^3 3% slower)
  if (n == 2048) return {16, 16, 8};
  std::vector<int> f;
  int m = n, a = 0;
  while (m % 2 == 0) { m /= 2; a++; }
  int n8 = a / 3, rem = a % 3;
  if (rem == 1 && n8 > 0) { n8--; f.push_back(4); f.push_back(4); }
  else if (rem == 1) f.push_back(2);
  else if (rem == 2) f.push_back(4);
  for (int i = 0; i < n8; i++) f.insert(f.begin(), 8);
  for (int p : {3, 5, 7, 11, 13})
    while (m % p == 0) { f.push_back(p); m /= p; }
  for (int p = 17; (long long)p * p <= m; p += 2)
    while (m % p == 0) { f.push_back(p); m /= p; }
  if (m > 1) f.push_back(m);
  return f;
}

static bool is_templated(int r) {
  return r == 2 || r == 3 || r == 4 || r == 5 || r == 7 || r == 8 || r == 11 || r == 13 || r == 16;
}

static std::mutex g_len_mu;
static std::map<int, Len1D> g_lens;
static std::vector<tw_t*> g_retired_tw;
static void prep_kernels(const Len1D& L);

std::vector<int> default_factors(int n) { return n > 1 ? factorize(n) : std::vector<int>{}; }

// plan of length n with the radix schedule f (stage s has radix f[s]):
// stage descriptors and the FP64-generated twiddle table
static Len1D build_len(int n, const std::vector<int>& f) {
  Len1D L;
  std::memset(&L.d, 0, sizeof(L.d));
  L.d.n = n;
  long long prod = 1;
  for (int r : f) prod *= r;
  if ((n > 1 && prod != n) || (n <= 1 && !f.empty())) throw Error{"radix schedule does not multiply to the length", -1};
  if ((int)f.size() > kMaxStages) throw Error{"transform length has too many factors", -1};
  std::vector<int2> idx;
  int ns = 1;
  for (size_t s = 0; s < f.size(); s++) {
    const int R = f[s];
    L.d.radix[s] = R;
    L.d.ns[s] = ns;
    L.d.twoff[s] = (int)idx.size();
    L.d.dns[s] = make_fdiv(ns);
    L.d.dnr[s] = make_fdiv(n / R);
    if (!is_templated(R)) L.has_generic = true;
    if (s > 0 && is_templated(R))
      for (int k = 0; k < ns; k++)
        for (int r = 1; r < R; r++) idx.push_back(make_int2(k * r, ns * R));
    ns *= R;
  }
  L.d.nst = (int)f.size();
  L.d.dn = make_fdiv(n);
  L.d.fulloff = (int)idx.size();
  if (L.has_generic)
    for (int e = 0; e < n; e++) idx.push_back(make_int2(e, n));
  L.d.twcount = (int)idx.size();
  if (n == 512) {
    // the wave kernels' table W_512^e
    std::vector<int2> fidx(n);
    for (int e = 0; e < n; e++) fidx[e] = make_int2(e, n);
    int2* didx = nullptr;
    check_hip(hipMalloc(&didx, n * sizeof(int2)), "hipMalloc(twiddle index)");
    check_hip(hipMemcpy(didx, fidx.data(), n * sizeof(int2), hipMemcpyHostToDevice), "hipMemcpy");
    check_hip(hipMalloc(&L.twf, n * sizeof(Tw)), "hipMalloc(twiddles)");
    k_gen_tw<Tw><<<(n + 255) / 256, 256>>>(didx, n, L.twf);
    check_hip(hipGetLastError(), "k_gen_tw launch");
    check_hip(hipDeviceSynchronize(), "k_gen_tw");
    check_hip(hipFree(didx), "hipFree");
  }
  if (!idx.empty()) {
    int2* didx = nullptr;
    check_hip(hipMalloc(&didx, idx.size() * sizeof(int2)), "hipMalloc(twiddle index)");
    check_hip(hipMemcpy(didx, idx.data(), idx.size() * sizeof(int2), hipMemcpyHostToDevice), "hipMemcpy");
    check_hip(hipMalloc(&L.tw, idx.size() * sizeof(Tw)), "hipMalloc(twiddles)");
    const int cnt = (int)idx.size();
    k_gen_tw<Tw><<<(cnt + 255) / 256, 256>>>(didx, cnt, L.tw);
    check_hip(hipGetLastError(), "k_gen_tw launch");
    check_hip(hipDeviceSynchronize(), "k_gen_tw");
    check_hip(hipFree(didx), "hipFree");
  }
  prep_kernels(L);
  return L;
}

const Len1D& get_len(int n) {
  std::lock_guard<std::mutex> lk(g_len_mu);
  auto it = g_lens.find(n);
  if (it != g_lens.end()) return it->second;
  auto res = g_lens.emplace(n, build_len(n, default_factors(n)));
  return res.first->second;
}

void set_len_plan(int n, const std::vector<int>& factors, bool ct_ok, int tile_bytes, int col_seg) {
  std::lock_guard<std::mutex> lk(g_len_mu);
  Len1D L = build_len(n, factors);
  L.ct_ok = ct_ok;
  L.tile_bytes = tile_bytes;
  L.col_seg = col_seg;
  auto it = g_lens.find(n);
  if (it == g_lens.end()) {
    g_lens.emplace(n, L);
    return;
  }
  // the old table is retired, not freed: launches captured earlier (graphs)
  // may still read it; retired tables are freed with the plans
  if (it->second.tw) g_retired_tw.push_back(it->second.tw);
  if (it->second.twf) g_retired_tw.push_back(it->second.twf);
  it->second = L;  // same map node: references to the entry stay valid
}

void free_lens() {
  std::lock_guard<std::mutex> lk(g_len_mu);
  (void)hipDeviceSynchronize();
  for (auto& kv : g_lens) {
    if (kv.second.tw) (void)hipFree(kv.second.tw);
    if (kv.second.twf) (void)hipFree(kv.second.twf);
  }
  for (tw_t* t : g_retired_tw) (void)hipFree(t);
  g_retired_tw.clear();
  g_lens.clear();
}

// Per-thread element budget of every stage (registers hold EPT complex values)
// and the workgroup sizes the kernels are built for.
static const int kEPT = 16;
static const int kNTs[] = {64, 128, 256, 512, 1024};
static constexpr size_t kLdsMax = 160 * 1024;
// data bytes per tile the tile-shape heuristics aim for (two tiles per CU)
static const size_t kTileBytes = 64 * 1024;

int max_len() { return 8192; }

// elements one thread can own in every stage of this length
static int len_cap_per_thread(const Len1D& L) {
  int c = kEPT;
  for (int s = 0; s < L.d.nst; s++) {
    const int R = L.d.radix[s];
    if (is_templated(R)) c = std::min(c, R * (kEPT / R));
  }
  return c;
}

static int pick_nt(long long elems, int cpt) {
  for (int nt : kNTs)
    if (elems <= (long long)nt * cpt) return nt;
  throw Error{"transform tile exceeds the 1024-thread workgroup capacity", -1};
}

// dynamic LDS up to the CU's 160 KiB less the kernel's static LDS
template <class K>
static void allow_big_lds(K kernel) {
  hipFuncAttributes fa;
  size_t stat = 0;
  if (hipFuncGetAttributes(&fa, (const void*)kernel) == hipSuccess) stat = fa.sharedSizeBytes;
  (void)hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(kLdsMax - stat));
  (void)hipGetLastError();  // a refused attribute leaves the default limit, not a pending error
}

// choose where twiddles live; returns LDS bytes
static size_t lds_layout(size_t data_bytes, const Len1D& L, int* tw_lds, int* tw_off) {
  size_t off = (data_bytes + 15) & ~(size_t)15;
  size_t twb = (size_t)L.d.twcount * sizeof(Tw);
  if (L.d.twcount > 0 && off + twb <= 96 * 1024) {
    *tw_lds = 1;
    *tw_off = (int)off;
    return off + twb;
  }
  *tw_lds = 0;
  *tw_off = 0;
  return data_bytes;
}

constexpr int kEPTc = 16;

// launch functors: one per kernel family, instantiated per workgroup size
struct GoXr2c {
  template <int NT>
  static void go(const XArgs& a, unsigned g, size_t b, hipStream_t s) {
    k_x_r2c<real_t, NT, kEPTc><<<g, NT, b, s>>>(a);
  }
  template <int NT>
  static void prep() { allow_big_lds(k_x_r2c<real_t, NT, kEPTc>); }
};
struct GoXc2r {
  template <int NT>
  static void go(const XArgs& a, unsigned g, size_t b, hipStream_t s) {
    k_x_c2r<real_t, NT, kEPTc><<<g, NT, b, s>>>(a);
  }
  template <int NT>
  static void prep() { allow_big_lds(k_x_c2r<real_t, NT, kEPTc>); }
};
template <bool FWD>
struct GoCol {
  template <int NT>
  static void go(const ColArgs& a, unsigned g, size_t b, hipStream_t s) {
    k_col<real_t, FWD, NT, kEPTc><<<g, NT, b, s>>>(a);
  }
  template <int NT>
  static void prep() { allow_big_lds(k_col<real_t, FWD, NT, kEPTc>); }
};

template <class F, class A>
static void dispatch_nt(int nt, const A& a, unsigned g, size_t b, hipStream_t s) {
  switch (nt) {
    case 64: F::template go<64>(a, g, b, s); break;
    case 128: F::template go<128>(a, g, b, s); break;
    case 256: F::template go<256>(a, g, b, s); break;
    case 512: F::template go<512>(a, g, b, s); break;
    default: F::template go<1024>(a, g, b, s); break;
  }
}

// ---- compile-time specialised lengths ----------------------------------
// Tuning constants of the specialised kernels (compile-time): complex values
// per thread, bytes of adjacent columns per row of a column tile, and the
// largest column workgroup.  The column passes are bound by the memory
// requests in flight per CU: over a 512^3 fwd+bwd, FP32 256 threads of 32
// values took 1.41 ms, 512 of 16 1.36 ms, 1024 of 8 1.32 ms; FP64 256 of 16
// 2.67 ms, 512 of 8 2.54 ms, 1024 of 4 2.54 ms.
#ifndef P3DFFT_CT_EPT
#define P3DFFT_CT_EPT 8
#endif
#ifndef P3DFFT_CT_COL_BYTES
#define P3DFFT_CT_COL_BYTES 128
#endif
#ifndef P3DFFT_CT_NT_MAX
#define P3DFFT_CT_NT_MAX (sizeof(real_t) == 4 ? 1024 : 512)
#endif
constexpr int kEPTct = P3DFFT_CT_EPT;
#ifndef P3DFFT_CT_LONG
#define P3DFFT_CT_LONG 2048
#endif
constexpr int pow2floor(int x) { int p = 1; while (p * 2 <= x) p *= 2; return p; }
// values per thread of the column kernels of length N: lengths of 2048 and
// more take twice the default, so that a tile still spans 8 columns
constexpr int ct_ept_col(int N) { return N >= P3DFFT_CT_LONG ? 2 * kEPTct : kEPTct; }
constexpr int ct_col_tx(int N) {
  int tx = P3DFFT_CT_COL_BYTES / (int)sizeof(cplx_t);
  int cap = pow2floor(P3DFFT_CT_NT_MAX * ct_ept_col(N) / N);
  if (cap < 1) cap = 1;
  return tx < cap ? tx : cap;
}
constexpr int ct_nt_ept(int elems, int ept) {
  int nt = (elems / ept + 63) / 64 * 64;
  return nt < 64 ? 64 : (nt > P3DFFT_CT_NT_MAX ? P3DFFT_CT_NT_MAX : nt);
}
constexpr int ct_nt(int elems) { return ct_nt_ept(elems, kEPTct); }
constexpr int ct_waves(long long lds_bytes, int nt) {
  long long wg = lds_bytes > 0 ? (160 * 1024) / lds_bytes : 8;
  if (wg < 1) wg = 1;
  const long long w = wg * (nt / 64) / 4;
  return w < 1 ? 1 : (w > 4 ? 4 : (int)w);  // at least 128 registers per lane
}
// X kernels: up to 16 row pairs per tile, 16 complex values per thread, at
// most 512 threads and an 80 KiB padded tile (two workgroups per CU)
#ifndef P3DFFT_X_EPT
#define P3DFFT_X_EPT 16
#endif
#ifndef P3DFFT_X_NT_MAX
#define P3DFFT_X_NT_MAX 512
#endif
constexpr int kEPTx = P3DFFT_X_EPT;
constexpr int ct_x_pairs(int N) {
  int b = 16;
  while (b > 1 && (b * N > P3DFFT_X_NT_MAX * kEPTx || (long long)(b + 1) * N * (long long)sizeof(cplx_t) > 80 * 1024))
    b /= 2;
  return b;
}
// a tile that gives fewer than P3DFFT_X_NT_MAX threads at kEPTx values each
// takes up to twice the threads at half the values: more waves per CU where
// the LDS allows few workgroups
constexpr int ct_nt_x(int elems) {
  int nt = (elems / kEPTx + 63) / 64 * 64;
  if (nt < P3DFFT_X_NT_MAX) nt = (elems / (kEPTx / 2) + 63) / 64 * 64;
  return nt < 64 ? 64 : (nt > P3DFFT_X_NT_MAX ? P3DFFT_X_NT_MAX : nt);
}

// Twiddles of a specialised length sit in LDS behind the data tile when the
// two fit in this many bytes (two workgroups per CU at the largest tiles).
constexpr long long kCtLdsTw = 80 * 1024;
constexpr long long round16(long long x) { return (x + 15) / 16 * 16; }

struct CtInfo {
  int TX, NTC, BX, NTX, nr;
  int radix[kMaxStages];
  bool twl_col, twl_x;  // twiddles staged in LDS (column / X kernels)
  bool twl_tab;         // the same for the table-addressed column kernel
  bool pipe;            // the column tile takes more than half the LDS: k_colio_pipe exists
  bool twl_pipe_tab;    // twiddles in LDS in the table-addressed k_colio_pipe
  long long twb;        // twiddle table bytes
  void (*col)(bool, const ColArgs&, unsigned, size_t, hipStream_t);
  void (*colio)(bool, bool, bool, const ColArgs&, unsigned, size_t, hipStream_t);
  void (*colpipe)(bool, bool, const ColArgs&, unsigned, size_t, hipStream_t);
  void (*xr2c)(const XArgs&, unsigned, size_t, hipStream_t);
  void (*xc2r)(const XArgs&, unsigned, size_t, hipStream_t);
  void (*zsolve)(const ColArgs&, double, int, int, unsigned, size_t, hipStream_t);
  void (*prep)();
  // grouped launches over many arrays of this length
  void (*xgrp)(bool, const XArgs&, const GrpArgs&, unsigned, size_t, hipStream_t);
  void (*cgrp)(bool, const ColArgs&, const GrpArgs&, unsigned, size_t, hipStream_t);
  void (*zgrp)(const ColArgs&, const GrpArgs&, unsigned, size_t, hipStream_t);
};

template <int... Rs>
constexpr int rmax_of() {
  int m = 0;
  for (int r : {Rs...}) m = r > m ? r : m;
  return m;
}

template <int... Rs>
struct CtLen {
  typedef RList<Rs...> RL;
  static constexpr int N = RProd<Rs...>::value;
  // values per thread of the column kernels: radix-16 stages hold 16
  static constexpr int RMAX = rmax_of<Rs...>();
  static constexpr int EPTC = RMAX >= 16 && ct_ept_col(N) < 16 ? 16 : ct_ept_col(N);
  static constexpr int TX = ct_col_tx(N);
  static constexpr int NTC = ct_nt_ept(TX * N, EPTC);
  static constexpr int BX = ct_x_pairs(N);
  static constexpr int NTX = ct_nt_x(BX * N);
  static constexpr long long TWB = (long long)TwCount<Rs...>::value() * sizeof(Tw);
  static constexpr long long DC = round16((long long)TX * N * sizeof(cplx_t));
  static constexpr long long DX = round16((long long)(BX + 1) * N * sizeof(cplx_t));
  // twiddles in LDS behind the column tile: when both fit in kCtLdsTw (two
  // workgroups per CU), or when the tile alone takes more than that (one
  // workgroup per CU either way) and both fit in the LDS
  static constexpr bool TWLC =
      TWB > 0 && (DC + TWB <= kCtLdsTw || (DC > kCtLdsTw && DC + TWB <= (long long)kLdsMax));
  // the table-addressed column kernel also keeps a table of N rows in LDS
  // (read from global memory instead, the tables cost more than the twiddles)
  static constexpr bool TWLT = TWLC && round16(DC + TWB) + N * (long long)sizeof(int2) <= (long long)kLdsMax;
  static constexpr bool TWLX = TWB > 0 && DX + TWB <= kCtLdsTw;
  // persistent column kernel: where the tile alone allows one workgroup per
  // CU; the table-addressed form keeps both row tables (2 N int2) in LDS
  static constexpr bool PIPE = DC > kCtLdsTw;
  static constexpr bool TWLP = TWLC && round16(DC + TWB) + 2 * N * (long long)sizeof(int2) <= (long long)kLdsMax;
  // waves per SIMD the LDS footprint allows: the register budget of each kernel
  static constexpr int WC = ct_waves(DC + (TWLC ? TWB : 0), NTC);
  static constexpr int WX = ct_waves(DX + (TWLX ? TWB : 0), NTX);
  template <class K>
  static void go(K k, unsigned g, int nt, size_t bytes, hipStream_t s, const ColArgs& a) {
    k<<<g, nt, bytes, s>>>(a);
  }
  template <class K>
  static void gox(K k, unsigned g, int nt, size_t bytes, hipStream_t s, const XArgs& a) {
    k<<<g, nt, bytes, s>>>(a);
  }
  static void col(bool fwd, const ColArgs& a, unsigned g, size_t bytes, hipStream_t s) {
    if (fwd) go(k_col_ct<real_t, true, TX, NTC, RL, WC, TWLC>, g, NTC, bytes, s, a);
    else go(k_col_ct<real_t, false, TX, NTC, RL, WC, TWLC>, g, NTC, bytes, s, a);
  }
  // pr: pruning maps; wide: 64-bit offsets (unpruned views only); a.ltab:
  // table-addressed rows
  static void colio(bool fwd, bool pr, bool wide, const ColArgs& a, unsigned g, size_t bytes, hipStream_t s) {
    typedef unsigned U32;
    typedef unsigned long long U64;
    if (a.ltab) {
      if (fwd) go(k_colio_ct<real_t, true, TX, NTC, RL, WC, TWLT, false, TabRows>, g, NTC, bytes, s, a);
      else go(k_colio_ct<real_t, false, TX, NTC, RL, WC, TWLT, false, TabRows>, g, NTC, bytes, s, a);
      return;
    }
    if (fwd) {
      if (pr) go(k_colio_ct<real_t, true, TX, NTC, RL, WC, TWLC, true, U32>, g, NTC, bytes, s, a);
      else if (wide) go(k_colio_ct<real_t, true, TX, NTC, RL, WC, TWLC, false, U64>, g, NTC, bytes, s, a);
      else go(k_colio_ct<real_t, true, TX, NTC, RL, WC, TWLC, false, U32>, g, NTC, bytes, s, a);
    } else {
      if (pr) go(k_colio_ct<real_t, false, TX, NTC, RL, WC, TWLC, true, U32>, g, NTC, bytes, s, a);
      else if (wide) go(k_colio_ct<real_t, false, TX, NTC, RL, WC, TWLC, false, U64>, g, NTC, bytes, s, a);
      else go(k_colio_ct<real_t, false, TX, NTC, RL, WC, TWLC, false, U32>, g, NTC, bytes, s, a);
    }
  }
  // tab: table-addressed rows (a.ltab), else unpruned rows at 32-bit offsets
  static void colpipe(bool fwd, bool tab, const ColArgs& a, unsigned g, size_t bytes, hipStream_t s) {
    if constexpr (PIPE) {
      if (tab) {
        if (fwd) go(k_colio_pipe<real_t, true, TX, NTC, RL, WC, TWLP, TabRows>, g, NTC, bytes, s, a);
        else go(k_colio_pipe<real_t, false, TX, NTC, RL, WC, TWLP, TabRows>, g, NTC, bytes, s, a);
      } else {
        if (fwd) go(k_colio_pipe<real_t, true, TX, NTC, RL, WC, TWLC, unsigned>, g, NTC, bytes, s, a);
        else go(k_colio_pipe<real_t, false, TX, NTC, RL, WC, TWLC, unsigned>, g, NTC, bytes, s, a);
      }
    }
  }
  static void xr2c(const XArgs& a, unsigned g, size_t bytes, hipStream_t s) {
    if (a.xt) gox(k_x_r2c_ct<real_t, BX, NTX, RL, WX, TWLX, true>, g, NTX, bytes, s, a);
    else gox(k_x_r2c_ct<real_t, BX, NTX, RL, WX, TWLX, false>, g, NTX, bytes, s, a);
  }
  static void xc2r(const XArgs& a, unsigned g, size_t bytes, hipStream_t s) {
    if (a.xt) gox(k_x_c2r_ct<real_t, BX, NTX, RL, WX, TWLX, true>, g, NTX, bytes, s, a);
    else gox(k_x_c2r_ct<real_t, BX, NTX, RL, WX, TWLX, false>, g, NTX, bytes, s, a);
  }
  static void zsolve(const ColArgs& a, double scale, int nxh, int ny, unsigned g, size_t bytes, hipStream_t s) {
    k_zsolve_ct<real_t, TX, NTC, RL, WC, TWLC><<<g, NTC, bytes, s>>>(a, scale, nxh, ny);
  }
  // raise the dynamic-LDS limit of every kernel of this length (host-side
  // attribute calls, made once when the length is planned, so that launches
  // alone remain for stream capture)
  template <bool FWD>
  static void prep_dir() {
    allow_big_lds(k_col_ct<real_t, FWD, TX, NTC, RL, WC, TWLC>);
    allow_big_lds(k_colio_ct<real_t, FWD, TX, NTC, RL, WC, TWLT, false, TabRows>);
    allow_big_lds(k_colio_ct<real_t, FWD, TX, NTC, RL, WC, TWLC, true, unsigned>);
    allow_big_lds(k_colio_ct<real_t, FWD, TX, NTC, RL, WC, TWLC, false, unsigned long long>);
    allow_big_lds(k_colio_ct<real_t, FWD, TX, NTC, RL, WC, TWLC, false, unsigned>);
    if constexpr (PIPE) {
      allow_big_lds(k_colio_pipe<real_t, FWD, TX, NTC, RL, WC, TWLP, TabRows>);
      allow_big_lds(k_colio_pipe<real_t, FWD, TX, NTC, RL, WC, TWLC, unsigned>);
    }
  }
  static void xgrp(bool r2c, const XArgs& a, const GrpArgs& ga, unsigned g, size_t bytes, hipStream_t s) {
    if (r2c) k_x_grp<real_t, true, BX, NTX, RL, WX, TWLX><<<g, NTX, bytes, s>>>(a, ga);
    else k_x_grp<real_t, false, BX, NTX, RL, WX, TWLX><<<g, NTX, bytes, s>>>(a, ga);
  }
  static void cgrp(bool fwd, const ColArgs& a, const GrpArgs& ga, unsigned g, size_t bytes, hipStream_t s) {
    if (fwd) k_col_grp<real_t, true, TX, NTC, RL, WC, TWLC><<<g, NTC, bytes, s>>>(a, ga);
    else k_col_grp<real_t, false, TX, NTC, RL, WC, TWLC><<<g, NTC, bytes, s>>>(a, ga);
  }
  static void zgrp(const ColArgs& a, const GrpArgs& ga, unsigned g, size_t bytes, hipStream_t s) {
    k_zs_grp<real_t, TX, NTC, RL, WC, TWLC><<<g, NTC, bytes, s>>>(a, ga);
  }
  static void prep() {
    prep_dir<true>();
    prep_dir<false>();
    allow_big_lds(k_x_grp<real_t, true, BX, NTX, RL, WX, TWLX>);
    allow_big_lds(k_x_grp<real_t, false, BX, NTX, RL, WX, TWLX>);
    allow_big_lds(k_col_grp<real_t, true, TX, NTC, RL, WC, TWLC>);
    allow_big_lds(k_col_grp<real_t, false, TX, NTC, RL, WC, TWLC>);
    allow_big_lds(k_zs_grp<real_t, TX, NTC, RL, WC, TWLC>);
    allow_big_lds(k_x_r2c_ct<real_t, BX, NTX, RL, WX, TWLX, true>);
    allow_big_lds(k_x_r2c_ct<real_t, BX, NTX, RL, WX, TWLX, false>);
    allow_big_lds(k_x_c2r_ct<real_t, BX, NTX, RL, WX, TWLX, true>);
    allow_big_lds(k_x_c2r_ct<real_t, BX, NTX, RL, WX, TWLX, false>);
    allow_big_lds(k_zsolve_ct<real_t, TX, NTC, RL, WC, TWLC>);
  }
  static CtInfo info() {
    CtInfo c;
    std::memset(&c, 0, sizeof(c));
    c.TX = TX; c.NTC = NTC; c.BX = BX; c.NTX = NTX;
    const int rs[] = {Rs...};
    c.nr = (int)sizeof...(Rs);
    for (int i = 0; i < c.nr; i++) c.radix[i] = rs[i];
    c.twl_col = TWLC; c.twl_x = TWLX; c.twl_tab = TWLT; c.twb = TWB;
    c.pipe = PIPE; c.twl_pipe_tab = TWLP;
    c.col = &col; c.colio = &colio; c.colpipe = &colpipe; c.xr2c = &xr2c; c.xc2r = &xc2r; c.zsolve = &zsolve; c.prep = &prep;
    c.xgrp = &xgrp; c.cgrp = &cgrp; c.zgrp = &zgrp;
    return c;
  }
};

// Specialised kernel set for a length, if one exists and its radix schedule
// matches the length plan (the twiddle tables are laid out by that schedule).
static const CtInfo* find_ct(const Len1D& L) {
  if (!L.ct_ok) return nullptr;
  static const std::map<int, CtInfo> tab = {
      {32, CtLen<8, 4>::info()},       {48, CtLen<4, 4, 3>::info()},    {64, CtLen<8, 8>::info()},
      {96, CtLen<8, 4, 3>::info()},    {128, CtLen<8, 4, 4>::info()},   {160, CtLen<8, 4, 5>::info()},
      {192, CtLen<8, 8, 3>::info()},   {256, CtLen<8, 8, 4>::info()},   {320, CtLen<8, 8, 5>::info()},
      {384, CtLen<8, 4, 4, 3>::info()}, {512, CtLen<8, 8, 8>::info()},  {1024, CtLen<8, 8, 4, 4>::info()},
      {2048, CtLen<16, 16, 8>::info()}};
  auto it = tab.find(L.d.n);
  if (it == tab.end()) return nullptr;
  const CtInfo& c = it->second;
  if (c.nr != L.d.nst) return nullptr;
  for (int i = 0; i < c.nr; i++)
    if (c.radix[i] != L.d.radix[i]) return nullptr;
  return &c;
}

bool len_has_ct(int n) {
  Len1D probe;
  std::memset(&probe.d, 0, sizeof(probe.d));
  const std::vector<int> f = default_factors(n);
  probe.d.n = n;
  probe.d.nst = (int)f.size();
  for (size_t i = 0; i < f.size() && i < (size_t)kMaxStages; i++) probe.d.radix[i] = f[i];
  return find_ct(probe) != nullptr;
}

std::string len_describe(int n) {
  const Len1D& L = get_len(n);
  std::string r = "radix ";
  for (int s = 0; s < L.d.nst; s++) r += (s ? "x" : "") + std::to_string(L.d.radix[s]);
  if (find_ct(L)) return r + ", specialised";
  return r + ", generic, tile " + std::to_string(L.tile_bytes / 1024) + " KiB, segment " +
         std::to_string(L.col_seg) + " B";
}

template <class F>
static void prep_generic_nt() {
  F::template prep<64>();
  F::template prep<128>();
  F::template prep<256>();
  F::template prep<512>();
  F::template prep<1024>();
}

// Kernel attributes (the dynamic-LDS limit) are set when a length is first
// planned, never at launch: launches are then the only device-side work of a
// transform, which is what stream capture into a graph requires.
static void prep_plane_kernels(int n);
static void prep_kernels(const Len1D& L) {
  static bool generic_done = false;
  if (!generic_done) {
    prep_generic_nt<GoXr2c>();
    prep_generic_nt<GoXc2r>();
    prep_generic_nt<GoCol<true>>();
    prep_generic_nt<GoCol<false>>();
    allow_big_lds(k_spectrum<real_t>);
    generic_done = true;
  }
  if (const CtInfo* c = find_ct(L)) c->prep();
  prep_plane_kernels(L.d.n);
}

static bool g_use_ct = true;
static bool g_alt_order = true;
static int g_flat_order = 2;  // 0 XCD-aware, 1 dispatch order, 2 by shape
static bool g_wide_offsets = false;
static bool g_dct_direct = true;
void set_dct_direct(bool on) { g_dct_direct = on; }
void set_wide_offsets(bool on) { g_wide_offsets = on; }
void set_flat_order(int mode) { g_flat_order = mode < 0 || mode > 2 ? 2 : mode; }
int flat_order() { return g_flat_order; }
// Order of the tiles of a column pass over one contiguous inner axis.  By
// shape: dispatch order when the row stride (one plane of the view) lies
// within 1/128 above 512 KiB, 1 MiB or 2 MiB.  With the XCD-aware order the
// 8 XCDs' concurrent tiles sit in 8 regions of each row, and rows that
// start a near power of two apart put those regions on few DRAM channels;
// in dispatch order the tiles running at once cover one contiguous range of
// each row.  Measured on MI355X (Z fwd / bwd, us, XCD-aware -> dispatch):
// FP32 512^3 248/213 -> 210/189, 512x256x512 113/114 -> 104/101,
// 1024x512x128 99/96 -> 93/87; FP64 512^3 448/419 -> 441/397,
// 512x256x512 249/206 -> 212/186, 256x512x512 362/319 -> 335/286; strides
// away from those powers lose with dispatch order (FP64 384^3 176/175 ->
// 211/198, 128x128x1024 54/52 -> 76/76), as do FP32 256x512x512 (+4%) and
// FP64 1024x512x128 (+10%, a 4 MiB stride).
static bool flat_dispatch_order(const ColSpec& c) {
  if (g_flat_order != 2) return g_flat_order == 1;
  const long long stride = std::max(c.sl_in, c.sl_out) * (long long)sizeof(cplx_t);
  for (long long p = 1LL << 19; p <= 1LL << 21; p <<= 1)
    if (stride >= p && stride - p <= p / 128) return true;
  return false;
}
void set_tile_order(bool alternate) { g_alt_order = alternate; }
bool tile_order_alternate() { return g_alt_order; }
void set_specialised(bool on) { g_use_ct = on; }

// row-pair count per X tile
static long long x_pairs_per_tile(const Len1D& L, int nx, long long pairs) {
  const long long tgt = (long long)(L.tile_bytes / sizeof(cplx_t));
  long long B = std::max<long long>(1, tgt / nx);
  // keep at least ~2048 workgroups in flight when the problem allows
  B = std::min<long long>(B, std::max<long long>(1, pairs / 2048));
  return B;
}

static void x_setup(XArgs& a, const Len1D& L, long long nrows, int nx, size_t* bytes, int* nt, unsigned* nwg) {
  const long long pairs = (nrows + 1) / 2;
  long long B = x_pairs_per_tile(L, nx, pairs);
  const int cpt = len_cap_per_thread(L);
  while (B > 1 && B * nx > 1024LL * cpt) B--;
  a.B = (int)B;
  a.L = L.d;
  a.tw = L.tw;
  const long long elems = B * nx;
  *bytes = lds_layout((size_t)elems * sizeof(cplx_t), L, &a.tw_lds, &a.tw_off);
  if (*bytes > kLdsMax) throw Error{"X transform does not fit in LDS", -1};
  *nt = pick_nt(elems, cpt);
  const long long g = (pairs + B - 1) / B;
  if (g > 0x7fffffffLL) throw Error{"grid too large", -1};
  *nwg = (unsigned)g;
}

// LDS bytes of a specialised tile of `data` bytes; sets the twiddle placement
static size_t ct_lds(size_t data, bool twl, const CtInfo& c, int* tw_lds, int* tw_off) {
  const size_t d16 = (size_t)round16((long long)data);
  *tw_lds = twl ? 1 : 0;
  *tw_off = twl ? (int)d16 : 0;
  return twl ? d16 + (size_t)c.twb : data;
}

// false when the 32-bit tile offsets of the specialised X kernels do not fit
static bool x_setup_ct(XArgs& a, const Len1D& L, const CtInfo& c, long long nrows, long long si_row,
                       long long so_row, size_t* bytes, unsigned* nwg) {
  const long long span = (2LL * c.BX + 1) * std::max(si_row, so_row) * (long long)sizeof(cplx_t);
  if (span >= (1LL << 32)) return false;
  const long long pairs = (nrows + 1) / 2;
  a.B = c.BX;
  a.L = L.d;
  a.tw = L.tw;
  *bytes = ct_lds((size_t)(c.BX + 1) * L.d.n * sizeof(cplx_t), c.twl_x, c, &a.tw_lds, &a.tw_off);
  const long long g = (pairs + c.BX - 1) / c.BX;
  if (g > 0x7fffffffLL) throw Error{"grid too large", -1};
  *nwg = (unsigned)g;
  return true;
}

static bool g_x_wave = false;
void set_x_wave(bool on) { g_x_wave = on; }

// the wave kernels (X rows of 512): four row pairs per workgroup
static bool launch_x_wave(bool r2c, const void* in, void* out, long long nrows, long long si_row, long long so_row,
                          int nk, const Len1D& L, bool rev, hipStream_t st) {
  if (!g_x_wave || !g_use_ct || L.d.n != 512 || !L.twf) return false;
  // workgroups walk the row pairs at a stride of the grid: at most 8 per CU
  const long long g = std::min<long long>(((nrows + 1) / 2 + 3) / 4, 256LL * 8);
  XArgs a;
  std::memset(&a, 0, sizeof(a));
  a.in = in; a.out = out; a.nrows = nrows; a.si_row = si_row; a.so_row = so_row;
  a.nxo = nk;
  a.rev = rev ? 1 : 0;
  const size_t lds = 512 * sizeof(Tw);
  if (r2c) k_x_r2c_w512<real_t><<<(unsigned)g, 256, lds, st>>>(a, L.twf);
  else k_x_c2r_w512<real_t><<<(unsigned)g, 256, lds, st>>>(a, L.twf);
  check_hip(hipGetLastError(), "k_x_w512 launch");
  return true;
}

void launch_x_r2c(const real_t* in, cplx_t* out, long long nrows, long long si_row, long long so_row, int nx,
                  int nxo, hipStream_t st, bool rev) {
  if (nrows <= 0 || nxo <= 0) return;
  if (nx > max_len()) {
    launch_x_long(true, in, out, nrows, si_row, so_row, nx, nxo, st);
    return;
  }
  const Len1D& L = get_len(nx);
  if (launch_x_wave(true, in, out, nrows, si_row, so_row, nxo, L, rev, st)) return;
  XArgs a;
  std::memset(&a, 0, sizeof(a));
  a.in = in; a.out = out; a.nrows = nrows; a.si_row = si_row; a.so_row = so_row;
  a.nxo = nxo; a.dnxo = make_fdiv(nxo);
  a.rev = rev ? 1 : 0;
  if (const CtInfo* c = g_use_ct ? find_ct(L) : nullptr) {
    size_t bytes; unsigned nwg;
    if (x_setup_ct(a, L, *c, nrows, si_row, so_row, &bytes, &nwg)) {
      c->xr2c(a, nwg, bytes, st);
      check_hip(hipGetLastError(), "k_x_r2c_ct launch");
      return;
    }
  }
  size_t bytes; int nt; unsigned nwg;
  x_setup(a, L, nrows, nx, &bytes, &nt, &nwg);
  dispatch_nt<GoXr2c>(nt, a, nwg, bytes, st);
  check_hip(hipGetLastError(), "k_x_r2c launch");
}

bool launch_x_r2c_tab(const real_t* in, cplx_t* out, long long nrows, int nx, int nxo, const int2* xt,
                      hipStream_t st) {
  if (nrows <= 0 || nxo <= 0) return true;
  if (nx > max_len()) return false;
  const Len1D& L = get_len(nx);
  const CtInfo* c = g_use_ct ? find_ct(L) : nullptr;
  if (!c) return false;
  XArgs a;
  std::memset(&a, 0, sizeof(a));
  a.in = in; a.out = out; a.nrows = nrows; a.si_row = nx; a.so_row = 0;
  a.nxo = nxo; a.dnxo = make_fdiv(nxo);
  a.xt = xt;
  size_t bytes; unsigned nwg;
  if (!x_setup_ct(a, L, *c, nrows, nx, 1, &bytes, &nwg)) return false;
  c->xr2c(a, nwg, bytes, st);
  check_hip(hipGetLastError(), "k_x_r2c_ct (table) launch");
  return true;
}

bool launch_x_c2r_tab(const cplx_t* in, real_t* out, long long nrows, int nx, int nxi, const int2* xt,
                      hipStream_t st) {
  if (nrows <= 0) return true;
  if (nx > max_len()) return false;
  const Len1D& L = get_len(nx);
  const CtInfo* c = g_use_ct ? find_ct(L) : nullptr;
  if (!c) return false;
  XArgs a;
  std::memset(&a, 0, sizeof(a));
  a.in = in; a.out = out; a.nrows = nrows; a.si_row = 0; a.so_row = nx;
  a.nxo = nxi; a.dnxo = make_fdiv(nxi > 0 ? nxi : 1);
  a.xt = xt;
  size_t bytes; unsigned nwg;
  if (!x_setup_ct(a, L, *c, nrows, 1, nx, &bytes, &nwg)) return false;
  c->xc2r(a, nwg, bytes, st);
  check_hip(hipGetLastError(), "k_x_c2r_ct (table) launch");
  return true;
}

void launch_x_c2r(const cplx_t* in, real_t* out, long long nrows, long long si_row, long long so_row, int nx,
                  int nxi, hipStream_t st, bool rev) {
  if (nrows <= 0) return;
  if (nx > max_len()) {
    launch_x_long(false, in, out, nrows, si_row, so_row, nx, nxi, st);
    return;
  }
  const Len1D& L = get_len(nx);
  if (launch_x_wave(false, in, out, nrows, si_row, so_row, nxi, L, rev, st)) return;
  XArgs a;
  std::memset(&a, 0, sizeof(a));
  a.in = in; a.out = out; a.nrows = nrows; a.si_row = si_row; a.so_row = so_row;
  a.nxo = nxi; a.dnxo = make_fdiv(nxi > 0 ? nxi : 1);
  a.rev = rev ? 1 : 0;
  if (const CtInfo* c = g_use_ct ? find_ct(L) : nullptr) {
    size_t bytes; unsigned nwg;
    if (x_setup_ct(a, L, *c, nrows, si_row, so_row, &bytes, &nwg)) {
      c->xc2r(a, nwg, bytes, st);
      check_hip(hipGetLastError(), "k_x_c2r_ct launch");
      return;
    }
  }
  size_t bytes; int nt; unsigned nwg;
  x_setup(a, L, nrows, nx, &bytes, &nt, &nwg);
  dispatch_nt<GoXc2r>(nt, a, nwg, bytes, st);
  check_hip(hipGetLastError(), "k_x_c2r launch");
}

// Plane kernels (k_xy_plane_fwd / k_yx_plane_bwd) for the (nx, ny) pairs
// instantiated below; mode 0 off, 1 wherever they apply, 2 (default) on the
// grids where they measured faster: up to 2^22 points, and at least 64 planes
// unless the planes are 64 x 64 or smaller (one workgroup per plane).
// Measured on MI355X, eager fwd+bwd, separate passes -> plane kernels (us):
// FP64 32^3 27.2 -> 14.0, 64^3 24.8 -> 19.8, 64x64x256 47.6 -> 33.9,
// 128x64x64 33.2 -> 29.8; FP32 64x64x256 33.5 -> 23.7, 128^3 48.1 -> 40.8,
// 128x128x64 32.9 -> 32.1, but 128x128x32 26.5 -> 28.2.
static int g_plane_mode = 2;
void set_plane_kernels(int mode) { g_plane_mode = mode < 0 || mode > 2 ? 2 : mode; }

template <class RLX, class RLY>
struct PlaneLen {
  typedef PlaneGeom<real_t, RLX, RLY> G;
  // four threads per butterfly of the first X stage (up to 1024): the loads,
  // the separation and the stores spread over all of them, the butterflies
  // over a quarter (64^3 FP64 fwd+bwd 22.4 -> 19.8 us against one thread
  // per butterfly)
  static constexpr int NT0 = G::BX * G::NX / 8 * 4;
  static constexpr int NT = NT0 < 64 ? 64 : (NT0 > 1024 ? 1024 : (NT0 + 63) / 64 * 64);
  static void prep() {
    allow_big_lds(k_xy_plane_fwd<real_t, RLX, RLY, NT>);
    allow_big_lds(k_yx_plane_bwd<real_t, RLX, RLY, NT>);
  }
  static void go(bool fwd, const PlaneArgs& a, int nz, size_t bytes, hipStream_t s) {
    if (fwd) k_xy_plane_fwd<real_t, RLX, RLY, NT><<<nz, NT, bytes, s>>>(a);
    else k_yx_plane_bwd<real_t, RLX, RLY, NT><<<nz, NT, bytes, s>>>(a);
  }
};

struct PlaneInfo {
  long long data;  // LDS bytes of the two tiles
  void (*prep)();
  void (*go)(bool, const PlaneArgs&, int, size_t, hipStream_t);
};
template <class RLX, class RLY>
static PlaneInfo plane_info() {
  typedef PlaneLen<RLX, RLY> P;
  return PlaneInfo{P::G::DATA, &P::prep, &P::go};
}

static const PlaneInfo* find_plane(int nx, int ny);
// LDS limits of the plane kernels, raised when a length they use is planned
// (host-side attribute calls outside any stream capture)
static void prep_plane_kernels(int n) {
  static std::once_flag once;
  if (n != 32 && n != 64 && n != 128) return;
  std::call_once(once, [] {
    for (int x : {32, 64, 128})
      for (int y : {32, 64, 128})
        if (const PlaneInfo* q = find_plane(x, y)) q->prep();
  });
}

static const PlaneInfo* find_plane(int nx, int ny) {
  typedef RList<8, 4> R32;
  typedef RList<8, 8> R64;
  typedef RList<8, 4, 4> R128;
  static const std::map<std::pair<int, int>, PlaneInfo> tab = {
      {{32, 32}, plane_info<R32, R32>()},   {{64, 64}, plane_info<R64, R64>()},
      {{32, 64}, plane_info<R32, R64>()},   {{64, 32}, plane_info<R64, R32>()},
      {{128, 64}, plane_info<R128, R64>()}, {{64, 128}, plane_info<R64, R128>()},
      {{128, 128}, plane_info<R128, R128>()}};
  auto it = tab.find({nx, ny});
  return it == tab.end() ? nullptr : &it->second;
}

bool launch_xy_plane(bool fwd, const void* in, void* out, int nx, int ny, int nz, hipStream_t st) {
  if (g_plane_mode == 0 || !g_use_ct || nz <= 0) return false;
  if (g_plane_mode == 2 &&
      ((long long)nx * ny * nz > (1LL << 22) || (nz < 64 && (long long)nx * ny > 64 * 64)))
    return false;
  const PlaneInfo* p = find_plane(nx, ny);
  if (!p) return false;
  const Len1D& Lx = get_len(nx);
  const Len1D& Ly = get_len(ny);
  // the kernels' radix schedules must be the plans' (the twiddle layout)
  if (!find_ct(Lx) || !find_ct(Ly)) return false;
  PlaneArgs a;
  a.in = in; a.out = out;
  a.twx = Lx.tw; a.twy = Ly.tw;
  a.ntwx = Lx.d.twcount; a.ntwy = nx == ny ? 0 : Ly.d.twcount;
  const size_t bytes = (size_t)p->data + (size_t)(a.ntwx + a.ntwy) * sizeof(Tw);
  if (bytes > kLdsMax) return false;
  p->go(fwd, a, nz, bytes, st);
  check_hip(hipGetLastError(), fwd ? "k_xy_plane_fwd launch" : "k_yx_plane_bwd launch");
  return true;
}

// The byte offsets of a column tile relative to its first column stay below
// 2^32: lanes span at most TX/n_inner + 2 outer indices and n_inner inner
// ones, plus `rows` rows.
static bool col_fits32(const ColSpec& c, int TX, int rows) {
  const long long nout = TX / std::max(c.n_inner, 1) + 2;
  auto span = [&](long long si, long long sl, long long so) {
    return (nout * so + (long long)c.n_inner * si + (long long)rows * sl + 1) * (long long)sizeof(cplx_t);
  };
  return span(c.si_in, c.sl_in, c.so_in) < (1LL << 32) && span(c.si_out, c.sl_out, c.so_out) < (1LL << 32);
}

// Persistent column kernels (k_colio_pipe) for the lengths whose tile takes
// more than half the LDS (2048 points): on (1, default) or off (0).  Measured
// on MI355X, share-mode rank of 2048^3 on 16x16, FP32, off -> on (us):
// ftran_y 101.0 / 100.3 -> 91.4 / 93.2, btran_y 108.6 / 109.1 -> 100.3 / 99.3,
// ftran_z 89.8 / 82.7 -> 80.9 / 81.1; fwd+bwd 0.569 / 0.561 -> 0.533 / 0.547 ms.
static int g_col_pipe = 1;
void set_col_pipe(int on) { g_col_pipe = on ? 1 : 0; }

// workgroups of a persistent launch: one per CU of the current device, and
// at least min(ntiles, 8), so that each XCD's share of the tiles has a
// workgroup (k_colio_pipe)
static unsigned pipe_grid(long long ntiles) {
  int dev = 0, ncu = 0;
  check_hip(hipGetDevice(&dev), "hipGetDevice");
  check_hip(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
  return (unsigned)std::max<long long>(1, std::min<long long>(ntiles, std::max(ncu, 8)));
}

bool launch_col_tab(const ColSpec& c, const int2* ltab, const int2* stab, hipStream_t st) {
  if (c.n_inner <= 0 || c.n_outer <= 0) return true;
  if (c.kind != COL_C2C || !g_use_ct || c.n > max_len()) return false;
  const Len1D& L = get_len(c.n);
  const CtInfo* ct = find_ct(L);
  if (!ct) return false;
  const long long ncol = (long long)c.n_inner * c.n_outer;
  if (ncol > 0x7fffffffLL) return false;
  ColArgs a;
  std::memset(&a, 0, sizeof(a));
  a.in = c.in; a.out = c.out;
  a.n_inner = c.n_inner; a.n_outer = c.n_outer;
  a.kind = COL_C2C;
  a.ltab = ltab; a.stab = stab;
  a.rev = c.rev ? 1 : 0;
  a.xcd = 1;  // table-addressed rows: tiles share lines across peer blocks
  const int TX = ct->TX;
  a.TX = TX;
  a.dTX = make_fdiv(TX);
  a.tiles = (c.n_inner + TX - 1) / TX;
  a.L = L.d;
  a.tw = L.tw;
  if (g_col_pipe && ct->pipe) {
    size_t pb = ct_lds((size_t)TX * c.n * sizeof(cplx_t), ct->twl_pipe_tab, *ct, &a.tw_lds, &a.tw_off);
    a.tab_off = (int)round16((long long)pb);
    pb = (size_t)a.tab_off + 2 * (size_t)c.n * sizeof(int2);
    if (pb <= kLdsMax) {
      ct->colpipe(c.fwd, true, a, pipe_grid((ncol + TX - 1) / TX), pb, st);
      check_hip(hipGetLastError(), "k_colio_pipe (table) launch");
      return true;
    }
  }
  size_t bytes = ct_lds((size_t)TX * c.n * sizeof(cplx_t), ct->twl_tab, *ct, &a.tw_lds, &a.tw_off);
  a.tab_off = (int)round16((long long)bytes);
  bytes = (size_t)a.tab_off + (size_t)c.n * sizeof(int2);
  if (bytes > kLdsMax) return false;
  ct->colio(c.fwd, false, false, a, (unsigned)((ncol + TX - 1) / TX), bytes, st);
  check_hip(hipGetLastError(), "k_colio_ct (table) launch");
  return true;
}

bool launch_col(const ColSpec& c, hipStream_t st) {
  if (c.n_inner <= 0 || c.n_outer <= 0 || c.nst <= 0) return c.cheby;
  ColArgs a;
  std::memset(&a, 0, sizeof(a));
  a.in = c.in; a.out = c.out;
  a.si_in = c.si_in; a.sl_in = c.sl_in; a.so_in = c.so_in;
  a.si_out = c.si_out; a.sl_out = c.sl_out; a.so_out = c.so_out;
  a.n_inner = c.n_inner; a.n_outer = c.n_outer;
  a.kind = c.kind;
  a.rev = c.rev ? 1 : 0;
  // one contiguous inner axis (the Z pass over the flattened (x, y) index):
  // tiles never share a line, so they may also run in dispatch order, which
  // lets the concurrently running tiles cover one contiguous range of each
  // row (see set_flat_order); views whose tiles share lines at their edges
  // (the Y pass over rows of nx/2+1 elements) keep the XCD-aware order
  const bool flat = c.n_outer == 1 && c.si_in == 1 && c.si_out == 1;
  a.xcd = (flat && flat_dispatch_order(c)) ? 0 : 1;
  a.nld = c.nld; a.ld_lo = c.ld_lo; a.ld_hi = c.ld_hi; a.ld_shift = c.ld_shift;
  a.nst = c.nst; a.st_lo = c.st_lo; a.st_shift = c.st_shift;
  int nfft;
  const Len1D* L = nullptr;
  if (c.kind == COL_COPY) nfft = c.nld;
  else {
    int m = c.n;
    if (c.kind == COL_DCT1) {
      if (c.n < 2) throw Error{"DCT-I needs at least 2 points", -1};
      m = 2 * (c.n - 1);
    } else if (c.kind == COL_DST1) m = 2 * (c.n + 1);
    if (m > max_len()) {
      launch_col_long(c, m, st);
      return false;
    }
    L = &get_len(m);
    nfft = m;
  }
  const int rows = std::max(nfft, c.nld + (c.kind == COL_DST1 ? 2 : 0));
  a.nrows_lds = rows;
  a.ld_lenfast = (c.sl_in == 1 && c.si_in != 1) ? 1 : 0;
  a.st_lenfast = (c.sl_out == 1 && c.si_out != 1) ? 1 : 0;
  a.dnld = make_fdiv(c.nld);
  a.dnst = make_fdiv(c.nst);
  // DCT-I and DST-I (REDFT00 / RODFT00) are the same real transforms in
  // ftran and btran: their extensions always go through the forward FFT
  const bool fwd = c.fwd || c.kind == COL_DCT1 || c.kind == COL_DST1;
  if (const CtInfo* ct = (L && g_use_ct) ? find_ct(*L) : nullptr) {
    const int TX = ct->TX;
    a.TX = TX;
    a.dTX = make_fdiv(TX);
    a.tiles = (c.n_inner + TX - 1) / TX;
    a.L = L->d;
    a.tw = L->tw;
    const size_t bytes = ct_lds((size_t)TX * rows * sizeof(cplx_t), ct->twl_col, *ct, &a.tw_lds, &a.tw_off);
    if (bytes > kLdsMax) throw Error{"column transform does not fit in LDS", -1};
    const long long ncol = (long long)c.n_inner * c.n_outer;
    if (ncol > 0x7fffffffLL) throw Error{"grid too large", -1};
    // identity row maps (no zero gap, no shift, all rows): no pruning
    const bool ld_id = c.nld == nfft && (c.ld_lo >= nfft || (c.ld_hi <= c.ld_lo && c.ld_shift == 0));
    const bool st_id = c.nst == nfft && (c.st_lo >= nfft || c.st_shift == 0);
    const bool pr = !(ld_id && st_id);
    const bool fits = col_fits32(c, TX, rows) && !g_wide_offsets;
    if (c.kind == COL_C2C && !a.ld_lenfast && !a.st_lenfast && fits && !pr && g_col_pipe && ct->pipe) {
      ct->colpipe(fwd, false, a, pipe_grid((ncol + TX - 1) / TX), bytes, st);
      check_hip(hipGetLastError(), "k_colio_pipe launch");
      return false;
    }
    if (c.kind == COL_C2C && !a.ld_lenfast && !a.st_lenfast && (fits || !pr)) {
      ct->colio(fwd, pr, !fits, a, (unsigned)((ncol + TX - 1) / TX), bytes, st);
      check_hip(hipGetLastError(), "k_colio_ct launch");
      return false;
    }
    // DCT-I without the fused Chebyshev recurrence: the direct-load kernel
    // reads the even extension's mirrored rows from the input itself (no LDS
    // load pass); the store map keeps the first nst rows
    const bool ld_all = c.nld == c.n && (c.ld_lo >= c.n || (c.ld_hi <= c.ld_lo && c.ld_shift == 0));
    if (c.kind == COL_DCT1 && !c.cheby && g_dct_direct && ld_all && !a.ld_lenfast && !a.st_lenfast && fits) {
      a.mirror = c.n;
      ct->colio(true, true, false, a, (unsigned)((ncol + TX - 1) / TX), bytes, st);
      check_hip(hipGetLastError(), "k_colio_ct (DCT-I) launch");
      return false;
    }
    const long long nwg = (long long)a.tiles * c.n_outer;
    if (nwg > 0x7fffffffLL) throw Error{"grid too large", -1};
    size_t b2 = bytes;
    if (c.cheby && c.kind == COL_DCT1 && 64 % TX == 0 && ct->NTC % 64 == 0 &&
        (c.nst + ct->NTC / TX - 1) / (ct->NTC / TX) <= kChebyRM) {
      // per-wave FP64 sums of the fused recurrence behind the tile
      a.cheby = 1;
      a.cheby_off = (int)round16((long long)bytes);
      a.cheby_scale = c.cheby_scale;
      a.cheby_L = c.cheby_L;
      b2 = (size_t)a.cheby_off + (size_t)(ct->NTC / 64) * TX * 4 * sizeof(double);
      if (b2 > kLdsMax) { a.cheby = 0; b2 = bytes; }
    }
    // DCT-I with the fused recurrence: the first stage reads the mirrored
    // rows from HBM instead of an LDS load pass and extension copy
    if (c.kind == COL_DCT1 && g_dct_direct && ld_all && !a.ld_lenfast && !a.st_lenfast && fits) a.mirror = c.n;
    ct->col(fwd, a, (unsigned)nwg, b2, st);
    check_hip(hipGetLastError(), "k_col_ct launch");
    return a.cheby != 0;
  }
  const int cpt = L ? len_cap_per_thread(*L) : kEPT;
  // tile width: one segment of adjacent columns (128 bytes unless the plan
  // says otherwise) per row where capacity allows
  int TX = std::max(1, (int)((L ? L->col_seg : 128) / (int)sizeof(cplx_t)));
  const long long tgt = (long long)((L ? (size_t)L->tile_bytes : kTileBytes) / sizeof(cplx_t));
  while (TX > 1 && (long long)TX * rows > std::max<long long>(tgt, rows)) TX /= 2;
  while (TX > 1 && (long long)TX * rows > 1024LL * cpt) TX /= 2;
  while (TX > 1 && TX / 2 >= c.n_inner) TX /= 2;
  a.TX = TX;
  a.dTX = make_fdiv(TX);
  a.tiles = (c.n_inner + TX - 1) / TX;
  const long long nwg = (long long)a.tiles * c.n_outer;
  if (nwg > 0x7fffffffLL) throw Error{"grid too large", -1};
  const long long elems = (long long)TX * rows;
  size_t bytes;
  if (L) {
    a.L = L->d;
    a.tw = L->tw;
    bytes = lds_layout((size_t)elems * sizeof(cplx_t), *L, &a.tw_lds, &a.tw_off);
  } else {
    a.L.n = nfft;
    bytes = (size_t)elems * sizeof(cplx_t);
  }
  if (bytes > kLdsMax) throw Error{"column transform does not fit in LDS", -1};
  const int nt = pick_nt(elems, cpt);
  if (fwd) dispatch_nt<GoCol<true>>(nt, a, (unsigned)nwg, bytes, st);
  else dispatch_nt<GoCol<false>>(nt, a, (unsigned)nwg, bytes, st);
  check_hip(hipGetLastError(), "k_col launch");
  return false;
}

bool launch_zsolve(cplx_t* spec, int nxh, int ny, int nz, double scale, hipStream_t st) {
  const Len1D& L = get_len(nz);
  const CtInfo* ct = g_use_ct ? find_ct(L) : nullptr;
  if (!ct) return false;
  ColArgs a;
  std::memset(&a, 0, sizeof(a));
  const long long ninner = (long long)nxh * ny;
  if (ninner > 0x7fffffffLL) return false;
  a.in = spec; a.out = spec;
  a.si_in = 1; a.sl_in = ninner; a.si_out = 1; a.sl_out = ninner;
  a.n_inner = (int)ninner; a.n_outer = 1;
  a.kind = COL_C2C;
  const int TX = ct->TX;
  a.TX = TX;
  a.tiles = (int)((ninner + TX - 1) / TX);
  a.L = L.d;
  a.tw = L.tw;
  if ((ninner + (long long)nz * ninner + TX) * (long long)sizeof(cplx_t) >= (1LL << 32)) return false;
  const size_t bytes = ct_lds((size_t)TX * nz * sizeof(cplx_t), ct->twl_col, *ct, &a.tw_lds, &a.tw_off);
  if (bytes > kLdsMax) return false;
  ct->zsolve(a, scale, nxh, ny, (unsigned)a.tiles, bytes, st);
  check_hip(hipGetLastError(), "k_zsolve_ct launch");
  return true;
}

// ---------------------------------------------------------------------------
// grouped Poisson batches: the five passes of every solve of a batch as one
// launch per pass and transform length
// ---------------------------------------------------------------------------
struct GrpStage {
  int kind;  // 0 X r2c, 1 Y forward, 2 Z solve, 3 Y backward, 4 X c2r
  const CtInfo* ct;
  XArgs xa;
  ColArgs ca;
  GrpArgs ga;
  unsigned nwg;
  size_t bytes;
};
struct GroupBatch {
  std::vector<GrpStage> st;
  void* dmem = nullptr;
};

static bool grp_eligible(int nx, int ny, int nz, const CtInfo** cx, const CtInfo** cy, const CtInfo** cz) {
  if (!g_use_ct || nx < 2 || (nx & 1)) return false;
  *cx = find_ct(get_len(nx));
  *cy = find_ct(get_len(ny));
  *cz = find_ct(get_len(nz));
  if (!*cx || !*cy || !*cz) return false;
  const long long nxh = nx / 2 + 1, plane = nxh * ny;
  // 32-bit tile offsets of every kernel (the single-array checks)
  if ((2LL * (*cx)->BX + 1) * nxh * (long long)sizeof(cplx_t) >= (1LL << 32)) return false;
  if ((plane * (nz + 2) + (*cz)->TX) * (long long)sizeof(cplx_t) >= (1LL << 32)) return false;
  if (nxh * nz > 0x7fffffffLL || plane > 0x7fffffffLL) return false;
  return true;
}

GroupBatch* group_batch_create(int mode, int n, const int* shapes, real_t* const* fields, void* const* works,
                               std::vector<int>* singles) {
  singles->clear();
  struct Req {
    int i, nx, ny, nz;
    const CtInfo *cx, *cy, *cz;
  };
  std::vector<Req> rq;
  for (int i = 0; i < n; i++) {
    Req r{i, shapes[3 * i], shapes[3 * i + 1], shapes[3 * i + 2], nullptr, nullptr, nullptr};
    if (grp_eligible(r.nx, r.ny, r.nz, &r.cx, &r.cy, &r.cz)) rq.push_back(r);
    else singles->push_back(i);
  }
  GroupBatch* B = new GroupBatch;
  // host images of every stage's descriptor and tile tables
  struct HostStage {
    GrpStage s;
    std::vector<GrpDesc> d;
    std::vector<int> treq, tfirst;
  };
  std::vector<HostStage> hs;
  // passes: 0 X r2c, 1 Y forward, 2 Z solve, 3 Y backward, 4 X c2r, 5 Z forward, 6 Z backward
  static const std::vector<int> kPasses[3] = {{0, 1, 2, 3, 4}, {0, 1, 5}, {6, 3, 4}};
  for (int kind : kPasses[mode]) {
    std::map<int, std::vector<const Req*>> by_len;
    for (const Req& r : rq)
      by_len[kind == 0 || kind == 4 ? r.nx : (kind == 1 || kind == 3 ? r.ny : r.nz)].push_back(&r);
    for (auto& kv : by_len) {
      const int len = kv.first;
      HostStage h;
      std::memset(&h.s.xa, 0, sizeof(h.s.xa));
      std::memset(&h.s.ca, 0, sizeof(h.s.ca));
      h.s.kind = kind;
      const Len1D& L = get_len(len);
      const CtInfo* ct = find_ct(L);
      h.s.ct = ct;
      int tiles_total = 0;
      for (const Req* r : kv.second) {
        const int nxh = r->nx / 2 + 1;
        GrpDesc d;
        std::memset(&d, 0, sizeof(d));
        long long tiles;
        if (kind == 0 || kind == 4) {
          d.in = kind == 0 ? (const void*)fields[r->i] : works[r->i];
          d.out = kind == 0 ? works[r->i] : (void*)fields[r->i];
          d.nrows = (long long)r->ny * r->nz;
          tiles = ((d.nrows + 1) / 2 + ct->BX - 1) / ct->BX;
        } else if (kind == 2 || kind == 5 || kind == 6) {
          d.in = works[r->i];
          d.out = works[r->i];
          d.n_inner = nxh * r->ny;
          d.nrows = 1;
          d.sl = (long long)nxh * r->ny;
          d.nxh = nxh;
          d.ny = r->ny;
          d.scale = 1.0 / ((double)r->nx * r->ny * r->nz);
          tiles = ((long long)d.n_inner + ct->TX - 1) / ct->TX;
        } else {
          d.in = works[r->i];
          d.out = works[r->i];
          d.n_inner = nxh;
          d.nrows = r->nz;
          d.sl = nxh;
          d.so = (long long)nxh * r->ny;
          tiles = ((long long)nxh * r->nz + ct->TX - 1) / ct->TX;
        }
        if (tiles_total + tiles > 0x7fffffffLL) throw Error{"grouped batch too large", -1};
        h.tfirst.push_back(tiles_total);
        for (long long t = 0; t < tiles; t++) h.treq.push_back((int)h.d.size());
        tiles_total += (int)tiles;
        h.d.push_back(d);
      }
      h.s.nwg = (unsigned)tiles_total;
      if (kind == 0 || kind == 4) {
        XArgs& a = h.s.xa;
        const int nxh = len / 2 + 1;
        a.si_row = kind == 0 ? len : nxh;
        a.so_row = kind == 0 ? nxh : len;
        a.nxo = nxh;
        a.dnxo = make_fdiv(nxh);
        a.B = ct->BX;
        a.L = L.d;
        a.tw = L.tw;
        h.s.bytes = ct_lds((size_t)(ct->BX + 1) * len * sizeof(cplx_t), ct->twl_x, *ct, &a.tw_lds, &a.tw_off);
      } else {
        ColArgs& a = h.s.ca;
        a.kind = COL_C2C;
        a.TX = ct->TX;
        a.dTX = make_fdiv(ct->TX);
        a.L = L.d;
        a.tw = L.tw;
        h.s.bytes = ct_lds((size_t)ct->TX * len * sizeof(cplx_t), ct->twl_col, *ct, &a.tw_lds, &a.tw_off);
      }
      if (h.s.bytes > kLdsMax) throw Error{"grouped pass does not fit in LDS", -1};
      hs.push_back(std::move(h));
    }
  }
  // one device allocation: per stage descriptors, tile -> array, first tiles
  size_t total = 0;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  for (auto& h : hs) total += al(h.d.size() * sizeof(GrpDesc)) + al(h.treq.size() * 4) + al(h.tfirst.size() * 4);
  if (total) check_hip(hipMalloc(&B->dmem, total), "hipMalloc(grouped batch)");
  size_t off = 0;
  for (auto& h : hs) {
    char* base = (char*)B->dmem;
    GrpDesc* dd = (GrpDesc*)(base + off);
    off += al(h.d.size() * sizeof(GrpDesc));
    int* tr = (int*)(base + off);
    off += al(h.treq.size() * 4);
    int* tf = (int*)(base + off);
    off += al(h.tfirst.size() * 4);
    check_hip(hipMemcpy(dd, h.d.data(), h.d.size() * sizeof(GrpDesc), hipMemcpyHostToDevice), "hipMemcpy");
    check_hip(hipMemcpy(tr, h.treq.data(), h.treq.size() * 4, hipMemcpyHostToDevice), "hipMemcpy");
    check_hip(hipMemcpy(tf, h.tfirst.data(), h.tfirst.size() * 4, hipMemcpyHostToDevice), "hipMemcpy");
    h.s.ga.d = dd;
    h.s.ga.tile_req = tr;
    h.s.ga.tile_first = tf;
    B->st.push_back(h.s);
  }
  return B;
}

void group_batch_launch(GroupBatch* B, hipStream_t st) {
  for (const GrpStage& s : B->st) {
    if (s.nwg == 0) continue;
    switch (s.kind) {
      case 0: s.ct->xgrp(true, s.xa, s.ga, s.nwg, s.bytes, st); break;
      case 1: s.ct->cgrp(true, s.ca, s.ga, s.nwg, s.bytes, st); break;
      case 2: s.ct->zgrp(s.ca, s.ga, s.nwg, s.bytes, st); break;
      case 3: s.ct->cgrp(false, s.ca, s.ga, s.nwg, s.bytes, st); break;
      case 5: s.ct->cgrp(true, s.ca, s.ga, s.nwg, s.bytes, st); break;
      case 6: s.ct->cgrp(false, s.ca, s.ga, s.nwg, s.bytes, st); break;
      default: s.ct->xgrp(false, s.xa, s.ga, s.nwg, s.bytes, st); break;
    }
    check_hip(hipGetLastError(), "grouped pass launch");
  }
}

int group_batch_launches(const GroupBatch* B) { return (int)B->st.size(); }

void group_batch_free(GroupBatch* B) {
  if (!B) return;
  if (B->dmem) (void)hipFree(B->dmem);
  delete B;
}

// workgroups per plane: at least ~2048 workgroups in all, up to 64 per plane
static int spectrum_chunks(int d2) { return std::max(1, std::min(64, (2048 + d2 - 1) / std::max(d2, 1))); }

size_t spectrum_partial_bytes(int d2, int kmax) {
  return (size_t)d2 * spectrum_chunks(d2) * (kmax + 1) * sizeof(double);
}

void launch_spectrum(const cplx_t* B, int d0, int d1, int d2, int f0, int f1, int f2, int kmax, double* partial,
                     double* E, hipStream_t st) {
  const size_t lds = 4 * (size_t)(kmax + 1) * sizeof(double);
  if (lds > kLdsMax) throw Error{"power spectrum: too many shells for one workgroup's LDS", -1};
  if ((long long)d0 * d1 >= (1LL << 31)) throw Error{"power spectrum: plane too large", -1};
  // k2 in 32 bits: every index, folded or not, is below 2^14
  if (std::max({d0, d1, d2, f0, f1, f2}) >= (1 << 14)) throw Error{"power spectrum: axis too long", -1};
  const int nchunk = spectrum_chunks(d2);
  const long long nwg = (long long)d2 * nchunk;
  if (nwg > 0x7fffffffLL) throw Error{"power spectrum: grid too large", -1};
  if (nwg > 0) {
    k_spectrum<real_t><<<(unsigned)nwg, 256, lds, st>>>(B, d0, d1, f0, f1, f2, kmax, nchunk, partial);
    check_hip(hipGetLastError(), "k_spectrum launch");
  }
  k_spectrum_sum<<<kmax + 1, 256, 0, st>>>(partial, (int)nwg, E);
  check_hip(hipGetLastError(), "k_spectrum_sum launch");
}

void launch_cheby(cplx_t* out, long long ncol, int nzc, long long cs, long long ks, double scale, double L,
                  hipStream_t st) {
  if (ncol <= 0) return;
  // 16 adjacent columns (256 bytes of FP64 complex per row) x 16 row chunks
  constexpr int TX = 16, P = 16;
  const long long nwg = (ncol + TX - 1) / TX;
  if (nwg > 0x7fffffffLL) throw Error{"cheby: grid too large", -1};
  k_cheby<real_t, TX, P><<<(unsigned)nwg, TX * P, 0, st>>>(out, ncol, nzc, cs, ks, scale, L);
  check_hip(hipGetLastError(), "k_cheby launch");
}

void launch_copy16(const void* in, void* out, long long n16, hipStream_t st) {
  if (n16 <= 0) return;
  const long long blocks = (n16 + 255) / 256;
  if (blocks > 0x7fffffffLL) throw Error{"copy too large", -1};
  k_copy16<<<(unsigned)blocks, 256, 0, st>>>((const copy_vec*)in, (copy_vec*)out, n16);
  check_hip(hipGetLastError(), "k_copy16 launch");
}

}  // namespace p3amd

