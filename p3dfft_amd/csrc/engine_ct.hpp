// engine_ct.hpp -- compile-time specialised Stockham engine.
//
// For the transform lengths the workloads use, the radix schedule, the tile
// width and the thread count are template parameters, so every index
// computation of a stage folds to shifts and constant multiplies and the
// per-thread register arrays have fixed size.  The LDS tile holds B sequences
// with element e of sequence b at lds[e*BP + b]; BP = B+1 pads each row so that
// both the sequence-fastest accesses of the stages and the element-fastest
// accesses of row loads avoid bank conflicts.  Stage twiddles use the same
// table layout as the run-time engine (stage 0 has none; stage s holds
// Ns*(R-1) entries W_{Ns R}^{k r}, k < Ns, r = 1..R-1).
#pragma once
#include <type_traits>
#include "fft_device.hpp"

namespace p3amd {

template <int... Rs>
struct RList {};

template <int A, int... Rs>
struct RProd {
  static constexpr int value = A * RProd<Rs...>::value;
};
template <int A>
struct RProd<A> {
  static constexpr int value = A;
};

// One stage: radix R, Ns = product of the earlier radices, twiddles at Off.
template <class T, int R, bool FWD, int N, int B, int BP, int NT, int Ns, int Off>
__device__ __forceinline__ void stage_ct(cpx<T>* lds, const typename TwT<T>::type* tw) {
  constexpr int nR = N / R;
  constexpr int total = B * nR;
  constexpr int IT = (total + NT - 1) / NT;
  cpx<T> v[IT][R];
#pragma unroll
  for (int m = 0; m < IT; m++) {
    const int it = threadIdx.x + m * NT;
    if ((total % NT) == 0 || it < total) {
      const int j = it / B, b = it % B;
#pragma unroll
      for (int r = 0; r < R; r++) v[m][r] = lds[(j + r * nR) * BP + b];
      if constexpr (Ns > 1) {
        const int k = j % Ns;
        const typename TwT<T>::type* w = tw + Off + k * (R - 1);
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
    if ((total % NT) == 0 || it < total) {
      const int j = it / B, b = it % B;
      const int k = j % Ns;
      const int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) lds[(base + r * Ns) * BP + b] = v[m][r];
    }
  }
  __syncthreads();
}

template <class T, bool FWD, int N, int B, int BP, int NT, int Ns, int Off, int R, int... Rest>
__device__ __forceinline__ void stages_ct(cpx<T>* lds, const typename TwT<T>::type* tw) {
  stage_ct<T, R, FWD, N, B, BP, NT, Ns, Off>(lds, tw);
  if constexpr (sizeof...(Rest) > 0)
    stages_ct<T, FWD, N, B, BP, NT, Ns * R, (Ns > 1 ? Off + Ns * (R - 1) : Off), Rest...>(lds, tw);
}

template <class T, bool FWD, int B, int BP, int NT, int... Rs>
__device__ __forceinline__ void run_fft_ct(cpx<T>* lds, const typename TwT<T>::type* tw, RList<Rs...>) {
  constexpr int N = RProd<Rs...>::value;
  stages_ct<T, FWD, N, B, BP, NT, 1, 0, Rs...>(lds, tw);
}

// ---------------------------------------------------------------------------
// Column transforms with the first stage fed straight from global memory and
// the last stage stored straight to global memory (one LDS round trip fewer
// at each end; every thread keeps IT*R independent loads in flight).
// Addresses are a wave-uniform 64-bit tile base plus 32-bit per-lane byte
// offsets (Off = unsigned), so an element access costs one 32-bit add; views
// whose offsets do not fit use 64-bit offsets (Off = unsigned long long).
// With PR the rows are
// remapped on load (pruning zero-fill) and on store (pruning selection);
// without it both maps are the identity.
// ---------------------------------------------------------------------------
struct ColIO {
  const char* in;             // input tile base
  char* out;                  // output tile base
  unsigned long long lane_in, lane_out;  // this lane's column, bytes from the tile bases
  unsigned long long sl_in, sl_out;      // row strides, bytes
  int nb;                     // valid columns of this tile
  int ld_lo, ld_hi, ld_shift; // PR: LDS row e <- input row e (e<ld_lo), e-ld_shift (e>=ld_hi), else 0
  int nst, st_lo, st_shift;   // PR: output row e <- FFT row e (e<st_lo) or e+st_shift
  int mirror;                 // PR, > 0: DCT-I even extension, row e <- input e (e<mirror) or 2(mirror-1)-e
  // table-addressed rows (Off = TabRows): row e of this lane's column at
  // element ltab[e].x + lane_o * ltab[e].y + lane_i of `in` (zero when
  // ltab[e].x < 0), FFT row p stored at stab[p] likewise (skipped when < 0)
  const int2* ltab;
  const int2* stab;
  unsigned lane_o, lane_i;    // this lane's outer and inner column index
};

// offset type selecting the table-addressed rows
struct TabRows {};

template <int R, int N, int B, int NT>
struct Stage0 {
  static constexpr int nR = N / R;
  static constexpr int total = B * nR;
  static constexpr int IT = (total + NT - 1) / NT;
};

template <class T, class Off>
__device__ __forceinline__ cpx<T> ld_at(const char* base, Off off) {
  return *(const cpx<T>*)(base + off);
}

template <class T, class Off>
__device__ __forceinline__ void st_at(char* base, Off off, cpx<T> v) {
  *(cpx<T>*)(base + off) = v;
}

// the first stage's values of one tile, held in registers between the loads
// and the butterflies
template <class T, int R, int N, int B, int NT>
struct Stage0Regs {
  cpx<T> v[Stage0<R, N, B, NT>::IT][R];
};

// first stage, loads: the tile's values of every butterfly of this thread
template <class T, int R, int N, int B, int NT, bool PR, class Off>
__device__ __forceinline__ void stage0_load(Stage0Regs<T, R, N, B, NT>& rg, const ColIO& io) {
  static_assert(NT % B == 0, "one column per thread");
  constexpr bool TB = std::is_same<Off, TabRows>::value;
  typedef Stage0<R, N, B, NT> S;
  const bool valid = (int)(threadIdx.x % B) < io.nb;
#pragma unroll
  for (int m = 0; m < S::IT; m++) {
    const int it = threadIdx.x + m * NT;
    if ((S::total % NT) == 0 || it < S::total) {
      const int j = it / B;
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int e = j + r * S::nR;
        // items outside the tile load its first element (always valid) and
        // are zeroed, so no load sits under a branch
        if constexpr (TB) {
          const int2 t = io.ltab[e];
          const bool ok = valid && t.x >= 0;
          const unsigned el = ok ? (unsigned)t.x + io.lane_o * (unsigned)t.y + io.lane_i : 0u;
          const cpx<T> u = ((const cpx<T>*)io.in)[el];
          rg.v[m][r] = ok ? u : mk<T>(0, 0);
        } else if constexpr (PR) {
          const int src = io.mirror ? (e < io.mirror ? e : 2 * (io.mirror - 1) - e)
                                    : (e < io.ld_lo ? e : (e >= io.ld_hi ? e - io.ld_shift : -1));
          const bool ok = valid && src >= 0;
          const cpx<T> u = ld_at<T, Off>(io.in, ok ? (Off)io.lane_in + (Off)src * (Off)io.sl_in : (Off)0);
          rg.v[m][r] = ok ? u : mk<T>(0, 0);
        } else {
          const cpx<T> u = ld_at<T, Off>(io.in, valid ? (Off)io.lane_in + (Off)e * (Off)io.sl_in : (Off)0);
          rg.v[m][r] = valid ? u : mk<T>(0, 0);
        }
      }
    }
  }
}

// first stage, butterflies (no twiddles) and results into the LDS tile
template <class T, int R, bool FWD, int N, int B, int NT>
__device__ __forceinline__ void stage0_put(cpx<T>* lds, Stage0Regs<T, R, N, B, NT>& rg) {
  typedef Stage0<R, N, B, NT> S;
#pragma unroll
  for (int m = 0; m < S::IT; m++) {
    const int it = threadIdx.x + m * NT;
    if ((S::total % NT) == 0 || it < S::total) {
      const int j = it / B, b = it % B;
      dft<R, FWD>(rg.v[m]);
#pragma unroll
      for (int r = 0; r < R; r++) lds[(j * R + r) * B + b] = rg.v[m][r];
    }
  }
  __syncthreads();
}

// first stage: loads of one tile into registers, butterflies (no twiddles),
// results into the LDS tile (the loads as stage0_load, written out here so
// that every load of the tile is issued back to back: composed of the two
// steps, the compiler interleaved them with address arithmetic and waited
// before the last one)
template <class T, int R, bool FWD, int N, int B, int NT, bool PR, class Off>
__device__ __forceinline__ void stage0_global(cpx<T>* lds, const ColIO& io) {
  static_assert(NT % B == 0, "one column per thread");
  constexpr bool TB = std::is_same<Off, TabRows>::value;
  typedef Stage0<R, N, B, NT> S;
  const bool valid = (int)(threadIdx.x % B) < io.nb;
  cpx<T> v[S::IT][R];
#pragma unroll
  for (int m = 0; m < S::IT; m++) {
    const int it = threadIdx.x + m * NT;
    if ((S::total % NT) == 0 || it < S::total) {
      const int j = it / B;
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int e = j + r * S::nR;
        // items outside the tile load its first element (always valid) and
        // are zeroed, so no load sits under a branch
        if constexpr (TB) {
          const int2 t = io.ltab[e];
          const bool ok = valid && t.x >= 0;
          const unsigned el = ok ? (unsigned)t.x + io.lane_o * (unsigned)t.y + io.lane_i : 0u;
          const cpx<T> u = ((const cpx<T>*)io.in)[el];
          v[m][r] = ok ? u : mk<T>(0, 0);
        } else if constexpr (PR) {
          const int src = io.mirror ? (e < io.mirror ? e : 2 * (io.mirror - 1) - e)
                                    : (e < io.ld_lo ? e : (e >= io.ld_hi ? e - io.ld_shift : -1));
          const bool ok = valid && src >= 0;
          const cpx<T> u = ld_at<T, Off>(io.in, ok ? (Off)io.lane_in + (Off)src * (Off)io.sl_in : (Off)0);
          v[m][r] = ok ? u : mk<T>(0, 0);
        } else {
          const cpx<T> u = ld_at<T, Off>(io.in, valid ? (Off)io.lane_in + (Off)e * (Off)io.sl_in : (Off)0);
          v[m][r] = valid ? u : mk<T>(0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int m = 0; m < S::IT; m++) {
    const int it = threadIdx.x + m * NT;
    if ((S::total % NT) == 0 || it < S::total) {
      const int j = it / B, b = it % B;
      dft<R, FWD>(v[m]);
#pragma unroll
      for (int r = 0; r < R; r++) lds[(j * R + r) * B + b] = v[m][r];
    }
  }
  __syncthreads();
}

// last stage: Ns * R == N, so every butterfly index j < Ns and k == j
template <class T, int R, bool FWD, int N, int B, int NT, int Ns, int TwOff, bool PR, class Off>
__device__ __forceinline__ void stageL_global(const cpx<T>* lds, const typename TwT<T>::type* tw,
                                              const ColIO& io) {
  static_assert(Ns * R == N, "last stage");
  static_assert(NT % B == 0, "one column per thread");
  constexpr bool TB = std::is_same<Off, TabRows>::value;
  constexpr int total = B * Ns;
  constexpr int IT = (total + NT - 1) / NT;
  const bool valid = (int)(threadIdx.x % B) < io.nb;
#pragma unroll
  for (int m = 0; m < IT; m++) {
    const int it = threadIdx.x + m * NT;
    if ((total % NT) == 0 || it < total) {
      const int j = it / B, b = it % B;
      cpx<T> v[R];
#pragma unroll
      for (int r = 0; r < R; r++) v[r] = lds[(j + r * Ns) * B + b];
      const typename TwT<T>::type* w = tw + TwOff + j * (R - 1);
#pragma unroll
      for (int r = 1; r < R; r++) v[r] = twmul<FWD>(v[r], w[r - 1]);
      dft<R, FWD>(v);
      // table-addressed rows: every entry and address first, so the table
      // reads issue together instead of one read and wait per store
      unsigned el[TB ? R : 1];
      bool st[TB ? R : 1];
      if constexpr (TB) {
#pragma unroll
        for (int r = 0; r < R; r++) {
          const int2 t = io.stab[j + r * Ns];
          el[r] = (unsigned)t.x + io.lane_o * (unsigned)t.y + io.lane_i;
          st[r] = t.x >= 0;
        }
      }
      if (valid) {
#pragma unroll
        for (int r = 0; r < R; r++) {
          const int p = j + r * Ns;
          if constexpr (TB) {
            if (st[r]) ((cpx<T>*)io.out)[el[r]] = v[r];
          } else if constexpr (PR) {
            const int e = p < io.st_lo ? p : p - io.st_shift;
            if (p < io.st_lo || (p >= io.st_lo + io.st_shift && e < io.nst))
              st_at<T, Off>(io.out, (Off)io.lane_out + (Off)e * (Off)io.sl_out, v[r]);
          } else {
            st_at<T, Off>(io.out, (Off)io.lane_out + (Off)p * (Off)io.sl_out, v[r]);
          }
        }
      }
    }
  }
}

template <class T, bool FWD, int N, int B, int NT, int Ns, int TwOff, bool PR, class Off, int R, int... Rest>
__device__ __forceinline__ void mid_and_last(cpx<T>* lds, const typename TwT<T>::type* tw, const ColIO& io) {
  if constexpr (sizeof...(Rest) == 0) {
    stageL_global<T, R, FWD, N, B, NT, Ns, TwOff, PR, Off>(lds, tw, io);
  } else {
    stage_ct<T, R, FWD, N, B, B, NT, Ns, TwOff>(lds, tw);
    mid_and_last<T, FWD, N, B, NT, Ns * R, (Ns > 1 ? TwOff + Ns * (R - 1) : TwOff), PR, Off, Rest...>(lds, tw, io);
  }
}

template <class T, bool FWD, int B, int NT, bool PR, class Off, int R0, int... Rs>
__device__ __forceinline__ void run_fft_io(cpx<T>* lds, const typename TwT<T>::type* tw, const ColIO& io,
                                           RList<R0, Rs...>) {
  static_assert(sizeof...(Rs) >= 1, "at least two stages");
  constexpr int N = R0 * RProd<Rs...>::value;
  stage0_global<T, R0, FWD, N, B, NT, PR, Off>(lds, io);
  mid_and_last<T, FWD, N, B, NT, R0, 0, PR, Off, Rs...>(lds, tw, io);
}

// run_fft_io over the tiles t0, t0 + dt, ... below t1 of one workgroup: the
// next tile's first-stage loads are issued before the current tile's later
// stages and its stores, so they are in flight while those run.  make_io(t)
// gives tile t's addressing.
template <class T, bool FWD, int B, int NT, bool PR, class Off, class MakeIO, int R0, int... Rs>
__device__ __forceinline__ void run_fft_io_pipe(cpx<T>* lds, const typename TwT<T>::type* tw, int t0, int t1,
                                                int dt, MakeIO make_io, RList<R0, Rs...>) {
  static_assert(sizeof...(Rs) >= 1, "at least two stages");
  constexpr int N = R0 * RProd<Rs...>::value;
  if (t0 >= t1) return;
  Stage0Regs<T, R0, N, B, NT> rg;
  ColIO io = make_io(t0);
  stage0_load<T, R0, N, B, NT, PR, Off>(rg, io);
  for (int t = t0; t < t1; t += dt) {
    stage0_put<T, R0, FWD, N, B, NT>(lds, rg);
    const ColIO cur = io;
    if (t + dt < t1) {
      io = make_io(t + dt);
      stage0_load<T, R0, N, B, NT, PR, Off>(rg, io);
    }
    mid_and_last<T, FWD, N, B, NT, R0, 0, PR, Off, Rs...>(lds, tw, cur);
    // the last stage's LDS reads before the next tile's first-stage writes
    __syncthreads();
  }
}

// Forward FFT of a tile whose first stage is fed from global memory (with
// the PR row maps, e.g. a DCT-I mirror) and whose result stays in the LDS
// tile (natural order, element e of column b at lds[e*B + b]).
template <class T, int B, int NT, class Off, int R0, int... Rs>
__device__ __forceinline__ void run_fft_in_io(cpx<T>* lds, const typename TwT<T>::type* tw, const ColIO& io,
                                              RList<R0, Rs...>) {
  static_assert(sizeof...(Rs) >= 1, "at least two stages");
  constexpr int N = R0 * RProd<Rs...>::value;
  stage0_global<T, R0, true, N, B, NT, true, Off>(lds, io);
  stages_ct<T, true, N, B, B, NT, R0, 0, Rs...>(lds, tw);
}

// Table-addressed rows (Off = TabRows) with the row tables staged in LDS at
// `tl` (n int2): the load table before the first stage, the store table
// after it, so every table read is an LDS read.
template <class T, bool FWD, int B, int NT, int R0, int... Rs>
__device__ __forceinline__ void run_fft_io_tab(cpx<T>* lds, int2* tl, const typename TwT<T>::type* tw, ColIO io,
                                               const int2* ltab, const int2* stab, RList<R0, Rs...>) {
  static_assert(sizeof...(Rs) >= 1, "at least two stages");
  constexpr int N = R0 * RProd<Rs...>::value;
  for (int i = threadIdx.x; i < N; i += NT) tl[i] = ltab[i];
  __syncthreads();
  io.ltab = tl;
  stage0_global<T, R0, FWD, N, B, NT, false, TabRows>(lds, io);
  for (int i = threadIdx.x; i < N; i += NT) tl[i] = stab[i];
  __syncthreads();
  io.stab = tl;
  mid_and_last<T, FWD, N, B, NT, R0, 0, false, TabRows, Rs...>(lds, tw, io);
}

// Spectral solve along the column: forward FFT, per-mode multiply by
// -scale/|k|^2 (k=0 -> 0), inverse FFT, all in one LDS tile.  Column i0 + b
// of the tile is kx + nxh * ky_index; the row is the kz index.
struct SolveCtx {
  int nxh, ny, nz;
  double scale;
};

template <class T, int B, int NT, int R0, int... Rs>
__device__ __forceinline__ void run_fft_solve(cpx<T>* lds, const typename TwT<T>::type* tw, const ColIO& io,
                                              long long i0, const SolveCtx& sc, RList<R0, Rs...>) {
  static_assert(sizeof...(Rs) >= 1, "at least two stages");
  constexpr int N = R0 * RProd<Rs...>::value;
  stage0_global<T, R0, true, N, B, NT, false, unsigned>(lds, io);
  stages_ct<T, true, N, B, B, NT, R0, 0, Rs...>(lds, tw);
  for (int q = threadIdx.x; q < B * N; q += NT) {
    const int e = q / B, b = q % B;
    const long long i = i0 + b;
    const int kx = (int)(i % sc.nxh);
    const int y = (int)(i / sc.nxh);
    const double ky = y <= sc.ny / 2 ? y : y - sc.ny;
    const double kz = e <= sc.nz / 2 ? e : e - sc.nz;
    const double k2 = (double)kx * kx + ky * ky + kz * kz;
    const double f = k2 > 0.0 ? -sc.scale / k2 : 0.0;
    const cpx<T> v = lds[q];
    lds[q] = mk<T>((T)(v.x * f), (T)(v.y * f));
  }
  __syncthreads();
  mid_and_last<T, false, N, B, NT, 1, 0, false, unsigned, R0, Rs...>(lds, tw, io);
}

// Twiddle table entries used by a radix list (all stages after the first).
template <int... Rs>
struct TwCount;
template <int R0, int... Rs>
struct TwCount<R0, Rs...> {
  template <int Ns, int R, int... Rest>
  static constexpr int tail() {
    if constexpr (sizeof...(Rest) == 0) return Ns * (R - 1);
    else return Ns * (R - 1) + tail<Ns * R, Rest...>();
  }
  static constexpr int value() {
    if constexpr (sizeof...(Rs) == 0) return 0;
    else return tail<R0, Rs...>();
  }
};

}  // namespace p3amd

namespace p3amd {
template <class RL>
struct RL_N;
template <int... Rs>
struct RL_N<RList<Rs...>> {
  static constexpr int value = RProd<Rs...>::value;
};
}  // namespace p3amd
