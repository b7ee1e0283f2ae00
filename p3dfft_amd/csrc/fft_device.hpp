// fft_device.hpp -- device-side building blocks of the gfx950 FFT engine.
//
// Complex numbers are (re, im) pairs of float or double.  A transform of length
// n = r_0 r_1 ... r_{S-1} is run as S Stockham autosort stages over data held
// in LDS; each stage reads R values per butterfly into registers, applies the
// stage twiddles, runs a length-R DFT in registers and writes the results to
// their autosorted position.  Twiddles are FP64-generated tables; the FP32
// engine stores each twiddle as a hi+lo pair of floats and multiplies with
// both (compensated twiddle multiply).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace p3amd {

template <class T>
struct alignas(2 * sizeof(T)) cpx {
  T x, y;
};

template <class T>
__device__ __forceinline__ cpx<T> mk(T a, T b) { cpx<T> r; r.x = a; r.y = b; return r; }
template <class T>
__device__ __forceinline__ cpx<T> add(cpx<T> a, cpx<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <class T>
__device__ __forceinline__ cpx<T> sub(cpx<T> a, cpx<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <class T>
__device__ __forceinline__ cpx<T> scl(cpx<T> a, T s) { return mk<T>(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (backward)
template <bool FWD, class T>
__device__ __forceinline__ cpx<T> rot(cpx<T> a) { return FWD ? mk<T>(a.y, -a.x) : mk<T>(-a.y, a.x); }

// ---------------------------------------------------------------------------
// twiddle storage
//   double : double2 w = exp(-2 pi i e / M)
//   float  : float4 (hi.re, hi.im, lo.re, lo.im) with hi = fl32(w), lo = fl32(w - hi)
// Backward transforms use the conjugate.
// ---------------------------------------------------------------------------
template <class T> struct TwT;
template <> struct TwT<double> { typedef double2 type; };
template <> struct TwT<float> { typedef float4 type; };

template <bool FWD>
__device__ __forceinline__ cpx<double> twmul(cpx<double> a, double2 w) {
  double wy = FWD ? w.y : -w.y;
  return mk<double>(fma(a.x, w.x, -a.y * wy), fma(a.x, wy, a.y * w.x));
}
template <bool FWD>
__device__ __forceinline__ cpx<float> twmul(cpx<float> a, float4 w) {
  float hy = FWD ? w.y : -w.y;
  float ly = FWD ? w.w : -w.w;
  float tr = fmaf(a.x, w.z, -a.y * ly);
  float ti = fmaf(a.x, ly, a.y * w.z);
  return mk<float>(fmaf(a.x, w.x, fmaf(-a.y, hy, tr)), fmaf(a.x, hy, fmaf(a.y, w.x, ti)));
}

// ---------------------------------------------------------------------------
// small DFTs in registers
// ---------------------------------------------------------------------------
template <bool FWD, class T>
__device__ __forceinline__ void dft2(cpx<T>* v) {
  cpx<T> a = v[0], b = v[1];
  v[0] = add(a, b);
  v[1] = sub(a, b);
}

template <bool FWD, class T>
__device__ __forceinline__ void dft4(cpx<T>* v) {
  cpx<T> t0 = add(v[0], v[2]), t1 = sub(v[0], v[2]);
  cpx<T> t2 = add(v[1], v[3]), t3 = rot<FWD>(sub(v[1], v[3]));
  v[0] = add(t0, t2);
  v[2] = sub(t0, t2);
  v[1] = add(t1, t3);
  v[3] = sub(t1, t3);
}

template <bool FWD, class T>
__device__ __forceinline__ void dft8(cpx<T>* v) {
  const T h = (T)0.70710678118654752440;
  cpx<T> e[4] = {v[0], v[2], v[4], v[6]};
  cpx<T> o[4] = {v[1], v[3], v[5], v[7]};
  dft4<FWD>(e);
  dft4<FWD>(o);
  // o[k] *= W8^k  (W8 = exp(-i pi/4) forward, conjugate backward)
  cpx<T> o1, o3;
  if (FWD) {
    o1 = mk<T>((o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h);
    o3 = mk<T>((o[3].y - o[3].x) * h, -(o[3].x + o[3].y) * h);
  } else {
    o1 = mk<T>((o[1].x - o[1].y) * h, (o[1].x + o[1].y) * h);
    o3 = mk<T>(-(o[3].x + o[3].y) * h, (o[3].x - o[3].y) * h);
  }
  cpx<T> o2 = rot<FWD>(o[2]);
  v[0] = add(e[0], o[0]);
  v[4] = sub(e[0], o[0]);
  v[1] = add(e[1], o1);
  v[5] = sub(e[1], o1);
  v[2] = add(e[2], o2);
  v[6] = sub(e[2], o2);
  v[3] = add(e[3], o3);
  v[7] = sub(e[3], o3);
}

// 16 points as 4 x 4: DFT4 over the stride-4 subsequences n1 + 4 n2, the
// twiddles W16^(n1 k1), DFT4 over n1; output in natural order
template <bool FWD, class T>
__device__ __forceinline__ void dft16(cpx<T>* v) {
  const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173;
  const T h = (T)0.70710678118654752440;
  cpx<T> y[4][4];
#pragma unroll
  for (int n1 = 0; n1 < 4; n1++) {
    cpx<T> t[4] = {v[n1], v[n1 + 4], v[n1 + 8], v[n1 + 12]};
    dft4<FWD>(t);
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) y[n1][k1] = t[k1];
  }
  // W16^m (forward exp(-2 pi i m / 16), backward its conjugate), m = n1 k1
  auto w = [&](cpx<T> a, int m) -> cpx<T> {
    T c, sn;
    switch (m) {
      case 1: c = c1; sn = s1; break;
      case 2: c = h; sn = h; break;
      case 3: c = s1; sn = c1; break;
      case 4: c = 0; sn = 1; break;
      case 6: c = -h; sn = h; break;
      default: c = -c1; sn = -s1; break;  // m = 9
    }
    const T sy = FWD ? -sn : sn;
    return mk<T>(a.x * c - a.y * sy, a.x * sy + a.y * c);
  };
  y[1][1] = w(y[1][1], 1);
  y[1][2] = w(y[1][2], 2);
  y[1][3] = w(y[1][3], 3);
  y[2][1] = w(y[2][1], 2);
  y[2][2] = rot<FWD>(y[2][2]);
  y[2][3] = w(y[2][3], 6);
  y[3][1] = w(y[3][1], 3);
  y[3][2] = w(y[3][2], 6);
  y[3][3] = w(y[3][3], 9);
#pragma unroll
  for (int k1 = 0; k1 < 4; k1++) {
    cpx<T> t[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
    dft4<FWD>(t);
#pragma unroll
    for (int k2 = 0; k2 < 4; k2++) v[k1 + 4 * k2] = t[k2];
  }
}

template <bool FWD, class T>
__device__ __forceinline__ void dft3(cpx<T>* v) {
  const T c3 = (T)0.86602540378443864676;
  cpx<T> s = add(v[1], v[2]);
  cpx<T> d = scl(rot<FWD>(sub(v[1], v[2])), c3);
  cpx<T> m = sub(v[0], scl(s, (T)0.5));
  v[0] = add(v[0], s);
  v[1] = add(m, d);
  v[2] = sub(m, d);
}

template <bool FWD, class T>
__device__ __forceinline__ void dft5(cpx<T>* v) {
  const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
  const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
  cpx<T> s14 = add(v[1], v[4]), d14 = sub(v[1], v[4]);
  cpx<T> s23 = add(v[2], v[3]), d23 = sub(v[2], v[3]);
  cpx<T> a0 = v[0];
  cpx<T> m1 = mk<T>(a0.x + c1 * s14.x + c2 * s23.x, a0.y + c1 * s14.y + c2 * s23.y);
  cpx<T> m2 = mk<T>(a0.x + c2 * s14.x + c1 * s23.x, a0.y + c2 * s14.y + c1 * s23.y);
  cpx<T> n1 = rot<FWD>(mk<T>(s1 * d14.x + s2 * d23.x, s1 * d14.y + s2 * d23.y));
  cpx<T> n2 = rot<FWD>(mk<T>(s2 * d14.x - s1 * d23.x, s2 * d14.y - s1 * d23.y));
  v[0] = add(a0, add(s14, s23));
  v[1] = add(m1, n1);
  v[4] = sub(m1, n1);
  v[2] = add(m2, n2);
  v[3] = sub(m2, n2);
}

// odd-prime DFTs by the pairwise (a_p +- a_{P-p}) form; tables hold
// cos(2 pi j/P), sin(2 pi j/P) for j = 1..(P-1)/2.
template <int P> struct PrimeTab;
template <> struct PrimeTab<7> {
  static constexpr double c[4] = {1.0, 0.6234898018587336, -0.22252093395631434, -0.900968867902419};
  static constexpr double s[4] = {0.0, 0.7818314824680298, 0.9749279121818236, 0.43388373911755823};
};
template <> struct PrimeTab<11> {
  static constexpr double c[6] = {1.0, 0.8412535328311812, 0.41541501300188644, -0.142314838273285,
                                  -0.654860733945285, -0.9594929736144974};
  static constexpr double s[6] = {0.0, 0.5406408174555976, 0.9096319953545183, 0.9898214418809328,
                                  0.7557495743542583, 0.28173255684142967};
};
template <> struct PrimeTab<13> {
  static constexpr double c[7] = {1.0, 0.8854560256532099, 0.5680647467311559, 0.120536680255323,
                                  -0.35460488704253545, -0.7485107481711012, -0.970941817426052};
  static constexpr double s[7] = {0.0, 0.4647231720437685, 0.8229838658936564, 0.992708874098054,
                                  0.9350162426854148, 0.6631226582407952, 0.23931566428755768};
};

template <int P, bool FWD, class T>
__device__ __forceinline__ void dftp(cpx<T>* v) {
  constexpr int H = (P - 1) / 2;
  cpx<T> sp[H + 1], dp[H + 1];
#pragma unroll
  for (int p = 1; p <= H; p++) {
    sp[p] = add(v[p], v[P - p]);
    dp[p] = sub(v[p], v[P - p]);
  }
  cpx<T> a0 = v[0];
  cpx<T> y0 = a0;
#pragma unroll
  for (int p = 1; p <= H; p++) y0 = add(y0, sp[p]);
#pragma unroll
  for (int q = 1; q <= H; q++) {
    cpx<T> m = a0, n = mk<T>(0, 0);
#pragma unroll
    for (int p = 1; p <= H; p++) {
      const int j = (p * q) % P;
      const T c = (T)PrimeTab<P>::c[j <= H ? j : P - j];
      const T s = j <= H ? (T)PrimeTab<P>::s[j] : -(T)PrimeTab<P>::s[P - j];
      m = mk<T>(m.x + c * sp[p].x, m.y + c * sp[p].y);
      n = mk<T>(n.x + s * dp[p].x, n.y + s * dp[p].y);
    }
    n = rot<FWD>(n);
    v[q] = add(m, n);
    v[P - q] = sub(m, n);
  }
  v[0] = y0;
}

template <int R, bool FWD, class T>
__device__ __forceinline__ void dft(cpx<T>* v) {
  if constexpr (R == 2) dft2<FWD>(v);
  else if constexpr (R == 3) dft3<FWD>(v);
  else if constexpr (R == 4) dft4<FWD>(v);
  else if constexpr (R == 5) dft5<FWD>(v);
  else if constexpr (R == 8) dft8<FWD>(v);
  else if constexpr (R == 16) dft16<FWD>(v);
  else dftp<R, FWD>(v);
}

// ---------------------------------------------------------------------------
// exact small-integer division: q = (x * m) >> 32 with m = ceil(2^32 / d),
// valid for x, d < 2^16 (all in-workgroup indices are below 2^16).
// ---------------------------------------------------------------------------
struct FDiv {
  uint64_t m;
  int d;
};
__device__ __forceinline__ int fdiv(int x, const FDiv& f) {
  return (int)(((uint64_t)(uint32_t)x * f.m) >> 32);
}

// ---------------------------------------------------------------------------
// length descriptor (host builds it; passed by value in kernel arguments)
// ---------------------------------------------------------------------------
constexpr int kMaxStages = 24;
struct LenDesc {
  int n;
  int nst;                 // number of stages
  int radix[kMaxStages];   // radix of each stage
  int ns[kMaxStages];      // product of the radices of earlier stages
  int twoff[kMaxStages];   // offset of each stage's twiddles in the table
  FDiv dns[kMaxStages];    // divide by ns
  FDiv dnr[kMaxStages];    // divide by n / radix
  FDiv dn;                 // divide by n
  int fulloff;             // offset of the full table W_n^e, e = 0..n-1
  int twcount;             // total entries (stage tables + full table)
};

}  // namespace p3amd
