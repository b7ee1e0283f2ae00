// engine_long.hip -- transforms longer than one workgroup's LDS tile holds:
// lengths above max_len(), and DCT-I / DST-I extensions above it.
//
// The reference plans any length through FFTW (build/init_plan.F90:100-128,
// 338-471; build/fft_spec.F90).  Here a length m above max_len() runs in
// several passes over a scratch buffer that holds a chunk of the view's
// columns row-major, column fastest (element (row e, column c) at e*ncc + c),
// so every pass is a coalesced column pass of the one-tile kernels:
//  - four-step, m = N1*N2 with N1, N2 <= max_len(): the N2-point FFTs of the
//    N1 subsequences x[n1 + N1*n2] (in place), the twiddle W_m^(n1*k2)
//    (evaluated in FP64), then the N1-point FFTs written in transposed order,
//    X[k2 + N2*k1];
//  - Bluestein's chirp-z form where m has no such split (or only one with a
//    large prime factor): X_k = c_k sum_n (x_n c_n) conj(c_(k-n)),
//    c_j = exp(-+ i pi j^2 / m), as a circular convolution of length M, the
//    smallest 2^a 3^b 5^c >= 2m-1, through two four-step FFTs of length M and
//    the filter's transform, computed once per length.
// Gather and scatter kernels apply the view's strides, its load/store row
// maps (pruning), the DCT-I / DST-I extensions, and on the X axis the
// two-for-one packing of two real rows into one complex row.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "engine.hpp"

namespace p3amd {

namespace {

template <class T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// exp(i pi num / den) in FP64 (num reduced exactly by the caller)
__device__ __forceinline__ cpx<double> expipi(long long num, long long den) {
  double s, c;
  sincospi((double)num / (double)den, &s, &c);
  return mk<double>(c, s);
}

template <class T>
__device__ __forceinline__ cpx<T> to_t(cpx<double> v) { return mk<T>((T)v.x, (T)v.y); }

// Bluestein chirp c_j = exp(sgn i pi j^2 / m), j^2 reduced mod 2m in integers
__device__ __forceinline__ cpx<double> chirp(long long j, int m, int sgn) {
  const long long r = (j * j) % (2LL * m);
  return expipi(sgn < 0 ? -r : r, m);
}

struct LgCol {
  const void* in;
  void* out;
  long long si, sl, so;  // strides of the view read (gather) or written (scatter)
  int n_inner;
  int kind;
  int nld, ld_lo, ld_hi, ld_shift;
  int nst, st_lo, st_shift;
  long long c0;  // first column of the chunk (flattened (outer, inner) index)
  long long ncc;  // columns in the chunk
  int m;          // FFT length (the extension's for DCT-I / DST-I)
};

// scratch row e of chunk column c <- the view's row e (load map, extension)
// I: the index type of the element and column counters and their divisions
// (unsigned where every index of the launch is below 2^31, else 64-bit)
template <class T, class I>
__global__ __launch_bounds__(256) void k_lg_gather(LgCol a, cpx<T>* A) {
  const cpx<T>* in = (const cpx<T>*)a.in;
  const I total = (I)((long long)a.m * a.ncc), ncc = (I)a.ncc, nin = (I)a.n_inner, c0 = (I)a.c0;
  for (I q = (I)blockIdx.x * 256 + threadIdx.x; q < total; q += (I)gridDim.x * 256) {
    const I e0 = q / ncc, c = q - e0 * ncc;
    const I col = c0 + c;
    const I o = col / nin, i = col - o * nin;
    const long long e = (long long)e0;
    const cpx<T>* base = in + (long long)o * a.so + (long long)i * a.si;
    auto load = [&](long long r) -> cpx<T> {
      if (r < 0 || r >= a.nld) return mk<T>(0, 0);
      const long long src = r < a.ld_lo ? r : (r >= a.ld_hi ? r - a.ld_shift : -1);
      return src >= 0 ? base[src * a.sl] : mk<T>(0, 0);
    };
    cpx<T> v;
    if (a.kind == COL_DCT1) {
      // even extension: row e <- row m-e
      v = e < a.nld ? load(e) : load(a.m - e);
    } else if (a.kind == COL_DST1) {
      // odd extension: rows 0 and nld+1 zero, row e <- input e-1, then -row m-e
      if (e == 0 || e == a.nld + 1) v = mk<T>(0, 0);
      else if (e <= a.nld) v = load(e - 1);
      else { const cpx<T> u = load(a.m - e - 1); v = mk<T>(-u.x, -u.y); }
    } else {
      v = load(e);
    }
    A[q] = v;
  }
}

// the view's output row e <- scratch row p(e) (store map; DST-I output is
// -i times row p+1 of the odd extension's transform)
template <class T, class I>
__global__ __launch_bounds__(256) void k_lg_scatter(LgCol a, const cpx<T>* R) {
  cpx<T>* out = (cpx<T>*)a.out;
  const I total = (I)((long long)a.nst * a.ncc), ncc = (I)a.ncc, nin = (I)a.n_inner, c0 = (I)a.c0;
  for (I q = (I)blockIdx.x * 256 + threadIdx.x; q < total; q += (I)gridDim.x * 256) {
    const I e0 = q / ncc, c = q - e0 * ncc;
    const I col = c0 + c;
    const I o = col / nin, i = col - o * nin;
    const long long e = (long long)e0;
    const long long p = e < a.st_lo ? e : e + a.st_shift;
    cpx<T> v;
    if (a.kind == COL_DST1) { const cpx<T> u = R[(p + 1) * a.ncc + c]; v = mk<T>(-u.y, u.x); }
    else v = R[p * a.ncc + c];
    out[(long long)o * a.so + (long long)i * a.si + e * a.sl] = v;
  }
}

// four-step twiddle: row n1 + N1*k2 of every column times W_m^(sgn n1 k2)
template <class T, class I>
__global__ __launch_bounds__(256) void k_lg_twiddle(cpx<T>* A, long long ncc, int N1, int m, int sgn) {
  const I total = (I)((long long)m * ncc), nc = (I)ncc;
  for (I q = (I)blockIdx.x * 256 + threadIdx.x; q < total; q += (I)gridDim.x * 256) {
    const unsigned row = (unsigned)(q / nc);  // < m
    const unsigned k2 = row / (unsigned)N1, n1 = row - k2 * (unsigned)N1;
    const long long r = (long long)n1 * k2;  // < m
    if (r == 0) continue;
    const cpx<double> w = expipi(sgn < 0 ? -2 * r : 2 * r, m);
    A[q] = cmul(A[q], to_t<T>(w));
  }
}

// Bluestein: P[j] = x_j c_j (j < m), 0 up to M
template <class T, class I>
__global__ __launch_bounds__(256) void k_lg_blu_pre(const cpx<T>* A, cpx<T>* P, long long ncc, int m, int M, int sgn) {
  const I total = (I)((long long)M * ncc), nc = (I)ncc;
  for (I q = (I)blockIdx.x * 256 + threadIdx.x; q < total; q += (I)gridDim.x * 256) {
    const long long j = (long long)(q / nc);
    P[q] = j < m ? cmul(A[q], to_t<T>(chirp(j, m, sgn))) : mk<T>(0, 0);
  }
}

// Bluestein: Q[j] *= H[j] (forward) or conj(H[(M-j) mod M]) (backward)
template <class T, class I>
__global__ __launch_bounds__(256) void k_lg_blu_mul(cpx<T>* Q, const cpx<T>* H, long long ncc, int M, int flip) {
  const I total = (I)((long long)M * ncc), nc = (I)ncc;
  for (I q = (I)blockIdx.x * 256 + threadIdx.x; q < total; q += (I)gridDim.x * 256) {
    const long long j = (long long)(q / nc);
    cpx<T> h;
    if (flip) { const cpx<T> g = H[j == 0 ? 0 : M - j]; h = mk<T>(g.x, -g.y); }
    else h = H[j];
    Q[q] = cmul(Q[q], h);
  }
}

// Bluestein: X_k = c_k Y_k / M (k < m)
template <class T, class I>
__global__ __launch_bounds__(256) void k_lg_blu_post(const cpx<T>* P, cpx<T>* A, long long ncc, int m, int M, int sgn) {
  const I total = (I)((long long)m * ncc), nc = (I)ncc;
  const double inv = 1.0 / (double)M;
  for (I q = (I)blockIdx.x * 256 + threadIdx.x; q < total; q += (I)gridDim.x * 256) {
    const long long k = (long long)(q / nc);
    const cpx<double> c = chirp(k, m, sgn);
    A[q] = cmul(P[q], to_t<T>(mk<double>(c.x * inv, c.y * inv)));
  }
}

// the Bluestein filter of length M: h_j = conj(c_j) of the forward chirp at
// j and M-j (0 < j < m), zero elsewhere
template <class T>
__global__ __launch_bounds__(256) void k_lg_blu_filter(cpx<T>* h, int m, int M) {
  for (long long j = blockIdx.x * 256LL + threadIdx.x; j < M; j += (long long)gridDim.x * 256) {
    cpx<T> v = mk<T>(0, 0);
    const long long jj = j < m ? j : (j > M - m ? M - j : -1);
    if (jj >= 0) v = to_t<T>(chirp(jj, m, +1));
    h[j] = v;
  }
}

// X axis, forward: scratch row e of pair p = (row 2p, row 2p+1) element e
template <class T, class I>
__global__ __launch_bounds__(256) void k_lg_x_gather(const T* in, cpx<T>* A, long long nrows, long long si_row,
                                                     int nx, long long p0, long long npc) {
  const I total = (I)((long long)nx * npc), np = (I)npc;
  for (I q = (I)blockIdx.x * 256 + threadIdx.x; q < total; q += (I)gridDim.x * 256) {
    const long long e = (long long)(q / np), p = (long long)(q - (q / np) * np);
    const long long r0 = 2 * (p0 + p);
    const T va = r0 < nrows ? in[r0 * si_row + e] : (T)0;
    const T vb = r0 + 1 < nrows ? in[(r0 + 1) * si_row + e] : (T)0;
    A[q] = mk<T>(va, vb);
  }
}

// X axis, forward: the two half spectra of each pair, modes k < nxo
template <class T, class I>
__global__ __launch_bounds__(256) void k_lg_x_separate(const cpx<T>* R, cpx<T>* out, long long nrows, long long so_row,
                                                       int nx, int nxo, long long p0, long long npc) {
  const I total = (I)((long long)nxo * npc), np = (I)npc;
  for (I q = (I)blockIdx.x * 256 + threadIdx.x; q < total; q += (I)gridDim.x * 256) {
    const long long k = (long long)(q / np), p = (long long)(q - (q / np) * np);
    const cpx<T> z = R[q];
    const cpx<T> zm = R[(k == 0 ? 0 : nx - k) * npc + p];
    const T h = (T)0.5;
    const long long r0 = 2 * (p0 + p);
    if (r0 < nrows) out[r0 * so_row + k] = mk<T>(h * (z.x + zm.x), h * (z.y - zm.y));
    if (r0 + 1 < nrows) out[(r0 + 1) * so_row + k] = mk<T>(h * (z.y + zm.y), h * (zm.x - z.x));
  }
}

// X axis, backward: the full spectrum of (row 2p) + i (row 2p+1) from the
// stored modes k < nxi (Hermitian fill above nx/2; Im of the DC and Nyquist
// modes ignored; modes at and above nxi zero)
template <class T, class I>
__global__ __launch_bounds__(256) void k_lg_x_merge(const cpx<T>* in, cpx<T>* A, long long nrows, long long si_row,
                                                    int nx, int nxi, long long p0, long long npc) {
  const I total = (I)((long long)nx * npc), np = (I)npc;
  const int half = nx / 2;
  for (I q = (I)blockIdx.x * 256 + threadIdx.x; q < total; q += (I)gridDim.x * 256) {
    const long long k = (long long)(q / np), p = (long long)(q - (q / np) * np);
    const bool up = k > half;
    const long long kk = up ? nx - k : k;
    const bool real_mode = kk == 0 || (nx % 2 == 0 && kk == half);
    const long long r0 = 2 * (p0 + p);
    auto get = [&](long long r) -> cpx<T> {
      if (r >= nrows || kk >= nxi) return mk<T>(0, 0);
      cpx<T> v = in[r * si_row + kk];
      if (real_mode) v.y = 0;
      if (up) v.y = -v.y;
      return v;
    };
    const cpx<T> va = get(r0), vb = get(r0 + 1);
    A[q] = mk<T>(va.x - vb.y, va.y + vb.x);
  }
}

// X axis, backward: real rows 2p, 2p+1 <- Re, Im of the inverse transform
template <class T, class I>
__global__ __launch_bounds__(256) void k_lg_x_store(const cpx<T>* R, T* out, long long nrows, long long so_row, int nx,
                                                    long long p0, long long npc) {
  const I total = (I)((long long)nx * npc), np = (I)npc;
  for (I q = (I)blockIdx.x * 256 + threadIdx.x; q < total; q += (I)gridDim.x * 256) {
    const long long e = (long long)(q / np), p = (long long)(q - (q / np) * np);
    const long long r0 = 2 * (p0 + p);
    const cpx<T> v = R[q];
    if (r0 < nrows) out[r0 * so_row + e] = v.x;
    if (r0 + 1 < nrows) out[(r0 + 1) * so_row + e] = v.y;
  }
}

// 32-bit index arithmetic where every index of a launch stays below 2^31
// (the grid stride then cannot wrap either)
bool idx32(long long n) { return n < (1LL << 31); }

unsigned grid_for(long long total) {
  const long long g = (total + 255) / 256;
  return (unsigned)std::max<long long>(1, std::min<long long>(g, 8192));
}

// ---------------------------------------------------------------------------
// plans
// ---------------------------------------------------------------------------
struct LongPlan {
  int m = 0;
  int N1 = 0, N2 = 0;  // four-step split of m (N1 = 0: Bluestein)
  int M = 0;           // Bluestein convolution length
  cplx_t* H = nullptr;  // transform of the Bluestein filter (M values)
};

std::recursive_mutex g_long_mu;
std::map<int, LongPlan> g_long;
// scratch: grown on demand outside stream capture; a replaced buffer is
// retired, not freed, so launches captured earlier stay valid
void* g_ls = nullptr;
size_t g_ls_bytes = 0;
std::vector<void*> g_ls_retired;

int max_prime(long long n) {
  int mp = 1;
  for (long long f = 2; f * f <= n; f++)
    while (n % f == 0) { mp = (int)std::max<long long>(mp, f); n /= f; }
  if (n > 1) mp = (int)std::max<long long>(mp, n);
  return mp;
}

// the most balanced split m = N1*N2 with both factors <= max_len() whose
// largest prime factor is at most `pmax` (false: none)
bool split_len(long long m, int pmax, int* n1, int* n2) {
  const long long lim = max_len();
  long long best = -1;
  for (long long d = 2; d <= lim && d * d <= m * 1; d++) {
    if (m % d) continue;
    const long long e = m / d;
    if (e > lim) continue;
    if (max_prime(d) > pmax || max_prime(e) > pmax) continue;
    best = d;  // increasing d: the last one found is the closest to sqrt(m)
  }
  if (best < 0) return false;
  *n1 = (int)best;  // the smaller factor
  *n2 = (int)(m / best);
  return true;
}

// smallest 2^a 3^b 5^c >= n
long long smooth235(long long n) {
  long long best = 1LL << 62;
  for (long long a = 1; a < 2 * n + 2; a *= 2)
    for (long long b = a; b < 2 * n + 2; b *= 3)
      for (long long c = b; c < 2 * n + 2; c *= 5)
        if (c >= n && c < best) best = c;
  return best;
}

void* long_scratch(size_t bytes, hipStream_t st) {
  std::lock_guard<std::recursive_mutex> lk(g_long_mu);
  if (bytes <= g_ls_bytes) return g_ls;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  check_hip(hipStreamIsCapturing(st, &cs), "hipStreamIsCapturing");
  if (cs != hipStreamCaptureStatusNone)
    throw Error{"transform lengths above " + std::to_string(max_len()) +
                    " need their scratch allocated outside stream capture: run the transform once before capturing it",
                -1};
  if (g_ls) g_ls_retired.push_back(g_ls);
  g_ls = nullptr;
  g_ls_bytes = 0;
  check_hip(hipMalloc(&g_ls, bytes), "hipMalloc(long-length scratch)");
  g_ls_bytes = bytes;
  return g_ls;
}

// scratch budget per chunk of columns (complex elements)
constexpr long long kChunkElems = (768LL << 20) / (long long)sizeof(cplx_t);

void flat_pass(cplx_t* in, cplx_t* out, long long so_in, long long sl_in, long long so_out, long long sl_out,
               long long ncc, int n_outer, int n, bool fwd, hipStream_t st) {
  ColSpec c;
  std::memset(&c, 0, sizeof(c));
  c.in = in; c.out = out;
  c.si_in = 1; c.sl_in = sl_in; c.so_in = so_in;
  c.si_out = 1; c.sl_out = sl_out; c.so_out = so_out;
  c.n_inner = (int)ncc; c.n_outer = n_outer;
  c.n = n; c.kind = COL_C2C; c.fwd = fwd;
  c.nld = n; c.ld_lo = n; c.ld_hi = n; c.ld_shift = 0;
  c.nst = n; c.st_lo = n; c.st_shift = 0;
  c.rev = false;
  (void)launch_col(c, st);
}

const LongPlan& get_long(int m);

// FFT of length m of ncc columns laid out row-major in A (column fastest),
// using B (split) or P, Q (Bluestein) as scratch; returns where the result is
cplx_t* fft_flat(cplx_t* A, cplx_t* B, cplx_t* P, cplx_t* Q, long long ncc, int m, bool fwd, hipStream_t st) {
  if (m <= max_len()) {
    flat_pass(A, A, 0, ncc, 0, ncc, ncc, 1, m, fwd, st);
    return A;
  }
  const LongPlan& L = get_long(m);
  const int sgn = fwd ? -1 : 1;
  if (L.N1) {
    const int N1 = L.N1, N2 = L.N2;
    // N2-point FFTs of the subsequences n1 (outer, stride ncc) over n2 (stride N1*ncc)
    flat_pass(A, A, ncc, (long long)N1 * ncc, ncc, (long long)N1 * ncc, ncc, N1, N2, fwd, st);
    if (idx32((long long)m * ncc))
      k_lg_twiddle<real_t, unsigned><<<grid_for((long long)m * ncc), 256, 0, st>>>(A, ncc, N1, m, sgn);
    else
      k_lg_twiddle<real_t, unsigned long long><<<grid_for((long long)m * ncc), 256, 0, st>>>(A, ncc, N1, m, sgn);
    check_hip(hipGetLastError(), "k_lg_twiddle launch");
    // N1-point FFTs over n1 (stride ncc) of each k2 (outer, stride N1*ncc),
    // stored at row k2 + N2*k1
    flat_pass(A, B, (long long)N1 * ncc, ncc, ncc, (long long)N2 * ncc, ncc, N2, N1, fwd, st);
    return B;
  }
  const int M = L.M;
  const bool i32 = idx32((long long)M * ncc);  // M > m
  if (i32) k_lg_blu_pre<real_t, unsigned><<<grid_for((long long)M * ncc), 256, 0, st>>>(A, P, ncc, m, M, sgn);
  else k_lg_blu_pre<real_t, unsigned long long><<<grid_for((long long)M * ncc), 256, 0, st>>>(A, P, ncc, m, M, sgn);
  check_hip(hipGetLastError(), "k_lg_blu_pre launch");
  cplx_t* R = fft_flat(P, Q, nullptr, nullptr, ncc, M, true, st);
  if (i32) k_lg_blu_mul<real_t, unsigned><<<grid_for((long long)M * ncc), 256, 0, st>>>(R, L.H, ncc, M, fwd ? 0 : 1);
  else k_lg_blu_mul<real_t, unsigned long long><<<grid_for((long long)M * ncc), 256, 0, st>>>(R, L.H, ncc, M, fwd ? 0 : 1);
  check_hip(hipGetLastError(), "k_lg_blu_mul launch");
  cplx_t* R2 = fft_flat(R, R == P ? Q : P, nullptr, nullptr, ncc, M, false, st);
  if (i32) k_lg_blu_post<real_t, unsigned><<<grid_for((long long)m * ncc), 256, 0, st>>>(R2, A, ncc, m, M, sgn);
  else k_lg_blu_post<real_t, unsigned long long><<<grid_for((long long)m * ncc), 256, 0, st>>>(R2, A, ncc, m, M, sgn);
  check_hip(hipGetLastError(), "k_lg_blu_post launch");
  return A;
}

// complex elements per column of a length-m transform's scratch
long long col_elems(int m) {
  if (m <= max_len()) return m;
  const LongPlan& L = get_long(m);
  return L.N1 ? 2LL * m : (long long)m + 2LL * L.M;
}

LongPlan build_long(int m) {
  LongPlan L;
  L.m = m;
  // four-step when both factors have small primes (the one-tile kernels run
  // a prime factor p above 13 at O(p) per element)
  if (split_len(m, 127, &L.N1, &L.N2)) {
    get_len(L.N1);
    get_len(L.N2);
    return L;
  }
  L.N1 = L.N2 = 0;
  const long long M = smooth235(2LL * m - 1);
  int a, b;
  if (M > (long long)max_len() * max_len() || !split_len(M, 5, &a, &b))
    throw Error{"transform length " + std::to_string(m) + " is not supported", -1};
  L.M = (int)M;
  // the convolution's own four-step plan (M is 5-smooth)
  const LongPlan& LM = get_long((int)M);
  if (!LM.N1) throw Error{"transform length " + std::to_string(m) + " is not supported", -1};
  a = LM.N1;
  b = LM.N2;
  // the filter's transform, once per length (FFT of length M in the build's
  // precision on the GPU)
  cplx_t* h = nullptr;
  cplx_t* t = nullptr;
  check_hip(hipMalloc(&h, (size_t)M * sizeof(cplx_t)), "hipMalloc(Bluestein filter)");
  check_hip(hipMalloc(&t, (size_t)M * sizeof(cplx_t)), "hipMalloc(Bluestein scratch)");
  k_lg_blu_filter<real_t><<<grid_for(M), 256, 0, 0>>>(h, m, (int)M);
  check_hip(hipGetLastError(), "k_lg_blu_filter launch");
  // one column, four-step (M is 5-smooth and splits)
  const int sgn = -1;
  flat_pass(h, h, 1, a, 1, a, 1, a, b, true, 0);
  k_lg_twiddle<real_t, unsigned><<<grid_for(M), 256, 0, 0>>>(h, 1, a, (int)M, sgn);  // M < 2^27
  check_hip(hipGetLastError(), "k_lg_twiddle launch");
  flat_pass(h, t, a, 1, 1, b, 1, b, a, true, 0);
  check_hip(hipDeviceSynchronize(), "Bluestein filter");
  check_hip(hipFree(h), "hipFree");
  L.H = t;
  return L;
}

const LongPlan& get_long(int m) {
  std::lock_guard<std::recursive_mutex> lk(g_long_mu);
  auto it = g_long.find(m);
  if (it != g_long.end()) return it->second;
  return g_long.emplace(m, build_long(m)).first->second;
}

// plans of lengths above max_len() are built (device allocations and
// synchronisation) on first use, which stream capture does not allow
void require_planned(int m, hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  check_hip(hipStreamIsCapturing(st, &cs), "hipStreamIsCapturing");
  if (cs == hipStreamCaptureStatusNone) return;
  std::lock_guard<std::recursive_mutex> lk(g_long_mu);
  auto it = g_long.find(m);
  if (it == g_long.end() || (it->second.M && g_long.find(it->second.M) == g_long.end()))
    throw Error{"transform length " + std::to_string(m) +
                    " is planned on first use: run the transform once before capturing it",
                -1};
}

}  // namespace

int max_len_any() { return 1 << 24; }

void plan_len(int n) {
  if (n <= max_len()) (void)get_len(n);
  else (void)get_long(n);
}

void launch_col_long(const ColSpec& c, int m, hipStream_t st) {
  require_planned(m, st);
  const long long ncol = (long long)c.n_inner * c.n_outer;
  const long long per = col_elems(m);
  const long long ncc = std::max<long long>(1, std::min<long long>(ncol, kChunkElems / per));
  cplx_t* A = (cplx_t*)long_scratch((size_t)(ncc * per) * sizeof(cplx_t), st);
  cplx_t* B = A + ncc * m;
  const LongPlan& L = get_long(m);
  cplx_t* P = L.N1 ? nullptr : A + ncc * m;
  cplx_t* Q = L.N1 ? nullptr : P + ncc * L.M;
  const bool fwd = c.fwd || c.kind == COL_DCT1 || c.kind == COL_DST1;
  for (long long c0 = 0; c0 < ncol; c0 += ncc) {
    LgCol g;
    std::memset(&g, 0, sizeof(g));
    g.in = c.in; g.out = c.out;
    g.si = c.si_in; g.sl = c.sl_in; g.so = c.so_in;
    g.n_inner = c.n_inner; g.kind = c.kind;
    g.nld = c.nld; g.ld_lo = c.ld_lo; g.ld_hi = c.ld_hi; g.ld_shift = c.ld_shift;
    g.nst = c.nst; g.st_lo = c.st_lo; g.st_shift = c.st_shift;
    g.c0 = c0;
    g.ncc = std::min(ncc, ncol - c0);
    g.m = m;
    // 32-bit counters: element index (rows x chunk columns) and flattened column
    const bool i32 = idx32((long long)std::max(m, c.nst) * g.ncc) && idx32(ncol);
    if (i32) k_lg_gather<real_t, unsigned><<<grid_for((long long)m * g.ncc), 256, 0, st>>>(g, A);
    else k_lg_gather<real_t, unsigned long long><<<grid_for((long long)m * g.ncc), 256, 0, st>>>(g, A);
    check_hip(hipGetLastError(), "k_lg_gather launch");
    const cplx_t* R = fft_flat(A, B, P, Q, g.ncc, m, fwd, st);
    g.si = c.si_out; g.sl = c.sl_out; g.so = c.so_out;
    if (i32) k_lg_scatter<real_t, unsigned><<<grid_for((long long)c.nst * g.ncc), 256, 0, st>>>(g, R);
    else k_lg_scatter<real_t, unsigned long long><<<grid_for((long long)c.nst * g.ncc), 256, 0, st>>>(g, R);
    check_hip(hipGetLastError(), "k_lg_scatter launch");
  }
}

void launch_x_long(bool r2c, const void* in, void* out, long long nrows, long long si_row, long long so_row, int nx,
                   int nk, hipStream_t st) {
  require_planned(nx, st);
  const long long pairs = (nrows + 1) / 2;
  const long long per = col_elems(nx);
  const long long npc = std::max<long long>(1, std::min<long long>(pairs, kChunkElems / per));
  cplx_t* A = (cplx_t*)long_scratch((size_t)(npc * per) * sizeof(cplx_t), st);
  cplx_t* B = A + npc * nx;
  const LongPlan& L = get_long(nx);
  cplx_t* P = L.N1 ? nullptr : A + npc * nx;
  cplx_t* Q = L.N1 ? nullptr : P + npc * L.M;
  for (long long p0 = 0; p0 < pairs; p0 += npc) {
    const long long n = std::min(npc, pairs - p0);
    const unsigned gx = grid_for((long long)nx * n);
    const bool i32 = idx32((long long)nx * n);  // nk <= nx
    if (r2c) {
      if (i32) k_lg_x_gather<real_t, unsigned><<<gx, 256, 0, st>>>((const real_t*)in, A, nrows, si_row, nx, p0, n);
      else k_lg_x_gather<real_t, unsigned long long><<<gx, 256, 0, st>>>((const real_t*)in, A, nrows, si_row, nx, p0, n);
      check_hip(hipGetLastError(), "k_lg_x_gather launch");
      const cplx_t* R = fft_flat(A, B, P, Q, n, nx, true, st);
      if (i32)
        k_lg_x_separate<real_t, unsigned><<<grid_for((long long)nk * n), 256, 0, st>>>(R, (cplx_t*)out, nrows, so_row,
                                                                                       nx, nk, p0, n);
      else
        k_lg_x_separate<real_t, unsigned long long><<<grid_for((long long)nk * n), 256, 0, st>>>(
            R, (cplx_t*)out, nrows, so_row, nx, nk, p0, n);
      check_hip(hipGetLastError(), "k_lg_x_separate launch");
    } else {
      if (i32) k_lg_x_merge<real_t, unsigned><<<gx, 256, 0, st>>>((const cplx_t*)in, A, nrows, si_row, nx, nk, p0, n);
      else k_lg_x_merge<real_t, unsigned long long><<<gx, 256, 0, st>>>((const cplx_t*)in, A, nrows, si_row, nx, nk, p0, n);
      check_hip(hipGetLastError(), "k_lg_x_merge launch");
      const cplx_t* R = fft_flat(A, B, P, Q, n, nx, false, st);
      if (i32) k_lg_x_store<real_t, unsigned><<<gx, 256, 0, st>>>(R, (real_t*)out, nrows, so_row, nx, p0, n);
      else k_lg_x_store<real_t, unsigned long long><<<gx, 256, 0, st>>>(R, (real_t*)out, nrows, so_row, nx, p0, n);
      check_hip(hipGetLastError(), "k_lg_x_store launch");
    }
  }
}

void free_long() {
  std::lock_guard<std::recursive_mutex> lk(g_long_mu);
  (void)hipDeviceSynchronize();
  for (auto& kv : g_long)
    if (kv.second.H) (void)hipFree(kv.second.H);
  g_long.clear();
  for (void* p : g_ls_retired) (void)hipFree(p);
  g_ls_retired.clear();
  if (g_ls) (void)hipFree(g_ls);
  g_ls = nullptr;
  g_ls_bytes = 0;
}

void free_long_scratch() {
  std::lock_guard<std::recursive_mutex> lk(g_long_mu);
  (void)hipDeviceSynchronize();
  for (void* p : g_ls_retired) (void)hipFree(p);
  g_ls_retired.clear();
  if (g_ls) (void)hipFree(g_ls);
  g_ls = nullptr;
  g_ls_bytes = 0;
}

}  // namespace p3amd
