// api_poisson.cpp -- spectral Poisson solves on resident HBM fields, the unit of
// work of the served / batched path (no counterpart in the reference; it
// composes the reference's forward and backward transforms, build/ftran.F90:
// 489-780 and build/btran.F90:396-679, around a per-mode multiply).
//
// A solve of -lap(u) = -f on the periodic box of an (nx, ny, nz) grid:
//   f_hat = r2c(f);  u_hat = -f_hat / |k|^2 / (nx ny nz), u_hat(k=0) = 0;
//   u = c2r(u_hat)
// in place on the field, with integer wavenumbers k = j (j <= n/2) or j - n.
// Solves take their work area from the caller, so any number of solves of
// any shapes can be in flight on different streams.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <string>
#include <vector>

#include "engine.hpp"
#include "p3dfft_amd_internal.hpp"

using namespace p3amd;

namespace p3amd {

template <class T>
__global__ void k_poisson(cpx<T>* spec, int nxhp, int ny, int nz, double scale) {
  const long long n = (long long)nxhp * ny * nz;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = (int)(i % nxhp);
  const long long r = i / nxhp;
  const int y = (int)(r % ny);
  const int z = (int)(r / ny);
  const double ky = y <= ny / 2 ? y : y - ny;
  const double kz = z <= nz / 2 ? z : z - nz;
  const double k2 = (double)x * x + ky * ky + kz * kz;
  const double f = k2 > 0.0 ? -scale / k2 : 0.0;
  const cpx<T> v = spec[i];
  spec[i] = mk<T>((T)(v.x * f), (T)(v.y * f));
}

// device RNG: counter-based (splitmix64 finaliser) uniform pairs, Box-Muller
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <class T>
__global__ void k_fill_normal(T* p, long long n, uint64_t seed) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t a = mix64(seed * 0x100000001B3ull + 2 * (uint64_t)i);
  const uint64_t b = mix64(seed * 0x100000001B3ull + 2 * (uint64_t)i + 1);
  const double u1 = ((a >> 11) + 1.0) * (1.0 / 9007199254740993.0);  // (0, 1]
  const double u2 = (b >> 11) * (1.0 / 9007199254740992.0);          // [0, 1)
  p[i] = (T)(sqrt(-2.0 * log(u1)) * cospi(2.0 * u2));
}

// fused Z forward / multiply / Z inverse kernel (default on); off composes
// the solve from the forward transform, a multiply pass and the backward one
static bool g_fused_solve = true;

// Poisson graphs alive (their captured launches hold plan twiddle tables)
std::atomic<int> g_live_graphs{0};

size_t poisson_work_bytes(int nx, int ny, int nz) {
  // the spectrum: every pass after X r2c runs in place on it
  return (size_t)(nx / 2 + 1) * ny * nz * sizeof(cplx_t);
}

}  // namespace p3amd

extern "C" {

// Build (once) the plan and twiddle table of a transform length, so that later
// calls of any entry point with that length launch without set-up work.
int p3dfft_amd_prepare_length(int n) {
  try {
    if (n <= 0 || n > max_len()) return -1;
    get_len(n);
    return 0;
  } catch (const Error& e) {
    set_error(e.msg);
    return -1;
  }
}

// 1 (default): fused Z-solve kernel; 0: forward transform, multiply pass,
// backward transform (the composition the fused kernel replaces).
void p3dfft_poisson_set_fused(int on) { g_fused_solve = (on != 0); }

// HBM work bytes one solve of this shape needs besides its field.
size_t p3dfft_poisson_work_bytes(int nx, int ny, int nz) { return poisson_work_bytes(nx, ny, nz); }

// Solve in place on `field` (real (nx, ny, nz), x fastest, in HBM) using
// `work` (at least p3dfft_poisson_work_bytes) on `stream`.  Asynchronous.
int p3dfft_poisson_solve(int nx, int ny, int nz, p3amd::real_t* field, void* work, size_t work_bytes,
                         void* stream) {
  try {
    if (nx <= 0 || ny <= 0 || nz <= 0 || nx > max_len() || ny > max_len() || nz > max_len()) {
      set_error("p3dfft_poisson_solve: invalid grid");
      return -1;
    }
    if (work_bytes < poisson_work_bytes(nx, ny, nz)) {
      set_error("p3dfft_poisson_solve: work area too small");
      return -1;
    }
    const int nxhp = nx / 2 + 1;
    const long long nxy = (long long)nxhp * ny;
    cplx_t* spec = (cplx_t*)work;
    hipStream_t st = (hipStream_t)stream;
    const double scale = 1.0 / ((double)nx * ny * nz);
    const bool alt = tile_order_alternate();
    // every pass after X r2c runs in place on the spectrum; consecutive
    // passes walk their tiles in opposite orders (the last-level cache keeps
    // what the previous pass read)
    auto col = [&](int axis, bool fwd, bool rev) {
      ColSpec c{};
      c.in = spec; c.out = spec;
      c.si_in = 1; c.si_out = 1;
      if (axis == 1) {
        c.sl_in = c.sl_out = nxhp; c.so_in = c.so_out = nxy; c.n_inner = nxhp; c.n_outer = nz; c.n = ny;
      } else {
        c.sl_in = c.sl_out = nxy; c.so_in = c.so_out = 0; c.n_inner = (int)nxy; c.n_outer = 1; c.n = nz;
      }
      c.kind = COL_C2C; c.fwd = fwd;
      c.nld = c.n; c.ld_lo = c.n; c.ld_hi = c.n; c.ld_shift = 0;
      c.nst = c.n; c.st_lo = c.n; c.st_shift = 0;
      c.rev = rev;
      launch_col(c, st);
    };
    launch_x_r2c(field, spec, (long long)ny * nz, nx, nxhp, nx, nxhp, st);
    col(1, true, alt);
    if (!(g_fused_solve && launch_zsolve(spec, nxhp, ny, nz, scale, st))) {
      // composed: Z forward, the multiply pass, Z inverse
      col(2, true, false);
      const long long n = nxy * nz;
      k_poisson<real_t><<<(unsigned)((n + 255) / 256), 256, 0, st>>>(spec, nxhp, ny, nz, scale);
      check_hip(hipGetLastError(), "k_poisson");
      col(2, false, alt);
    }
    col(1, false, false);
    launch_x_c2r(spec, field, (long long)ny * nz, nxhp, nx, nx, nxhp, st, alt);
    return 0;
  } catch (const Error& e) {
    set_error(e.msg);
    return -1;
  }
}

// The per-mode multiply of a solve on its own: spectrum (nx/2+1, ny, nz) in
// HBM times -1/(|k|^2 nx ny nz), k=0 -> 0, for callers that compose a solve
// from p3dfft_ftran_r2c_dev and p3dfft_btran_c2r_dev.  Asynchronous.
int p3dfft_poisson_multiply(p3amd::real_t* spec, int nx, int ny, int nz, void* stream) {
  if (nx <= 0 || ny <= 0 || nz <= 0) {
    set_error("p3dfft_poisson_multiply: invalid grid");
    return -1;
  }
  const int nxhp = nx / 2 + 1;
  const long long n = (long long)nxhp * ny * nz;
  k_poisson<real_t><<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(
      (cplx_t*)spec, nxhp, ny, nz, 1.0 / ((double)nx * ny * nz));
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// A fixed batch of solves captured once into a HIP graph and replayed: the
// launches of every solve (five to seven kernels each) are issued by one
// graph launch.  The solves are spread over `nstreams` forked branches
// (largest first, each to the branch with the least work so far), so
// independent solves overlap on the device as they do on separate streams.
struct PoissonGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int nsolves = 0;
};

int p3dfft_poisson_graph_create(int n, const int* shapes, p3amd::real_t* const* fields, void* const* works,
                                const size_t* work_bytes, int nstreams, void** graph_out) {
  if (!graph_out || n <= 0 || !shapes || !fields || !works || !work_bytes) {
    set_error("p3dfft_poisson_graph_create: invalid arguments");
    return -1;
  }
  *graph_out = nullptr;
  for (int i = 0; i < n; i++) {
    const int nx = shapes[3 * i], ny = shapes[3 * i + 1], nz = shapes[3 * i + 2];
    if (nx <= 0 || ny <= 0 || nz <= 0 || nx > max_len() || ny > max_len() || nz > max_len()) {
      set_error("p3dfft_poisson_graph_create: invalid grid");
      return -1;
    }
    if (!fields[i] || !works[i] || work_bytes[i] < poisson_work_bytes(nx, ny, nz)) {
      set_error("p3dfft_poisson_graph_create: missing field or work area too small");
      return -1;
    }
  }
  const int nb = std::max(1, std::min(nstreams, n));
  // plans, twiddle tables and kernel attributes first: capture records launches only
  for (int i = 0; i < 3 * n; i++)
    if (p3dfft_amd_prepare_length(shapes[i]) != 0) return -1;
  // largest-first assignment to the least-loaded branch
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  auto pts = [&](int i) { return (double)shapes[3 * i] * shapes[3 * i + 1] * shapes[3 * i + 2]; };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return pts(a) > pts(b); });
  std::vector<double> load(nb, 0.0);
  std::vector<std::vector<int>> branch(nb);
  for (int i : order) {
    const int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    branch[b].push_back(i);
    load[b] += pts(i);
  }
  hipStream_t cap = nullptr;
  std::vector<hipStream_t> side(nb, nullptr);
  std::vector<hipEvent_t> done(nb, nullptr);
  hipEvent_t fork = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool capturing = false;
  int rc = -1;
  try {
    check_hip(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "hipStreamCreate");
    check_hip(hipEventCreateWithFlags(&fork, hipEventDisableTiming), "hipEventCreate");
    for (int b = 0; b < nb; b++) {
      check_hip(hipStreamCreateWithFlags(&side[b], hipStreamNonBlocking), "hipStreamCreate");
      check_hip(hipEventCreateWithFlags(&done[b], hipEventDisableTiming), "hipEventCreate");
    }
    check_hip(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    capturing = true;
    check_hip(hipEventRecord(fork, cap), "hipEventRecord");
    for (int b = 0; b < nb; b++) {
      check_hip(hipStreamWaitEvent(side[b], fork, 0), "hipStreamWaitEvent");
      for (int i : branch[b])
        if (p3dfft_poisson_solve(shapes[3 * i], shapes[3 * i + 1], shapes[3 * i + 2], fields[i], works[i],
                                 work_bytes[i], (void*)side[b]) != 0)
          throw Error{std::string(), -1};  // the solve set the message
      check_hip(hipEventRecord(done[b], side[b]), "hipEventRecord");
      check_hip(hipStreamWaitEvent(cap, done[b], 0), "hipStreamWaitEvent");
    }
    capturing = false;
    check_hip(hipStreamEndCapture(cap, &graph), "hipStreamEndCapture");
    check_hip(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    PoissonGraph* g = new PoissonGraph;
    g->graph = graph;
    g->exec = exec;
    g->nsolves = n;
    *graph_out = g;
    g_live_graphs++;
    graph = nullptr;
    exec = nullptr;
    rc = 0;
  } catch (const Error& e) {
    if (!e.msg.empty()) set_error(e.msg);
  }
  if (capturing) {
    hipGraph_t partial = nullptr;
    (void)hipStreamEndCapture(cap, &partial);
    if (partial) (void)hipGraphDestroy(partial);
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  for (int b = 0; b < nb; b++) {
    if (done[b]) (void)hipEventDestroy(done[b]);
    if (side[b]) (void)hipStreamDestroy(side[b]);
  }
  if (fork) (void)hipEventDestroy(fork);
  if (cap) (void)hipStreamDestroy(cap);
  return rc;
}

// A fixed batch of solves run as grouped launches: every pass of every
// solve whose lengths have specialised kernels is issued as one launch per
// pass and transform length (five to a few tens of launches per batch instead
// of five to seven per solve), each workgroup finding its field in a tile
// table; the other solves run one by one after them.  Results are
// bit-identical to p3dfft_poisson_solve of each field.
struct PoissonBatch {
  GroupBatch* gb = nullptr;
  std::vector<int> singles, shapes;
  std::vector<p3amd::real_t*> fields;
  std::vector<void*> works;
  std::vector<size_t> wbytes;
};

int p3dfft_poisson_batch_create(int n, const int* shapes, p3amd::real_t* const* fields, void* const* works,
                                const size_t* work_bytes, void** batch_out) {
  if (!batch_out || n <= 0 || !shapes || !fields || !works || !work_bytes) {
    set_error("p3dfft_poisson_batch_create: invalid arguments");
    return -1;
  }
  *batch_out = nullptr;
  for (int i = 0; i < n; i++) {
    const int nx = shapes[3 * i], ny = shapes[3 * i + 1], nz = shapes[3 * i + 2];
    if (nx <= 0 || ny <= 0 || nz <= 0 || nx > max_len() || ny > max_len() || nz > max_len()) {
      set_error("p3dfft_poisson_batch_create: invalid grid");
      return -1;
    }
    if (!fields[i] || !works[i] || work_bytes[i] < poisson_work_bytes(nx, ny, nz)) {
      set_error("p3dfft_poisson_batch_create: missing field or work area too small");
      return -1;
    }
  }
  for (int i = 0; i < 3 * n; i++)
    if (p3dfft_amd_prepare_length(shapes[i]) != 0) return -1;
  PoissonBatch* b = new PoissonBatch;
  try {
    b->gb = group_batch_create(0, n, shapes, fields, works, &b->singles);
  } catch (const Error& e) {
    delete b;
    set_error(e.msg);
    return -1;
  }
  b->shapes.assign(shapes, shapes + 3 * n);
  b->fields.assign(fields, fields + n);
  b->works.assign(works, works + n);
  b->wbytes.assign(work_bytes, work_bytes + n);
  g_live_graphs++;
  *batch_out = b;
  return 0;
}

// Issue the batch on `stream` (asynchronous; in place on the fields).
int p3dfft_poisson_batch_launch(void* batch, void* stream) {
  if (!batch) {
    set_error("p3dfft_poisson_batch_launch: null batch");
    return -1;
  }
  PoissonBatch* b = (PoissonBatch*)batch;
  try {
    group_batch_launch(b->gb, (hipStream_t)stream);
  } catch (const Error& e) {
    set_error(e.msg);
    return -1;
  }
  for (int i : b->singles)
    if (p3dfft_poisson_solve(b->shapes[3 * i], b->shapes[3 * i + 1], b->shapes[3 * i + 2], b->fields[i],
                             b->works[i], b->wbytes[i], stream) != 0)
      return -1;
  return 0;
}

// Launches one p3dfft_poisson_batch_launch issues for the grouped solves,
// and how many solves run one by one after them.
int p3dfft_poisson_batch_info(void* batch, int* grouped_launches, int* single_solves) {
  if (!batch) return -1;
  PoissonBatch* b = (PoissonBatch*)batch;
  if (grouped_launches) *grouped_launches = group_batch_launches(b->gb);
  if (single_solves) *single_solves = (int)b->singles.size();
  return 0;
}

int p3dfft_poisson_batch_destroy(void* batch) {
  if (!batch) return 0;
  PoissonBatch* b = (PoissonBatch*)batch;
  group_batch_free(b->gb);
  delete b;
  g_live_graphs--;
  return 0;
}

// Batched 3D transforms of mixed shapes (no setup needed): direction 1 is
// r2c 'fft' of reals[i] (nx, ny, nz) into spectra[i] (nx/2+1, ny, nz),
// unnormalised, as p3dfft_ftran_r2c_dev computes it; direction -1 is c2r
// 'tff' of spectra[i] (overwritten) into reals[i].  One launch per pass and
// transform length; every length must have a specialised kernel.
int p3dfft_amd_batch_create(int direction, int n, const int* shapes, p3amd::real_t* const* reals,
                            p3amd::real_t* const* spectra, void** batch_out) {
  if (!batch_out || n <= 0 || !shapes || !reals || !spectra || (direction != 1 && direction != -1)) {
    set_error("p3dfft_amd_batch_create: invalid arguments");
    return -1;
  }
  *batch_out = nullptr;
  for (int i = 0; i < n; i++) {
    const int nx = shapes[3 * i], ny = shapes[3 * i + 1], nz = shapes[3 * i + 2];
    if (nx <= 0 || ny <= 0 || nz <= 0 || nx > max_len() || ny > max_len() || nz > max_len() || !reals[i] ||
        !spectra[i]) {
      set_error("p3dfft_amd_batch_create: invalid grid or array");
      return -1;
    }
  }
  for (int i = 0; i < 3 * n; i++)
    if (p3dfft_amd_prepare_length(shapes[i]) != 0) return -1;
  PoissonBatch* b = new PoissonBatch;
  try {
    b->gb = group_batch_create(direction == 1 ? 1 : 2, n, shapes, reals, (void* const*)spectra, &b->singles);
  } catch (const Error& e) {
    delete b;
    set_error(e.msg);
    return -1;
  }
  if (!b->singles.empty()) {
    group_batch_free(b->gb);
    delete b;
    set_error("p3dfft_amd_batch_create: a length has no specialised kernel");
    return -1;
  }
  g_live_graphs++;
  *batch_out = b;
  return 0;
}

int p3dfft_amd_batch_launch(void* batch, void* stream) {
  if (!batch) {
    set_error("p3dfft_amd_batch_launch: null batch");
    return -1;
  }
  try {
    group_batch_launch(((PoissonBatch*)batch)->gb, (hipStream_t)stream);
  } catch (const Error& e) {
    set_error(e.msg);
    return -1;
  }
  return 0;
}

int p3dfft_amd_batch_destroy(void* batch) { return p3dfft_poisson_batch_destroy(batch); }

// Replay every captured solve on `stream` (asynchronous; the fields are
// solved in place again, so the caller refreshes them between replays).
int p3dfft_poisson_graph_launch(void* graph, void* stream) {
  if (!graph) {
    set_error("p3dfft_poisson_graph_launch: null graph");
    return -1;
  }
  const hipError_t e = hipGraphLaunch(((PoissonGraph*)graph)->exec, (hipStream_t)stream);
  if (e != hipSuccess) {
    set_error(std::string("hipGraphLaunch: ") + hipGetErrorString(e));
    return -1;
  }
  return 0;
}

int p3dfft_poisson_graph_destroy(void* graph) {
  if (!graph) return 0;
  PoissonGraph* g = (PoissonGraph*)graph;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  g_live_graphs--;
  return 0;
}

// Fill n values with i.i.d. N(0,1) from a counter-based generator keyed by seed.
int p3dfft_amd_fill_normal(p3amd::real_t* p, long long n, unsigned long long seed, void* stream) {
  if (n <= 0) return 0;
  k_fill_normal<real_t><<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(p, n, seed);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Streams for callers that run work concurrently (default priority).
void* p3dfft_amd_stream_create(void) {
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
  return (void*)s;
}
int p3dfft_amd_stream_destroy(void* s) { return hipStreamDestroy((hipStream_t)s) == hipSuccess ? 0 : -1; }
int p3dfft_amd_stream_synchronize(void* s) {
  return hipStreamSynchronize((hipStream_t)s) == hipSuccess ? 0 : -1;
}
// 1 when every operation before the event has completed, 0 otherwise.
int p3dfft_amd_event_query(void* e) { return hipEventQuery((hipEvent_t)e) == hipSuccess ? 1 : 0; }

}  // extern "C"
