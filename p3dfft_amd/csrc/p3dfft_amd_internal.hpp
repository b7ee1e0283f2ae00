// p3dfft_amd_internal.hpp -- module state shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <string>
#include <vector>

#include "engine.hpp"
#include "share.hpp"

namespace p3amd {

struct Pending {
  int slot;
  hipEvent_t a, b;
};

// The reference's module variables (build/module.F90:101-176) for one setup.
struct State {
  bool set = false;
  int nx = 0, ny = 0, nz = 0, nxc = 0, nyc = 0, nzc = 0;
  bool ow = true;
  bool stride1 = false, stride1_next = false;
  int p1 = 1, p2 = 1, ipid = 0, jpid = 0;
  int nxh = 0, nxhp = 0, nxhc = 0, nxhpc = 0, nyh = 0, nzh = 0, nyhc = 0, nzhc = 0, nycph = 0, nzcph = 0;
  int iistart = 0, iiend = 0, iisize = 0, jistart = 0, jiend = 0, jisize = 0;
  int jjstart = 0, jjend = 0, jjsize = 0, kjstart = 0, kjend = 0, kjsize = 0;
  std::vector<int> iist, iisz, jist, jisz, jjst, jjsz, kjst, kjsz;
  int memsize[3] = {0, 0, 0};
  // HBM work buffers: w1 (nxhpc, ny, nz), w2 (nxhpc, nyc, nz)
  cplx_t* w1 = nullptr;
  cplx_t* w2 = nullptr;
  int nv_alloc = 0;  // variables the work buffers hold (grown by the _many calls)
  // per-plane partial shell sums of the power spectrum (grown on demand)
  double* spec_partial = nullptr;
  size_t spec_partial_bytes = 0;
  // share mode (one rank of a P1 x P2 grid): Z pencil work buffer and splits
  bool share = false;
  cplx_t* w3 = nullptr;
  ShareGeom sg;
  // staging of host arrays for the host-pointer entry points
  void* stage_a = nullptr;
  void* stage_b = nullptr;
  size_t stage_bytes = 0;
  // X/Y chunk size in bytes of intermediate, 0 = whole grid in one pass
  // (measured on MI355X at 512^3: chunks of 16-128 MiB were 10-110% slower)
  size_t chunk_bytes = 0;
  hipStream_t stream = nullptr;
  // 12 timer slots (seconds)
  double timers[12] = {0};
  bool timing = true;
  hipEvent_t open[12] = {};
  std::vector<Pending> pend;
  std::vector<hipEvent_t> evpool;
};

extern State g;
// Poisson graphs and captured transform graphs alive (plans may not be
// reset while any exists)
extern std::atomic<int> g_live_graphs;
// captured transform graphs alive: they hold the setup's work buffers, so
// p3dfft_clean and a growth of those buffers are refused while any exists
extern std::atomic<int> g_live_captures;

bool fatal(const std::string& msg);
void notice(const std::string& msg);
void set_error(const std::string& msg);
bool require_setup();
void tstart(State& S, int slot, hipStream_t s);
void tstop(State& S, int slot, hipStream_t s);
void timers_resolve();
void map_data_to_proc(int data, int proc, int* st, int* en, int* sz);
bool compute_geometry(State& s, const int dims[2], int ipid, int jpid);
size_t real_elems();
size_t spec_elems();
void ensure_staging();
// cheby_done non-null: fuse the Chebyshev recurrence (L = cheby_L) into the
// DCT-I Z pass where the kernel allows; *cheby_done tells whether it was
bool forward_dev(State& S, const real_t* in, cplx_t* out, const unsigned char* op, hipStream_t s,
                 double cheby_L = 0, bool* cheby_done = nullptr);
bool backward_dev(State& S, cplx_t* in, real_t* out, const unsigned char* op, hipStream_t s);
bool cheby_dev(State& S, const real_t* in, cplx_t* out, double Lz, hipStream_t s);
void release_state();
bool zop_len_ok(int nz, int kind);
bool ensure_many(State& S, int nv);
int chunk_planes(State& S);
void share_release();

}  // namespace p3amd
