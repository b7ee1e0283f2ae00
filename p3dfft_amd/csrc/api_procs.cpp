// api_procs.cpp -- the processor-grid queries of the reference's module for
// the current setup: every rank's pencil extents (the proc_dims and
// proc_id2coords arrays that build/setup.F90:224-229, 551-574 gather with
// MPI_Allgather) and get_proc_parts (build/module.F90:888-1054), which splits
// a box of the global grid into the pieces held by each rank.  Integer host
// work only (no device calls): the extents of every rank follow from the
// setup's MapDataToProc splits, which the share mode and the 1x1 setup both
// keep for the whole grid.
#include <algorithm>
#include <vector>

#include "p3dfft_amd_internal.hpp"

using namespace p3amd;

namespace {

// Rank ids of the reference's Cartesian communicator (build/setup.F90:195-229,
// the default non-DIMS_C order): MPI_Cart_create over (jproc, iproc) ranks
// row-major, ipid = cartid(2), jpid = cartid(1), so id = jpid * iproc + ipid.
int rank_of(int ip, int jp) { return jp * g.p1 + ip; }

// p3dfft_get_dims (build/module.F90:225-273) of rank (ip, jp): start, end and
// size (3 each, 1-based) of configuration conf (1: real X pencil, 2: complex
// Z pencil, (z, y, x) under STRIDE1)
void dims_of(int ip, int jp, int conf, int st[3], int sz[3]) {
  if (conf == 1) {
    st[0] = 1; sz[0] = g.nx;
    st[1] = g.jist[ip]; sz[1] = g.jisz[ip];
    st[2] = g.kjst[jp]; sz[2] = g.kjsz[jp];
  } else if (g.stride1) {
    st[0] = 1; sz[0] = g.nzc;
    st[1] = g.jjst[jp]; sz[1] = g.jjsz[jp];
    st[2] = g.iist[ip]; sz[2] = g.iisz[ip];
  } else {
    st[0] = g.iist[ip]; sz[0] = g.iisz[ip];
    st[1] = g.jjst[jp]; sz[1] = g.jjsz[jp];
    st[2] = 1; sz[2] = g.nzc;
  }
}

// one axis of the box: [b, b + n) split at the boundaries of the ranks along
// that axis of the grid (coordinate c = 0..P-1 of the rank owning start
// starts[c] with size sizes[c]); appends (coordinate, base, size) pieces in
// order.  False when the box leaves the axis.
struct Piece {
  int c, base, size;
};
bool split_axis(int b, int n, const std::vector<int>& starts, const std::vector<int>& sizes,
                std::vector<Piece>* out) {
  const int P = (int)starts.size();
  const int lo = starts[0], hi = starts[P - 1] + sizes[P - 1];  // [lo, hi)
  if (n < 1 || b < lo || b + n > hi) return false;
  int c = 0;
  while (b >= starts[c] + sizes[c]) c++;  // the rank that holds the base point
  while (n > 0) {
    const int take = std::min(n, starts[c] + sizes[c] - b);
    out->push_back(Piece{c, b, take});
    b += take;
    n -= take;
    c++;
  }
  return true;
}

}  // namespace

extern "C" {

// proc_dims[(c-1) + 2*((f-1) + 9*p)]: conf c = 1, 2; field f = 1..9 (start x,
// y, z, end x, y, z, size x, y, z as p3dfft_get_dims gives them); rank p.
// id2coords[2p], id2coords[2p+1]: (ipid, jpid) of rank p.  Null arrays are
// skipped.  Returns the number of ranks P, or -1 before setup.
int p3dfft_amd_proc_dims(int* proc_dims, int* id2coords) {
  if (!g.set) {
    set_error(" P3DFFT error: call setup before other routines");
    return -1;
  }
  const int P = g.p1 * g.p2;
  for (int jp = 0; jp < g.p2; jp++)
    for (int ip = 0; ip < g.p1; ip++) {
      const int p = rank_of(ip, jp);
      if (id2coords) {
        id2coords[2 * p] = ip;
        id2coords[2 * p + 1] = jp;
      }
      if (!proc_dims) continue;
      for (int conf = 1; conf <= 2; conf++) {
        int st[3], sz[3];
        dims_of(ip, jp, conf, st, sz);
        for (int a = 0; a < 3; a++) {
          proc_dims[(conf - 1) + 2 * (a + 9 * p)] = st[a];
          proc_dims[(conf - 1) + 2 * (a + 3 + 9 * p)] = st[a] + sz[a] - 1;
          proc_dims[(conf - 1) + 2 * (a + 6 + 9 * p)] = sz[a];
        }
      }
    }
  return P;
}

// build/module.F90:888-1054: the pieces of the box with 1-based base point
// (base_x, base_y, base_z) and sizes (size_x, size_y, size_z) of
// configuration conf held by each rank, in the reference's order (pieces
// along the rank-grid's i axis outermost, then along j; conf 1: i = y,
// j = z; conf 2: i = x, j = y).
// parts is the reference's proc_parts(P, 7) (Fortran order,
// parts[(col-1)*P + (q-1)] for piece q): column 1 the rank id, 2-4 the base
// point, 5-7 the sizes (x, y, z order); rows past the last piece are -1.
// ierr: 0; 1 for a conf other than 1 or 2; -1 when the box is not inside the
// grid (the reference finds no rank for a base point beyond the grid and
// does not check the others).  Returns the number of pieces (or -1).
int p3dfft_amd_get_proc_parts(int base_x, int base_y, int base_z, int size_x, int size_y, int size_z, int conf,
                              int* parts, int* ierr) {
  if (!g.set) {
    set_error(" P3DFFT error: call setup before other routines");
    if (ierr) *ierr = -1;
    return -1;
  }
  const int P = g.p1 * g.p2;
  if (parts)
    for (int i = 0; i < 7 * P; i++) parts[i] = -1;
  if (conf != 1 && conf != 2) {
    if (ierr) *ierr = 1;
    return -1;
  }
  const int base[3] = {base_x, base_y, base_z}, size[3] = {size_x, size_y, size_z};
  // axes (0-based, in the configuration's index order) split over the rank
  // grid's i and j directions, and the one each pencil holds whole (the
  // STRIDE1 complex layout is (z, y, x): x is split over i there)
  int ai = 0, aj = 1, ak = 2;
  if (conf == 1) { ai = 1; aj = 2; ak = 0; }
  else if (g.stride1) { ai = 2; aj = 1; ak = 0; }
  std::vector<int> si(g.p1), ni(g.p1), sj(g.p2), nj(g.p2);
  int st[3], sz[3];
  for (int ip = 0; ip < g.p1; ip++) {
    dims_of(ip, 0, conf, st, sz);
    si[ip] = st[ai];
    ni[ip] = sz[ai];
  }
  for (int jp = 0; jp < g.p2; jp++) {
    dims_of(0, jp, conf, st, sz);
    sj[jp] = st[aj];
    nj[jp] = sz[aj];
  }
  dims_of(0, 0, conf, st, sz);
  std::vector<Piece> pi, pj;
  if (!split_axis(base[ai], size[ai], si, ni, &pi) || !split_axis(base[aj], size[aj], sj, nj, &pj) ||
      size[ak] < 1 || base[ak] < st[ak] || base[ak] + size[ak] > st[ak] + sz[ak]) {
    if (ierr) *ierr = -1;
    return -1;
  }
  int q = 0;
  for (const Piece& a : pi)
    for (const Piece& b : pj) {
      if (parts) {
        int pb[3], ps[3];
        pb[ai] = a.base; ps[ai] = a.size;
        pb[aj] = b.base; ps[aj] = b.size;
        pb[ak] = base[ak]; ps[ak] = size[ak];
        parts[0 * P + q] = rank_of(a.c, b.c);
        for (int d = 0; d < 3; d++) {
          parts[(1 + d) * P + q] = pb[d];
          parts[(4 + d) * P + q] = ps[d];
        }
      }
      q++;
    }
  if (ierr) *ierr = 0;
  return q;
}

}  // extern "C"
