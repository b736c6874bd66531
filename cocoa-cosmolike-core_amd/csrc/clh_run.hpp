// Per-GL-node run of the ln P cotangent (the accumulator of the Limber
// adjoint kernels) and the two places a run lands: its node's row of the
// run slab (ss / gs nodes) or a band of (k cell, z column) cells shared by a
// tile's gg nodes.  Host and device: the host test suite compiles this header
// with the system C++ compiler and checks it against direct sums.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CLH_RUN_HD __device__ __forceinline__
#define CLH_ATOMIC_ADD(p, v) atomicAdd((p), (v))
// workgroup scope for LDS: also keeps the compiler from merging it with a
// global add into one flat atomic
#define CLH_LDS_ADD(p, v) __hip_atomic_fetch_add((p), (v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
#else
#define CLH_RUN_HD inline
#define CLH_ATOMIC_ADD(p, v) (*(p) += (v))
#define CLH_LDS_ADD(p, v) (*(p) += (v))
#endif

// ----------------------------------------------------------------------------
// ln P cotangent of one GL node (a fixed z column j) over the rows of a tile,
// kept as a run along the k index: cell i receives the sum over rows with
// cell i of g (1 - dx) plus the sum over rows with cell i - 1 of g dx.  Each
// finished cell goes to the sink S as put(i, v, shared).
// Precondition: along the rows of a node, tile after tile, the cell index i
// never decreases.  It holds because the layout's ell rows are
// non-decreasing (checked on the host), log10 is monotone and the cell is
// the bisection index of log10 k.  Then a cell strictly inside a tile's index
// range is reached by that tile alone; the first two and the last two cells
// of the range can be shared with the neighbouring tiles (shared = true).
// ----------------------------------------------------------------------------
template <class S>
struct ClhLnPRunT {
  S out;
  int ci = -1, first = -1;
  double a0 = 0.0, a1 = 0.0;
  CLH_RUN_HD void put(int i, double v, bool edge) {
    if (v == 0.0) return;
    out.put(i, v, edge || i <= first + 1);
  }
  CLH_RUN_HD void add(int i, double dx, double g) {
    if (i != ci) {
      if (ci < 0) {
        first = i;
      } else {
        put(ci, a0, false);
        if (i == ci + 1) {
          a0 = a1;
        } else {
          put(ci + 1, a1, false);
          a0 = 0.0;
        }
        a1 = 0.0;
      }
      ci = i;
    }
    a0 += g * (1 - dx);
    a1 += g * dx;
  }
  CLH_RUN_HD void done() {
    if (ci < 0) return;
    put(ci, a0, true);
    put(ci + 1, a1, true);
  }
};

// The node's own row R of the run slab (zeroed before the launch; a gather
// folds the z weights in afterwards): plain stores for cells no other tile
// reaches, atomics for the shared ones.
struct ClhSlab {
  double* R = nullptr;
  CLH_RUN_HD void put(int i, double v, bool shared) {
    if (shared) CLH_ATOMIC_ADD(R + i, v);
    else R[i] = v;
  }
};
using ClhLnPRun = ClhLnPRunT<ClhSlab>;

// A node with z column j and weight dy, writing into the band of its block:
// cell i of column j takes v (1 - dy), of column j + 1 v dy.  Column c of
// the node (c = 0: j, 1: j + 1) holds cells lo[c] .. hi[c] at A[b[c] + i];
// a cell outside that range (or a column without room: lo > hi) goes to the
// draw's ln P cotangent G[i nz + column] with an atomic add.  A is zeroed
// before and added into G after the rows (every node of the block adds into
// it, hence atomics there too).
struct ClhBand {
  double* A = nullptr;
  double* G = nullptr;
  int nz = 0, j = 0;
  double dy = 0.0;
  int lo[2] = {1, 1}, hi[2] = {0, 0}, b[2] = {0, 0};
  CLH_RUN_HD void cell(int c, int i, double v) {
    if (v == 0.0) return;
    if (i >= lo[c] && i <= hi[c]) CLH_LDS_ADD(A + b[c] + i, v);
    else CLH_ATOMIC_ADD(G + (size_t)i * nz + j + c, v);
  }
  CLH_RUN_HD void put(int i, double v, bool) {
    cell(0, i, v * (1 - dy));
    cell(1, i, v * dy);
  }
};

// Column layout of a band: columns c = 0 .. n - 1 with cell ranges
// lo[c] .. hi[c] (empty when hi < lo) packed in order from offset 0; a
// column that would pass cap doubles, and every one after it, gets no room
// (start -1).  Returns the doubles used.  (The device kernels compute the
// same offsets with a wave-wide scan.)
CLH_RUN_HD int clh_band_layout(const int* lo, const int* hi, int n, int cap, int* start) {
  int used = 0;
  bool full = false;
  for (int c = 0; c < n; c++) {
    const int w = hi[c] >= lo[c] ? hi[c] - lo[c] + 1 : 0;
    if (full || used + w > cap) {
      full = true;
      start[c] = -1;
    } else {
      start[c] = used;
      used += w;
    }
  }
  return used;
}
