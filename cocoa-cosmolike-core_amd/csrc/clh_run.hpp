// Per-GL-node run of the ln P cotangent (the accumulator of the Limber
// adjoint kernels).  Host and device: the host test suite compiles this
// header with the system C++ compiler and checks it against direct sums.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CLH_RUN_HD __device__ __forceinline__
#define CLH_ATOMIC_ADD(p, v) atomicAdd((p), (v))
#else
#define CLH_RUN_HD inline
#define CLH_ATOMIC_ADD(p, v) (*(p) += (v))
#endif

// ----------------------------------------------------------------------------
// ln P cotangent of one GL node (a fixed z column j) over the rows of a tile,
// kept as a run along the k index: R[i] = sum over rows with cell i of
// g (1 - dx) plus sum over rows with cell i - 1 of g dx, where R is the node's
// own row of the run slab (zeroed before the launch); k_lnp_gather folds the
// z weights in.
// Precondition: along the rows of a node, tile after tile, the cell index i
// never decreases.  It holds because the layout's ell rows are
// non-decreasing (checked on the host), log10 is monotone and the cell is
// the bisection index of log10 k.  Then a cell strictly inside a tile's index
// range is reached by that tile alone and takes a plain store; the first two
// and the last two cells of the range can be shared with the neighbouring
// tiles and take atomics.
// ----------------------------------------------------------------------------
struct ClhLnPRun {
  double* R;
  int ci = -1, first = -1;
  double a0 = 0.0, a1 = 0.0;
  CLH_RUN_HD void put(int i, double v, bool edge) {
    if (v == 0.0) return;
    if (edge || i <= first + 1) CLH_ATOMIC_ADD(R + i, v);
    else R[i] = v;
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
