// Index searches over strictly increasing tables and the XCD-contiguous
// block order of the Limber launches.  Host and device: the host test suite
// compiles this header with the system C++ compiler.
#pragma once

#if defined(__HIPCC__)
#define CLH_HD __host__ __device__ __forceinline__
#else
#define CLH_HD inline
#endif

// ilo/ihi bisection used by every table lookup of cosmo3D.c (e.g. :46-57)
CLH_HD int clh_bs(const double* __restrict__ x, int n, double v) {
  int ilo = 0, ihi = n - 1;
  while (ihi > ilo + 1) {
    const int ll = (ihi + ilo) >> 1;
    if (x[ll] > v) ihi = ll; else ilo = ll;
  }
  return ilo;
}

// The same index as clh_bs for a strictly increasing table, found by walking
// from a previous index j (cheap when successive queries move monotonically).
//
// Galloping: one or two steps cost the same loads as a linear walk; a long
// move (the first rows of a low-ell tile move chi_1 far from chi_0) doubles
// its stride, then bisects the last bracket, so its dependent loads grow
// with log2 of the distance.  The result is the clh_bs index: the largest
// j <= n-2 with x[j] <= v, or 0.
CLH_HD int clh_walk(const double* __restrict__ x, int n, double v, int j) {
  if (j > 0 && x[j] > v) {  // down: x[hi] > v throughout
    int hi = j, lo = j - 1, step = 1;
    while (lo > 0 && x[lo] > v) {
      hi = lo;
      step <<= 1;
      lo = hi - step > 0 ? hi - step : 0;
    }
    if (x[lo] > v) return 0;
    while (hi - lo > 1) {
      const int m = (lo + hi) >> 1;
      if (x[m] <= v) lo = m; else hi = m;
    }
    return lo;
  }
  if (j < n - 2 && x[j + 1] <= v) {  // up: x[lo] <= v; hi = n-1 or x[hi] > v
    int lo = j + 1, hi = j + 2, step = 1;
    while (hi <= n - 2 && x[hi] <= v) {
      lo = hi;
      step <<= 1;
      hi = lo + step < n - 1 ? lo + step : n - 1;
    }
    while (hi - lo > 1) {
      const int m = (lo + hi) >> 1;
      if (x[m] <= v) lo = m; else hi = m;
    }
    return lo;
  }
  return j;
}

// Work item of block b in a launch of n blocks.  The hardware hands blocks to
// the 8 XCDs round robin (block b on XCD b % 8); this bijection gives each
// XCD a contiguous run of the work list, so the tiles of one draw (its P(k,z)
// table, node block and cotangent rows) share one XCD's L2.  Correctness does
// not depend on the placement, only the locality does.
CLH_HD int clh_xcd_item(int b, int n) {
  const int x = b & 7, q = b >> 3, per = n >> 3, rem = n & 7;
  return x * per + (x < rem ? x : rem) + q;
}
