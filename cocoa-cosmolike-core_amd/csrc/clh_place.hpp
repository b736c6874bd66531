// Data-vector block placement: block starts of an order, the check that they
// tile the vector, and the Fourier-space placement of C_ell rows with shear
// calibration and the lmax_shear cut.  Plain C++ (no HIP), so the host checks
// build it with a host compiler and AddressSanitizer.
#pragma once

#include <utility>
#include <vector>

// compute_data_vector_Mx2pt_N_starts (generic_interface.hpp:770-791) for M
// blocks: indices = stable ascending sort of ord, start(i) = sum of
// sizes(indices(j)) over j < indices(i)
inline void clh_block_starts(const int* sizes, const int* ord, int M, int* start) {
  int idx[8];
  for (int i = 0; i < M; i++) idx[i] = i;
  for (int i = 1; i < M; i++)
    for (int j = i; j > 0 && ord[idx[j - 1]] > ord[idx[j]]; j--) std::swap(idx[j], idx[j - 1]);
  for (int i = 0; i < M; i++) {
    start[i] = 0;
    for (int j = 0; j < idx[i]; j++) start[i] += sizes[idx[j]];
  }
}

// ord is a permutation of 0..M-1 whose block starts (above) put every block
// inside the vector with no two blocks overlapping.  The formula tiles the
// vector only for involutive orders ((0,1,2), (1,0,2), (2,1,0), (0,2,1) for
// M = 3); a 3-cycle such as (1,2,0) makes blocks overlap or run past the end.
// start receives the starts (also when the answer is false).
inline bool clh_blocks_tile(const int* sizes, const int* ord, int M, int* start) {
  if (!ord || M < 1 || M > 8) return false;
  for (int i = 0; i < M; i++) {
    if (ord[i] < 0 || ord[i] >= M) return false;
    for (int j = 0; j < i; j++)
      if (ord[i] == ord[j]) return false;
  }
  clh_block_starts(sizes, ord, M, start);
  long long n = 0;
  for (int i = 0; i < M; i++) n += sizes[i];
  for (int i = 0; i < M; i++) {
    if (start[i] < 0 || (long long)start[i] + sizes[i] > n) return false;
    for (int j = 0; j < i; j++)
      if (sizes[i] > 0 && sizes[j] > 0 && start[i] < start[j] + sizes[j] && start[j] < start[i] + sizes[i])
        return false;
  }
  return true;
}

// The Fourier-space 3x2pt vector (compute_Mx2pt_N_masked<1,3>,
// generic_interface.cpp): blocks ss (EE, zero at ell >= lmax_shear), gs, gg
// of ncl ells each, ss pairs in N_shear order (z1 <= z2), gs columns
// zl * ns + zs at data-vector row gs_row[zl * ns + zs] (-1: an excluded pair,
// not placed; null: every pair, row = column) with calibration (1 + m_zs),
// ss with (1 + m_z1)(1 + m_z2).  cl: [n_ss + ns nl + nl][ncl] rows.  Every
// write is checked against n (returns -1 and writes nothing if a block would
// leave out[0, n)); returns the number of entries written.
inline int clh_place_fourier(int ns, int nl, int ncl, const double* ell, double lmax_shear, const double* m,
                             const int* gs_row, const int start[3], const double* cl, double* out, int n) {
  const int nss = ns * (ns + 1) / 2, ngs = ns * nl;
  // the included pairs' rows are 0 .. ngs_dv - 1, each once
  int ngs_dv = 0;
  std::vector<char> seen(ngs > 0 ? ngs : 1, 0);
  for (int q = 0; q < ngs; q++) {
    const int r = gs_row ? gs_row[q] : q;
    if (r < -1 || r >= ngs) return -1;
    if (r >= 0) {
      if (seen[r]) return -1;
      seen[r] = 1;
      ngs_dv++;
    }
  }
  for (int r = 0; r < ngs_dv; r++)
    if (!seen[r]) return -1;
  const int rows[3] = {nss, ngs_dv, nl};
  for (int k = 0; k < 3; k++)
    if (start[k] < 0 || (long long)start[k] + (long long)rows[k] * ncl > n) return -1;
  int w = 0;
  for (int q = 0; q < nss; q++) {
    int z1 = 0, z2 = q;
    while (z2 >= ns - z1) {
      z2 -= ns - z1;
      z1++;
    }
    z2 += z1;
    for (int i = 0; i < ncl; i++, w++)
      out[start[0] + ncl * q + i] =
          ell[i] < lmax_shear ? cl[(long long)q * ncl + i] * (1.0 + m[z1]) * (1.0 + m[z2]) : 0.0;
  }
  for (int q = 0; q < ngs; q++) {
    const int r = gs_row ? gs_row[q] : q;
    if (r < 0) continue;
    for (int i = 0; i < ncl; i++, w++)
      out[start[1] + ncl * r + i] = cl[(long long)(nss + q) * ncl + i] * (1.0 + m[q % ns]);
  }
  for (int q = 0; q < nl; q++)
    for (int i = 0; i < ncl; i++, w++) out[start[2] + ncl * q + i] = cl[(long long)(nss + ngs + q) * ncl + i];
  return w;
}
