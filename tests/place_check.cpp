// Host check of the data-vector placement (csrc/clh_place.hpp), built with
// AddressSanitizer by the test that runs it.
//  - Block starts: for every order of 3 and of 6 blocks with pairwise
//    distinct sizes, the starts tile the vector exactly (each entry covered
//    once) if and only if the order is an involution; the non-tiling orders
//    are listed by whether a block runs past the end of the vector (an
//    out-of-range write into a buffer of the vector's size) or only overlaps.
//  - Fourier placement: for every tiling order, ns = 1..10 source bins,
//    nl = 0..6 lens bins, 1, 2, 3 or 20 ells (some above lmax_shear) and
//    none or a random set of excluded (lens, source) pairs, the vector is
//    written into a buffer of exactly its size: every entry once, with the
//    value of a direct restatement (ss pairs enumerated z1 <= z2, included gs
//    pairs in lens-major order, calibration factors, the lmax_shear cut); a
//    row map that repeats or skips a row is refused before any write.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "clh_place.hpp"

static bool involution(const int* o, int M) {
  for (int i = 0; i < M; i++)
    if (o[o[i]] != i) return false;
  return true;
}

int main() {
  long bad = 0, cases = 0, overrun = 0, overlap = 0;
  // block starts, M = 3 and M = 6, distinct sizes
  for (int M : {3, 6}) {
    int ord[6];
    for (int i = 0; i < M; i++) ord[i] = i;
    const int szs[3][6] = {{7, 3, 11, 2, 5, 13}, {1, 40, 9, 4, 17, 6}, {300, 20, 1, 33, 8, 2}};
    do {
      for (const auto& sz : szs) {
        int start[6];
        const bool t = clh_blocks_tile(sz, ord, M, start);
        int n = 0;
        for (int i = 0; i < M; i++) n += sz[i];
        std::vector<int> cov(n, 0);
        bool past = false, ok = true;
        for (int i = 0; i < M; i++)
          for (int k = 0; k < sz[i]; k++) {
            const int e = start[i] + k;
            if (e < 0 || e >= n) past = true;
            else cov[e]++;
          }
        for (int e = 0; e < n; e++) ok = ok && cov[e] == 1;
        ok = ok && !past;
        cases++;
        if (t != ok || t != involution(ord, M)) bad++;
        if (!t) (past ? overrun : overlap)++;
      }
    } while (std::next_permutation(ord, ord + M));
  }
  // malformed orders are refused
  {
    const int sz[3] = {4, 5, 6};
    const int badord[5][3] = {{0, 0, 1}, {0, 1, 3}, {-1, 0, 1}, {2, 2, 2}, {1, 2, 0}};
    for (const auto& o : badord) {
      int start[3];
      cases++;
      if (clh_blocks_tile(sz, o, 3, start)) bad++;
    }
  }
  // Fourier placement into a buffer of exactly the vector's size
  const int ords[4][3] = {{0, 1, 2}, {1, 0, 2}, {2, 1, 0}, {0, 2, 1}};
  std::mt19937 rng(7);
  for (int ns = 1; ns <= 10; ns++)
    for (int nl = 0; nl <= 6; nl++)
      for (int ncl : {1, 2, 3, 20})
        for (const auto& o : ords)
          for (int ex = 0; ex < 2; ex++) {
            const int nss = ns * (ns + 1) / 2, ngs = ns * nl;
            // excluded pairs (ex = 1: each with probability 1/4), row map in lens-major order
            std::vector<int> row(ngs);
            int ngs_dv = 0;
            for (int q = 0; q < ngs; q++) row[q] = (ex && rng() % 4 == 0) ? -1 : ngs_dv++;
            const int sizes[3] = {ncl * nss, ncl * ngs_dv, ncl * nl};
            const int n = sizes[0] + sizes[1] + sizes[2];
            int start[3];
            cases++;
            if (!clh_blocks_tile(sizes, o, 3, start)) {
              bad++;
              continue;
            }
            std::vector<double> ell(ncl), m(ns), cl((size_t)(nss + ngs + nl) * ncl);
            for (int i = 0; i < ncl; i++) ell[i] = 30.0 * std::pow(1.4, i);
            const double lmax_shear = ncl > 1 ? ell[ncl / 2] : 1e6;
            for (int b = 0; b < ns; b++) m[b] = 0.01 * (b + 1);
            for (size_t k = 0; k < cl.size(); k++) cl[k] = 1.0 + (double)k;
            std::vector<double> out(n, NAN);  // exactly n: ASan flags any write past it
            const int w = clh_place_fourier(ns, nl, ncl, ell.data(), lmax_shear, m.data(), ex ? row.data() : nullptr,
                                            start, cl.data(), out.data(), n);
            if (w != n) {
              bad++;
              continue;
            }
            // direct restatement
            int q = 0;
            for (int z1 = 0; z1 < ns; z1++)
              for (int z2 = z1; z2 < ns; z2++, q++)
                for (int i = 0; i < ncl; i++) {
                  const double ref =
                      ell[i] < lmax_shear ? cl[(size_t)q * ncl + i] * (1 + m[z1]) * (1 + m[z2]) : 0.0;
                  if (!(out[start[0] + ncl * q + i] == ref)) bad++;
                }
            for (int zl = 0, r = 0; zl < nl; zl++)
              for (int zs = 0; zs < ns; zs++) {
                const int col = zl * ns + zs;
                if (row[col] < 0) continue;
                for (int i = 0; i < ncl; i++)
                  if (!(out[start[1] + ncl * r + i] == cl[(size_t)(nss + col) * ncl + i] * (1 + m[zs]))) bad++;
                r++;
              }
            for (int zl = 0; zl < nl; zl++)
              for (int i = 0; i < ncl; i++)
                if (!(out[start[2] + ncl * zl + i] == cl[(size_t)(nss + ngs + zl) * ncl + i])) bad++;
            for (int e = 0; e < n; e++)
              if (std::isnan(out[e])) bad++;
          }
  // a block that would leave the buffer is refused before any write
  {
    const int start[3] = {0, 10, 3};
    std::vector<double> ell{100.0}, m{0.0, 0.0}, cl(3 + 2 + 1, 1.0), out(6, -1.0);
    cases++;
    if (clh_place_fourier(2, 1, 1, ell.data(), 1e6, m.data(), nullptr, start, cl.data(), out.data(), 6) != -1) bad++;
    for (double v : out)
      if (v != -1.0) bad++;
  }
  // row maps that repeat or skip a row
  {
    const int start[3] = {0, 3, 5};
    std::vector<double> ell{100.0}, m{0.0, 0.0}, cl(3 + 2 + 1, 1.0), out(6, -1.0);
    const int dup[2] = {0, 0}, skip[2] = {1, -1}, big[2] = {0, 2};
    for (const int* rm : {dup, skip, big}) {
      cases++;
      if (clh_place_fourier(2, 1, 1, ell.data(), 1e6, m.data(), rm, start, cl.data(), out.data(), 6) != -1) bad++;
    }
    for (double v : out)
      if (v != -1.0) bad++;
  }
  std::printf("place: cases %ld non-tiling orders: past the end %ld, overlapping %ld bad %ld\n", cases, overrun,
              overlap, bad);
  return bad ? 1 : 0;
}
