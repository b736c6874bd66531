// Host check of the ln P run accumulator (csrc/clh_run.hpp).
// Slab: random node sequences of (cell i, weight dx, value g) over rows in
// non-decreasing cell order are cut into tiles; every tile is one run over
// the same slab row, as the ss / gs adjoint does it.  The slab must equal the
// direct sums
//   R[i] = sum_{rows in cell i} g (1 - dx) + sum_{rows in cell i-1} g dx
// whatever order the tiles run in (plain stores of one tile must never hit
// a cell that another tile adds to), including jumps of many cells and
// one-row tiles.
// Band: blocks of nodes with z columns j (and j + 1, weights 1 - dy, dy) in
// a few neighbouring columns, tile by tile as the gg adjoint does it: the
// band's column ranges from each node's first and last cell of the tile
// (clh_band_layout), the nodes' runs into the band, the band added into the
// ln P cotangent G[i][j].  G must equal the direct bilinear sums, also with
// a band too small for every column (cap) and with reported ranges one cell
// short (cells outside go to G directly).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "clh_run.hpp"

int main() {
  std::mt19937_64 rng(11);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  long bad = 0, cases = 0;
  const int nk = 400;
  for (int trial = 0; trial < 4000; trial++) {
    const int nrow = 1 + (int)(u(rng) * 120);
    std::vector<int> ci(nrow);
    std::vector<double> dx(nrow), g(nrow);
    int c = (int)(u(rng) * 60);
    for (int r = 0; r < nrow; r++) {
      const double t = u(rng);
      c += t < 0.4 ? 0 : (t < 0.85 ? 1 : (int)(u(rng) * 20));  // stay, step, jump
      c = std::min(c, nk - 2);
      ci[r] = c;
      dx[r] = u(rng);
      g[r] = u(rng) < 0.1 ? 0.0 : u(rng) - 0.3;
    }
    std::vector<double> ref(nk, 0.0);
    for (int r = 0; r < nrow; r++) {
      ref[ci[r]] += g[r] * (1 - dx[r]);
      ref[ci[r] + 1] += g[r] * dx[r];
    }
    // tiles
    std::vector<int> cut = {0};
    for (int r = 1; r < nrow; r++)
      if (u(rng) < (trial % 4 == 0 ? 0.6 : 0.15)) cut.push_back(r);
    cut.push_back(nrow);
    for (int order = 0; order < 2; order++) {
      std::vector<double> R(nk, 0.0);
      const int nt = (int)cut.size() - 1;
      for (int tt = 0; tt < nt; tt++) {
        const int t = order ? nt - 1 - tt : tt;
        ClhLnPRun run;
        run.out.R = R.data();
        for (int r = cut[t]; r < cut[t + 1]; r++) run.add(ci[r], dx[r], g[r]);
        run.done();
      }
      cases++;
      for (int i = 0; i < nk; i++)
        if (std::fabs(R[i] - ref[i]) > 1e-12 * (1 + std::fabs(ref[i]))) {
          bad++;
          break;
        }
    }
  }
  // band
  const int nz = 30;
  long bcases = 0, bbad = 0, spill = 0;
  for (int trial = 0; trial < 3000; trial++) {
    const int nn = 1 + (int)(u(rng) * 40), nrow = 1 + (int)(u(rng) * 90);
    const int jbase = (int)(u(rng) * 20);
    std::vector<int> nj(nn);
    std::vector<double> ndy(nn);
    std::vector<std::vector<int>> ci(nn, std::vector<int>(nrow));
    std::vector<std::vector<double>> dx(nn, std::vector<double>(nrow)), g(nn, std::vector<double>(nrow));
    std::vector<double> ref((size_t)nk * nz, 0.0);
    for (int q = 0; q < nn; q++) {
      nj[q] = jbase + (int)(u(rng) * 8);
      ndy[q] = u(rng) < 0.1 ? 0.0 : u(rng);
      int c = (int)(u(rng) * 100);
      for (int r = 0; r < nrow; r++) {
        const double t = u(rng);
        c = std::min(c + (t < 0.5 ? 0 : (t < 0.9 ? 1 : (int)(u(rng) * 15))), nk - 2);
        ci[q][r] = c;
        dx[q][r] = u(rng);
        g[q][r] = u(rng) < 0.1 ? 0.0 : u(rng) - 0.3;
        const double a = g[q][r] * (1 - dx[q][r]), b = g[q][r] * dx[q][r];
        ref[(size_t)c * nz + nj[q]] += a * (1 - ndy[q]);
        ref[(size_t)(c + 1) * nz + nj[q]] += b * (1 - ndy[q]);
        ref[(size_t)c * nz + nj[q] + 1] += a * ndy[q];
        ref[(size_t)(c + 1) * nz + nj[q] + 1] += b * ndy[q];
      }
    }
    std::vector<int> cut = {0};
    for (int r = 1; r < nrow; r++)
      if (u(rng) < 0.1) cut.push_back(r);
    cut.push_back(nrow);
    const int cap = trial % 3 == 0 ? 1 + (int)(u(rng) * 60) : 1024;
    const bool short_range = trial % 5 == 0;
    std::vector<double> G((size_t)nk * nz, 0.0);
    for (int t = 0; t + 1 < (int)cut.size(); t++) {
      const int r0 = cut[t], r1 = cut[t + 1] - 1;
      int jlo = nz;
      for (int q = 0; q < nn; q++) jlo = std::min(jlo, nj[q]);
      const int ncol = 10;
      std::vector<int> lo(ncol, 1 << 30), hi(ncol, -1), start(ncol);
      for (int q = 0; q < nn; q++)
        for (int c = nj[q] - jlo; c <= nj[q] - jlo + 1; c++) {
          lo[c] = std::min(lo[c], ci[q][r0] + (short_range ? 1 : 0));
          hi[c] = std::max(hi[c], ci[q][r1] + 1);
        }
      const int used = clh_band_layout(lo.data(), hi.data(), ncol, cap, start.data());
      std::vector<double> A(std::max(used, 1), 0.0);
      for (int q = 0; q < nn; q++) {
        ClhLnPRunT<ClhBand> run;
        run.out.A = A.data();
        run.out.G = G.data();
        run.out.nz = nz;
        run.out.j = nj[q];
        run.out.dy = ndy[q];
        for (int c = 0; c < 2; c++) {
          const int cc = nj[q] - jlo + c;
          if (start[cc] >= 0) {
            run.out.lo[c] = lo[cc];
            run.out.hi[c] = hi[cc];
            run.out.b[c] = start[cc] - lo[cc];
          } else {
            spill++;
          }
        }
        for (int r = r0; r <= r1; r++) run.add(ci[q][r], dx[q][r], g[q][r]);
        run.done();
      }
      for (int c = 0; c < ncol; c++)
        if (start[c] >= 0)
          for (int i = lo[c]; i <= hi[c]; i++) G[(size_t)i * nz + jlo + c] += A[start[c] + i - lo[c]];
    }
    bcases++;
    for (size_t k = 0; k < G.size(); k++)
      if (std::fabs(G[k] - ref[k]) > 1e-12 * (1 + std::fabs(ref[k]))) {
        bbad++;
        break;
      }
  }
  std::printf("cases %ld bad %ld band cases %ld bad %ld (columns without room %ld)\n", cases, bad, bcases, bbad,
              spill);
  return bad != 0 || bbad != 0 || spill == 0;
}
