// Host check of the ln P run accumulator (csrc/clh_run.hpp).  Random node
// sequences of (cell i, weight dx, value g) over rows in non-decreasing
// cell order are cut into tiles; every tile is one run over the same slab
// row, as the adjoint kernels do it.  The slab must equal the direct sums
//   R[i] = sum_{rows in cell i} g (1 - dx) + sum_{rows in cell i-1} g dx
// whatever order the tiles run in (plain stores of one tile must never hit
// a cell that another tile adds to), including jumps of many cells and
// one-row tiles.
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
        run.R = R.data();
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
  std::printf("cases %ld bad %ld\n", cases, bad);
  return bad != 0;
}
