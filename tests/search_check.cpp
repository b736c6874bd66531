// Host check of the table searches in clh_search.hpp: for random strictly
// increasing tables, every start index and queries below, inside, on and
// above the knots, clh_walk must return the clh_bs index, and clh_bs the
// largest j <= n-2 with x[j] <= v (0 when v < x[0]).  clh_xcd_item must be a
// bijection of 0..n-1 whose image on each XCD is one contiguous run.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "clh_search.hpp"

static int brute(const std::vector<double>& x, double v) {
  int j = 0;
  for (int i = 0; i + 1 < (int)x.size(); i++)
    if (x[i] <= v) j = i;
  return j;
}

int main() {
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  long checks = 0, bad = 0;
  for (int n : {2, 3, 4, 5, 9, 17, 64, 401}) {
    std::vector<double> x(n);
    double acc = -1.0;
    for (int i = 0; i < n; i++) x[i] = (acc += 0.01 + u(rng));
    std::vector<double> q;
    for (int i = 0; i < n; i++) {
      q.push_back(x[i]);
      q.push_back(std::nextafter(x[i], -1e300));
      q.push_back(std::nextafter(x[i], 1e300));
    }
    for (int t = 0; t < 200; t++) q.push_back(x[0] - 2.0 + (x[n - 1] - x[0] + 4.0) * u(rng));
    q.push_back(-1e300);
    q.push_back(1e300);
    for (double v : q) {
      const int want = brute(x, v);
      if (clh_bs(x.data(), n, v) != want) bad++;
      for (int j = 0; j <= n - 2; j++) {
        if (clh_walk(x.data(), n, v, j) != want) bad++;
        checks++;
      }
    }
  }
  for (int n = 1; n < 300; n++) {
    std::vector<int> seen(n, 0), xcd_of(n, -1);
    for (int b = 0; b < n; b++) {
      const int it = clh_xcd_item(b, n);
      if (it < 0 || it >= n || seen[it]++) { bad++; continue; }
      xcd_of[it] = b & 7;
    }
    for (int it = 1; it < n; it++)  // runs are contiguous and in XCD order
      if (xcd_of[it] < xcd_of[it - 1]) bad++;
  }
  std::printf("checks %ld bad %ld\n", checks, bad);
  return bad != 0;
}
