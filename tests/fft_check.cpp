// Host check of the mixed-radix Stockham FFT and the c2r packing
// (csrc/clh_fft.hpp): complex inverse transforms of lengths with factors
// 2, 3, 5 (incl. the 450 and 750 of the non-Limber transforms) against
// the direct sum, and c2r transforms of length 2M of random Hermitian
// spectra (imaginary parts of X_0 and X_M dropped, as a c2r transform
// ignores them) against x[n] = X_0 + X_M (-1)^n + 2 sum_m Re(X_m e^{2 pi i m n / N}),
// and r2c transforms of random real x against the direct sum.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "clh_fft.hpp"

int main() {
  std::mt19937_64 rng(5);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  long bad = 0, cases = 0;
  double worst = 0.0;
  for (int M : {1, 2, 3, 5, 4, 6, 10, 15, 30, 60, 450, 750, 375, 1000}) {
    if (!clh_fft_ok(M)) { bad++; continue; }
    std::vector<int> rad;
    for (int m = M; m > 1; m /= clh_fft_radix(m)) rad.push_back(clh_fft_radix(m));
    auto tw = [M](int t) {
      const long double a = 2.0L * 3.14159265358979323846264338327950288L * t / M;
      return clh_c2{(double)std::cos(a), (double)std::sin(a)};
    };
    for (int trial = 0; trial < 3; trial++) {
      std::vector<clh_c2> a(M), b(M);
      for (auto& v : a) v = clh_c2{u(rng), u(rng)};
      const std::vector<clh_c2> in = a;
      int Ns = 1;
      for (int R : rad) {
        for (int j = 0; j < M / R; j++) clh_fft_bfly(a.data(), b.data(), clh_fft_stage(M, R, Ns), j, tw);
        Ns *= R;
        a.swap(b);
      }
      double scale = 0.0, err = 0.0;
      for (int n = 0; n < M; n++) {
        long double re = 0, im = 0;
        for (int k = 0; k < M; k++) {
          const long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)((long)k * n % M) / M;
          re += in[k].re * cosl(ang) - in[k].im * sinl(ang);
          im += in[k].re * sinl(ang) + in[k].im * cosl(ang);
        }
        scale = std::fmax(scale, std::hypot((double)re, (double)im));
        err = std::fmax(err, std::hypot(a[n].re - (double)re, a[n].im - (double)im));
      }
      cases++;
      worst = std::fmax(worst, err / scale);
      if (err > 1e-13 * scale) bad++;
    }
    // c2r of length N = 2M
    const int N = 2 * M;
    std::vector<clh_c2> X(M + 1);
    for (auto& v : X) v = clh_c2{u(rng), u(rng)};
    std::vector<clh_c2> a(M), b(M);
    auto twN = [N](int t) {
      const long double ang = 2.0L * 3.14159265358979323846264338327950288L * t / N;
      return clh_c2{(double)std::cos(ang), (double)std::sin(ang)};
    };
    clh_c2 X0 = X[0], XM = X[M];
    X0.im = 0.0;
    XM.im = 0.0;
    for (int k = 0; k < M; k++) {
      const clh_c2 xk = k == 0 ? X0 : X[k];
      const clh_c2 xmk = k == 0 ? XM : X[M - k];
      a[k] = clh_c2r_pack(xk, xmk, twN(k));
    }
    int Ns = 1;
    for (int R : rad) {
      for (int j = 0; j < M / R; j++) clh_fft_bfly(a.data(), b.data(), clh_fft_stage(M, R, Ns), j, tw);
      Ns *= R;
      a.swap(b);
    }
    double scale = 0.0, err = 0.0;
    for (int n = 0; n < N; n++) {
      long double s = X0.re + XM.re * ((n % 2) ? -1.0L : 1.0L);
      for (int m = 1; m < M; m++) {
        const long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)((long)m * n % N) / N;
        s += 2.0L * (X[m].re * cosl(ang) - X[m].im * sinl(ang));
      }
      const double got = (n % 2) ? a[n / 2].im : a[n / 2].re;
      scale = std::fmax(scale, std::fabs((double)s));
      err = std::fmax(err, std::fabs(got - (double)s));
    }
    cases++;
    worst = std::fmax(worst, err / scale);
    if (err > 1e-13 * scale) bad++;
  }
  // r2c: real x of length N = 2M through one length-M transform and the unpacking
  for (int M : {1, 3, 10, 450, 750}) {
    const int N = 2 * M;
    std::vector<int> rad;
    for (int m = M; m > 1; m /= clh_fft_radix(m)) rad.push_back(clh_fft_radix(m));
    auto tw = [M](int t) {
      const long double ang = 2.0L * 3.14159265358979323846264338327950288L * t / M;
      return clh_c2{(double)std::cos(ang), (double)std::sin(ang)};
    };
    std::vector<double> xr(N);
    for (auto& v : xr) v = u(rng);
    std::vector<clh_c2> a(M), b(M);
    for (int n = 0; n < M; n++) a[n] = clh_c2{xr[2 * n], xr[2 * n + 1]};
    int Ns = 1;
    for (int R : rad) {
      for (int j = 0; j < M / R; j++) clh_fft_bfly(a.data(), b.data(), clh_fft_stage(M, R, Ns), j, tw);
      Ns *= R;
      a.swap(b);
    }
    double scale = 0.0, err = 0.0;
    for (int k = 0; k <= M; k++) {
      const long double ang = 2.0L * 3.14159265358979323846264338327950288L * k / N;
      const clh_c2 e{(double)cosl(ang), (double)sinl(ang)};
      const clh_c2 y = clh_r2c_unpack(a[k % M], a[(M - k) % M], e);
      long double re = 0, im = 0;
      for (int n = 0; n < N; n++) {
        const long double an = 2.0L * 3.14159265358979323846264338327950288L * (long double)((long)k * n % N) / N;
        re += xr[n] * cosl(an);
        im += xr[n] * sinl(an);
      }
      scale = std::fmax(scale, std::hypot((double)re, (double)im));
      err = std::fmax(err, std::hypot(y.re - (double)re, y.im - (double)im));
    }
    cases++;
    worst = std::fmax(worst, err / scale);
    if (err > 1e-13 * scale) bad++;
  }
  cases++;
  if (clh_fft_ok(11 * 3) || clh_fft_ok(7 * 10)) bad++;  // 7 and 11 are not supported factors
  std::printf("fft: cases %ld worst relative error %.3e bad %ld\n", cases, worst, bad);
  return bad ? 1 : 0;
}
