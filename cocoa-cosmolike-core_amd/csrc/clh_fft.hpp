// Mixed-radix (2, 3, 5) Stockham FFT pieces for the non-Limber C_ell: the
// inverse real transforms of FFTLog (cfftlog_ells_p2, cosmolike/cosmo2D.c:
// 4492-4524, FFTW c2r there).  A c2r transform of even length N = 2M runs as
// one complex transform of length M (the even / odd outputs as real /
// imaginary parts).  Plain C++ with __host__ __device__ under hipcc, so the
// host checks build it with a host compiler.
#pragma once

#ifdef __HIPCC__
#define CLH_FFT_HD __host__ __device__
#else
#define CLH_FFT_HD
#endif

struct clh_c2 {
  double re, im;
};

CLH_FFT_HD inline clh_c2 clh_c2_mul(clh_c2 a, clh_c2 b) {
  return clh_c2{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// The stage radices of a length-M transform: 5s, then 3s, 2s (the next
// radix of the still unfactored part m; 0 when m has another factor)
CLH_FFT_HD inline int clh_fft_radix(int m) {
  return m % 5 == 0 ? 5 : m % 3 == 0 ? 3 : m % 2 == 0 ? 2 : 0;
}
CLH_FFT_HD inline bool clh_fft_ok(int M) {
  int m = M;
  while (m > 1) {
    const int R = clh_fft_radix(m);
    if (!R) return false;
    m /= R;
  }
  return M >= 1;
}

// v[k] <- sum_r v[r] e^{+2 pi i r k / R} for R = 2, 3, 5 (a template
// argument, so that v stays in registers)
template <int R>
CLH_FFT_HD inline void clh_dft_small(clh_c2* v) {
  if (R == 2) {
    const clh_c2 a = v[0], b = v[1];
    v[0] = clh_c2{a.re + b.re, a.im + b.im};
    v[1] = clh_c2{a.re - b.re, a.im - b.im};
  } else if (R == 3) {
    const double s = 0.86602540378443864676;  // sin(2 pi / 3)
    const clh_c2 t1{v[1].re + v[2].re, v[1].im + v[2].im}, t2{v[1].re - v[2].re, v[1].im - v[2].im};
    const clh_c2 m{v[0].re - 0.5 * t1.re, v[0].im - 0.5 * t1.im};
    v[0] = clh_c2{v[0].re + t1.re, v[0].im + t1.im};
    v[1] = clh_c2{m.re - s * t2.im, m.im + s * t2.re};
    v[2] = clh_c2{m.re + s * t2.im, m.im - s * t2.re};
  } else {  // R == 5
    const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;  // cos(2 pi / 5), cos(4 pi / 5)
    const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;   // sin(2 pi / 5), sin(4 pi / 5)
    const clh_c2 a1{v[1].re + v[4].re, v[1].im + v[4].im}, b1{v[1].re - v[4].re, v[1].im - v[4].im};
    const clh_c2 a2{v[2].re + v[3].re, v[2].im + v[3].im}, b2{v[2].re - v[3].re, v[2].im - v[3].im};
    const clh_c2 v0 = v[0];
    const clh_c2 m1{v0.re + c1 * a1.re + c2 * a2.re, v0.im + c1 * a1.im + c2 * a2.im};
    const clh_c2 m2{v0.re + c2 * a1.re + c1 * a2.re, v0.im + c2 * a1.im + c1 * a2.im};
    const clh_c2 n1{s1 * b1.re + s2 * b2.re, s1 * b1.im + s2 * b2.im};  // times i below
    const clh_c2 n2{s2 * b1.re - s1 * b2.re, s2 * b1.im - s1 * b2.im};
    v[0] = clh_c2{v0.re + a1.re + a2.re, v0.im + a1.im + a2.im};
    v[1] = clh_c2{m1.re - n1.im, m1.im + n1.re};
    v[4] = clh_c2{m1.re + n1.im, m1.im - n1.re};
    v[2] = clh_c2{m2.re - n2.im, m2.im + n2.re};
    v[3] = clh_c2{m2.re + n2.im, m2.im - n2.re};
  }
}

// One Stockham stage: radix R after stages whose radices multiply to Ns, its
// uniform constants computed once (no integer division per butterfly)
struct ClhFftStage {
  int M, R, Ns, stride, tsc;  // stride = M / R, tsc = M / (Ns R)
  float rNs;                  // 1 / Ns
};
CLH_FFT_HD inline ClhFftStage clh_fft_stage(int M, int R, int Ns) {
  return ClhFftStage{M, R, Ns, M / R, M / (Ns * R), 1.0f / (float)Ns};
}

// Butterfly j (0 <= j < M / R) of a stage: inputs in[j + r M / R] times
// e^{+2 pi i r (j mod Ns) / (Ns R)}, an R-point DFT, outputs to
// out[(j / Ns) Ns R + (j mod Ns) + r Ns].  After every stage, out holds the
// inverse DFT sum_k in[k] e^{+2 pi i k n / M} in natural order.
template <int R, class TW>
CLH_FFT_HD inline void clh_fft_bfly_r(const clh_c2* in, clh_c2* out, const ClhFftStage& st, int j, const TW& tw) {
  clh_c2 v[R];
  int jq = (int)((float)j * st.rNs);  // j / Ns, exact after one correction either way (j < 2^20)
  if (jq * st.Ns > j) jq--;
  else if ((jq + 1) * st.Ns <= j) jq++;
  const int jm = j - jq * st.Ns, tstep = st.tsc * jm;  // tstep r < M: no reduction
#pragma unroll
  for (int r = 0; r < R; r++) {
    clh_c2 a = in[j + r * st.stride];
    if (r > 0 && jm > 0) a = clh_c2_mul(a, tw(tstep * r));
    v[r] = a;
  }
  clh_dft_small<R>(v);
  const int o = jq * st.Ns * R + jm;
#pragma unroll
  for (int r = 0; r < R; r++) out[o + r * st.Ns] = v[r];
}
template <class TW>
CLH_FFT_HD inline void clh_fft_bfly(const clh_c2* in, clh_c2* out, const ClhFftStage& st, int j, const TW& tw) {
  switch (st.R) {
    case 2: clh_fft_bfly_r<2>(in, out, st, j, tw); break;
    case 3: clh_fft_bfly_r<3>(in, out, st, j, tw); break;
    default: clh_fft_bfly_r<5>(in, out, st, j, tw); break;
  }
}

// Input k (0 <= k < M) of the length-M complex transform whose output z_n
// holds the c2r inverse transform of length N = 2M: z_n = x[2n] + i x[2n+1],
// x[n'] = sum_{m=0}^{N-1} X_m e^{2 pi i m n' / N} for a Hermitian X given by
// X_0 .. X_M.  e = e^{+2 pi i k / N}.
CLH_FFT_HD inline clh_c2 clh_c2r_pack(clh_c2 Xk, clh_c2 XMk, clh_c2 e) {
  const clh_c2 s{Xk.re + XMk.re, Xk.im - XMk.im};  // X_k + conj(X_{M-k})
  const clh_c2 d{Xk.re - XMk.re, Xk.im + XMk.im};  // X_k - conj(X_{M-k})
  const clh_c2 ed = clh_c2_mul(e, d);
  return clh_c2{s.re - ed.im, s.im + ed.re};        // s + i e d
}

// The r2c counterpart: Z = sum_n z_n e^{+2 pi i k n / M} of z_n = x[2n] +
// i x[2n+1] (real x of length N = 2M) gives Y_k = sum_{n'} x[n'] e^{+2 pi i
// k n' / N} for 0 <= k <= M as E_k + e O_k, E_k = (Z_k + conj(Z_{M-k})) / 2,
// O_k = (Z_k - conj(Z_{M-k})) / (2 i), e = e^{+2 pi i k / N} (Z_M = Z_0).
CLH_FFT_HD inline clh_c2 clh_r2c_unpack(clh_c2 Zk, clh_c2 ZMk, clh_c2 e) {
  const clh_c2 E{0.5 * (Zk.re + ZMk.re), 0.5 * (Zk.im - ZMk.im)};
  const clh_c2 O{0.5 * (Zk.im + ZMk.im), -0.5 * (Zk.re - ZMk.re)};  // (Z_k - conj(Z_{M-k})) / (2 i)
  const clh_c2 eo = clh_c2_mul(e, O);
  return clh_c2{E.re + eo.re, E.im + eo.im};
}
