// Non-Limber clustering C_ell below LMAX_NOLIMBER = 150 and the w(theta) it
// feeds: C_cl_tomo (cosmolike/cosmo2D.c:4528-4769) with its FFTLog passes
// cfftlog_ells_p1 / p2 (cosmo2D.c:4047-4526), restated for the GPU.
//
// Per draw and lens bin i the reference transforms three functions of
// x = chi (Mpc) on NL_Nchi = 500 log-spaced nodes (density b1 n D H/H0 chi,
// RSD -f n D H/H0 chi, magnification W_mag D / chi / (c/H0)^2), each by one
// real FFT of length N_j, then for every ell one complex multiply by
// g_l(nu + i eta_m) e^{-i eta_m ln base_l} and one inverse real FFT, of which
// the 500 outputs at the original nodes are kept.  Here:
//   k_nl_prep  FFTLog inputs f(x) / x^nu                     (thread per node)
//   k_nl_fwd   the forward DFT T_ij[m], windowed             (thread per m)
//   k_nl_gl    G_j[l][m] = g_l e^{-i eta_m ln base_l}         (thread per m)
//   k_nl_idft  the inverse DFT restricted to the kept outputs as an FP64 MFMA
//              product: Out[l][q] = sum_m w_m Re(conj(T G) e^{2 pi i m n / N}),
//              fused with the C_ell integrand F^2 k^3 P_lin(k) and its sum
//              over a 16-node slice of q; first for the ten block-end ells
//              (15, 31, ..., 143, 149) only, then for the 16-ell tiles below
//              each bin's LMAX
//   k_nl_lmax  the reference's block-convergence switch to Limber: LMAX is the
//              end of the first 16-ell block whose last C_ell is within 1% of
//              the Limber C_ell (cosmo2D.c:4740-4750)
//   k_nl_cl    C_ell = 2/pi dln k sum + C_Limber(P_nl) - C_Limber(P_lin) and
//              the change of w(theta) against the Limber C_ell the folded
//              Legendre kernel already summed:
//              dw = sum_l Pw[theta][l] (C_used - C_Limber)[l]
// As in the reference, no ell at or above a bin's LMAX is transformed.
#include "clh_device.hpp"

typedef double clh_v4d __attribute__((ext_vector_type(4)));

// FFTLog configurations of C_cl_tomo (cosmo2D.c:4566-4580): nu, derivative
__constant__ double c_nl_nu[3] = {1.0, 1.01, 1.0};
__constant__ int c_nl_deriv[3] = {0, 2, 0};

// Lanczos approximation, g = 7, 9 terms (the coefficient table
// cfftlog_ells_p2 uses, cosmo2D.c:4160-4168; Godfrey's published set)
__constant__ double c_lanczos[9] = {0.99999999999980993227684700473478, 676.520368121885098567009190444019,
                                    -1259.13921672240287047156078755283, 771.3234287776530788486528258894,
                                    -176.61502916214059906584551354, 12.507343278686904814458936853,
                                    -0.13857109526572011689554707, 9.984369578019570859563e-6,
                                    1.50563273514931155834e-7};

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
  const double s = 1.0 / (b.x * b.x + b.y * b.y);
  return make_double2((a.x * b.x + a.y * b.y) * s, (a.y * b.x - a.x * b.y) * s);
}
__device__ __forceinline__ double2 clog2d(double2 a) { return make_double2(log(hypot(a.x, a.y)), atan2(a.y, a.x)); }
__device__ __forceinline__ double2 cexp2d(double2 a) {
  double s, c;
  sincos(a.y, &s, &c);
  const double e = exp(a.x);
  return make_double2(e * c, e * s);
}

// ln Gamma(zz + 1), Re zz > -1/2
__device__ double2 clh_lngamma1(double2 zz) {
  double2 s = make_double2(c_lanczos[0], 0.0);
  for (int w = 1; w < 9; w++) s = cadd(s, cdiv(make_double2(c_lanczos[w], 0.0), make_double2(zz.x + w, zz.y)));
  const double2 t = make_double2(zz.x + 7.5, zz.y);
  const double2 lt = clog2d(t);
  const double2 a = cmul(make_double2(zz.x + 0.5, zz.y), lt);
  const double2 ls = clog2d(s);
  return make_double2(0.5 * log(2.0 * M_PI) + a.x - t.x + ls.x, a.y - t.y + ls.y);
}

// ln Gamma(a) (imaginary part modulo 2 pi), reflection below Re a = 1/2
__device__ double2 clh_lngamma(double2 a) {
  if (a.x < 0.5) {
    double sp, cp;
    sincospi(a.x, &sp, &cp);
    const double2 sn = make_double2(sp * cosh(M_PI * a.y), cp * sinh(M_PI * a.y));  // sin(pi a)
    const double2 ls = clog2d(sn);
    const double2 g = clh_lngamma1(make_double2(-a.x, -a.y));
    return make_double2(log(M_PI) - ls.x - g.x, -ls.y - g.y);
  }
  return clh_lngamma1(make_double2(a.x - 1.0, a.y));
}

// g_l(z) of cfftlog_ells_p2 (cosmo2D.c:4253-4469): derivative 0
// 2^z Gamma((l+z)/2) / Gamma((3+l-z)/2); derivative 2
// (z-1)(z-2) 2^(z-2) Gamma((l+z-2)/2) / Gamma((5+l-z)/2)
__device__ double2 clh_nl_gl(int deriv, int l, double2 z) {
  const double ln2 = log(2.0);
  if (deriv == 0) {
    const double2 g1 = clh_lngamma(make_double2(0.5 * (l + z.x), 0.5 * z.y));
    const double2 g2 = clh_lngamma(make_double2(0.5 * (3 + l - z.x), -0.5 * z.y));
    return cexp2d(make_double2(z.x * ln2 + g1.x - g2.x, z.y * ln2 + g1.y - g2.y));
  }
  const double2 g1 = clh_lngamma(make_double2(0.5 * (l + z.x - 2), 0.5 * z.y));
  const double2 g2 = clh_lngamma(make_double2(0.5 * (5 + l - z.x), -0.5 * z.y));
  const double2 e = cexp2d(make_double2((z.x - 2) * ln2 + g1.x - g2.x, z.y * ln2 + g1.y - g2.y));
  return cmul(cmul(make_double2(z.x - 1, z.y), make_double2(z.x - 2, z.y)), e);
}

// g_l / g_{l-2}: the recurrence the reference applies inside a 16-ell block
__device__ __forceinline__ double2 clh_nl_step(int deriv, int l, double2 z) {
  if (deriv == 0) return cdiv(make_double2(l - 2 + z.x, z.y), make_double2(l + 1 - z.x, -z.y));
  return cdiv(make_double2(l - 4 + z.x, z.y), make_double2(l + 3 - z.x, -z.y));
}

__device__ __forceinline__ bool clh_nl_mag(const DrawDesc& d) {
  bool m = false;
  for (int i = 0; i < d.nl; i++) m = m || fabs(d.par->bmag[i]) > 1.e-12;
  return m;
}

// p_lin(k, a = 1) (cosmo3D.c:306-358), k in H0/c units
__device__ __forceinline__ double clh_plin0(const DrawDesc& d, double k) {
  const double log10k = log10(k / CLH_COVERH0);
  const int nz = d.nzL;
  const int i = clh_bs(d.lkL, d.nkL, log10k);
  const int j = clh_bs(d.pzL, nz, 0.0);
  const double dx = (log10k - d.lkL[i]) / (d.lkL[i + 1] - d.lkL[i]);
  const double dy = (0.0 - d.pzL[j]) / (d.pzL[j + 1] - d.pzL[j]);
  const double* P = d.lnPL + (size_t)i * nz + j;
  const double out = (1 - dx) * (1 - dy) * P[0] + (1 - dx) * dy * P[1] + dx * (1 - dy) * P[nz] + dx * dy * P[nz + 1];
  return exp(out) / (CLH_COVERH0 * CLH_COVERH0 * CLH_COVERH0);
}

// ----------------------------------------------------------------------------
// k_nl_prep: chi nodes and FFTLog inputs (cosmo2D.c:4582-4628).  Work item
// (draw, 256-node block); one thread per node.
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_nl_prep(const DrawDesc* __restrict__ D, const int2* __restrict__ work) {
  const int2 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const ClhNlOff off = clh_nl_off(d.nl);
  double* W = d.nlw;
  const int q = w.y * 256 + threadIdx.x;
  const double rc = CLH_COVERH0 / d.h0;
  const double chi_min = clh_chi(d, 1. / (1.0 + 0.002)) * rc;
  const double chi_max = clh_chi(d, 1. / (1.0 + 4.0)) * rc;
  const double dlnchi = log(chi_max / chi_min) / ((double)CLH_NL_NCHI - 1.0);
  if (q == 0) {
    W[off.scal + 0] = chi_min;
    W[off.scal + 1] = dlnchi;
    W[off.scal + 2] = rc;
  }
  if (q >= 512) return;
  double* fb = W + off.fb;
  if (q >= CLH_NL_NCHI) {
    W[off.x + q] = 0.0;
    for (int i = 0; i < d.nl; i++)
      for (int j = 0; j < 3; j++) fb[(i * 3 + j) * 512 + q] = 0.0;
    return;
  }
  const double x = chi_min * exp(dlnchi * q);
  W[off.x + q] = x;
  const double chi = x / rc;
  const double a = clh_a_chi(d, chi);
  const double z = 1. / a - 1.;
  double cc, dc;
  clh_chi_all(d, a, cc, dc);
  const double hh = 1.0 / ((a * a) * dc);
  const double Dg = clh_growfac(d, a), f = clh_f_growth(d, z);
  const double xnu1 = pow(x, 1.01);
  for (int i = 0; i < d.nl; i++) {
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    if (!(z < d.l_zmin[i] || z > d.l_zmax[i])) {
      const double pf = clh_nz_lens(d, z, i);
      const double WM = 1.5 * d.omega_m * chi / a * clh_g_lens(d, a, i);
      v0 = chi * pf * Dg * hh * d.par->b1[i];
      v1 = -chi * pf * Dg * hh * f;
      v2 = (WM / chi / (rc * rc)) * Dg;
    }
    fb[(i * 3 + 0) * 512 + q] = v0 / x;
    fb[(i * 3 + 1) * 512 + q] = v1 / xnu1;
    fb[(i * 3 + 2) * 512 + q] = v2 / x;
  }
}

// ----------------------------------------------------------------------------
// k_nl_fwd: T_ij[m] = sum_q fb[q] e^{-2 pi i m (N0 + q) / N} (the r2c FFT of
// the zero-padded input, cosmo2D.c:4085-4098) and the c-window
// (cosmo2D.c:4100-4116).  Work item (draw, bin, configuration, 256-m block).
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_nl_fwd(const DrawDesc* __restrict__ D, const int4* __restrict__ work) {
  __shared__ double tw[2 * 1500];
  const int4 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const int i = w.y, j = w.z;
  const ClhNlOff off = clh_nl_off(d.nl);
  const int N = d.nl_N[j], N0 = d.nl_N0[j], nh = N / 2 + 1;
  const int tw0 = j == 0 ? 0 : (j == 1 ? 2 * d.nl_N[0] : 2 * (d.nl_N[0] + d.nl_N[1]));
  for (int t = threadIdx.x; t < 2 * N; t += 256) tw[t] = d.nl_tw[tw0 + t];
  __syncthreads();
  const int m = w.w * 256 + threadIdx.x;
  if (m >= CLH_NL_NH) return;
  double2* T = reinterpret_cast<double2*>(d.nlw + off.T) + (size_t)(i * 3 + j) * CLH_NL_NH;
  if (m >= nh || (j == 2 && !clh_nl_mag(d))) {
    T[m] = make_double2(0.0, 0.0);
    return;
  }
  const double* fb = d.nlw + off.fb + (i * 3 + j) * 512;
  int t = (m * N0) % N;
  double re = 0.0, im = 0.0;
  for (int q = 0; q < CLH_NL_NCHI; q++) {
    const double v = fb[q];
    re += v * tw[2 * t];
    im -= v * tw[2 * t + 1];
    t += m;
    if (t >= N) t -= N;
  }
  const int halfN = N / 2;
  const int kmax = (int)(halfN * 0.25);
  if (m >= halfN - kmax) {
    const int k = halfN - m;
    const double Wc = (double)k / kmax - sin(2. * M_PI * k / kmax) / (2. * M_PI);
    re *= Wc;
    im *= Wc;
  }
  T[m] = make_double2(re, im);
}

// ----------------------------------------------------------------------------
// k_nl_gl: G_j[l][m] = g_l(nu + i eta_m) e^{-i eta_m ln base_l} with
// eta_m = 2 pi m / (dln x N), ln base_l = ln(x_0 e^{-2 N0 dln x} (l+1) / x_last)
// (cosmo2D.c:4120-4128, 4216-4231, 4484-4490); g at the first two ells of each
// 16-ell block from ln Gamma, the rest by the two-step recurrence.  Work item
// (draw, configuration, 256-m block).
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_nl_gl(const DrawDesc* __restrict__ D, const int4* __restrict__ work) {
  const int4 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const int j = w.y;
  const ClhNlOff off = clh_nl_off(d.nl);
  const int N = d.nl_N[j], N0 = d.nl_N0[j], nh = N / 2 + 1;
  const int m = w.z * 256 + threadIdx.x;
  if (m >= CLH_NL_NH) return;
  double2* G = reinterpret_cast<double2*>(d.nlw + off.G) + ((size_t)j * CLH_NL_NH + m) * CLH_NL_LT;
  if (m >= nh || (j == 2 && !clh_nl_mag(d))) {
    for (int l = 0; l < CLH_NL_LT; l++) G[l] = make_double2(0.0, 0.0);
    return;
  }
  const double* x = d.nlw + off.x;
  const double dlnx = log(x[1] / x[0]);
  const double eta = (2.0 * M_PI / (dlnx * N)) * m;
  const double2 z = make_double2(c_nl_nu[j], eta);
  const int deriv = c_nl_deriv[j];
  const double base = x[0] / exp(2 * N0 * dlnx);
  double2 g2 = make_double2(0.0, 0.0), g1 = make_double2(0.0, 0.0);
  for (int l = 0; l < CLH_NL_L; l++) {
    const int ls = l & ~15;
    const double2 g = (l < ls + 2) ? clh_nl_gl(deriv, l, z) : cmul(g2, clh_nl_step(deriv, l, z));
    g2 = g1;
    g1 = g;
    double s, c;
    sincos(-eta * log(base * (l + 1.) / x[CLH_NL_NCHI - 1]), &s, &c);
    G[l] = cmul(g, make_double2(c, s));
  }
  for (int l = CLH_NL_L; l < CLH_NL_LT; l++) G[l] = make_double2(0.0, 0.0);
}

// ----------------------------------------------------------------------------
// k_nl_idft: the inverse real transforms of every (bin, ell) kept at the 500
// original nodes, as an FP64 MFMA product over m, fused with the C_ell
// integrand (cosmo2D.c:4492-4524 and 4702-4734).  Work item (draw, bin,
// 16-ell tile, 256-node tile).
// Lane layout of v_mfma_f64_16x16x4: A[r][k] on lane r + 16 k, B[k][n] on
// lane n + 16 k, D[4 v + l / 16][l % 16] in value v of lane l.  A block of
// four waves covers 256 columns, each wave four 16-column tiles that share
// its A operand (four independent accumulation chains).  K index
// 4 s + k -> m = 2 s + k / 2, real (k even) or imaginary (k odd) part:
//   A[l][.] = (Re Z, Im Z) with Z = T[m] G[l][m],
//   B[.][q] = w_m (cos, sin)(2 pi m n / N), n = N0 + q, w_m = 1 at m = 0 and
// N/2, 2 elsewhere (the c2r sum over the Hermitian half), so that
// sum A B = sum_m w_m Re(conj(Z) e^{2 pi i m n / N}).
// CHECK: the tile's rows are the ten block-end ells (row r -> min(16 r + 15,
// 149), rows 10..15 empty) and the sums go to partc; otherwise the tile is
// ells 16 w.z .. 16 w.z + 15 and is skipped at or above the bin's LMAX.
// ----------------------------------------------------------------------------
__device__ __forceinline__ int clh_nl_row(bool check, int l0, int r) {
  if (!check) return l0 + r;
  return r < 10 ? (16 * r + 15 < CLH_NL_L ? 16 * r + 15 : CLH_NL_L - 1) : CLH_NL_LT - 1;
}

template <bool CHECK>
__global__ __launch_bounds__(256) void k_nl_idft(const DrawDesc* __restrict__ D, const int4* __restrict__ work) {
  constexpr int SUB = 4;  // 16-column output tiles per wave, sharing the A operand
  __shared__ double tw[2 * 1500];
  const int4 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const int i = w.y, l0 = 16 * w.z;
  const ClhNlOff off = clh_nl_off(d.nl);
  if (!CHECK && l0 >= (int)d.nlw[off.lmx + i]) return;  // whole block: ells the reference never transforms
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r = lane & 15, kq = lane >> 4, mo = kq >> 1, cpart = kq & 1;
  const int lrow = clh_nl_row(CHECK, l0, r);
  const int q0 = 256 * w.w + 64 * wv;  // the wave's columns q0 .. q0 + 63
  const bool mag = clh_nl_mag(d);
  const double2* Tall = reinterpret_cast<const double2*>(d.nlw + off.T);
  const double2* Gall = reinterpret_cast<const double2*>(d.nlw + off.G);
  clh_v4d acc[3][SUB];
  for (int j = 0; j < 3; j++)
    for (int u = 0; u < SUB; u++) acc[j][u] = (clh_v4d){0.0, 0.0, 0.0, 0.0};
  for (int j = 0; j < 3; j++) {
    if (j == 2 && !mag) break;
    const int N = d.nl_N[j], N0 = d.nl_N0[j], nh = N / 2 + 1;
    const int tw0 = j == 0 ? 0 : (j == 1 ? 2 * d.nl_N[0] : 2 * (d.nl_N[0] + d.nl_N[1]));
    __syncthreads();
    for (int t = threadIdx.x; t < 2 * N; t += 256) tw[t] = d.nl_tw[tw0 + t];
    __syncthreads();
    const double2* T = Tall + (size_t)(i * 3 + j) * CLH_NL_NH + mo;
    const double2* G = Gall + ((size_t)j * CLH_NL_NH + mo) * CLH_NL_LT + lrow;
    int t[SUB], dt[SUB];
    double bm[SUB];
    for (int u = 0; u < SUB; u++) {
      const int q = q0 + 16 * u + r;
      const int n = N0 + (q < CLH_NL_NCHI ? q : 0);
      bm[u] = q < CLH_NL_NCHI ? 2.0 : 0.0;  // w_m = 2 except at m = 0 and N/2
      t[u] = (mo * n) % N;
      dt[u] = (2 * n) % N;
    }
    const int nsteps = (nh + 1) / 2;
    clh_v4d a4[SUB];
    for (int u = 0; u < SUB; u++) a4[u] = acc[j][u];
    double2 Tn = T[0], Gn = G[0];
    for (int s = 0; s < nsteps; s++) {
      const double2 Z = cmul(Tn, Gn);
      if (s + 1 < nsteps) {  // next step's operands, loaded ahead of this step's MFMAs
        Tn = T[2 * (s + 1)];
        Gn = G[(size_t)2 * (s + 1) * CLH_NL_LT];
      }
      const int m = 2 * s + mo;
      const double a = cpart ? Z.y : Z.x;
      const double half = (m == 0 || 2 * m == N) ? 0.5 : 1.0;
#pragma unroll
      for (int u = 0; u < SUB; u++) {
        const double b = bm[u] * half * tw[2 * t[u] + cpart];
        a4[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, a4[u], 0, 0, 0);
        t[u] += dt[u];
        if (t[u] >= N) t[u] -= N;
      }
    }
    for (int u = 0; u < SUB; u++) acc[j][u] = a4[u];
  }
  // epilogue: F = sum_j Fy_j (magnification with b_mag l(l+1) / y^2),
  // integrand F^2 k^3 P_lin(k, a = 1) at k = y c/H0, y = (l + 1) / x[499 - q]
  const double* x = d.nlw + off.x;
  const double rc = d.nlw[off.scal + 2];
  const double bmag = d.par->bmag[i];
  double* part = d.nlw + (CHECK ? off.partc : off.part);
  for (int v = 0; v < 4; v++) {
    const int row = 4 * v + kq;
    const int l = clh_nl_row(CHECK, l0, row);
    for (int u = 0; u < SUB; u++) {
      const int q = q0 + 16 * u + r;
      double val = 0.0;
      if (l < CLH_NL_L && q < CLH_NL_NCHI) {
        const double xq = x[CLH_NL_NCHI - 1 - q];
        const double ty = (l + 1.) / xq;
        double F = 0.0;
        for (int j = 0; j < 3; j++) {
          if (j == 2 && !mag) break;
          const double nu = c_nl_nu[j];
          const double pre = sqrt(M_PI) / (4.0 * d.nl_N[j] * pow(l + 1., nu)) * pow(xq, nu);
          const double fy = acc[j][u][v] * pre;
          F += j == 2 ? bmag * (l * (l + 1.)) * fy / (ty * ty) : fy;
        }
        const double k1 = ty * rc;
        val = F * F * (k1 * k1 * k1) * clh_plin0(d, k1);
      }
      val += __shfl_xor(val, 1);
      val += __shfl_xor(val, 2);
      val += __shfl_xor(val, 4);
      val += __shfl_xor(val, 8);
      if (r == 0)
        part[((size_t)i * (CHECK ? 16 : CLH_NL_LT) + (CHECK ? row : l)) * CLH_NL_QT + 16 * w.w + 4 * wv + u] = val;
    }
  }
}

// ----------------------------------------------------------------------------
// k_nl_lmax: per (draw, lens bin), one wave: the non-Limber C_ell at the ten
// block ends and the first block whose end agrees with Limber to 1%
// (cosmo2D.c:4740-4750; LMAX = 150 if none does)
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_nl_lmax(const DrawDesc* __restrict__ D, const int2* __restrict__ work) {
  __shared__ double dev[16];
  const int2 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const int i = w.y, r = threadIdx.x;
  const ClhNlOff off = clh_nl_off(d.nl);
  if (r < 10) {
    const int L = clh_nl_row(true, 0, r);
    const double* part = d.nlw + off.partc + ((size_t)i * 16 + r) * CLH_NL_QT;
    double s = 0.0;
    for (int t = 0; t < CLH_NL_QT; t++) s += part[t];
    const double lnl = d.nl_lnl[(size_t)i * (CLH_NL_L - 1) + L - 1];
    const double llin = d.nl_llin[(size_t)i * (CLH_NL_L - 1) + L - 1];
    const double cl = s * d.nlw[off.scal + 1] * 2. / M_PI + lnl - llin;
    double dv = INFINITY;
    if (fabs(lnl) > 1e-300) dv = cl / lnl - 1.0;
    dev[r] = (isfinite(dv) && fabs(dv) < 0.01) ? 1.0 : 0.0;
  }
  __syncthreads();
  if (r == 0) {
    int lmax = CLH_NL_L;
    for (int b = 0; b < 10; b++)
      if (dev[b] != 0.0) {
        lmax = 16 * (b + 1) < CLH_NL_L ? 16 * (b + 1) : CLH_NL_L;
        break;
      }
    d.nlw[off.lmx + i] = (double)lmax;
  }
}

// ----------------------------------------------------------------------------
// k_nl_cl: per (draw, lens bin): the non-Limber C_ell, the convergence switch
// and the w(theta) change (cosmo2D.c:4695-4767; w_gg_tomo, cosmo2D.c:843-905).
// The folded Legendre kernel has already summed Pw against the Limber C_ell
// C_L (direct rows below LMIN_tab = 20, the ln-ell table fill above); this adds
// sum_l Pw[theta][l] (C_used[l] - C_L[l]) for 1 <= l < 150, where C_used is the
// non-Limber value below the bin's LMAX and the reference's Limber choice
// (table above LMIN_tab, direct at or below) from there on.
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_nl_cl(const DrawDesc* __restrict__ D, const int2* __restrict__ work) {
  __shared__ double NL[CLH_NL_L], DL[CLH_NL_L];
  __shared__ int s_lmax;
  const int2 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const int i = w.y, l = threadIdx.x;
  const ClhNlOff off = clh_nl_off(d.nl);
  const double dlnk = d.nlw[off.scal + 1];
  const double* lnl = d.nl_lnl + (size_t)i * (CLH_NL_L - 1);
  const double* llin = d.nl_llin + (size_t)i * (CLH_NL_L - 1);
  if (l < CLH_NL_L && l < (int)d.nlw[off.lmx + i]) {  // the transformed ells
    const double* part = d.nlw + off.part + ((size_t)i * CLH_NL_LT + l) * CLH_NL_QT;
    double s = 0.0;
    for (int t = 0; t < CLH_NL_QT; t++) s += part[t];
    NL[l] = s * dlnk * 2. / M_PI + (l >= 1 ? lnl[l - 1] - llin[l - 1] : 0.0);
  }
  if (l == 0) s_lmax = (int)d.nlw[off.lmx + i];
  __syncthreads();
  if (l < CLH_NL_L) {
    const double* tab = d.cl + (size_t)(d.n_ss + d.n_gs + i) * d.nrow;
    double interp = 0.0;
    if (l >= CLH_LMIN_TAB) {
      const double r = (log((double)l) - d.lim_gx[0]) * (1.0 / d.lim_gx[2]);
      const int ii = (int)floor(r);
      const int ic = ii < 0 ? 0 : (ii >= d.n_ell - 1 ? d.n_ell - 2 : ii);
      const double t = r - ic;
      interp = tab[CLH_NLOW + ic] * (1.0 - t) + tab[CLH_NLOW + ic + 1] * t;
    }
    const double direct = l >= 1 ? lnl[l - 1] : 0.0;
    const double limber_path = l < CLH_LMIN_TAB ? (l >= 1 ? tab[l - 1] : 0.0) : interp;
    const double used = l < s_lmax ? NL[l] : (l > CLH_LMIN_TAB ? interp : direct);
    d.nlw[off.used + (size_t)i * CLH_NL_L + l] = used;
    DL[l] = l >= 1 ? used - limber_path : 0.0;
  }
  __syncthreads();
  if (l < d.ntheta && d.dv) {
    const double* pw = d.nl_pw + (size_t)l * CLH_NL_L;
    double dw = 0.0;
    for (int k = 1; k < CLH_NL_L; k++) dw += pw[k] * DL[k];
    const int nt = d.ntheta;
    const int canon = nt * (2 * d.n_ss + d.n_gs_dv) + nt * i + l;
    if (!clh_masked(d, canon)) d.dv[d.dv_start[2] + nt * i + l] += dw;
  }
}

hipError_t clh_launch_nl_forward(const DrawDesc* D, const int2* wprep, int nprep, const int4* wfwd, int nfwd,
                                 const int4* wgl, int ngl, const int4* wchk, int nchk, const int2* wbin, int nbin,
                                 const int4* widft, int nidft, hipStream_t s) {
  if (nprep <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_nl_prep, dim3(nprep), dim3(256), 0, s, D, wprep);
  hipLaunchKernelGGL(k_nl_fwd, dim3(nfwd), dim3(256), 0, s, D, wfwd);
  hipLaunchKernelGGL(k_nl_gl, dim3(ngl), dim3(256), 0, s, D, wgl);
  hipLaunchKernelGGL(k_nl_idft<true>, dim3(nchk), dim3(256), 0, s, D, wchk);
  hipLaunchKernelGGL(k_nl_lmax, dim3(nbin), dim3(64), 0, s, D, wbin);
  hipLaunchKernelGGL(k_nl_idft<false>, dim3(nidft), dim3(256), 0, s, D, widft);
  return hipGetLastError();
}

hipError_t clh_launch_nl_cl(const DrawDesc* D, const int2* work, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_nl_cl, dim3(n), dim3(256), 0, s, D, work);
  return hipGetLastError();
}
