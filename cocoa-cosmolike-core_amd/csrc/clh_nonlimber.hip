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
//   k_nl_fwd   the forward r2c FFT T_ij[m], windowed         (one per bin, j)
//   k_nl_gl    G_j[l][m] = g_l e^{-i eta_m ln base_l}         (thread per m)
//   k_nl_fft   per (draw, ell) and lens bin: the inverse real transforms
//              Out[n] = sum_m w_m Re(conj(T G) e^{2 pi i m n / N}) as one
//              mixed-radix Stockham FFT of length N/2 in LDS each
//              (csrc/clh_fft.hpp), fused with the C_ell integrand
//              F^2 k^3 P_lin(k) at the 500 kept outputs and its sum; first
//              for the ten block-end ells (15, 31, ..., 143, 149) only, then
//              for every ell below each bin's LMAX
//   k_nl_lmax  the reference's block-convergence switch to Limber: LMAX is the
//              end of the first 16-ell block whose last C_ell is within 1% of
//              the Limber C_ell (cosmo2D.c:4740-4750)
//   k_nl_cl    C_ell = 2/pi dln k sum + C_Limber(P_nl) - C_Limber(P_lin) and
//              the change of w(theta) against the Limber C_ell the folded
//              Legendre kernel already summed:
//              dw = sum_l Pw[theta][l] (C_used - C_Limber)[l]
// As in the reference, no ell at or above a bin's LMAX is transformed.
#include "clh_device.hpp"
#include "clh_fft.hpp"

// e^{2 pi i t / M} as the product of two LDS tables: lo[t % S], hi[t / S]
struct ClhTwid {
  const clh_c2* lo;
  const clh_c2* hi;
  int S;
  __device__ __forceinline__ clh_c2 operator()(int t) const { return clh_c2_mul(hi[t / S], lo[t % S]); }
};
#define CLH_NL_TWS 28  // S: S * S >= the largest M (750)


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
    W[off.xnu + q] = 0.0;
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
  W[off.xnu + q] = xnu1;
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
// (cosmo2D.c:4100-4116).  Work item (draw, bin, configuration): one r2c
// transform of length N as a complex Stockham FFT of length N / 2 in LDS
// (conj of the + sign transform of the real input).
// ----------------------------------------------------------------------------
// TAN: the VJP's tangent inputs (DrawDesc::nlt, ClhNlTan), work item (draw,
// bin, set): sets 0-2 in configurations 0-2, set 3 in configuration 0.
template <bool TAN>
__global__ __launch_bounds__(256) void k_nl_fwd(const DrawDesc* __restrict__ D, const int4* __restrict__ work) {
  __shared__ clh_c2 buf[2][CLH_NL_NH - 2];
  __shared__ clh_c2 twl[CLH_NL_TWS], twh[CLH_NL_TWS];
  const int4 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const int i = w.y, set = w.z, j = TAN ? (set < 3 ? set : 0) : set, tid = threadIdx.x;
  const ClhNlOff off = clh_nl_off(d.nl);
  const ClhNlTan ot = clh_nl_tan(d.nl);
  const int N = d.nl_N[j], N0 = d.nl_N0[j], M = N / 2, nh = N / 2 + 1;
  const double* tw = d.nl_tw + (j == 0 ? 0 : (j == 1 ? 2 * d.nl_N[0] : 2 * (d.nl_N[0] + d.nl_N[1])));
  clh_c2* T = TAN ? reinterpret_cast<clh_c2*>(d.nlt + ot.T) + (size_t)(i * 4 + set) * CLH_NL_NH
                  : reinterpret_cast<clh_c2*>(d.nlw + off.T) + (size_t)(i * 3 + j) * CLH_NL_NH;
  if (j == 2 && !clh_nl_mag(d)) {
    for (int m = tid; m < CLH_NL_NH; m += 256) T[m] = clh_c2{0.0, 0.0};
    return;
  }
  for (int u = tid; u < 2 * CLH_NL_TWS; u += 256) {  // e^{2 pi i t / M} = entry 2 t of the N table
    const int hi = u / CLH_NL_TWS, t = u % CLH_NL_TWS;
    const int tt = (hi ? t * CLH_NL_TWS : t) % M;
    (hi ? twh : twl)[t] = clh_c2{tw[4 * tt], tw[4 * tt + 1]};
  }
  const double* fb = TAN ? d.nlt + ot.fb + (i * 4 + set) * 512 : d.nlw + off.fb + (i * 3 + j) * 512;
  for (int n = tid; n < M; n += 256) {
    const int a = 2 * n - N0, b = 2 * n + 1 - N0;
    buf[0][n] = clh_c2{(a >= 0 && a < CLH_NL_NCHI) ? fb[a] : 0.0, (b >= 0 && b < CLH_NL_NCHI) ? fb[b] : 0.0};
  }
  __syncthreads();
  const ClhTwid twM{twl, twh, CLH_NL_TWS};
  int src = 0, Ns = 1;
  for (int m = M; m > 1;) {  // N / 2 factors into 2, 3, 5 (checked on the host)
    const int R = clh_fft_radix(m);
    const ClhFftStage st = clh_fft_stage(M, R, Ns);
    for (int jb = tid; jb < st.stride; jb += 256) clh_fft_bfly(buf[src], buf[src ^ 1], st, jb, twM);
    __syncthreads();
    src ^= 1;
    Ns *= R;
    m /= R;
  }
  const int halfN = N / 2;
  const int kmax = (int)(halfN * 0.25);
  for (int m = tid; m < CLH_NL_NH; m += 256) {
    if (m >= nh) {
      T[m] = clh_c2{0.0, 0.0};
      continue;
    }
    const clh_c2 e = m == M ? clh_c2{-1.0, 0.0} : clh_c2{tw[2 * m], tw[2 * m + 1]};  // e^{2 pi i m / N}
    const clh_c2 y = clh_r2c_unpack(buf[src][m % M], buf[src][(M - m) % M], e);
    double re = y.re, im = -y.im;  // e^{-2 pi i ...}: the conjugate for real input
    if (m >= halfN - kmax) {
      const int k = halfN - m;
      const double Wc = (double)k / kmax - sin(2. * M_PI * k / kmax) / (2. * M_PI);
      re *= Wc;
      im *= Wc;
    }
    T[m] = clh_c2{re, im};
  }
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
  // G_j stored [l][m] (m fastest): these stores and the transforms' reads coalesce
  double2* G = reinterpret_cast<double2*>(d.nlw + off.G) + (size_t)j * CLH_NL_LT * CLH_NL_NH + m;
  if (m >= nh || (j == 2 && !clh_nl_mag(d))) {
    for (int l = 0; l < CLH_NL_LT; l++) G[(size_t)l * CLH_NL_NH] = make_double2(0.0, 0.0);
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
    G[(size_t)l * CLH_NL_NH] = cmul(g, make_double2(c, s));
  }
  for (int l = CLH_NL_L; l < CLH_NL_LT; l++) G[(size_t)l * CLH_NL_NH] = make_double2(0.0, 0.0);
}

// ----------------------------------------------------------------------------
// k_nl_fft: the inverse real transforms of one ell for every lens bin below
// its LMAX, at the 500 original nodes, fused with the C_ell integrand
// (cosmo2D.c:4492-4524 and 4702-4734).  Work item (draw, ell) or, with
// CHECK, (draw, r) for the block-end ell of row r (ten rows).  Per bin and
// configuration j: X_m = conj(T_j[m] G_j[l][m]) for m = 0 .. N/2 (the
// imaginary parts at m = 0 and N/2 dropped, as the c2r transform does),
// packed into one complex transform of length M = N/2 whose outputs z_n
// hold Out[2n] + i Out[2n+1] (Stockham stages ping-ponging between two LDS
// buffers, twiddles from the layout's cos / sin table of 2 pi t / N);
// Out[N0 + q] then enters F(q) = sum_j Fy_j (magnification with
// b_mag l(l+1) / y^2), and the bin's sum_q F^2 k^3 P_lin(k) at k = y c/H0,
// y = (l + 1) / x[499 - q], goes to part (partc with CHECK).
// ----------------------------------------------------------------------------
__device__ __forceinline__ int clh_nl_row(bool check, int l0, int r) {
  if (!check) return l0 + r;
  return r < 10 ? (16 * r + 15 < CLH_NL_L ? 16 * r + 15 : CLH_NL_L - 1) : CLH_NL_LT - 1;
}

// number of Stockham stages of a length-M transform
__device__ __forceinline__ int clh_fft_nst(int M) {
  int n = 0;
  for (int m = M; m > 1; m /= clh_fft_radix(m)) n++;
  return n;
}

// Two halves of 128 threads run the density (j = 0) and RSD (j = 1)
// transforms of a bin side by side, with barriers shared; the magnification
// transform (j = 2) follows on the second half.
// TAN (the VJP): per bin whose ell carries a cotangent sbar of the sum,
// three passes: F from the forward's transforms, then F_z from the lens
// shift's tangent transforms (sets 0-2) and F_b from the b1 tangent (set 3,
// density only); sbar sum_q 2 F F_z k^3 P_lin goes to g_lens, sbar sum_q
// 2 F F_b k^3 P_lin to g_b1 (the sum is linear in each transformed input).
template <bool CHECK, bool TAN = false>
__global__ __launch_bounds__(256) void k_nl_fft(const DrawDesc* __restrict__ D, const int4* __restrict__ work) {
  static_assert(!(CHECK && TAN), "the tangent passes run at the transformed ells");
  __shared__ clh_c2 fbuf[2 * (CLH_NL_NH - 2) + 2 * 450];  // per half two buffers of its M (M0 + M1 <= 1200)
  __shared__ clh_c2 twl[2][CLH_NL_TWS], twh[2][CLH_NL_TWS];  // per M (configuration 0; 1 and 2)
  __shared__ double kP[CLH_NL_NCHI];
  __shared__ double Fh[2][512];  // the linear P column at z = 0 first, then F(q) per half
  __shared__ double red[4];
  __shared__ double Fk[TAN ? 512 : 1];  // TAN: F of the first pass
  const int4 w = work[clh_xcd_item(blockIdx.x, gridDim.x)];
  const DrawDesc& d = D[w.x];
  const int r = w.y, l = CHECK ? clh_nl_row(true, 0, r) : r;
  const ClhNlOff off = clh_nl_off(d.nl);
  const ClhNlTan ot = clh_nl_tan(d.nl);
  const int tid = threadIdx.x, h = tid >> 7, ht = tid & 127;
  if (!CHECK) {  // nothing to do when every bin switched to Limber at or below l
    bool any = false;
    for (int i = 0; i < d.nl; i++) any = any || l < (int)d.nlw[off.lmx + i];
    if (!any) return;
  }
  // twiddle tables of M = N_j / 2 (N_1 = N_2) from the layout's cos / sin of 2 pi t / N_j
  for (int u = tid; u < 4 * CLH_NL_TWS; u += 256) {
    const int c = u / (2 * CLH_NL_TWS), hi = (u / CLH_NL_TWS) & 1, t = u % CLH_NL_TWS;
    const int N = d.nl_N[c], M = N / 2;
    const double* tw = d.nl_tw + (c == 0 ? 0 : 2 * d.nl_N[0]);
    const int tt = (hi ? t * CLH_NL_TWS : t) % M;  // e^{2 pi i tt / M} = entry 2 tt of the N table
    (hi ? twh : twl)[c][t] = clh_c2{tw[4 * tt], tw[4 * tt + 1]};
  }
  // k^3 P_lin(k, z = 0) at k = (l + 1) / x[499 - q] c/H0: the z = 0 column
  // of ln P_lin staged in LDS (interpolation in k of the z-interpolated
  // column is the bilinear value of p_lin, cosmo3D.c:306-358)
  const double* x = d.nlw + off.x;
  const double* xnu = d.nlw + off.xnu;
  const double rc = d.nlw[off.scal + 2];
  const bool stage = d.nkL <= 512;
  double* s_lk = Fh[0];
  double* s_col = Fh[1];
  if (stage) {
    const int nz = d.nzL, j0 = clh_bs(d.pzL, nz, 0.0);
    const double dy = (0.0 - d.pzL[j0]) / (d.pzL[j0 + 1] - d.pzL[j0]);
    for (int k = tid; k < d.nkL; k += 256) {
      s_lk[k] = d.lkL[k];
      const double* P = d.lnPL + (size_t)k * nz + j0;
      s_col[k] = (1 - dy) * P[0] + dy * P[1];
    }
  }
  __syncthreads();
  for (int q = tid; q < CLH_NL_NCHI; q += 256) {
    const double xq = x[CLH_NL_NCHI - 1 - q];
    const double k1 = (l + 1.) / xq * rc;
    double pl;
    if (stage) {
      const double log10k = log10(k1 / CLH_COVERH0);
      const int i = clh_bs(s_lk, d.nkL, log10k);
      const double dx = (log10k - s_lk[i]) / (s_lk[i + 1] - s_lk[i]);
      pl = exp((1 - dx) * s_col[i] + dx * s_col[i + 1]) / (CLH_COVERH0 * CLH_COVERH0 * CLH_COVERH0);
    } else {
      pl = clh_plin0(d, k1);
    }
    kP[q] = (k1 * k1 * k1) * pl;
  }
  const bool mag = clh_nl_mag(d);
  const clh_c2* Tall = reinterpret_cast<const clh_c2*>(d.nlw + off.T);
  const clh_c2* Gall = reinterpret_cast<const clh_c2*>(d.nlw + off.G);
  const int M0 = d.nl_N[0] / 2;
  clh_c2* mybuf = h == 0 ? fbuf : fbuf + 2 * M0;  // this half's two buffers
  for (int i = 0; i < d.nl; i++) {
    if (!CHECK && l >= (int)d.nlw[off.lmx + i]) continue;  // uniform over the block
    const double sbar = TAN ? d.nlt[ot.sbar + (size_t)i * CLH_NL_LT + l] : 0.0;
    if (TAN && sbar == 0.0) continue;  // (uniform)
    for (int pass = 0; pass < (TAN ? 3 : 1); pass++) {
    __syncthreads();  // kP / the previous bin's (pass's) sum done with Fh
    for (int q = ht; q < 512; q += 128) Fh[h][q] = 0.0;
    const double bmag = d.par->bmag[i];
    for (int round = 0; round < (pass == 2 ? 1 : (mag ? 2 : 1)); round++) {
      const int j = round == 0 ? h : 2;
      const bool act = pass == 2 ? h == 0 : (round == 0 || h == 1);
      const int N = d.nl_N[j], N0 = d.nl_N0[j], M = N / 2, c = j == 0 ? 0 : 1;
      const double* tw = d.nl_tw + (j == 0 ? 0 : (j == 1 ? 2 * d.nl_N[0] : 2 * (d.nl_N[0] + d.nl_N[1])));
      clh_c2* b0 = mybuf;
      clh_c2* b1 = mybuf + M;
      if (act) {
        const clh_c2* T = pass == 0 ? Tall + (size_t)(i * 3 + j) * CLH_NL_NH
                                    : reinterpret_cast<const clh_c2*>(d.nlt + ot.T) +
                                          (size_t)(i * 4 + (pass == 1 ? j : 3)) * CLH_NL_NH;
        const clh_c2* G = Gall + ((size_t)j * CLH_NL_LT + l) * CLH_NL_NH;
        for (int k = ht; k < M; k += 128) {
          const clh_c2 a = clh_c2_mul(T[k], G[k]), b = clh_c2_mul(T[M - k], G[M - k]);
          clh_c2 Xk{a.re, -a.im}, XMk{b.re, -b.im};
          if (k == 0) Xk.im = XMk.im = 0.0;
          b0[k] = clh_c2r_pack(Xk, XMk, clh_c2{tw[2 * k], tw[2 * k + 1]});
        }
      }
      __syncthreads();
      // the stage loop runs the larger stage count of the active halves;
      // N / 2 factors into 2, 3, 5 (nl_fft_size, checked on the host)
      const int nst = round == 0 ? max(clh_fft_nst(M0), clh_fft_nst(d.nl_N[1] / 2)) : clh_fft_nst(M);
      const ClhTwid twM{twl[c], twh[c], CLH_NL_TWS};
      clh_c2* src = b0;
      clh_c2* dst = b1;
      int Ns = 1, m = M;
      for (int s = 0; s < nst; s++) {
        if (act && m > 1) {
          const int R = clh_fft_radix(m);
          const ClhFftStage st = clh_fft_stage(M, R, Ns);
          for (int jb = ht; jb < st.stride; jb += 128) clh_fft_bfly(src, dst, st, jb, twM);
          Ns *= R;
          m /= R;
          clh_c2* t = src;
          src = dst;
          dst = t;
        }
        __syncthreads();
      }
      if (act) {
        const double nu = c_nl_nu[j];
        const double pre = sqrt(M_PI) / (4.0 * N * (nu == 1.0 ? l + 1. : pow(l + 1., nu)));
        for (int q = ht; q < CLH_NL_NCHI; q += 128) {
          const int n = N0 + q;
          const double o = (n & 1) ? src[n >> 1].im : src[n >> 1].re;
          const double xq = x[CLH_NL_NCHI - 1 - q];
          const double fy = o * pre * (nu == 1.0 ? xq : xnu[CLH_NL_NCHI - 1 - q]);
          if (j == 2) {
            const double ty = (l + 1.) / xq;
            Fh[h][q] += bmag * (l * (l + 1.)) * fy / (ty * ty);
          } else {
            Fh[h][q] += fy;
          }
        }
      }
      __syncthreads();  // this round's readers done with the buffers, F complete
    }
    if constexpr (TAN) {
      if (pass == 0) {
        for (int q = tid; q < CLH_NL_NCHI; q += 256) Fk[q] = Fh[0][q] + Fh[1][q];
      } else {  // (each thread reads the Fk entries it wrote)
        double s = 0.0;
        for (int q = tid; q < CLH_NL_NCHI; q += 256) s += 2.0 * Fk[q] * (Fh[0][q] + Fh[1][q]) * kP[q];
        s = clh_block_sum(s, red);
        if (tid == 0 && s != 0.0) atomicAdd(pass == 1 ? d.g_lens + i : d.g_b1 + i, sbar * s);
      }
    } else {
      double s = 0.0;
      for (int q = tid; q < CLH_NL_NCHI; q += 256) {
        const double F = Fh[0][q] + Fh[1][q];
        s += F * F * kP[q];
      }
      s = clh_block_sum(s, red);
      if (tid == 0) {
        double* part = d.nlw + (CHECK ? off.partc : off.part);
        part[(size_t)i * (CHECK ? 16 : CLH_NL_LT) + (CHECK ? r : l)] = s;
      }
    }
    }  // pass
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
    const double s = d.nlw[off.partc + (size_t)i * 16 + r];
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
    const double s = d.nlw[off.part + (size_t)i * CLH_NL_LT + l];
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

// ----------------------------------------------------------------------------
// Adjoint (VJP) of the non-Limber w.
// k_nl_prep_tan: the tangent FFTLog inputs (ClhNlTan sets) at the chi nodes
// of k_nl_prep: d/d dz of the density b1 n D H/H0 chi, RSD -f n D H/H0 chi
// and magnification W_mag D / chi / (c/H0)^2 inputs (n(z - dz) and g_lens
// through their shift tangents), and d/d b1 of the density input.  The bin's
// z range (l_zmin .. l_zmax) does not move with dz.
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_nl_prep_tan(const DrawDesc* __restrict__ D, const int2* __restrict__ work) {
  const int2 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const ClhNlOff off = clh_nl_off(d.nl);
  const ClhNlTan ot = clh_nl_tan(d.nl);
  double* fb = d.nlt + ot.fb;
  const int q = w.y * 256 + threadIdx.x;
  if (q >= 512) return;
  if (q >= CLH_NL_NCHI) {
    for (int i = 0; i < d.nl; i++)
      for (int j = 0; j < 4; j++) fb[(i * 4 + j) * 512 + q] = 0.0;
    return;
  }
  const double rc = d.nlw[off.scal + 2];
  const double x = d.nlw[off.x + q], xnu1 = d.nlw[off.xnu + q];
  const double chi = x / rc;
  const double a = clh_a_chi(d, chi);
  const double z = 1. / a - 1.;
  double cc, dc;
  clh_chi_all(d, a, cc, dc);
  const double hh = 1.0 / ((a * a) * dc);
  const double Dg = clh_growfac(d, a), f = clh_f_growth(d, z);
  for (int i = 0; i < d.nl; i++) {
    double t0 = 0.0, t1 = 0.0, t2 = 0.0, tb = 0.0;
    if (!(z < d.l_zmin[i] || z > d.l_zmax[i])) {
      const double pf = clh_nz_lens(d, z, i), dpf = clh_dnz_lens(d, z, i);
      const double dWM = 1.5 * d.omega_m * chi / a * clh_dg_lens(d, a, i);
      t0 = chi * dpf * Dg * hh * d.par->b1[i];
      t1 = -chi * dpf * Dg * hh * f;
      t2 = (dWM / chi / (rc * rc)) * Dg;
      tb = chi * pf * Dg * hh;
    }
    fb[(i * 4 + 0) * 512 + q] = t0 / x;
    fb[(i * 4 + 1) * 512 + q] = t1 / xnu1;
    fb[(i * 4 + 2) * 512 + q] = t2 / x;
    fb[(i * 4 + 3) * 512 + q] = tb / x;
  }
}

// ----------------------------------------------------------------------------
// k_nl_cl_adj: adjoint of k_nl_cl per (draw, lens bin), thread = ell.  With u
// the masked w cotangent of the bin, dL[l] = sum_theta Pw[theta][l] u[theta]
// is the cotangent of DL = C_used - C_L:
//  - below LMAX: sbar[l] = dL 2/pi dln k (the FFTLog sum, k_nl_fft<TAN>),
//    +dL on the P_nl view's row l and -dL on the P_lin view's;
//  - from LMAX on: +dL on the table rows of the interpolated value (l >
//    LMIN_tab) or on the P_nl view's row (l <= LMIN_tab);
//  - -dL on what the folded kernel summed: the direct row l - 1 (l <
//    LMIN_tab) or the table rows (l >= LMIN_tab) of the gg column.
// The LMAX switch is piecewise constant in the parameters.  Runs after
// k_leg_adjoint (which stored the gg column's cotangent) and before the gg
// adjoints read it.
// ----------------------------------------------------------------------------
__device__ __forceinline__ void clh_nl_interp_adj(const DrawDesc& d, double* tab, int l, double v) {
  const double r = (log((double)l) - d.lim_gx[0]) * (1.0 / d.lim_gx[2]);
  const int ii = (int)floor(r);
  const int ic = ii < 0 ? 0 : (ii >= d.n_ell - 1 ? d.n_ell - 2 : ii);
  const double t = r - ic;
  atomicAdd(tab + CLH_NLOW + ic, v * (1.0 - t));
  atomicAdd(tab + CLH_NLOW + ic + 1, v * t);
}

__global__ __launch_bounds__(256) void k_nl_cl_adj(const DrawDesc* __restrict__ D, const int2* __restrict__ work) {
  __shared__ double u[32];
  const int2 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const int i = w.y, l = threadIdx.x, nt = d.ntheta;
  const ClhNlOff off = clh_nl_off(d.nl);
  const ClhNlTan ot = clh_nl_tan(d.nl);
  if (l < nt) {
    const int canon = nt * (2 * d.n_ss + d.n_gs_dv) + nt * i + l;
    u[l] = clh_masked(d, canon) ? 0.0 : d.dvbar[d.dv_start[2] + nt * i + l];
  }
  __syncthreads();
  if (l >= CLH_NL_LT) return;
  double* sbar = d.nlt + ot.sbar + (size_t)i * CLH_NL_LT;
  if (l == 0 || l >= CLH_NL_L) {
    sbar[l] = 0.0;
    return;
  }
  double x = 0.0;
  for (int th = 0; th < nt; th++) x += d.nl_pw[(size_t)th * CLH_NL_L + l] * u[th];
  const int lmax = (int)d.nlw[off.lmx + i];
  const int ncol = d.n_ss + d.n_gs + d.nl, col = d.n_ss + d.n_gs + i;
  double* vnl = d.nlc_bar + (size_t)col * (CLH_NL_L - 1);
  double* vlin = d.nlc_bar + (size_t)(ncol + col) * (CLH_NL_L - 1);
  double* tab = d.cbar + (size_t)col * d.nrow;
  sbar[l] = l < lmax ? x * d.nlw[off.scal + 1] * 2. / M_PI : 0.0;
  vnl[l - 1] = (l < lmax || l <= CLH_LMIN_TAB) ? x : 0.0;
  vlin[l - 1] = l < lmax ? -x : 0.0;
  if (x == 0.0) return;
  if (l >= lmax && l > CLH_LMIN_TAB) clh_nl_interp_adj(d, tab, l, x);  // C_used: the table value
  if (l < CLH_LMIN_TAB) atomicAdd(tab + (l - 1), -x);                  // C_L: the direct row
  else clh_nl_interp_adj(d, tab, l, -x);                               // C_L: the table value
}

hipError_t clh_launch_nl_cl_adj(const DrawDesc* D, const int2* wbin, int nbin, hipStream_t s) {
  if (nbin <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_nl_cl_adj, dim3(nbin), dim3(256), 0, s, D, wbin);
  return hipGetLastError();
}

// the tangent inputs, their transforms and the tangent passes (g_b1, g_lens)
hipError_t clh_launch_nl_tangent(const DrawDesc* D, const int2* wprep, int nprep, const int4* wfwd4, int nfwd4,
                                 const int4* widft, int nidft, hipStream_t s) {
  if (nprep <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_nl_prep_tan, dim3(nprep), dim3(256), 0, s, D, wprep);
  hipLaunchKernelGGL(k_nl_fwd<true>, dim3(nfwd4), dim3(256), 0, s, D, wfwd4);
  hipLaunchKernelGGL((k_nl_fft<false, true>), dim3(nidft), dim3(256), 0, s, D, widft);
  return hipGetLastError();
}

hipError_t clh_launch_nl_forward(const DrawDesc* D, const int2* wprep, int nprep, const int4* wfwd, int nfwd,
                                 const int4* wgl, int ngl, const int4* wchk, int nchk, const int2* wbin, int nbin,
                                 const int4* widft, int nidft, hipStream_t s) {
  if (nprep <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_nl_prep, dim3(nprep), dim3(256), 0, s, D, wprep);
  hipLaunchKernelGGL(k_nl_fwd<false>, dim3(nfwd), dim3(256), 0, s, D, wfwd);
  hipLaunchKernelGGL(k_nl_gl, dim3(ngl), dim3(256), 0, s, D, wgl);
  hipLaunchKernelGGL(k_nl_fft<true>, dim3(nchk), dim3(256), 0, s, D, wchk);
  hipLaunchKernelGGL(k_nl_lmax, dim3(nbin), dim3(64), 0, s, D, wbin);
  hipLaunchKernelGGL(k_nl_fft<false>, dim3(nidft), dim3(256), 0, s, D, widft);
  return hipGetLastError();
}

hipError_t clh_launch_nl_cl(const DrawDesc* D, const int2* work, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_nl_cl, dim3(n), dim3(256), 0, s, D, work);
  return hipGetLastError();
}
