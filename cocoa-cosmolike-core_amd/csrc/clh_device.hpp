// Device-side restatements of the per-point physics the projection path
// evaluates: background interpolation (cosmolike/cosmo3D.c), P(k,z) lookup,
// n(z) splines (cosmolike/redshift_spline.c), lens efficiencies and radial
// weights (cosmolike/radial_weights.c), NLA amplitude (cosmolike/IA.c).
#pragma once
#include <hip/hip_runtime.h>
#include "clh_internal.hpp"
#include "clh_search.hpp"

// chi_all: cosmo3D.c:40-92 (chi and dchi/da, in c/H0 units).  The table is
// validated on the host to extend two nodes beyond every queried z.
__device__ __forceinline__ void clh_chi_all(const DrawDesc& d, double a, double& chi,
                                            double& dchida) {
  const double z = 1.0 / a - 1.0;
  const double* __restrict__ zz = d.chi_z;
  const double* __restrict__ cc = d.chi_v;
  int j = clh_bs(zz, d.chi_n, z);
  if (j > d.chi_n - 3) j = d.chi_n - 3;  // host guarantees this never triggers
  const double dy = (z - zz[j]) / (zz[j + 1] - zz[j]);
  const double out0 = cc[j] + dy * (cc[j + 1] - cc[j]);
  const double up = (cc[j + 2] - cc[j]) / (zz[j + 2] - zz[j]);
  double down;
  if (j > 0) down = (cc[j + 1] - cc[j - 1]) / (zz[j + 1] - zz[j - 1]);
  else down = (cc[j + 1] - cc[j]) / (zz[j + 1] - zz[j]);
  double out1 = down + dy * (up - down);
  out1 = out1 / CLH_COVERH0;
  out1 = out1 / (a * a);
  chi = out0 / CLH_COVERH0;
  dchida = out1;
}

// chi(a) of clh_chi_all with its interval found by walking from j0 (the
// index of its bisection, whatever the start; j0 < 0: the bisection)
__device__ __forceinline__ double clh_chi_near(const DrawDesc& d, double a, int j0) {
  const double z = 1.0 / a - 1.0;
  const double* __restrict__ zz = d.chi_z;
  const double* __restrict__ cc = d.chi_v;
  int j = j0 < 0 ? clh_bs(zz, d.chi_n, z) : clh_walk(zz, d.chi_n, z, j0);
  if (j > d.chi_n - 3) j = d.chi_n - 3;
  const double dy = (z - zz[j]) / (zz[j + 1] - zz[j]);
  return (cc[j] + dy * (cc[j + 1] - cc[j])) / CLH_COVERH0;
}

__device__ __forceinline__ double clh_chi(const DrawDesc& d, double a) {
  double c, dc;
  clh_chi_all(d, a, c, dc);
  return c;
}

// 1 / x for the per-evaluation divisions of the node helpers: the hardware
// reciprocal and two Newton steps (each squares the relative error), within
// an ulp or two of the rounded quotient, about half the cost of the IEEE
// division sequence; x is a normal, finite number there.  FP32: the division.
__device__ __forceinline__ double clh_rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.0), r, r);
  return fma(fma(-x, r, 1.0), r, r);
}
__device__ __forceinline__ float clh_rcp(float x) { return 1.0f / x; }

// hoverh0v2: cosmo3D.c:104-107
__device__ __forceinline__ double clh_hoverh0v2(double a, double dchida) {
  return 1.0 / ((a * a) * dchida);
}

// a_chi: cosmo3D.c:109-133
__device__ __forceinline__ double clh_a_chi(const DrawDesc& d, double io_chi) {
  const double chi = io_chi * CLH_COVERH0;
  const int j = clh_bs(d.chi_v, d.chi_n, chi);
  const double dy = (chi - d.chi_v[j]) / (d.chi_v[j + 1] - d.chi_v[j]);
  const double z = d.chi_z[j] + dy * (d.chi_z[j + 1] - d.chi_z[j]);
  return 1.0 / (1.0 + z);
}

// norm_growfac(a, true): cosmo3D.c:148-198 (G(0) precomputed per draw)
__device__ __forceinline__ double clh_growfac(const DrawDesc& d, double a) {
  const double z = 1.0 / a - 1.0;
  const int j = clh_bs(d.G_z, d.G_n, z);
  const double dy = (z - d.G_z[j]) / (d.G_z[j + 1] - d.G_z[j]);
  const double G = d.G_v[j] + dy * (d.G_v[j + 1] - d.G_v[j]);
  return (G * a) / d.G0;
}

// f_growth: cosmo3D.c:200-228
__device__ __forceinline__ double clh_f_growth(const DrawDesc& d, double z) {
  const int j = clh_bs(d.G_z, d.G_n, z);
  const double dy = (z - d.G_z[j]) / (d.G_z[j + 1] - d.G_z[j]);
  const double G = d.G_v[j] + dy * (d.G_v[j + 1] - d.G_v[j]);
  const double dlnGdlnz = ((d.G_v[j + 1] - d.G_v[j]) / (d.G_z[j + 1] - d.G_z[j])) * z / G;
  const double dlnGdlna = -dlnGdlnz * (1 + z) / z;
  return 1 + dlnGdlna;
}

// p_nonlin without baryons: cosmo3D.c:362-417 (bilinear in log10 k, z of ln P;
// linear extrapolation outside the table)
__device__ __forceinline__ double clh_p_nonlin(const DrawDesc& d, double k, double a) {
  const double log10k = log10(k / CLH_COVERH0);
  const double z = 1.0 / a - 1.0;
  const int nz = d.nzp;
  const int i = clh_bs(d.lk, d.nk, log10k);
  const int j = clh_bs(d.pz, nz, z);
  const double dx = (log10k - d.lk[i]) / (d.lk[i + 1] - d.lk[i]);
  const double dy = (z - d.pz[j]) / (d.pz[j + 1] - d.pz[j]);
  const double* __restrict__ P = d.lnP + (size_t)i * nz + j;
  const double out = (1 - dx) * (1 - dy) * P[0] + (1 - dx) * dy * P[1] + dx * (1 - dy) * P[nz] +
                     dx * dy * P[nz + 1];
  return exp(out) / (CLH_COVERH0 * CLH_COVERH0 * CLH_COVERH0);
}

// natural cubic spline through histogram centres (gsl cspline evaluation)
__device__ __forceinline__ double clh_spline(const double* __restrict__ zc,
                                             const double* __restrict__ coef, int rows, double x) {
  int i = clh_bs(zc, rows, x);
  if (i > rows - 2) i = rows - 2;
  const double* c = coef + 4 * i;
  const double dx = x - zc[i];
  return c[0] + dx * (c[1] + dx * (c[2] + dx * c[3]));
}

// nz_source_photoz evaluation: redshift_spline.c:513-537
__device__ __forceinline__ double clh_nz_src(const DrawDesc& d, double zz, int b) {
  zz = zz - d.par->src_dz[b];
  const int n = d.s_rows;
  if (zz < d.s_zc[0] || zz > d.s_zc[n - 1]) return 0.0;
  return clh_spline(d.s_zc, d.s_coef + (size_t)b * 4 * (n - 1), n, zz);
}

// nz_lens_photoz evaluation: redshift_spline.c:736-762
__device__ __forceinline__ double clh_nz_lens(const DrawDesc& d, double zz, int b) {
  zz = (zz - d.par->lens_dz[b] - d.l_zmean[b]) / d.par->lens_st[b] + d.l_zmean[b];
  const int n = d.l_rows;
  if (zz <= d.l_zc[0] || zz >= d.l_zc[n - 1]) return 0.0;
  const double r = clh_spline(d.l_zc, d.l_coef + (size_t)b * 4 * (n - 1), n, zz);
  return r / d.par->lens_st[b];
}

// interpol1d: basics.c:457-480 (constant extrapolation)
__device__ __forceinline__ double clh_interpol1d(const double* __restrict__ f, int n, double a,
                                                 double dx, double x) {
  if (x < a) return f[0];
  const double r = (x - a) / dx;
  const int i = (int)floor(r);
  if (i + 1 >= n) return f[n - 1];
  return (r - i) * (f[i + 1] - f[i]) + f[i];
}

// g_tomo evaluation: redshift_spline.c:919-922
__device__ __forceinline__ double clh_g_tomo(const DrawDesc& d, double a, int b) {
  const double amin = 1.0 / (d.s_zmax_all + 1.0), amax = 0.999999;
  const double dac = (amax - amin) / ((double)d.n_a - 1.0);
  return (a <= amin || a > 1.0 - dac) ? 0.0
         : clh_interpol1d(d.gtab + (size_t)b * d.n_a, d.n_a, amin, dac, a);
}

// g_lens evaluation: redshift_spline.c:1106-1108
__device__ __forceinline__ double clh_g_lens(const DrawDesc& d, double a, int b) {
  const double amin = 1.0 / (d.l_zmax_all + 1.0), amax = 0.999999;
  const double dac = (amax - amin) / ((double)d.n_a - 1.0);
  return (a < amin || a > 1.0 - dac) ? 0.0
         : clh_interpol1d(d.gtab + (size_t)(d.ns + b) * d.n_a, d.n_a, amin, dac, a);
}

// IA_A1_Z1Z2 amplitude times Omega_m c1 rho_crit / D: IA.c:238-296
__device__ __forceinline__ double clh_ia_c1(const DrawDesc& d, double a, double growfac_a, int b) {
  double A = 0.0;
  if (d.ia_mode == 2) A = d.par->ia_bin[b];
  else if (d.ia_mode == 3) A = d.par->ia_A * pow((1.0 / a) / 1.62, d.par->ia_eta);
  return A * (d.omega_m * 0.01389 / growfac_a);
}

// IA_A2_Z1 (IA.c:305-368): A2 Omega_m c1rhocrit / D^2; redshift evolution
// A2 (1 / (a (1+z0)))^eta2
__device__ __forceinline__ double clh_ia_c2(const DrawDesc& d, double a, double growfac_a, int b) {
  double A = 0.0;
  if (d.ia_mode == 2) A = d.par->ia2_bin[b];
  else if (d.ia_mode == 3) A = d.par->ia2_A * pow(1.0 / (a * 1.62), d.par->ia2_eta);
  return A * (d.omega_m * 0.01389 / (growfac_a * growfac_a));
}

// IA_BTA_Z1 (IA.c:370-420): b_TA per bin or one value; 1 without IA
__device__ __forceinline__ double clh_ia_bta(const DrawDesc& d, int b) {
  if (d.ia_mode == 2) return d.par->bta_bin[b];
  if (d.ia_mode == 3) return d.par->bta_A;
  return 1.0;
}

// A FAST-PT table at ln k (cosmo2D.c:1414-1447, 2566-2593): uniform in ln k
// with the step (ln kmax - ln kmin) / N, zero outside the range, the last
// entry at its top end; linear between entries (interpol1d, basics.c:457-480)
__device__ __forceinline__ double clh_ptab(const double* __restrict__ tab, int n, double l0, double l1,
                                           double dl, double lnk) {
  if (lnk < l0 || lnk > l1) return 0.0;
  const double r = (lnk - l0) / dl;
  const int i = (int)floor(r);
  if (i + 1 >= n) return tab[n - 1];
  const double w = r - i;
  return w * (tab[i + 1] - tab[i]) + tab[i];
}
// FAST-PT IA table t (set_IA_PS)
__device__ __forceinline__ double clh_fpt(const DrawDesc& d, int t, double lnk) {
  return clh_ptab(d.fpt + (size_t)t * d.fpt_n, d.fpt_n, d.fpt_l0, d.fpt_l1, d.fpt_dl, lnk);
}
// FAST-PT bias table t (set_bias_PS)
__device__ __forceinline__ double clh_fptb(const DrawDesc& d, int t, double lnk) {
  return clh_ptab(d.fptb + (size_t)t * d.fptb_n, d.fptb_n, d.fptb_l0, d.fptb_l1, d.fptb_dl, lnk);
}

// one-loop bias term of the gg integrand without W_gal^2 (cosmo2D.c:2560-2603):
// g4 fac (b1 b2 d1d2 + b2^2/4 d2d2 + b1 bs2 d1s2 + b2 bs2/2 d2s2 + bs2^2/4 s2s2
// + b1 b3 d1p3) + 2 b1 bK k^2 P fac, with d2d2, d2s2, s2s2 less 2, 4/3, 8/9
// sigma4 inside the table range
__device__ __forceinline__ double clh_oneloop_gg(const DrawDesc& d, int zl, double lnk, double k, double g4fac,
                                                 double pfac) {
  const bool in = !(lnk < d.fptb_l0 || lnk > d.fptb_l1);
  const double s4 = d.fptb_s4;
  const double d1d2 = clh_fptb(d, 0, lnk);
  const double d2d2 = in ? clh_fptb(d, 1, lnk) - 2. * s4 : 0.0;
  const double d1s2 = clh_fptb(d, 2, lnk);
  const double d2s2 = in ? clh_fptb(d, 3, lnk) - 4. / 3. * s4 : 0.0;
  const double s2s2 = in ? clh_fptb(d, 4, lnk) - 8. / 9. * s4 : 0.0;
  const double d1p3 = clh_fptb(d, 5, lnk);
  const double b1 = d.par->b1[zl], b2 = d.par->b2[zl], bs2 = d.par->bs2[zl], b3 = d.par->b3[zl],
               bk = d.par->bk[zl];
  return g4fac * (b1 * b2 * d1d2 + 0.25 * b2 * b2 * d2d2 + b1 * bs2 * d1s2 + 0.5 * b2 * bs2 * d2s2 +
                  0.25 * bs2 * bs2 * s2s2 + b1 * b3 * d1p3) +
         2 * b1 * bk * k * k * pfac;
}

// The gg one-loop term (clh_oneloop_gg) split for the adjoint: tot, its
// P-proportional part 2 b1 bK k^2 P fac (the part with a ln P cotangent) and
// its derivative with respect to b1 at fixed bs2.
__device__ __forceinline__ void clh_oneloop_gg_parts(const DrawDesc& d, int zl, double lnk, double k, double g4fac,
                                                     double pfac, double& tot, double& pp, double& db1) {
  const bool in = !(lnk < d.fptb_l0 || lnk > d.fptb_l1);
  const double s4 = d.fptb_s4;
  const double d1d2 = clh_fptb(d, 0, lnk);
  const double d2d2 = in ? clh_fptb(d, 1, lnk) - 2. * s4 : 0.0;
  const double d1s2 = clh_fptb(d, 2, lnk);
  const double d2s2 = in ? clh_fptb(d, 3, lnk) - 4. / 3. * s4 : 0.0;
  const double s2s2 = in ? clh_fptb(d, 4, lnk) - 8. / 9. * s4 : 0.0;
  const double d1p3 = clh_fptb(d, 5, lnk);
  const double b1 = d.par->b1[zl], b2 = d.par->b2[zl], bs2 = d.par->bs2[zl], b3 = d.par->b3[zl],
               bk = d.par->bk[zl];
  pp = 2 * b1 * bk * k * k * pfac;
  tot = g4fac * (b1 * b2 * d1d2 + 0.25 * b2 * b2 * d2d2 + b1 * bs2 * d1s2 + 0.5 * b2 * bs2 * d2s2 +
                 0.25 * bs2 * bs2 * s2s2 + b1 * b3 * d1p3) +
        pp;
  db1 = g4fac * (b2 * d1d2 + bs2 * d1s2 + b3 * d1p3) + 2 * bk * k * k * pfac;
}

// one-loop bias term of the gs integrand times fac (cosmo2D.c:2027-2057):
// 0.5 g4 fac (b2 d1d2 + bs2 d1s2 + b3 d1p3) + bK k^2 P fac
__device__ __forceinline__ double clh_oneloop_gs(const DrawDesc& d, int zl, double lnk, double k, double g4fac,
                                                 double pfac) {
  return 0.5 * g4fac *
             (d.par->b2[zl] * clh_fptb(d, 0, lnk) + d.par->bs2[zl] * clh_fptb(d, 2, lnk) +
              d.par->b3[zl] * clh_fptb(d, 5, lnk)) +
         d.par->bk[zl] * k * k * pfac;
}

// W_RSD: radial_weights.c:87-105 (hoverh0 via dchi_dz, f via f_growth)
__device__ __forceinline__ double clh_w_rsd(const DrawDesc& d, double l, double a0, double a1,
                                            int ni) {
  double c0, dc0, c1, dc1;
  clh_chi_all(d, a0, c0, dc0);
  clh_chi_all(d, a1, c1, dc1);
  const double h0 = 1.0 / ((a0 * a0) * dc0);
  const double h1 = 1.0 / ((a1 * a1) * dc1);
  double w = (1 + 8. * l) / ((2. * l + 1.) * (2. * l + 1.)) * clh_nz_lens(d, 1. / a0 - 1., ni) * h0 *
             clh_f_growth(d, 1.0 / a0 - 1.0);
  w -= 4. / (2 * l + 3.) * sqrt((2 * l + 1.) / (2 * l + 3.)) * clh_nz_lens(d, 1. / a1 - 1., ni) *
       h1 * clh_f_growth(d, 1.0 / a1 - 1.0);
  return w;
}

__device__ __forceinline__ void clh_ss_pair(int ns, int q, int& i, int& j) {
  i = 0;
  while (q >= ns - i) { q -= ns - i; i++; }
  j = i + q;
}

// sum of v over the block (blockDim.x <= 256: four waves); lds holds 4 doubles
__device__ __forceinline__ double clh_block_sum(double v, double* lds) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  const double s = lds[0] + lds[1] + lds[2] + lds[3];
  __syncthreads();
  return s;
}

// amin_lens / amax_lens: redshift_spline.c:50-79
// |dz| of the lens limits (lim_eps: see DrawDesc)
__device__ __forceinline__ double clh_lim_adz(const DrawDesc& d, int ni) {
  const double dz = d.par->lens_dz[ni];
  return fabs(dz) + (dz != 0.0 ? d.lim_eps : 0.0);
}
__device__ __forceinline__ double clh_amin_lens(const DrawDesc& d, int ni) {
  const double zmax = (d.l_zmax[ni] - d.l_zmean[ni]) * d.par->lens_st[ni] + d.l_zmean[ni];
  return 1. / (1 + zmax + 2. * clh_lim_adz(d, ni));
}
__device__ __forceinline__ double clh_amax_lens(const DrawDesc& d, int ni) {
  const double zmin = (d.l_zmin[ni] - d.l_zmean[ni]) * d.par->lens_st[ni] + d.l_zmean[ni];
  if (d.par->bmag[ni] != 0) return 1. / (1. + fmax(d.s_zmin_all, 0.001));
  return 1. / (1 + fmax(zmin - 2. * clh_lim_adz(d, ni), 0.001));
}

// ------------------------------------------------ walked-index evaluations
// The same values as clh_chi_all (chi only), clh_nz_src / clh_nz_lens and
// their shift derivatives, for a thread that evaluates them along monotone z:
// the table index is walked from the previous call (start with j = -1).
// chi at z along monotone z (chi_all's linear chi, cosmo3D.c:40-92): the
// interval's knots and reciprocal width are kept until the walk moves
struct ClhChiWalk {
  int j = -1, jj = 0;
  double z0 = 0.0, c0 = 0.0, dc = 0.0, inv = 0.0;
  double lo = 0.0, hi = 0.0;  // chi_z[j], chi_z[j + 1]: the walk test without loads
  __device__ __forceinline__ double chi(const DrawDesc& d, double z) {
    const double* __restrict__ zz = d.chi_z;
    const int n = d.chi_n;
    // clh_walk keeps j exactly when this test holds
    const bool stay = j >= 0 && (z >= lo || j == 0) && (z < hi || j == n - 2);
    if (!stay) {
      j = (j < 0) ? clh_bs(zz, n, z) : clh_walk(zz, n, z, j);
      lo = zz[j];
      hi = zz[j + 1];
      jj = j > n - 3 ? n - 3 : j;
      const double* __restrict__ cc = d.chi_v;
      z0 = zz[jj];
      inv = 1.0 / (zz[jj + 1] - z0);
      c0 = cc[jj];
      dc = cc[jj + 1] - c0;
    }
    return (c0 + (z - z0) * inv * dc) * (1.0 / CLH_COVERH0);
  }
};

// ---------------------------------------------------------------- derivatives
// d/dx of the spline piece: b + dx (2c + 3 dx d)
__device__ __forceinline__ double clh_spline_d(const double* __restrict__ zc,
                                               const double* __restrict__ coef, int rows, double x) {
  int i = clh_bs(zc, rows, x);
  if (i > rows - 2) i = rows - 2;
  const double* c = coef + 4 * i;
  const double dx = x - zc[i];
  return c[1] + dx * (2.0 * c[2] + 3.0 * dx * c[3]);
}

// d n_src(z - dz_b) / d dz_b (zero where the forward value is clamped to zero)
__device__ __forceinline__ double clh_dnz_src(const DrawDesc& d, double zz, int b) {
  zz = zz - d.par->src_dz[b];
  const int n = d.s_rows;
  if (zz < d.s_zc[0] || zz > d.s_zc[n - 1]) return 0.0;
  return -clh_spline_d(d.s_zc, d.s_coef + (size_t)b * 4 * (n - 1), n, zz);
}

// d n_lens / d dz_b of nz_lens_photoz: -spline'(z') / s^2
__device__ __forceinline__ double clh_dnz_lens(const DrawDesc& d, double zz, int b) {
  const double s = d.par->lens_st[b];
  zz = (zz - d.par->lens_dz[b] - d.l_zmean[b]) / s + d.l_zmean[b];
  const int n = d.l_rows;
  if (zz <= d.l_zc[0] || zz >= d.l_zc[n - 1]) return 0.0;
  return -clh_spline_d(d.l_zc, d.l_coef + (size_t)b * 4 * (n - 1), n, zz) / (s * s);
}

// d g / d dz of the bin's own shift: same interpolation over the tangent table
__device__ __forceinline__ double clh_dg_tomo(const DrawDesc& d, double a, int b) {
  const double amin = 1.0 / (d.s_zmax_all + 1.0), amax = 0.999999;
  const double dac = (amax - amin) / ((double)d.n_a - 1.0);
  return (a <= amin || a > 1.0 - dac) ? 0.0
         : clh_interpol1d(d.dgtab + (size_t)b * d.n_a, d.n_a, amin, dac, a);
}
__device__ __forceinline__ double clh_dg_lens(const DrawDesc& d, double a, int b) {
  const double amin = 1.0 / (d.l_zmax_all + 1.0), amax = 0.999999;
  const double dac = (amax - amin) / ((double)d.n_a - 1.0);
  return (a < amin || a > 1.0 - dac) ? 0.0
         : clh_interpol1d(d.dgtab + (size_t)(d.ns + b) * d.n_a, d.n_a, amin, dac, a);
}

// ------------------------------------------------- data-vector placement
// point-mass term of the gamma_t pair (lens zl, source zs), without its
// 1/theta^2 (PointMass::get_pm, generic_interface.cpp:2626-2641; a_lens from
// the lens bin's zmean); TANGENT: its derivative with respect to the source
// bin's photo-z shift (through g_tomo)
template <bool TANGENT = false>
__device__ __forceinline__ double clh_pm_fac(const DrawDesc& d, int zl, int zs) {
  const double pm = d.par->pm[zl];
  if (pm == 0.0) return 0.0;
  const double a = 1.0 / (1.0 + d.l_zmean[zl]);
  const double g = TANGENT ? clh_dg_tomo(d, a, zs) : clh_g_tomo(d, a, zs);
  return 4.0 * M_PI * 1.6e-23 * pm * 1.e13 * g / (clh_chi(d, a) * a * a * a);
}

// 0 where the likelihood mask removes data-vector entry `canon` (index in the
// (0,1,2) block order)
__device__ __forceinline__ bool clh_masked(const DrawDesc& d, int canon) {
  return d.apply_calib && d.mask && d.mask[canon] == 0;
}

// d C1 / d A_IA for the redshift-evolution mode (C1 is linear in A)
__device__ __forceinline__ double clh_ia_c1_unit(const DrawDesc& d, double a, double growfac_a) {
  if (d.ia_mode != 3) return 0.0;
  return pow((1.0 / a) / 1.62, d.par->ia_eta) * (d.omega_m * 0.01389 / growfac_a);
}

// p_nonlin plus the table cell and bilinear weights it read (for the adjoint)
__device__ __forceinline__ double clh_p_nonlin_cells(const DrawDesc& d, double k, double a, int& cell,
                                                     double w[4]) {
  const double log10k = log10(k / CLH_COVERH0);
  const double z = 1.0 / a - 1.0;
  const int nz = d.nzp;
  const int i = clh_bs(d.lk, d.nk, log10k);
  const int j = clh_bs(d.pz, nz, z);
  const double dx = (log10k - d.lk[i]) / (d.lk[i + 1] - d.lk[i]);
  const double dy = (z - d.pz[j]) / (d.pz[j + 1] - d.pz[j]);
  cell = i * nz + j;
  w[0] = (1 - dx) * (1 - dy);
  w[1] = (1 - dx) * dy;
  w[2] = dx * (1 - dy);
  w[3] = dx * dy;
  const double* __restrict__ P = d.lnP + cell;
  const double out = w[0] * P[0] + w[1] * P[1] + w[2] * P[nz] + w[3] * P[nz + 1];
  return exp(out) / (CLH_COVERH0 * CLH_COVERH0 * CLH_COVERH0);
}

// index of the ss pair (i, j), i <= j, in the row-major upper triangle
__device__ __forceinline__ int clh_ss_index(int ns, int i, int j) {
  if (i > j) { const int t = i; i = j; j = t; }
  return i * ns - i * (i - 1) / 2 + (j - i);
}

// ----------------------------------------------------------- block helpers
__device__ __forceinline__ double clh_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// ------------------------------------------------------------- node blocks
// node block of an ss (kind CLH_W_SS) or gs (CLH_W_GS, lens bin zl) tile and
// the offset (in units of CLH_NP_SS) of its search-result columns
__device__ __forceinline__ const double* clh_sx_nodes(const DrawDesc& d, int kind, int zl) {
  return kind == CLH_W_SS ? d.nodes + d.node_ss : d.nodes + d.node_gs + (size_t)zl * d.node_gs_stride;
}
__device__ __forceinline__ int clh_sx_lut_col(const DrawDesc& d, int kind) {
  return kind == CLH_W_SS ? 3 + d.ns : 5 + d.ns;
}
__device__ __forceinline__ const double* clh_gg_nodes(const DrawDesc& d, int zl) {
  return d.nodes + d.node_gg + (size_t)zl * d.node_gg_stride;
}

// Row of the per-node ln P cotangent runs of ss GL node p (null when the
// caller did not ask for the ln P gradient).  The gs and gg nodes fold
// theirs in LDS bands (k_limber_sx_adj<..., GS>, k_limber_gg_adj).
__device__ __forceinline__ double* clh_run_row(const DrawDesc& d, int kind, int zl, int p) {
  (void)kind;
  (void)zl;
  if (!d.run) return nullptr;
  return d.run + (size_t)p * d.nk;
}

// stage a table of n doubles in LDS when it fits (cap entries), else use it
// in place; every thread of the block calls this, then the block syncs
__device__ __forceinline__ const double* clh_stage(const double* __restrict__ g, int n, double* lds,
                                                   int cap) {
  if (n > cap) return g;
  for (int i = threadIdx.x; i < n; i += blockDim.x) lds[i] = g[i];
  return lds;
}

// ------------------------------------------------------------- gg integrand
// Per-row constants of the gg integrand (ell = l + 1/2): log10(ell),
// chi_1/chi = (ell+1)/ell, the two W_RSD coefficients (radial_weights.c:98-103)
// and the magnification prefactor l(l+1)/ell^2.
struct ClhGgRow {
  double l10, rr1, c0, c1, pref, ell;
};

__device__ __forceinline__ ClhGgRow clh_gg_row(double l) {
  const double ell = l + 0.5;
  ClhGgRow R;
  R.l10 = log10(ell);
  R.rr1 = (ell + 1.0) / ell;
  R.c0 = (1 + 8. * ell) / ((2. * ell + 1.) * (2. * ell + 1.));
  R.c1 = 4. / (2 * ell + 3.) * sqrt((2 * ell + 1.) / (2 * ell + 3.));
  R.pref = l * (l + 1.) / (ell * ell);
  R.ell = ell;
  return R;
}

// One gg GL node evaluating the integrand (W_gal b1 + W_mag pref bmag +
// W_RSD)^2 P(ell/chi, a) fac (cosmo2D.c:2490-2666) row after row in
// increasing ell.  The a_0 half of W_RSD depends only on the node (chi_0 =
// chi) and is evaluated once; the a_1 half (chi_1 = chi (ell+1)/ell) moves
// monotonically toward the node, so every table it reads -- chi(z) for
// a_chi and the dchi/dz stencil (cosmo3D.c:40-133), G(z) for f_growth
// (cosmo3D.c:200-228), the lens n(z) spline (redshift_spline.c:736-762) and
// log10 k of P(k,z) (cosmo3D.c:362-417) -- is found by a walk from the
// previous row, with the reciprocal spacing of the interval recomputed only
// when the interval changes.  ADJ additionally returns d W_RSD / d dz_lens.
// The per-draw tables the gg integrand walks (chi(z), G(z), the lens n(z)
// spline of one bin, log10 k): in LDS when the block staged them, else the
// draw's own arrays.
struct ClhGgTabs {
  const double *cz, *cv, *gz, *gv, *lzc, *lcf, *lk;
};

// Stages the gg tables of lens bin zl into dyn (2 chi_n + 2 G_n + l_rows +
// 4 (l_rows - 1) doubles) when stage != 0.  Every thread of the block calls
// it; the block syncs before use.
__device__ __forceinline__ ClhGgTabs clh_gg_tabs(const DrawDesc& d, int zl, double* dyn, int stage,
                                                 double* s_lk) {
  ClhGgTabs T;
  const int nr = d.l_rows;
  const double* lcf = d.l_coef + (size_t)zl * 4 * (nr - 1);
  T.lk = clh_stage(d.lk, d.nk, s_lk, CLH_LDS_LK);
  if (!stage) {
    T.cz = d.chi_z; T.cv = d.chi_v; T.gz = d.G_z; T.gv = d.G_v; T.lzc = d.l_zc; T.lcf = lcf;
    return T;
  }
  double* p = dyn;
  auto cp = [&](const double* src, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = src[i];
    const double* r = p;
    p += n;
    return r;
  };
  T.cz = cp(d.chi_z, d.chi_n);
  T.cv = cp(d.chi_v, d.chi_n);
  T.gz = cp(d.G_z, d.G_n);
  T.gv = cp(d.G_v, d.G_n);
  T.lzc = cp(d.l_zc, nr);
  T.lcf = cp(lcf, 4 * (nr - 1));
  return T;
}

// The P(k,z) lookup of one GL node (fixed z column j, weight dy) along
// increasing log10 k (p_nonlin, cosmo3D.c:362-417): the bracketing log10 k
// values, the reciprocal spacing and the four ln P values of the current cell
// stay in registers; memory is read only when log10 k leaves the cell (the
// index is the one clh_bs would return).  Successive rows mostly move one
// cell up, so with AHEAD the log10 k knot and the ln P pair of row i + 2 are
// loaded one move ahead: a one-cell step then needs no load to complete (the
// values are the same ones set_cell would read).  The look-ahead costs 6
// VGPRs, so kernels near an occupancy step leave it off.
// TAN: the cell's ln P differences along z at its two k knots, M0 = ln P[i][j+1]
// - ln P[i][j] and M1 = ln P[i+1][j+1] - ln P[i+1][j], are kept as well, so
// that d ln P / d dy = (1 - dx) M0 + dx M1 (and d ln P / d log10 k =
// (L1 - L0) inv_k) is at hand for the tangent of the lookup.
// R: the arithmetic type of the walk and the interpolation (float in the
// FP32 integrands; the table stays FP64 in memory).
template <bool AHEAD = true, bool TAN = false, typename R = double>
struct ClhPkNodeT {
  int i, j, nk, nz;
  R M0, M1;
  // cell state: lo / hi bound the walk test (-inf / +inf at the table ends,
  // where clh_bs extrapolates), x0 = lk[i], inv_k = 1 / (lk[i+1] - lk[i]),
  // L0 / L1 the ln P interpolated in z at lk[i] / lk[i+1], so that
  // (1-dx)(1-dy)p00 + (1-dx)dy p01 + dx(1-dy)p10 + dx dy p11 = L0 + dx (L1 - L0)
  R dy, lo, hi, x0, inv_k, L0, L1;
  R nhi, n0, n1;  // lk[i + 2], ln P[i + 2][j], ln P[i + 2][j + 1] (i + 2 <= nk - 1)
  const double* __restrict__ lnP;

  __device__ __forceinline__ R zint(const double* __restrict__ P) const {
    return (1 - dy) * (R)P[0] + dy * (R)P[1];
  }
  __device__ __forceinline__ void ahead(const double* __restrict__ lk) {
    if (AHEAD && i + 2 <= nk - 1) {
      // used only at the next step, so their latency overlaps the rows between
      nhi = lk[i + 2];
      const double* __restrict__ P = lnP + (size_t)(i + 2) * nz + j;
      n0 = P[0];
      n1 = P[1];
    }
  }
  __device__ __forceinline__ void bounds(const double* __restrict__ lk) {
    x0 = lk[i];
    const R x1 = lk[i + 1];
    inv_k = clh_rcp(x1 - x0);
    lo = i == 0 ? -INFINITY : x0;
    hi = i == nk - 2 ? INFINITY : x1;
  }

  __device__ __forceinline__ void init(const DrawDesc& d, int j_, R dy_) {
    i = -1;
    lo = INFINITY;  // empty cell: the first lookup searches
    hi = -INFINITY;
    j = j_;
    dy = dy_;
    nk = __builtin_amdgcn_readfirstlane(d.nk);  // the draw's: equal in every lane
    nz = __builtin_amdgcn_readfirstlane(d.nzp);
    lnP = d.lnP;
  }
  __device__ __forceinline__ void set_cell(const double* __restrict__ lk) {
    bounds(lk);
    const double* __restrict__ P = lnP + (size_t)i * nz + j;
    L0 = zint(P);
    L1 = zint(P + nz);
    if (TAN) {
      M0 = P[1] - P[0];
      M1 = P[nz + 1] - P[nz];
    }
    ahead(lk);
  }
  // ln P at log10 k; dx: the log10 k weight of the cell (i, j)
  __device__ __forceinline__ R lnp(const double* __restrict__ lk, R log10k, R& dx) {
    if (!(log10k >= lo && log10k < hi)) {
      if (i < 0) {
        i = clh_bs(lk, nk, log10k);
        set_cell(lk);
      } else if (AHEAD && log10k >= hi && (i + 1 == nk - 2 || log10k < nhi)) {
        // one cell up: the clh_walk index is i + 1
        i++;
        x0 = hi;  // finite: the old cell was not the last
        lo = x0;
        hi = i == nk - 2 ? INFINITY : nhi;
        inv_k = clh_rcp(nhi - x0);
        L0 = L1;
        L1 = (1 - dy) * n0 + dy * n1;
        if (TAN) {
          M0 = M1;
          M1 = n1 - n0;
        }
        ahead(lk);
      } else {
        i = clh_walk(lk, nk, log10k, i);
        set_cell(lk);
      }
    }
    dx = (log10k - x0) * inv_k;
    return L0 + dx * (L1 - L0);
  }
  // TAN: d ln P / d log10 k and d ln P / d dy at the last lookup (weight dx)
  __device__ __forceinline__ R dlnp_dk() const { return (L1 - L0) * inv_k; }
  __device__ __forceinline__ R dlnp_dy(R dx) const { return (1 - dx) * M0 + dx * M1; }
};

typedef ClhPkNodeT<true> ClhPkNode;

// The log10 k cell of a P(k,z) lookup along increasing k without the ln P
// values: the same index and weight dx as ClhPkNode::lnp.
struct ClhKCell {
  int i = -1, nk;
  double lo, hi, inv_k;
  __device__ __forceinline__ double dx(const double* __restrict__ lk, double log10k) {
    if (i < 0) {
      i = clh_bs(lk, nk, log10k);
      set(lk);
    } else if (!((log10k >= lo || i == 0) && (log10k < hi || i == nk - 2))) {
      i = clh_walk(lk, nk, log10k, i);
      set(lk);
    }
    return (log10k - lo) * inv_k;
  }
  __device__ __forceinline__ void set(const double* __restrict__ lk) {
    lo = lk[i];
    hi = lk[i + 1];
    inv_k = 1.0 / (hi - lo);
  }
};

// One bracketed walk over a strictly increasing table: the index clh_bs
// would return for v, found from the previous one; true when it moved.
struct ClhBracket {
  int j;
  double lo, hi;
  __device__ __forceinline__ void set(const double* __restrict__ x, int j_) {
    j = j_;
    lo = x[j];
    hi = x[j + 1];
  }
  __device__ __forceinline__ bool move(const double* __restrict__ x, int n, double v) {
    if ((v >= lo || j == 0) && (v < hi || j == n - 2)) return false;
    set(x, clh_walk(x, n, v, j));
    return true;
  }
};

// n and d n / d dz of one source or lens bin along a run of z (the fine grid
// of k_gtab): the knot interval's bounds and spline coefficients stay in
// registers and are read again only when z leaves the interval (the index
// clh_bs would return).  Source bins: the spline at z - dz; lens bins: at
// (z - dz - zmean) / stretch + zmean, divided by the stretch (clh_nz_lens),
// the stretch's reciprocal taken once; zero outside the knots.
template <bool DER>
struct ClhNzRun {
  const double* __restrict__ zc;
  const double* __restrict__ cf;
  int rows;
  bool lens, have = false;
  double dz, is, zm;  // shift, 1 / stretch, zmean
  ClhBracket br;
  double c0 = 0.0, c1 = 0.0, c2 = 0.0, c3 = 0.0;
  __device__ __forceinline__ void init(const DrawDesc& d, bool src, int b) {
    lens = !src;
    if (src) {
      zc = d.s_zc;
      rows = d.s_rows;
      cf = d.s_coef + (size_t)b * 4 * (rows - 1);
      dz = d.par->src_dz[b];
      is = 1.0;
      zm = 0.0;
    } else {
      zc = d.l_zc;
      rows = d.l_rows;
      cf = d.l_coef + (size_t)b * 4 * (rows - 1);
      dz = d.par->lens_dz[b];
      is = 1.0 / d.par->lens_st[b];
      zm = d.l_zmean[b];
    }
  }
  __device__ __forceinline__ void coef() {
    const double* __restrict__ c = cf + 4 * br.j;
    c0 = c[0];
    c1 = c[1];
    c2 = c[2];
    c3 = c[3];
  }
  __device__ __forceinline__ void eval(double z, double& n, double& dn) {
    n = dn = 0.0;
    double zz;
    if (lens) {
      zz = (z - dz - zm) * is + zm;
      if (zz <= zc[0] || zz >= zc[rows - 1]) return;
    } else {
      zz = z - dz;
      if (zz < zc[0] || zz > zc[rows - 1]) return;
    }
    if (!have) {
      br.set(zc, clh_bs(zc, rows, zz));
      have = true;
      coef();
    } else if (br.move(zc, rows, zz)) {
      coef();
    }
    const double x = zz - br.lo;
    if (lens) {
      n = (c0 + x * (c1 + x * (c2 + x * c3))) * is;
      if (DER) dn = -(c1 + x * (2.0 * c2 + 3.0 * x * c3)) * (is * is);
    } else {
      n = c0 + x * (c1 + x * (c2 + x * c3));
      if (DER) dn = -(c1 + x * (2.0 * c2 + 3.0 * x * c3));
    }
  }
};

// TAN (with ADJ): eval also leaves in dhn the derivative of the a_1 half's
// hf1 n1 (W_RSD = c0 t0 - c1 hf1 n1) with respect to the node's chi at the
// row, on the row's table intervals, and the P(k,z) walk keeps its z
// differences: the pieces of the integrand's derivative along the node's a.
// R: the arithmetic type of the integrand (float in the FP32 mode's gg
// kernels: tables and records stay FP64 in memory, the walks compare in R).
template <bool ADJ, bool AHEAD = true, bool TAN = false, typename R = double>
struct ClhGgNode {
  static_assert(ADJ || !TAN, "the tangent uses the adjoint's n(z) derivative");
  // row-independent values of the node
  R fK, wgb, wmb, t0, t0d, c3fac, l10fk, chi_cut;  // chi_cut = chi(a_min) in Mpc
  R inv_st, zs0, zlo, zhi;  // lens n(z) argument zs = z1 inv_st + zs0; first and last knot
  // The a_1 side walks over the chi(z) table (a_chi), its z column (the
  // dchi/dz stencil), the G(z) table and the lens n(z) knots.  Per table the
  // current interval j, the bounds of the walk test (-inf / +inf at the table
  // ends, where the reference's bisection extrapolates) and the interval's
  // coefficients, read from the draw's interval records (DrawDesc::crec,
  // grec) when the interval changes.  Successive rows move chi_1 down, so a
  // change is almost always one interval down: one record load, no search.
  int jc, jg, jn, chi_n, G_n, l_rows, same;
  bool alive;  // chi_1 <= chi(a_min) on the current chi interval
  const double* __restrict__ crec;
  const double* __restrict__ grec;
  const double* __restrict__ nrec;  // this lens bin's n(z) interval records
  R clo, chi_, czc, zA, zS;         // chi interval: z1 = zA + c1v zS; czc = z at its lower end
  R D0, D1;                         // z interval (the chi interval's): dchi/dz = D0 + z1 D1
  R glo, ghi, G0, slG;              // G interval (own bounds only off the chi grid): G = G0 + z1 slG
  R nlo, nhi_, cf0, cf1, cf2, cf3;  // lens spline interval: knots, coefficients / stretch
  R dhn;                            // TAN: d(hf1 n1) / d chi of the last row
  ClhPkNodeT<AHEAD, TAN, R> pk;

  __device__ __forceinline__ void rec_z(const double* __restrict__ r) {  // the z interval's coefficients
    D0 = r[4];
    D1 = r[5];
    if (same) {
      G0 = r[6];
      slG = r[7];
    }
  }
  // chi interval jc from the records.  Its z interval, for the dchi/dz
  // stencil and (same grid) G(z), has the same index: z1 interpolated on the
  // chi interval lies between the interval's z knots, the reference's second
  // bisection (on z1) can differ only when z1 rounds onto a knot, where the
  // stencil is continuous.  The upper bound is capped at the chi(a_min) cut
  // (cosmo2D.c:2512-2514), beyond which the integrand is zero.
  __device__ __forceinline__ void set_c(R c1v) {
    if (c1v > chi_cut) {  // beyond the cut until chi_1 falls below it
      alive = false;
      clo = chi_cut;
      chi_ = INFINITY;
      return;
    }
    alive = true;
    const double* __restrict__ r = crec + 8 * (size_t)jc;
    clo = jc == 0 ? -INFINITY : r[0];
    chi_ = fmin(jc == chi_n - 2 ? (R)INFINITY : (R)r[8], chi_cut);
    czc = r[1];
    zA = r[2];
    zS = r[3];
    rec_z(r);
  }
  __device__ __forceinline__ void move_c(const ClhGgTabs& T, R c1v) {
    if (alive && c1v < clo && jc > 0) {  // one interval down: the upper bound is the old lower one
      const double* __restrict__ r = crec + 8 * (size_t)(jc - 1);
      const R lo1 = r[0];
      if (jc == 1 || c1v >= lo1) {
        chi_ = clo;
        jc--;
        clo = jc == 0 ? -INFINITY : lo1;
        czc = r[1];
        zA = r[2];
        zS = r[3];
        rec_z(r);
        return;
      }
    }
    jc = clh_walk(T.cv, chi_n, c1v, jc);
    set_c(c1v);
  }
  __device__ __forceinline__ void set_g() {
    const double* __restrict__ r = grec + 3 * (size_t)jg;
    glo = jg == 0 ? -INFINITY : r[0];
    ghi = jg == G_n - 2 ? INFINITY : r[3];
    G0 = r[1];
    slG = r[2];
  }
  __device__ __forceinline__ void rec_n(const double* __restrict__ r) {
    cf0 = r[1];
    cf1 = r[2];
    cf2 = r[3];
    cf3 = r[4];
  }
  // the lens n(z) interval of zs: a spline piece inside the knots, zero
  // (and an interval reaching to -inf / +inf) outside them
  __device__ __forceinline__ void set_n(const ClhGgTabs& T, R zs) {
    if (zs <= zlo || zs >= zhi) {
      nlo = zs <= zlo ? -INFINITY : zhi;
      nhi_ = zs <= zlo ? zlo : INFINITY;
      cf0 = cf1 = cf2 = cf3 = 0.0;
      jn = l_rows - 1;  // above the knots, one step down is the last interval
      return;
    }
    jn = clh_walk(T.lzc, l_rows, zs, jn);
    const double* __restrict__ r = nrec + 5 * (size_t)jn;
    nlo = r[0];
    nhi_ = r[5];
    rec_n(r);
  }
  __device__ __forceinline__ void move_n(const ClhGgTabs& T, R zs) {
    if (zs < nlo && nlo > zlo) {  // one knot down inside the knots (zs falls as chi_1 does)
      const double* __restrict__ r = nrec + 5 * (size_t)(jn - 1);
      const R lo1 = r[0];
      if (zs >= lo1 && zs > zlo) {
        nhi_ = nlo;
        nlo = lo1;
        jn--;
        rec_n(r);
        return;
      }
    }
    set_n(T, zs);
  }

  // o: the node block (a, fK, fac, W_gal, W_mag, then the search results of
  // k_nodes: P(k,z) z index and weight, log10(chi c/H0), the a_0 side of W_RSD
  // and the table indices at a_0 from which the walks start)
  __device__ __forceinline__ void init(const DrawDesc& d, const ClhGgTabs& T, const double* __restrict__ o,
                                      int p, int zl) {
    constexpr int np = CLH_NP_GG;
    constexpr double c3 = 1.0 / (CLH_COVERH0 * CLH_COVERH0 * CLH_COVERH0);
    chi_n = d.chi_n;
    G_n = d.G_n;
    l_rows = d.l_rows;
    same = d.zg_same;
    crec = d.crec;
    grec = d.grec;
    nrec = d.nrec + (size_t)zl * 5 * d.l_rows;
    fK = o[np + p];
    c3fac = c3 * o[2 * np + p];
    wgb = o[3 * np + p] * d.par->b1[zl];
    wmb = o[4 * np + p] * d.par->bmag[zl];
    const double* lut = o + 5 * np;
    pk.init(d, (int)lut[p], lut[np + p]);
    l10fk = lut[2 * np + p];
    jc = (int)lut[3 * np + p];
    jg = (int)lut[5 * np + p];
    jn = (int)lut[6 * np + p];
    t0 = lut[7 * np + p];
    t0d = (ADJ && !TAN) ? lut[8 * np + p] : 0.0;
    inv_st = 1.0 / d.par->lens_st[zl];
    const R zm = d.l_zmean[zl];
    zs0 = (-d.par->lens_dz[zl] - zm) * inv_st + zm;
    zlo = d.l_zc[0];
    zhi = d.l_zc[l_rows - 1];
    chi_cut = d.chi_amin * CLH_COVERH0;
    alive = false;   // empty intervals: the first row walks from the a_0 indices
    clo = INFINITY;
    chi_ = -INFINITY;
    nlo = -INFINITY;  // (not a one-knot step either)
    nhi_ = -INFINITY;
    czc = zA = zS = D1 = slG = 0.0;  // finite until a row lies inside the cut
    D0 = G0 = 1.0;
    cf0 = cf1 = cf2 = cf3 = 0.0;
    if (!same) set_g();
  }

  // false: chi_1 beyond chi(a_min), integrand zero (cosmo2D.c:2512-2514).
  // t: the bracket; pf = P fac; dx: log10 k weight of the P cell (pk.i, pk.j).
  // T: the draw's tables (staged in LDS when they fit).  F32 (R = double):
  // P(k,z) as expf of the FP64 bilinear value; R = float: all in FP32.  Each interval is the one the reference's
  // bisection finds; inside it the interpolations are the interval's linear
  // (n(z): cubic) pieces.
  template <bool F32 = false>
  __device__ __forceinline__ bool eval(const DrawDesc& d, const ClhGgTabs& T, const ClhGgRow& row,
                                       R& t, R& pf, R& dwr, R& dx) {
    constexpr R C = CLH_COVERH0;
    // a_chi(chi_1) (cosmo3D.c:109-133), chi_1 = chi (ell+1)/ell
    const R c1v = fK * (R)row.rr1 * C;
    if (!(c1v >= clo && c1v < chi_)) move_c(T, c1v);
    // dchi/dz at z1 (cosmo3D.c:64-81 stencils) on the same interval
    const R z1 = zA + c1v * zS;
    // G(z1) and f_growth (cosmo3D.c:200-228): hoverh0 f_growth =
    // (c/H0) / (dchi/dz) (1 - G'(1+z)/G), where G - G'(1+z) = G0 - slG on the interval
    if (!same && !(z1 >= glo && z1 < ghi)) {
      jg = clh_walk(T.gz, G_n, z1, jg);
      set_g();
    }
    const R DG = D0 + z1 * D1, GG = G0 + z1 * slG;
    R hf1, inv = 0;
    if (TAN) {
      inv = clh_rcp(DG * GG);
      hf1 = C * (G0 - slG) * inv;
    } else {
      hf1 = C * (G0 - slG) * clh_rcp(DG * GG);
    }
    // nz_lens(z1) (redshift_spline.c:736-762) and its shift derivative
    const R zs = z1 * inv_st + zs0;
    if (!(zs >= nlo && zs < nhi_)) move_n(T, zs);
    const R x = nlo > -INFINITY ? zs - nlo : R(0);  // below the knots: zero coefficients
    const R n1 = cf0 + x * (cf1 + x * (cf2 + x * cf3));
    R dn1 = 0;
    if (ADJ) dn1 = -(cf1 + x * (R(2) * cf2 + R(3) * x * cf3)) * inv_st;
    if (TAN) {
      // d(hf1 n1)/d z1 = hf1 (n1' - n1 (D1 / DG + slG / GG)), n1' = -dn1;
      // z1 = zA + chi rr1 (c/H0) zS on the chi interval
      dhn = hf1 * (-dn1 - n1 * (D1 * GG + slG * DG) * inv) * (zS * (R)row.rr1 * C);
    }
    const R c0 = row.c0, c1 = row.c1;
    const R wr = c0 * t0 - c1 * hf1 * n1;
    // (TAN: without the a_0 half c0 t0d, which the caller adds per node)
    if (ADJ) dwr = (TAN ? R(0) : c0 * t0d) - c1 * hf1 * dn1;
    t = wgb + (R)row.pref * wmb + wr;
    const R out = pk.lnp(T.lk, (R)row.l10 - l10fk, dx);
    if constexpr (sizeof(R) == 4) pf = expf(out) * c3fac;
    else if (F32) pf = (R)(expf((float)out) * (float)c3fac);
    else pf = exp(out) * c3fac;
    return alive;
  }
};
