/*
 * cl_oracle.c -- TEST INFRASTRUCTURE ONLY (parity oracle + CPU baseline).
 *
 * FP64 CPU restatement of CosmoLike's tomographic two-point projection path.
 * Only tests/, the smoke check and the benchmark's CPU-baseline leg load it;
 * the HIP library never does.  Each routine cites the reference routine whose
 * algorithm it restates.  Third-party arithmetic the reference delegates to
 * GSL (version not pinned anywhere in the reference tree) is restated from the
 * published algorithms:
 *   - gsl_integration_glfixed: n-point Gauss-Legendre on [a,b]; nodes here come
 *     from Newton iteration on P_n (pinned against numpy.leggauss and
 *     polynomial exactness in the tests);
 *   - gsl_interp_cspline: natural cubic spline (pinned against
 *     scipy CubicSpline(bc_type='natural'));
 *   - gsl_sf_legendre_Pl_deriv_array: Bonnet upward recurrence for P_l and
 *     (1-x^2) P'_l = l (P_{l-1} - x P_l) (pinned against mpmath).
 */
#include "cl_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OC_MAXB 20
#define OC_LMIN_TAB 20                 /* structs.c:44 */
#define OC_A_MIN (1.0 / (1.0 + 40.0))  /* structs.c:41 */
#define OC_COVERH0 2997.92458          /* structs.c:119 */

struct oc_ctx {
  int n_ell, n_a, lmax, hdi;
  int ntheta;
  double vtmin, vtmax;
  double omega_m;
  int chi_nz, G_nz, nk, nzp;
  double *chi_z, *chi_v, *G_z, *G_v, *lk, *pz, *lnP;
  /* source sample: histogram + spline through bin centres */
  int ns, s_rows;
  double *s_tab;  /* [ns+1][rows]; row ns = z column */
  double s_zmin_all, s_zmax_all, s_zmin[OC_MAXB], s_zmax[OC_MAXB];
  double *s_zc, *s_y, *s_c; /* centres [rows]; y,c [ns+1][rows] */
  /* lens sample */
  int nl, l_rows;
  double *l_tab;
  double l_zmin_all, l_zmax_all, l_zmin[OC_MAXB], l_zmax[OC_MAXB], l_zmean[OC_MAXB];
  double *l_zc, *l_y, *l_c;
  /* nuisance */
  double src_dz[OC_MAXB], lens_dz[OC_MAXB], lens_st[OC_MAXB], b1[OC_MAXB],
      bmag[OC_MAXB], shear_m[OC_MAXB];
  double pm[OC_MAXB]; /* point-mass amplitudes per lens bin (PointMass::pm_) */
  int ggl_ex[2 * OC_MAXB * OC_MAXB]; /* excluded (lens, source) pairs (tomo.ggl_exclude) */
  int nlb_set;  /* 1: the one-loop bias term follows nlb_on (the integrands' nonlinear_bias
                   argument), 0: has_b2_galaxies as the nointerp integrals pass it */
  int nlb_on;
  int n_ggl_ex;
  int* mask;          /* likelihood mask in the (0,1,2) block order, or NULL */
  int mask_n;
  int ia_mode;
  double ia_bin[OC_MAXB], ia_A, ia_eta, oneplusz0, c1rhocrit;
  /* TATT (IA_MODEL = 1): nuisance.ia[1][*] (A2), ia[2][*] (b_TA) and the
     FAST-PT IA tables of set_IA_PS (FPTIA.tab[0..9], k in c/H0 units) */
  int ia_model;
  double ia2_bin[OC_MAXB], ia2_A, ia2_eta, bta_bin[OC_MAXB], bta_A;
  double* fpt; /* [10][fpt_n] */
  int fpt_n;
  double fpt_kmin, fpt_kmax;
  /* non-Limber w (C_cl_tomo): h0, linear P(k,z) (set_linear_power_spectrum)
     and like.adopt_limber_gg (1 here by default; the reference default 0) */
  double h0;
  int nkL, nzL;
  double *lkL, *pzL, *lnPL;
  int limber_gg;
  int nl_lmax[OC_MAXB]; /* LMAX[i] of the last C_cl_tomo run */
  /* one-loop galaxy bias: nuisance.gb[1..3,5] (b2, bs2, b3, bK) and the
     FAST-PT bias tables of set_bias_PS (FPTbias.tab[0..5], k in c/H0 units) */
  double b2[OC_MAXB], bs2[OC_MAXB], b3[OC_MAXB], bk[OC_MAXB];
  double* fptb; /* [6][fptb_n] */
  int fptb_n;
  double fptb_kmin, fptb_kmax, fptb_s4;
  /* CMB lensing cross-correlations (init_cmb_cross_correlation): beam FWHM
     (rad), ell range of the beam filter, HEALPix pixel window (optional) */
  double cmb_fwhm;
  int cmb_lmin, cmb_lmax, cmb_nwin;
  double* cmb_win;
  /* kk band powers (init_cmb_auto_bandpower): binning matrix
     [nbp][lmax - lmin + 1] and theory offset [nbp]; nbp = 0: C_kk at like.ell */
  int kk_nbp, kk_lmin, kk_lmax;
  double *kk_M, *kk_off;
  /* caches */
  int g_valid;
  double *gt_tab, *gl_tab; /* [ns][N_a], [nl][N_a] */
  int K_valid;
  int K_ld;  /* 1: kernels evaluated in long double (default), 0: in double */
  double *K; /* [4][ntheta][lmax] */
  int status;
  char err[256];
};

static void oc_set_err(oc_ctx* c, const char* fmt, ...) {
  if (c->status == 0) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c->err, sizeof c->err, fmt, ap);
    va_end(ap);
  }
  c->status = -1;
}

const char* oc_last_error(oc_ctx* c) { return c->err; }

oc_ctx* oc_create(void) {
  oc_ctx* c = (oc_ctx*)calloc(1, sizeof(oc_ctx));
  if (!c) return NULL;
  /* defaults: structs.c:262-279, reset_nuisance_struct (structs.c:207-237) */
  c->n_ell = 300;
  c->n_a = 225;
  c->lmax = 50000;
  c->hdi = 0;
  c->c1rhocrit = 0.01389;
  c->K_ld = 1;
  c->oneplusz0 = 1.62; /* set_nuisance_IA, generic_interface.cpp:1756 */
  c->h0 = 0.7;
  c->limber_gg = 1;
  for (int i = 0; i < OC_MAXB; i++) c->lens_st[i] = 1.0;
  return c;
}

void oc_free(oc_ctx* c) {
  if (!c) return;
  free(c->chi_z); free(c->chi_v); free(c->G_z); free(c->G_v);
  free(c->lk); free(c->pz); free(c->lnP);
  free(c->s_tab); free(c->s_zc); free(c->s_y); free(c->s_c);
  free(c->l_tab); free(c->l_zc); free(c->l_y); free(c->l_c);
  free(c->gt_tab); free(c->gl_tab); free(c->K);
  free(c->mask);
  free(c->fpt);
  free(c->lkL); free(c->pzL); free(c->lnPL);
  free(c->fptb);
  free(c->cmb_win);
  free(c->kk_M); free(c->kk_off);
  free(c);
}

void oc_touch(oc_ctx* c) { c->g_valid = 0; }

static double* dup(const double* x, int n) {
  double* y = (double*)malloc(sizeof(double) * (size_t)n);
  memcpy(y, x, sizeof(double) * (size_t)n);
  return y;
}

static int bsearch_lo(const double* x, int n, double v) {
  /* the ilo/ihi bisection used throughout cosmo3D.c (e.g. :46-57) */
  size_t ilo = 0, ihi = (size_t)n - 1;
  while (ihi > ilo + 1) {
    size_t ll = (ihi + ilo) / 2;
    if (x[ll] > v) ihi = ll; else ilo = ll;
  }
  return (int)ilo;
}

/* ---------------------------------------------------------------------- */
/* Gauss-Legendre (restates the published gsl_integration_glfixed rule)    */
/* ---------------------------------------------------------------------- */
int oc_gl_nodes(int n, double* x, double* w) {
  const int m = (n + 1) / 2;
  for (int i = 0; i < m; i++) {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5));
    double pp = 0.0;
    for (int it = 0; it < 100; it++) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; j++) {
        double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      double z1 = z;
      z = z1 - p1 / pp;
      if (fabs(z - z1) < 1e-16) break;
    }
    /* recompute derivative at the converged node */
    {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; j++) {
        double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
    }
    const double wi = 2.0 / ((1.0 - z * z) * pp * pp);
    x[i] = -z;          /* ascending order */
    x[n - 1 - i] = z;
    w[i] = wi;
    w[n - 1 - i] = wi;
  }
  if (n % 2 == 1) x[n / 2] = 0.0;
  return 0;
}

typedef struct { int n; double x[1024]; double w[1024]; } gl_rule;

static void gl_make(gl_rule* r, int n) { r->n = n; oc_gl_nodes(n, r->x, r->w); }

static void gl_point(const gl_rule* r, double a, double b, int p, double* xi, double* wi) {
  const double A = 0.5 * (b - a), B = 0.5 * (a + b);
  *xi = B + A * r->x[p];
  *wi = A * r->w[p];
}

/* GL order selection: cosmo2D.c:1577-1581 (ss/gs) and 2634-2638 (gg) */
static int gl_order_ssgs(int hdi) {
  hdi = abs(hdi);
  return hdi == 0 ? 64 : hdi == 1 ? 96 : hdi == 2 ? 128 : hdi == 3 ? 256 : 512;
}
static int gl_order_gg(int hdi) {
  hdi = abs(hdi);
  return hdi == 0 ? 96 : hdi == 1 ? 128 : hdi == 2 ? 256 : hdi == 3 ? 512 : 1024;
}

/* ---------------------------------------------------------------------- */
/* natural cubic spline (restates gsl_interp_cspline)                      */
/* ---------------------------------------------------------------------- */
static void spline_init(int n, const double* x, const double* y, double* c) {
  /* c = second derivative / 2; natural: c[0] = c[n-1] = 0 */
  c[0] = 0.0;
  c[n - 1] = 0.0;
  const int sys = n - 2;
  if (sys <= 0) return;
  double* diag = (double*)malloc(sizeof(double) * sys);
  double* off = (double*)malloc(sizeof(double) * sys);
  double* g = (double*)malloc(sizeof(double) * sys);
  for (int i = 0; i < sys; i++) {
    const double h_i = x[i + 1] - x[i];
    const double h_ip1 = x[i + 2] - x[i + 1];
    const double ydiff_i = y[i + 1] - y[i];
    const double ydiff_ip1 = y[i + 2] - y[i + 1];
    const double g_i = (h_i != 0.0) ? 1.0 / h_i : 0.0;
    const double g_ip1 = (h_ip1 != 0.0) ? 1.0 / h_ip1 : 0.0;
    off[i] = h_ip1;
    diag[i] = 2.0 * (h_ip1 + h_i);
    g[i] = 3.0 * (ydiff_ip1 * g_ip1 - ydiff_i * g_i);
  }
  /* symmetric tridiagonal solve (Thomas) */
  for (int i = 1; i < sys; i++) {
    const double m = off[i - 1] / diag[i - 1];
    diag[i] -= m * off[i - 1];
    g[i] -= m * g[i - 1];
  }
  c[sys] = g[sys - 1] / diag[sys - 1];
  for (int i = sys - 2; i >= 0; i--) c[i + 1] = (g[i] - off[i] * c[i + 2]) / diag[i];
  free(diag); free(off); free(g);
}

static double spline_eval(int n, const double* x, const double* y, const double* c, double xx) {
  int i = bsearch_lo(x, n, xx);
  if (i > n - 2) i = n - 2;
  const double h = x[i + 1] - x[i];
  const double dy = y[i + 1] - y[i];
  const double dx = xx - x[i];
  const double b = dy / h - h * (c[i + 1] + 2.0 * c[i]) / 3.0;
  const double d = (c[i + 1] - c[i]) / (3.0 * h);
  return y[i] + dx * (b + dx * (c[i] + dx * d));
}

/* ---------------------------------------------------------------------- */
/* background (cosmolike/cosmo3D.c)                                        */
/* ---------------------------------------------------------------------- */
/* chi_all: cosmo3D.c:40-92 */
static void chi_all(oc_ctx* c, double a, double* chi, double* dchida) {
  const double z = 1.0 / a - 1.0;
  const double* zz = c->chi_z;
  const double* cc = c->chi_v;
  const int j = bsearch_lo(zz, c->chi_nz, z);
  if (j + 2 > c->chi_nz - 1) { /* the reference reads chi[j+2] (cosmo3D.c:66) */
    oc_set_err(c, "chi_all: z=%g needs chi table beyond its last two nodes", z);
    *chi = NAN; *dchida = NAN; return;
  }
  const double dy = (z - zz[j]) / (zz[j + 1] - zz[j]);
  const double out0 = cc[j] + dy * (cc[j + 1] - cc[j]);
  double out1;
  const double up = (cc[j + 2] - cc[j]) / (zz[j + 2] - zz[j]);
  if (j > 0) {
    const double down = (cc[j + 1] - cc[j - 1]) / (zz[j + 1] - zz[j - 1]);
    out1 = down + dy * (up - down);
  } else {
    const double down = (cc[j + 1] - cc[j]) / (zz[j + 1] - zz[j]);
    out1 = down + dy * (up - down);
  }
  out1 = out1 / OC_COVERH0;
  out1 = out1 / (a * a);
  *chi = out0 / OC_COVERH0;
  *dchida = out1;
}

static double chi_of_a(oc_ctx* c, double a) { double x, d; chi_all(c, a, &x, &d); return x; }

/* hoverh0v2: cosmo3D.c:104-107 */
static double hoverh0v2(double a, double dchida) { return 1.0 / ((a * a) * dchida); }
/* hoverh0: cosmo3D.c:99-102 via dchi_dz (:94-97) */
static double hoverh0(oc_ctx* c, double a) {
  double x, d;
  chi_all(c, a, &x, &d);
  return 1.0 / ((a * a) * d);
}

/* a_chi: cosmo3D.c:109-133 */
static double a_chi(oc_ctx* c, double io_chi) {
  const double chi = io_chi * OC_COVERH0;
  const int j = bsearch_lo(c->chi_v, c->chi_nz, chi);
  const double dy = (chi - c->chi_v[j]) / (c->chi_v[j + 1] - c->chi_v[j]);
  const double z = c->chi_z[j] + dy * (c->chi_z[j + 1] - c->chi_z[j]);
  return 1.0 / (1.0 + z);
}

/* norm_growfac(a, true): cosmo3D.c:148-198 */
static double growfac(oc_ctx* c, double a) {
  int j = bsearch_lo(c->G_z, c->G_nz, 0.0);
  double dy = (0.0 - c->G_z[j]) / (c->G_z[j + 1] - c->G_z[j]);
  const double g1 = c->G_v[j] + dy * (c->G_v[j + 1] - c->G_v[j]);
  const double z = 1.0 / a - 1.0;
  j = bsearch_lo(c->G_z, c->G_nz, z);
  dy = (z - c->G_z[j]) / (c->G_z[j + 1] - c->G_z[j]);
  const double G = c->G_v[j] + dy * (c->G_v[j + 1] - c->G_v[j]);
  return (G * a) / g1;
}

/* f_growth: cosmo3D.c:200-228 */
static double f_growth(oc_ctx* c, double z) {
  const int j = bsearch_lo(c->G_z, c->G_nz, z);
  const double dy = (z - c->G_z[j]) / (c->G_z[j + 1] - c->G_z[j]);
  const double G = c->G_v[j] + dy * (c->G_v[j + 1] - c->G_v[j]);
  const double dlnGdlnz = ((c->G_v[j + 1] - c->G_v[j]) / (c->G_z[j + 1] - c->G_z[j])) * z / G;
  const double dlnGdlna = -dlnGdlnz * (1 + z) / z;
  return 1 + dlnGdlna;
}

/* p_nonlin (no baryons): cosmo3D.c:362-417; Pdelta -> p_nonlin (:424-445) */
static double p_nonlin(oc_ctx* c, double k, double a) {
  const double log10k = log10(k / OC_COVERH0);
  const double z = 1.0 / a - 1.0;
  const int nk = c->nk, nz = c->nzp;
  const int i = bsearch_lo(c->lk, nk, log10k);
  const int j = bsearch_lo(c->pz, nz, z);
  const double dx = (log10k - c->lk[i]) / (c->lk[i + 1] - c->lk[i]);
  const double dy = (z - c->pz[j]) / (c->pz[j + 1] - c->pz[j]);
  const double* P = c->lnP;
  const double out = (1 - dx) * (1 - dy) * P[i * nz + j] + (1 - dx) * dy * P[i * nz + j + 1] +
                     dx * (1 - dy) * P[(i + 1) * nz + j] + dx * dy * P[(i + 1) * nz + j + 1];
  return exp(out) / (OC_COVERH0 * OC_COVERH0 * OC_COVERH0);
}

/* ---------------------------------------------------------------------- */
/* redshift distributions (cosmolike/redshift_spline.c)                    */
/* ---------------------------------------------------------------------- */
/* zdistr_histo_n / pf_histo_n: redshift_spline.c:383-410, 597-628 */
static double histo_n(const double* tab, int ntomo, int nzbins, double zmin_all,
                      double zmax_all, double z, int ni) {
  double res = 0.0;
  if ((z >= zmin_all) && (z < zmax_all)) {
    const double* z_v = tab + (size_t)ntomo * nzbins;
    const double dz_histo = (z_v[nzbins - 1] - z_v[0]) / ((double)nzbins - 1.);
    const int nj = (int)floor((z - z_v[0]) / dz_histo);
    if (nj < 0 || nj > nzbins - 1) return NAN;
    res = tab[(size_t)ni * nzbins + nj];
  }
  return res;
}

/* table + spline build of nz_source_photoz / nz_lens_photoz
 * (redshift_spline.c:446-512 and 665-731) */
static int build_nz_splines(oc_ctx* c, const double* tab, int ntomo, int nzbins, double zmin,
                            double zmax, double** zc_out, double** y_out, double** c_out) {
  double* zc = (double*)malloc(sizeof(double) * nzbins);
  double* y = (double*)calloc((size_t)(ntomo + 1) * nzbins, sizeof(double));
  double* cc = (double*)calloc((size_t)(ntomo + 1) * nzbins, sizeof(double));
  const double dz_histo = (zmax - zmin) / ((double)nzbins);
  for (int k = 0; k < nzbins; k++) zc[k] = zmin + (k + 0.5) * dz_histo;
  double NORM[OC_MAXB];
  double norm = 0;
  for (int i = 0; i < ntomo; i++) {
    NORM[i] = 0.0;
    for (int k = 0; k < nzbins; k++)
      NORM[i] += histo_n(tab, ntomo, nzbins, zmin, zmax, zc[k], i) * dz_histo;
    if (!(NORM[i] > 0.0)) {
      oc_set_err(c, "zero/negative n(z) normalization for bin %d", i);
      free(zc); free(y); free(cc);
      return -1;
    }
    norm += NORM[i];
  }
  for (int k = 0; k < nzbins; k++) {
    y[k] = 0;
    for (int i = 0; i < ntomo; i++) {
      y[(size_t)(i + 1) * nzbins + k] = histo_n(tab, ntomo, nzbins, zmin, zmax, zc[k], i) / NORM[i];
      y[k] += y[(size_t)(i + 1) * nzbins + k] * NORM[i] / norm;
    }
  }
  for (int i = 0; i < ntomo + 1; i++)
    spline_init(nzbins, zc, y + (size_t)i * nzbins, cc + (size_t)i * nzbins);
  *zc_out = zc; *y_out = y; *c_out = cc;
  return 0;
}

/* nz_source_photoz evaluation: redshift_spline.c:513-537 */
static double nz_source(oc_ctx* c, double zz, int nj) {
  const int n = c->s_rows;
  zz = zz - c->src_dz[nj];
  if (zz < c->s_zc[0] || zz > c->s_zc[n - 1]) return 0.0;
  return spline_eval(n, c->s_zc, c->s_y + (size_t)(nj + 1) * n, c->s_c + (size_t)(nj + 1) * n, zz);
}

/* nz_lens_photoz evaluation: redshift_spline.c:736-762 */
static double nz_lens(oc_ctx* c, double zz, int nj) {
  const int n = c->l_rows;
  zz = (zz - c->lens_dz[nj] - c->l_zmean[nj]) / c->lens_st[nj] + c->l_zmean[nj];
  if (zz <= c->l_zc[0] || zz >= c->l_zc[n - 1]) return 0.0;
  const double r = spline_eval(n, c->l_zc, c->l_y + (size_t)(nj + 1) * n, c->l_c + (size_t)(nj + 1) * n, zz);
  return r / c->lens_st[nj];
}

/* bin limits from a histogram column: set_source_sample / set_lens_sample
 * (generic_interface.cpp:1862-1867, 1966-1971) */
static void bin_limits(const double* tab, int ntomo, int nrows, int k, double* zlo, double* zhi) {
  const double* z_v = tab + (size_t)ntomo * nrows;
  const double* col = tab + (size_t)k * nrows;
  double mx = col[0];
  for (int i = 1; i < nrows; i++) if (col[i] > mx) mx = col[i];
  int first = -1, last = -1;
  for (int i = 0; i < nrows; i++)
    if (col[i] > 0.999e-8 * mx) { if (first < 0) first = i; last = i; }
  if (first < 0) { first = 0; last = nrows - 1; }
  *zlo = z_v[first];
  *zhi = z_v[last];
}

int oc_set_source_sample(oc_ctx* c, const double* table, int nrows, int ntomo) {
  c->status = 0;
  if (ntomo <= 0 || ntomo > OC_MAXB || nrows < 5) { oc_set_err(c, "bad source sample size"); return -1; }
  for (int i = 1; i < nrows; i++)
    if (!(table[(size_t)i * (ntomo + 1)] > table[(size_t)(i - 1) * (ntomo + 1)])) {
      oc_set_err(c, "bad n(z) file (z vector not monotonic)"); return -1;
    }
  free(c->s_tab);
  c->s_tab = (double*)malloc(sizeof(double) * (size_t)(ntomo + 1) * nrows);
  for (int i = 0; i < nrows; i++) {
    c->s_tab[(size_t)ntomo * nrows + i] = table[(size_t)i * (ntomo + 1)];
    for (int k = 0; k < ntomo; k++) c->s_tab[(size_t)k * nrows + i] = table[(size_t)i * (ntomo + 1) + k + 1];
  }
  c->ns = ntomo;
  c->s_rows = nrows;
  const double* z_v = c->s_tab + (size_t)ntomo * nrows;
  c->s_zmin_all = fmax(z_v[0], 1.e-5);
  c->s_zmax_all = z_v[nrows - 1] + (z_v[nrows - 1] - z_v[0]) / ((double)nrows - 1.);
  for (int k = 0; k < ntomo; k++) {
    double lo, hi;
    bin_limits(c->s_tab, ntomo, nrows, k, &lo, &hi);
    c->s_zmin[k] = fmax(lo, 1.001e-5);
    c->s_zmax[k] = hi;
  }
  if (c->s_zmax_all < c->s_zmax[ntomo - 1] || c->s_zmin_all > c->s_zmin[0]) {
    oc_set_err(c, "source n(z) range inconsistent"); return -1;
  }
  free(c->s_zc); free(c->s_y); free(c->s_c);
  c->s_zc = c->s_y = c->s_c = NULL;
  if (build_nz_splines(c, c->s_tab, ntomo, nrows, c->s_zmin_all, c->s_zmax_all, &c->s_zc, &c->s_y, &c->s_c))
    return -1;
  c->g_valid = 0;
  return 0;
}

/* zmean (redshift_spline.c:794-844): 256-point GL of z n(z) / n(z) over the
 * bin's support, evaluated when set_lens_sample runs (zero shift, unit stretch,
 * zmean still 0 -- generic_interface.cpp:1874-1876) */
static void compute_lens_zmean(oc_ctx* c) {
  static gl_rule r; /* 256/512/1024 points by accuracy level (:813-815) */
  const int hd = abs(c->hdi);
  gl_make(&r, hd == 0 ? 256 : hd == 1 ? 512 : 1024);
  double save_dz[OC_MAXB], save_st[OC_MAXB], save_zm[OC_MAXB];
  memcpy(save_dz, c->lens_dz, sizeof save_dz);
  memcpy(save_st, c->lens_st, sizeof save_st);
  memcpy(save_zm, c->l_zmean, sizeof save_zm);
  for (int i = 0; i < OC_MAXB; i++) { c->lens_dz[i] = 0.0; c->lens_st[i] = 1.0; c->l_zmean[i] = 0.0; }
  double zm[OC_MAXB];
  for (int i = 0; i < c->nl; i++) {
    double num = 0.0, den = 0.0;
    for (int p = 0; p < r.n; p++) {
      double x, w;
      gl_point(&r, c->l_zmin[i], c->l_zmax[i], p, &x, &w);
      const double n = nz_lens(c, x, i);
      num += w * x * n;
      den += w * n;
    }
    zm[i] = num / den;
  }
  memcpy(c->lens_dz, save_dz, sizeof save_dz);
  memcpy(c->lens_st, save_st, sizeof save_st);
  for (int i = 0; i < c->nl; i++) c->l_zmean[i] = zm[i];
  (void)save_zm;
}

int oc_set_lens_sample(oc_ctx* c, const double* table, int nrows, int ntomo) {
  c->status = 0;
  if (ntomo <= 0 || ntomo > OC_MAXB || nrows < 5) { oc_set_err(c, "bad lens sample size"); return -1; }
  for (int i = 1; i < nrows; i++)
    if (!(table[(size_t)i * (ntomo + 1)] > table[(size_t)(i - 1) * (ntomo + 1)])) {
      oc_set_err(c, "bad n(z) file (z vector not monotonic)"); return -1;
    }
  free(c->l_tab);
  c->l_tab = (double*)malloc(sizeof(double) * (size_t)(ntomo + 1) * nrows);
  for (int i = 0; i < nrows; i++) {
    c->l_tab[(size_t)ntomo * nrows + i] = table[(size_t)i * (ntomo + 1)];
    for (int k = 0; k < ntomo; k++) c->l_tab[(size_t)k * nrows + i] = table[(size_t)i * (ntomo + 1) + k + 1];
  }
  c->nl = ntomo;
  c->l_rows = nrows;
  const double* z_v = c->l_tab + (size_t)ntomo * nrows;
  c->l_zmin_all = fmax(z_v[0], 1.e-5);
  c->l_zmax_all = z_v[nrows - 1] + (z_v[nrows - 1] - z_v[0]) / ((double)nrows - 1.);
  for (int k = 0; k < ntomo; k++) bin_limits(c->l_tab, ntomo, nrows, k, &c->l_zmin[k], &c->l_zmax[k]);
  free(c->l_zc); free(c->l_y); free(c->l_c);
  c->l_zc = c->l_y = c->l_c = NULL;
  if (build_nz_splines(c, c->l_tab, ntomo, nrows, c->l_zmin_all, c->l_zmax_all, &c->l_zc, &c->l_y, &c->l_c))
    return -1;
  compute_lens_zmean(c);
  c->g_valid = 0;
  return 0;
}

/* ---------------------------------------------------------------------- */
/* lens efficiencies: g_tomo (redshift_spline.c:852-923), g_lens (1014-1109) */
/* ---------------------------------------------------------------------- */
static void build_g_tables(oc_ctx* c) {
  const int x = 25 * (1 + abs(c->hdi));
  const int Na = x * (c->n_a - 1) + 1;
  const double amax = 0.999999;
  free(c->gt_tab); free(c->gl_tab);
  c->gt_tab = (double*)calloc((size_t)c->ns * c->n_a, sizeof(double));
  c->gl_tab = c->nl > 0 ? (double*)calloc((size_t)c->nl * c->n_a, sizeof(double)) : NULL;
  {
    const double amin = 1.0 / (c->s_zmax_all + 1.0);
    const double da = (amax - amin) / ((double)Na - 1.0);
    #pragma omp parallel for schedule(static)
    for (int j = 0; j < c->ns; j++) {
      double* Pint = (double*)malloc(sizeof(double) * Na);
      double* Qint = (double*)malloc(sizeof(double) * Na);
      for (int i = 0; i < Na; i++) {
        const double a = amin + i * da;
        Pint[i] = nz_source(c, 1. / a - 1., j) / (a * a);
        const double ch = chi_of_a(c, amin + i * da);
        Qint[i] = Pint[i] / ch;
      }
      double P = 0.0, Q = 0.0;
      double* tab = c->gt_tab + (size_t)j * c->n_a;
      tab[0] = 0.0;
      for (int i = 1; i < Na; i++) {
        P += 0.5 * da * (Pint[i - 1] + Pint[i]);
        Q += 0.5 * da * (Qint[i - 1] + Qint[i]);
        if (i % x == 0) tab[i / x] = P - chi_of_a(c, amin + i * da) * Q;
      }
      free(Pint); free(Qint);
    }
  }
  if (c->nl > 0) {
    const double amin = 1.0 / (c->l_zmax_all + 1.0);
    const double amin_shear = 1.0 / (c->s_zmax_all + 1.0);
    const double da = (amax - amin) / ((double)Na - 1.0);
    const int Na0 = (amin_shear < amin) ? (int)ceil((amin - amin_shear) / da) + 1 : 1;
    const double da0 = (Na0 > 1) ? (amin - amin_shear) / ((double)Na0 - 1.) : 0.0;
    #pragma omp parallel for schedule(static)
    for (int j = 0; j < c->nl; j++) {
      double P = 0.0, Q = 0.0;
      for (int i = 1; i < Na0; i++) {
        const double apl = amin_shear + (i - 1) * da0;
        const double aph = amin_shear + i * da0;
        const double fl = nz_lens(c, 1.0 / apl - 1.0, j) / (apl * apl);
        const double fh = nz_lens(c, 1.0 / aph - 1.0, j) / (aph * aph);
        P += 0.5 * da0 * (fl + fh);
        Q += 0.5 * da0 * (fl / chi_of_a(c, apl) + fh / chi_of_a(c, aph));
      }
      double* Pint = (double*)malloc(sizeof(double) * Na);
      double* Qint = (double*)malloc(sizeof(double) * Na);
      for (int i = 0; i < Na; i++) {
        const double ap = amin + i * da;
        Pint[i] = nz_lens(c, 1.0 / ap - 1.0, j) / (ap * ap);
        Qint[i] = Pint[i] / chi_of_a(c, amin + i * da);
      }
      double* tab = c->gl_tab + (size_t)j * c->n_a;
      tab[0] = P - chi_of_a(c, amin) * Q;
      for (int i = 1; i < Na; i++) {
        P += 0.5 * da * (Pint[i - 1] + Pint[i]);
        Q += 0.5 * da * (Qint[i - 1] + Qint[i]);
        if (i % x == 0) tab[i / x] = P - chi_of_a(c, amin + i * da) * Q;
      }
      free(Pint); free(Qint);
    }
  }
  c->g_valid = 1;
}

/* interpol1d: basics.c:457-480 */
static double interpol1d(const double* f, int n, double a, double b, double dx, double x) {
  (void)b;
  if (x < a) return f[0];
  const double r = (x - a) / dx;
  const int i = (int)floor(r);
  if (i + 1 >= n) return f[n - 1];
  return (r - i) * (f[i + 1] - f[i]) + f[i];
}

static double g_tomo(oc_ctx* c, double a, int ni) {
  const double amin = 1.0 / (c->s_zmax_all + 1.0), amax = 0.999999;
  const double dac = (amax - amin) / ((double)c->n_a - 1.0);
  return (a <= amin || a > 1.0 - dac) ? 0.0
         : interpol1d(c->gt_tab + (size_t)ni * c->n_a, c->n_a, amin, amax, dac, a);
}

static double g_lens(oc_ctx* c, double a, int ni) {
  const double amin = 1.0 / (c->l_zmax_all + 1.0), amax = 0.999999;
  const double dac = (amax - amin) / ((double)c->n_a - 1.0);
  return (a < amin || a > 1.0 - dac) ? 0.0
         : interpol1d(c->gl_tab + (size_t)ni * c->n_a, c->n_a, amin, amax, dac, a);
}

/* ---------------------------------------------------------------------- */
/* radial weights (radial_weights.c:15-105), IA (IA.c:238-303)             */
/* ---------------------------------------------------------------------- */
static double W_kappa(oc_ctx* c, double a, double fK, int nz) {
  return (1.5 * c->omega_m * fK / a) * g_tomo(c, a, nz);
}
static double W_mag(oc_ctx* c, double a, double fK, int nz) {
  return (1.5 * c->omega_m * fK / a) * g_lens(c, a, nz);
}
static double W_gal(oc_ctx* c, double a, int ni, double hh) {
  const double z = 1. / a - 1;
  return nz_lens(c, z, ni) * hh;
}
static double W_source(oc_ctx* c, double a, int ni, double hh) {
  const double z = 1.0 / a - 1.0;
  return nz_source(c, z, ni) * hh;
}
static double f_rsd(oc_ctx* c, double a) { return f_growth(c, 1.0 / a - 1.0); }
static double W_RSD(oc_ctx* c, double l, double a0, double a1, int ni) {
  double w = (1 + 8. * l) / ((2. * l + 1.) * (2. * l + 1.)) * nz_lens(c, 1. / a0 - 1., ni) *
             hoverh0(c, a0) * f_rsd(c, a0);
  w -= 4. / (2 * l + 3.) * sqrt((2 * l + 1.) / (2 * l + 3.)) * nz_lens(c, 1. / a1 - 1., ni) *
       hoverh0(c, a1) * f_rsd(c, a1);
  return w;
}

static void IA_A1_Z1Z2(oc_ctx* c, double a, double growfac_a, int n1, int n2, double res[2]) {
  double A1 = 0.0, A2 = 0.0;
  switch (c->ia_mode) {
    case 0: break;
    case 2: A1 = c->ia_bin[n1]; A2 = c->ia_bin[n2]; break;
    case 3: A1 = c->ia_A * pow((1.0 / a) / c->oneplusz0, c->ia_eta); A2 = A1; break;
    default: oc_set_err(c, "IA mode %d not supported", c->ia_mode);
  }
  const double x = c->omega_m * c->c1rhocrit / growfac_a;
  res[0] = A1 * x;
  res[1] = A2 * x;
}

/* IA_A2_Z1Z2 (IA.c:305-361): A2 (Omega_m c1rhocrit / D^2), same evolution
   form as A1 with its own amplitude and exponent */
static void IA_A2_Z1Z2(oc_ctx* c, double a, double growfac_a, int n1, int n2, double res[2]) {
  double A1 = 0.0, A2 = 0.0;
  switch (c->ia_mode) {
    case 0: break;
    case 2: A1 = c->ia2_bin[n1]; A2 = c->ia2_bin[n2]; break;
    case 3: A1 = c->ia2_A * pow(1.0 / (a * c->oneplusz0), c->ia2_eta); A2 = A1; break;
    default: oc_set_err(c, "IA mode %d not supported", c->ia_mode);
  }
  const double x = c->omega_m * c->c1rhocrit / (growfac_a * growfac_a);
  res[0] = A1 * x;
  res[1] = A2 * x;
}

/* IA_BTA_Z1Z2 (IA.c:370-415): b_TA per bin or one value; 1 without IA */
static void IA_BTA_Z1Z2(oc_ctx* c, int n1, int n2, double res[2]) {
  switch (c->ia_mode) {
    case 0: res[0] = res[1] = 1.0; break;
    case 2: res[0] = c->bta_bin[n1]; res[1] = c->bta_bin[n2]; break;
    case 3: res[0] = res[1] = c->bta_A; break;
    default: oc_set_err(c, "IA mode %d not supported", c->ia_mode); res[0] = res[1] = 0.0;
  }
}

/* FAST-PT IA table t at k (cosmo2D.c:1414-1447): uniform in ln k with step
   (ln kmax - ln kmin) / N, zero outside [kmin, kmax], last entry at the top */
static double fpt_at(oc_ctx* c, int t, double k, double g4) {
  const double lnk = log(k);
  const double l0 = log(c->fpt_kmin), l1 = log(c->fpt_kmax), dl = (l1 - l0) / c->fpt_n;
  if (lnk < l0 || lnk > l1) return 0.0;
  const double r = (lnk - l0) / dl;
  const int i = (int)floor(r);
  const double* tab = c->fpt + (size_t)t * c->fpt_n;
  if (i + 1 >= c->fpt_n) return g4 * tab[c->fpt_n - 1];
  const double w = r - i;
  return g4 * (w * (tab[i + 1] - tab[i]) + tab[i]);
}

double oc_fpt_at(oc_ctx* c, int t, double k, double g4) { return fpt_at(c, t, k, g4); }

/* amin_lens / amax_lens: redshift_spline.c:50-79 */
static double amin_lens(oc_ctx* c, int ni) {
  const double zmax = (c->l_zmax[ni] - c->l_zmean[ni]) * c->lens_st[ni] + c->l_zmean[ni];
  return 1. / (1 + zmax + 2. * fabs(c->lens_dz[ni]));
}
static double amax_lens(oc_ctx* c, int ni) {
  const double zmin = (c->l_zmin[ni] - c->l_zmean[ni]) * c->lens_st[ni] + c->l_zmean[ni];
  if (c->bmag[ni] != 0) return 1. / (1. + fmax(c->s_zmin_all, 0.001));
  return 1. / (1 + fmax(zmin - 2. * fabs(c->lens_dz[ni]), 0.001));
}

/* pair maps: Z1/Z2 (redshift_spline.c:229-275), ZL/ZS (150-193, no exclusions) */
static void ss_pair(int ns, int q, int* z1, int* z2) {
  *z1 = 0; *z2 = 0;
  int n = 0;
  for (int i = 0; i < ns; i++)
    for (int j = i; j < ns; j++) { if (n == q) { *z1 = i; *z2 = j; return; } n++; }
}
int oc_npairs_ss(oc_ctx* c) { return c->ns * (c->ns + 1) / 2; }
int oc_npairs_gs(oc_ctx* c) { return c->nl * c->ns; }

/* test_zoverlap (redshift_spline.c:115-145): 0 for a listed (lens, source)
   pair of tomo.ggl_exclude, else 1 */
int oc_test_zoverlap(oc_ctx* c, int zl, int zs) {
  for (int k = 0; k < c->n_ggl_ex; k++)
    if (c->ggl_ex[2 * k] == zl && c->ggl_ex[2 * k + 1] == zs) return 0;
  return 1;
}

/* tomo.ggl_Npowerspectra (init_ntomo_powerspectra, generic_interface.cpp:775-802):
   the included pairs, which the gamma_t block holds in N_ggl order (lens
   major, source minor; redshift_spline.c:199-223) */
int oc_npairs_ggl(oc_ctx* c) {
  int n = 0;
  for (int zl = 0; zl < c->nl; zl++)
    for (int zs = 0; zs < c->ns; zs++) n += oc_test_zoverlap(c, zl, zs);
  return n;
}

/* init_ggl_exclude (generic_interface.cpp:874-896): n entries, pairs
   (lens, source) */
int oc_set_ggl_exclude(oc_ctx* c, const int* pairs, int n) {
  if (n < 0 || n / 2 > OC_MAXB * OC_MAXB || (n > 0 && !pairs)) return -1;
  c->n_ggl_ex = n / 2;
  for (int k = 0; k < 2 * c->n_ggl_ex; k++) c->ggl_ex[k] = pairs[k];
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Limber integrands                                                       */
/* ---------------------------------------------------------------------- */
/* int_for_C_ss_tomo_limber_core: cosmo2D.c:1370-1515 (TATT branch
   :1395-1487, NLA :1488-1504); k = (l + 1/2) / f_K */
static double ss_core(oc_ctx* c, double a, double fK, double PK, double growfac_a,
                      double dchida, double ell_prefactor, int n1, int n2, double WK1,
                      double WK2, double WS1, double WS2, int EE, double k) {
  double ans;
  if (c->ia_model == 1) {
    const double g4 = growfac_a * growfac_a * growfac_a * growfac_a;
    double AX[2];
    IA_A1_Z1Z2(c, a, growfac_a, n1, n2, AX);
    const double C11 = AX[0], C12 = AX[1];
    IA_A2_Z1Z2(c, a, growfac_a, n1, n2, AX);
    const double C21 = AX[0], C22 = AX[1];
    IA_BTA_Z1Z2(c, n1, n2, AX);
    const double bta1 = AX[0], bta2 = AX[1];
    if (1 == EE) {
      const double tt = fpt_at(c, 0, k, g4), ta_dE1 = fpt_at(c, 2, k, g4), ta_dE2 = fpt_at(c, 3, k, g4);
      const double ta = fpt_at(c, 4, k, g4), mixA = fpt_at(c, 6, k, g4), mixB = fpt_at(c, 7, k, g4);
      const double mixEE = fpt_at(c, 8, k, g4);
      ans = WK1 * WK2 * PK
          - WS1 * WK2 * (C11 * PK + C11 * bta1 * (ta_dE1 + ta_dE2) - 5 * C21 * (mixA + mixB))
          - WS2 * WK1 * (C12 * PK + C12 * bta2 * (ta_dE1 + ta_dE2) - 5 * C22 * (mixA + mixB))
          + WS1 * WS2 * (C11 * C12 * PK
                         + C11 * C12 * (bta1 * bta2 * ta + (bta1 + bta2) * (ta_dE1 + ta_dE2))
                         - 5. * (C11 * C22 + C12 * C21) * (mixA + mixB)
                         - 5. * (C11 * bta1 * C22 + C12 * bta2 * C21) * mixEE
                         + 25. * C21 * C22 * tt);
    } else {
      const double tt = fpt_at(c, 1, k, g4), ta = fpt_at(c, 5, k, g4), mix = fpt_at(c, 9, k, g4);
      ans = WS1 * WS2 * (C11 * C12 * bta1 * bta2 * ta - 5. * (C11 * bta1 * C22 + C12 * bta2 * C21) * mix +
                         25. * C21 * C22 * tt);
    }
    return ans * (dchida / (fK * fK)) * ell_prefactor;
  }
  if (1 == EE) {
    double IA[2];
    IA_A1_Z1Z2(c, a, growfac_a, n1, n2, IA);
    const double C11 = IA[0], C12 = IA[1];
    ans = WK1 * WK2 * PK - WS1 * WK2 * C11 * PK - WS2 * WK1 * C12 * PK + WS1 * WS2 * C11 * C12 * PK;
  } else {
    ans = 0.0;
  }
  return ans * (dchida / (fK * fK)) * ell_prefactor;
}

/* int_for_C_ss_tomo_limber: cosmo2D.c:1521-1555 */
static double int_ss(oc_ctx* c, double a, int n1, int n2, double l, int EE) {
  const double ell = l + 0.5;
  double fK, dchida;
  chi_all(c, a, &fK, &dchida);
  const double growfac_a = growfac(c, a);
  const double hh = hoverh0v2(a, dchida);
  const double k = ell / fK;
  const double PK = p_nonlin(c, k, a);
  const double ell4 = ell * ell * ell * ell;
  const double ell_prefactor = l * (l - 1.) * (l + 1.) * (l + 2.) / ell4;
  const double WK1 = W_kappa(c, a, fK, n1), WK2 = W_kappa(c, a, fK, n2);
  const double WS1 = W_source(c, a, n1, hh), WS2 = W_source(c, a, n2, hh);
  return ss_core(c, a, fK, PK, growfac_a, dchida, ell_prefactor, n1, n2, WK1, WK2, WS1, WS2, EE, k);
}

static int has_b2(oc_ctx* c);
static double fptb_at(oc_ctx* c, int t, double lnk);
/* the integrands' nonlinear_bias argument: has_b2_galaxies() in the C_ell
   integrals (cosmo2D.c:2154, 2643, 2988), the caller's value in an integrand
   evaluated on its own */
static int nlb(oc_ctx* c) { return c->nlb_set ? c->nlb_on : has_b2(c); }

/* one-loop bias term of the gs integrand (cosmo2D.c:1970-1992, 2027-2057):
   0.5 g4 (b2 d1d2 + bs2 d1s2 + b3 d1p3) + bK k^2 P when has_b2_galaxies */
static double gs_oneloop(oc_ctx* c, int nl, double k, double PK, double growfac_a) {
  if (!nlb(c)) return 0.0;
  if (!c->fptb) { oc_set_err(c, "one-loop bias needs the FAST-PT bias tables (set_bias_PS)"); return NAN; }
  const double lnk = log(k);
  const double g4 = growfac_a * growfac_a * growfac_a * growfac_a;
  return 0.5 * g4 * (c->b2[nl] * fptb_at(c, 0, lnk) + c->bs2[nl] * fptb_at(c, 2, lnk) + c->b3[nl] * fptb_at(c, 5, lnk)) +
         c->bk[nl] * k * k * PK;
}

/* int_for_C_gs_tomo_limber_core: cosmo2D.c:1890-2073 (TATT branch
   :1919-2006 and NLA :2007-2065, without RSD and one-loop bias, which are
   off by default); k = (l + 1/2) / f_K */
static double gs_core(oc_ctx* c, double a, double fK, double PK, double growfac_a,
                      double dchida, double ell_prefactor, double ell_prefactor2, int nl,
                      int ns, double WK, double WS, double WGAL, double WMAG, double k) {
  const double b1 = c->b1[nl];
  const double bmag = c->bmag[nl];
  if (c->ia_model == 1) {
    const double g4 = growfac_a * growfac_a * growfac_a * growfac_a;
    const double mixA = fpt_at(c, 6, k, g4), mixB = fpt_at(c, 7, k, g4);
    const double ta_dE1 = fpt_at(c, 2, k, g4), ta_dE2 = fpt_at(c, 3, k, g4);
    double AX[2];
    IA_A1_Z1Z2(c, a, growfac_a, ns, ns, AX);
    const double C1ZS = AX[0];
    IA_BTA_Z1Z2(c, ns, ns, AX);
    const double btazs = AX[0];
    IA_A2_Z1Z2(c, a, growfac_a, ns, ns, AX);
    const double C2ZS = AX[0];
    const double ans = WK * ((WGAL * b1 + WMAG * ell_prefactor * bmag + 0.0) * PK + WGAL * gs_oneloop(c, nl, k, PK, growfac_a))
                     - WS * (WGAL * b1 + WMAG * ell_prefactor * bmag) *
                           (C1ZS * PK + C1ZS * btazs * (ta_dE1 + ta_dE2) - 5 * C2ZS * (mixA + mixB));
    return ans * (dchida / (fK * fK)) * ell_prefactor2;
  }
  double IA[2];
  IA_A1_Z1Z2(c, a, growfac_a, ns, ns, IA);
  const double C1ZS = IA[0];
  const double ans = (WK - WS * C1ZS) *
                     ((WGAL * b1 + WMAG * ell_prefactor * bmag + 0.0) * PK + WGAL * gs_oneloop(c, nl, k, PK, growfac_a));
  return ans * (dchida / (fK * fK)) * ell_prefactor2;
}

/* int_for_C_gs_tomo_limber: cosmo2D.c:2079-2119 */
static double int_gs(oc_ctx* c, double a, int nl, int ns, double l) {
  const double growfac_a = growfac(c, a);
  double fK, dchida;
  chi_all(c, a, &fK, &dchida);
  const double hh = hoverh0v2(a, dchida);
  const double ell = l + 0.5;
  const double k = ell / fK;
  const double PK = p_nonlin(c, k, a);
  const double ell_prefactor = l * (l + 1.) / (ell * ell);
  const double tmp = (l - 1.) * l * (l + 1.) * (l + 2.);
  const double ell_prefactor2 = (tmp > 0) ? sqrt(tmp) / (ell * ell) : 0.0;
  const double WK = W_kappa(c, a, fK, ns);
  const double WS = W_source(c, a, ns, hh);
  const double WGAL = W_gal(c, a, nl, hh);
  const double WMAG = W_mag(c, a, fK, nl);
  return gs_core(c, a, fK, PK, growfac_a, dchida, ell_prefactor, ell_prefactor2, nl, ns, WK, WS, WGAL, WMAG, k);
}

/* int_for_C_gg_tomo_limber (linear bias, RSD on, HOD off): cosmo2D.c:2490-2608 */
/* p_lin: cosmo3D.c:306-358 (bilinear in log10 k, z of ln P_lin) */
static double p_lin(oc_ctx* c, double k, double a) {
  const double log10k = log10(k / OC_COVERH0);
  const double z = 1.0 / a - 1.0;
  const int nk = c->nkL, nz = c->nzL;
  const int i = bsearch_lo(c->lkL, nk, log10k);
  const int j = bsearch_lo(c->pzL, nz, z);
  const double dx = (log10k - c->lkL[i]) / (c->lkL[i + 1] - c->lkL[i]);
  const double dy = (z - c->pzL[j]) / (c->pzL[j + 1] - c->pzL[j]);
  const double* P = c->lnPL;
  const double out = (1 - dx) * (1 - dy) * P[i * nz + j] + (1 - dx) * dy * P[i * nz + j + 1] +
                     dx * (1 - dy) * P[(i + 1) * nz + j] + dx * dy * P[(i + 1) * nz + j + 1];
  return exp(out) / (OC_COVERH0 * OC_COVERH0 * OC_COVERH0);
}

/* has_b2_galaxies (cosmo2D.c:282-288) */
static int has_b2(oc_ctx* c) {
  for (int i = 0; i < c->nl; i++)
    if (c->b2[i] != 0.0) return 1;
  return 0;
}

/* FPTbias.tab[t] at ln k (cosmo2D.c:2566-2593): uniform in ln k with step
   (ln kmax - ln kmin) / N, zero outside the range */
static double fptb_at(oc_ctx* c, int t, double lnk) {
  const double l0 = log(c->fptb_kmin), l1 = log(c->fptb_kmax), dl = (l1 - l0) / c->fptb_n;
  if (lnk < l0 || lnk > l1) return 0.0;
  return interpol1d(c->fptb + (size_t)t * c->fptb_n, c->fptb_n, l0, l1, dl, lnk);
}

/* int_for_C_gg_tomo_limber: cosmo2D.c:2489-2606 (RSD on, no HOD); use_lin
   selects p_lin over Pdelta; the one-loop bias term when has_b2_galaxies and
   not use_lin (:2560-2603) */
static double int_gg_ps(oc_ctx* c, double a, int ni, double l, int use_lin) {
  double fK, dchida;
  chi_all(c, a, &fK, &dchida);
  const double hh = hoverh0v2(a, dchida);
  const double ell = l + 0.5;
  const double k = ell / fK;
  const double b1i = c->b1[ni];
  const double bmagi = c->bmag[ni];
  const double WGALi = W_gal(c, a, ni, hh);
  const double WMAGi = W_mag(c, a, fK, ni);
  const double ell_prefactor = l * (l + 1.) / (ell * ell);
  double res = 1.0;
  const double chi_a_min = chi_of_a(c, OC_A_MIN);
  const double chi_0 = ell / k;
  const double chi_1 = (ell + 1.0) / k;
  if (chi_1 > chi_a_min) return 0;
  const double a_0 = a_chi(c, chi_0);
  const double a_1 = a_chi(c, chi_1);
  const double WRSDi = W_RSD(c, ell, a_0, a_1, ni);
  res *= (WGALi * b1i + WMAGi * ell_prefactor * bmagi + WRSDi);
  res *= (WGALi * b1i + WMAGi * ell_prefactor * bmagi + WRSDi);
  const double PK = use_lin ? p_lin(c, k, a) : p_nonlin(c, k, a);
  res *= PK;
  double oneloop = 0.0;
  if (nlb(c) && !use_lin) {
    if (!c->fptb) { oc_set_err(c, "one-loop bias needs the FAST-PT bias tables (set_bias_PS)"); return NAN; }
    const double lnk = log(k);
    const double l0 = log(c->fptb_kmin), l1 = log(c->fptb_kmax);
    const int in = !(lnk < l0 || lnk > l1);
    const double s4 = c->fptb_s4;
    const double d1d2 = fptb_at(c, 0, lnk);
    const double d2d2 = in ? fptb_at(c, 1, lnk) - 2. * s4 : 0.0;
    const double d1s2 = fptb_at(c, 2, lnk);
    const double d2s2 = in ? fptb_at(c, 3, lnk) - 4. / 3. * s4 : 0.0;
    const double s2s2 = in ? fptb_at(c, 4, lnk) - 8. / 9. * s4 : 0.0;
    const double d1p3 = fptb_at(c, 5, lnk);
    const double g = growfac(c, a), g4 = g * g * g * g;
    const double b2 = c->b2[ni], bs2 = c->bs2[ni], b3 = c->b3[ni], bk = c->bk[ni];
    oneloop = WGALi * WGALi *
              (g4 * (b1i * b2 * d1d2 + 0.25 * b2 * b2 * d2d2 + b1i * bs2 * d1s2 + 0.5 * b2 * bs2 * d2s2 +
                     0.25 * bs2 * bs2 * s2s2 + b1i * b3 * d1p3) +
               (2 * b1i * bk * k * k * PK));
  }
  return (res + oneloop) * dchida / (fK * fK);
}

static int ensure_g(oc_ctx* c) {
  if (!c->chi_z || !c->G_z || !c->lnP || !c->s_tab) { oc_set_err(c, "inputs not set"); return -1; }
  if (!c->g_valid) build_g_tables(c);
  return c->status;
}

/* C_ss_tomo_limber_nointerp: cosmo2D.c:1562-1604 (init=0) */
static double C_ss_nointerp(oc_ctx* c, const gl_rule* r, double l, int ni, int nj, int EE) {
  const double amin = 1. / (c->s_zmax_all + 1.);
  const double amax = 1. / (1. + fmax(c->s_zmin_all, 1e-6));
  double s = 0.0;
  for (int p = 0; p < r->n; p++) {
    double x, w;
    gl_point(r, amin, amax, p, &x, &w);
    s += w * int_ss(c, x, ni, nj, l, EE);
  }
  return s;
}

/* C_gs_tomo_limber_nointerp: cosmo2D.c:2125-2182 */
static double C_gs_nointerp(oc_ctx* c, const gl_rule* r, double l, int nl, int ns) {
  const double amin = amin_lens(c, nl), amax = amax_lens(c, nl);
  if (!(amin > 0) || !(amin < 1) || !(amax > 0) || !(amax < 1) || !(amin < amax)) {
    oc_set_err(c, "0 < amin < amax < 1 not true");
    return NAN;
  }
  double s = 0.0;
  for (int p = 0; p < r->n; p++) {
    double x, w;
    gl_point(r, amin, amax, p, &x, &w);
    s += w * int_gs(c, x, nl, ns, l);
  }
  return s;
}

/* C_gg_tomo_limber_linpsopt_nointerp: cosmo2D.c:2614-2666 */
static double C_gg_nointerp_ps(oc_ctx* c, const gl_rule* r, double l, int ni, int use_lin);
static double C_gg_nointerp(oc_ctx* c, const gl_rule* r, double l, int ni) {
  return C_gg_nointerp_ps(c, r, l, ni, 0);
}
static double C_gg_nointerp_ps(oc_ctx* c, const gl_rule* r, double l, int ni, int use_lin) {
  const double amin = amin_lens(c, ni), amax = amax_lens(c, ni);
  if (!(amin > 0) || !(amin < 1) || !(amax > 0) || !(amax < 1) || !(amin < amax)) {
    oc_set_err(c, "0 < amin < amax < 1 not true");
    return NAN;
  }
  double s = 0.0;
  for (int p = 0; p < r->n; p++) {
    double x, w;
    gl_point(r, amin, amax, p, &x, &w);
    s += w * int_gg_ps(c, x, ni, l, use_lin);
  }
  return s;
}

/* ln-ell grids of the three tables (cosmo2D.c:1631-1634, 2204-2207, 2703-2706) */
static void ss_lim(oc_ctx* c, double lim[3]) {
  lim[0] = log(fmax(OC_LMIN_TAB - 1., 1.0));
  lim[1] = log(c->lmax + 1);
  lim[2] = (lim[1] - lim[0]) / ((double)c->n_ell - 1.);
}
static void gx_lim(oc_ctx* c, double lim[3]) {
  lim[0] = log(fmax(OC_LMIN_TAB, 1.0));
  lim[1] = log(c->lmax + 1);
  lim[2] = (lim[1] - lim[0]) / ((double)c->n_ell - 1.0);
}

/* C_ss_tomo_limber table: cosmo2D.c:1655-1747 */
static void table_ss(oc_ctx* c, double* tab /* [npairs][nell] */, int EE) {
  const int nell = c->n_ell, npts = gl_order_ssgs(c->hdi), ns = c->ns;
  const int np = oc_npairs_ss(c);
  double lim[3];
  ss_lim(c, lim);
  gl_rule r;
  gl_make(&r, npts);
  const double amin = 1. / (c->s_zmax_all + 1.);
  const double amax = 1. / (1. + fmax(c->s_zmin_all, 1e-6));
  double *A = malloc(sizeof(double) * npts), *WT = malloc(sizeof(double) * npts),
         *FK = malloc(sizeof(double) * npts), *GF = malloc(sizeof(double) * npts),
         *HH = malloc(sizeof(double) * npts), *DC = malloc(sizeof(double) * npts);
  for (int p = 0; p < npts; p++) { /* create_cosmo_nodes: cosmo2D.c:1331-1356 */
    gl_point(&r, amin, amax, p, &A[p], &WT[p]);
    chi_all(c, A[p], &FK[p], &DC[p]);
    GF[p] = growfac(c, A[p]);
    HH[p] = hoverh0v2(A[p], DC[p]);
  }
  double* lx = malloc(sizeof(double) * nell);
  double* pref = malloc(sizeof(double) * nell);
  for (int i = 0; i < nell; i++) {
    lx[i] = exp(lim[0] + i * lim[2]);
    const double ell = lx[i] + 0.5, ell4 = ell * ell * ell * ell;
    pref[i] = lx[i] * (lx[i] - 1.) * (lx[i] + 1.) * (lx[i] + 2.) / ell4;
  }
  double* PK = malloc(sizeof(double) * (size_t)nell * npts);
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < nell; i++)
    for (int p = 0; p < npts; p++) PK[(size_t)i * npts + p] = p_nonlin(c, (lx[i] + 0.5) / FK[p], A[p]);
  double* W = malloc(sizeof(double) * 2 * ns * npts);
  for (int b = 0; b < ns; b++)
    for (int p = 0; p < npts; p++) {
      W[(0 * ns + b) * npts + p] = W_kappa(c, A[p], FK[p], b);
      W[(1 * ns + b) * npts + p] = W_source(c, A[p], b, HH[p]);
    }
  #pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < nell; i++)
    for (int k = 0; k < np; k++) {
      int z1, z2;
      ss_pair(ns, k, &z1, &z2);
      double sum = 0.0;
      for (int p = 0; p < npts; p++)
        sum += ss_core(c, A[p], FK[p], PK[(size_t)i * npts + p], GF[p], DC[p], pref[i], z1, z2,
                       W[z1 * npts + p], W[z2 * npts + p], W[(ns + z1) * npts + p],
                       W[(ns + z2) * npts + p], EE, (lx[i] + 0.5) / FK[p]) * WT[p];
      tab[(size_t)k * nell + i] = sum;
    }
  free(A); free(WT); free(FK); free(GF); free(HH); free(DC);
  free(lx); free(pref); free(PK); free(W);
}

/* C_gs_tomo_limber table: cosmo2D.c:2229-2348 */
static void table_gs(oc_ctx* c, double* tab /* [nl*ns][nell] */) {
  const int nell = c->n_ell, npts = gl_order_ssgs(c->hdi), ns = c->ns, nl = c->nl;
  double lim[3];
  gx_lim(c, lim);
  gl_rule r;
  gl_make(&r, npts);
  double* lx = malloc(sizeof(double) * nell);
  double* p1 = malloc(sizeof(double) * nell);
  double* p2 = malloc(sizeof(double) * nell);
  for (int i = 0; i < nell; i++) {
    lx[i] = exp(lim[0] + i * lim[2]);
    const double ell = lx[i] + 0.5;
    p1[i] = lx[i] * (lx[i] + 1.) / (ell * ell);
    const double tmp = (lx[i] - 1.) * lx[i] * (lx[i] + 1.) * (lx[i] + 2.);
    p2[i] = (tmp > 0) ? sqrt(tmp) / (ell * ell) : 0.0;
  }
  for (int zl = 0; zl < nl; zl++) {
    const double amin = amin_lens(c, zl), amax = amax_lens(c, zl);
    if (!(amin > 0) || !(amin < 1) || !(amax > 0) || !(amax < 1) || !(amin < amax)) {
      oc_set_err(c, "0 < amin < amax < 1 not true"); break;
    }
    double A[1024], WT[1024], FK[1024], GF[1024], HH[1024], DC[1024];
    for (int p = 0; p < npts; p++) {
      gl_point(&r, amin, amax, p, &A[p], &WT[p]);
      chi_all(c, A[p], &FK[p], &DC[p]);
      GF[p] = growfac(c, A[p]);
      HH[p] = hoverh0v2(A[p], DC[p]);
    }
    double* PK = malloc(sizeof(double) * (size_t)nell * npts);
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < nell; i++)
      for (int p = 0; p < npts; p++) PK[(size_t)i * npts + p] = p_nonlin(c, (lx[i] + 0.5) / FK[p], A[p]);
    double WG[1024], WM[1024];
    double* WK = malloc(sizeof(double) * ns * npts);
    double* WS = malloc(sizeof(double) * ns * npts);
    for (int p = 0; p < npts; p++) {
      WG[p] = W_gal(c, A[p], zl, HH[p]);
      WM[p] = W_mag(c, A[p], FK[p], zl);
      for (int zs = 0; zs < ns; zs++) {
        WK[zs * npts + p] = W_kappa(c, A[p], FK[p], zs);
        WS[zs * npts + p] = W_source(c, A[p], zs, HH[p]);
      }
    }
    #pragma omp parallel for collapse(2) schedule(static)
    for (int zs = 0; zs < ns; zs++)
      for (int i = 0; i < nell; i++) {
        double sum = 0.0;
        for (int p = 0; p < npts; p++)
          sum += gs_core(c, A[p], FK[p], PK[(size_t)i * npts + p], GF[p], DC[p], p1[i], p2[i], zl,
                         zs, WK[zs * npts + p], WS[zs * npts + p], WG[p], WM[p], (lx[i] + 0.5) / FK[p]) *
                 WT[p];
        tab[(size_t)(zl * ns + zs) * nell + i] = sum;
      }
    free(PK); free(WK); free(WS);
  }
  free(lx); free(p1); free(p2);
}

/* C_gg_tomo_limber table: cosmo2D.c:2709-2732 */
static void table_gg(oc_ctx* c, double* tab /* [nl][nell] */) {
  const int nell = c->n_ell, nl = c->nl;
  double lim[3];
  gx_lim(c, lim);
  gl_rule r;
  gl_make(&r, gl_order_gg(c->hdi));
  #pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nl; k++)
    for (int i = 0; i < nell; i++) {
      const double lx = exp(lim[0] + i * lim[2]);
      tab[(size_t)k * nell + i] = C_gg_nointerp(c, &r, lx, k);
    }
}

/* *_fill: cosmo2D.c:1784-1807 (scalar branch) */
static void fill(const double* tab, const double lim[3], int n, int lmin, int lmax, double* out) {
  const double inv_dx = 1.0 / lim[2];
  const double a = lim[0];
  for (int l = lmin; l < lmax; l++) {
    const double r = (log((double)l) - a) * inv_dx;
    const int i = (int)floor(r);
    const int ic = i < 0 ? 0 : (i >= n - 1 ? n - 2 : i);
    const double t = r - ic;
    out[l] = tab[ic] + t * (tab[ic + 1] - tab[ic]);
  }
}

int oc_cl_tables(oc_ctx* c, double* ss, double* gs, double* gg) {
  c->status = 0;
  if (ensure_g(c)) return -1;
  if (ss) table_ss(c, ss, 1);
  if (gs && c->nl > 0) table_gs(c, gs);
  if (gg && c->nl > 0) table_gg(c, gg);
  return c->status;
}

int oc_cl_ss_nointerp(oc_ctx* c, const double* l, int nl, double* out) {
  c->status = 0;
  if (ensure_g(c)) return -1;
  gl_rule r;
  gl_make(&r, gl_order_ssgs(c->hdi));
  const int np = oc_npairs_ss(c);
  #pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < np; k++)
    for (int i = 0; i < nl; i++) {
      int z1, z2;
      ss_pair(c->ns, k, &z1, &z2);
      out[(size_t)k * nl + i] = C_ss_nointerp(c, &r, l[i], z1, z2, 1);
    }
  return c->status;
}

/* B-mode C_ss (C_ss_tomo_limber(l, ni, nj, EE = 0)): table at the table ells
   and direct GL at arbitrary ell */
int oc_cl_tables_bb(oc_ctx* c, double* ss_bb) {
  c->status = 0;
  if (ensure_g(c)) return -1;
  table_ss(c, ss_bb, 0);
  return c->status;
}
int oc_cl_ss_bb_nointerp(oc_ctx* c, const double* l, int nl, double* out) {
  c->status = 0;
  if (ensure_g(c)) return -1;
  gl_rule r;
  gl_make(&r, gl_order_ssgs(c->hdi));
  const int np = oc_npairs_ss(c);
  #pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < np; k++)
    for (int i = 0; i < nl; i++) {
      int z1, z2;
      ss_pair(c->ns, k, &z1, &z2);
      out[(size_t)k * nl + i] = C_ss_nointerp(c, &r, l[i], z1, z2, 0);
    }
  return c->status;
}

int oc_cl_gs_nointerp(oc_ctx* c, const double* l, int nl, double* out) {
  c->status = 0;
  if (ensure_g(c)) return -1;
  gl_rule r;
  gl_make(&r, gl_order_ssgs(c->hdi));
  const int np = oc_npairs_gs(c);
  #pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < np; k++)
    for (int i = 0; i < nl; i++) out[(size_t)k * nl + i] = C_gs_nointerp(c, &r, l[i], k / c->ns, k % c->ns);
  return c->status;
}

int oc_cl_gg_nointerp(oc_ctx* c, const double* l, int nl, double* out) {
  c->status = 0;
  if (ensure_g(c)) return -1;
  gl_rule r;
  gl_make(&r, gl_order_gg(c->hdi));
  #pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < c->nl; k++)
    for (int i = 0; i < nl; i++) out[(size_t)k * nl + i] = C_gg_nointerp(c, &r, l[i], k);
  return c->status;
}

/* Cl arrays exactly as xi_pm_tomo / w_gammat_tomo / w_gg_tomo(limber=1) build
 * them: l=0 -> 0, 1 <= l < LMIN_tab direct GL, LMIN_tab <= l < LMAX fill
 * (cosmo2D.c:468-496, 655-674, 831-849) */
/* C_ss at every integer ell < LMAX as xi_pm_tomo builds it (EE = 1: E mode,
   0: B mode, nonzero only under TATT) */
static void full_ss(oc_ctx* c, const gl_rule* r64, double* ss, int EE) {
  const int L = c->lmax, nell = c->n_ell, np = oc_npairs_ss(c);
  double* tab = malloc(sizeof(double) * (size_t)np * nell);
  table_ss(c, tab, EE);
  double lim[3];
  ss_lim(c, lim);
  #pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < np; k++) {
    int z1, z2;
    ss_pair(c->ns, k, &z1, &z2);
    double* out = ss + (size_t)k * L;
    out[0] = 0.0;
    for (int l = 1; l < OC_LMIN_TAB; l++) out[l] = C_ss_nointerp(c, r64, (double)l, z1, z2, EE);
    fill(tab + (size_t)k * nell, lim, nell, OC_LMIN_TAB, L, out);
  }
  free(tab);
}

int oc_cl_full(oc_ctx* c, double* ss, double* gs, double* gg) {
  c->status = 0;
  if (ensure_g(c)) return -1;
  const int L = c->lmax, nell = c->n_ell;
  gl_rule r64, r96;
  gl_make(&r64, gl_order_ssgs(c->hdi));
  gl_make(&r96, gl_order_gg(c->hdi));
  if (ss) full_ss(c, &r64, ss, 1);
  if (gs && c->nl > 0) {
    const int np = oc_npairs_gs(c);
    double* tab = malloc(sizeof(double) * (size_t)np * nell);
    table_gs(c, tab);
    double lim[3];
    gx_lim(c, lim);
    #pragma omp parallel for schedule(dynamic)
    for (int k = 0; k < np; k++) {
      double* out = gs + (size_t)k * L;
      out[0] = 0.0;
      for (int l = 1; l < OC_LMIN_TAB; l++) out[l] = C_gs_nointerp(c, &r64, (double)l, k / c->ns, k % c->ns);
      fill(tab + (size_t)k * nell, lim, nell, OC_LMIN_TAB, L, out);
    }
    free(tab);
  }
  if (gg && c->nl > 0) {
    double* tab = malloc(sizeof(double) * (size_t)c->nl * nell);
    table_gg(c, tab);
    double lim[3];
    gx_lim(c, lim);
    #pragma omp parallel for schedule(dynamic)
    for (int k = 0; k < c->nl; k++) {
      double* out = gg + (size_t)k * L;
      out[0] = 0.0;
      for (int l = 1; l < OC_LMIN_TAB; l++) out[l] = C_gg_nointerp(c, &r96, (double)l, k);
      fill(tab + (size_t)k * nell, lim, nell, OC_LMIN_TAB, L, out);
    }
    free(tab);
  }
  return c->status;
}

/* ---------------------------------------------------------------------- */
/* Legendre stage                                                          */
/* ---------------------------------------------------------------------- */
/* gsl_sf_legendre_Pl_deriv_array restated (called at basics.c:426,431) */
int oc_legendre_pl_dpl(int lmax, double x, double* P, double* dP) {
  if (lmax < 0 || x < -1.0 || x > 1.0) return -1;
  P[0] = 1.0;
  if (lmax >= 1) P[1] = x;
  {
    double pm2 = 1.0, pm1 = x;
    for (int l = 2; l <= lmax; l++) {
      const double p = (x * (2 * l - 1) * pm1 - (l - 1) * pm2) / l;
      pm2 = pm1;
      pm1 = p;
      P[l] = p;
    }
  }
  dP[0] = 0.0;
  if (lmax >= 1) dP[1] = 1.0;
  const double sqeps = 1.4901161193847656e-08;
  if (fabs(x - 1.0) * (lmax + 1.0) * (lmax + 1.0) < sqeps) {
    for (int l = 2; l <= lmax; l++) {
      const double pre = 0.5 * l * (l + 1.0);
      dP[l] = pre * (1.0 - 0.25 * (1.0 - x) * (l + 2.0) * (l - 1.0));
    }
  } else if (fabs(x + 1.0) * (lmax + 1.0) * (lmax + 1.0) < sqeps) {
    for (int l = 2; l <= lmax; l++) {
      const double sgn = (l & 1) ? 1.0 : -1.0;
      const double pre = sgn * 0.5 * l * (l + 1.0);
      dP[l] = pre * (1.0 - 0.25 * (1.0 + x) * (l + 2.0) * (l - 1.0));
    }
  } else {
    const double da = 1.0 + x, db = 1.0 - x;
    for (int l = 2; l <= lmax; l++) dP[l] = -l * (x * P[l] - P[l - 1]) / (da * db);
  }
  return 0;
}

/* set_bin_average (basics.c:382-455) + G builds (cosmo2D.c:377-460,
 * 581-644, 758-823).  The formulas are evaluated in REAL: the reference's
 * double (K_ld = 0), or long double (K_ld = 1, default) because the G+-
 * kernels are small differences of O(l^2) terms divided by x_min - x_max
 * (~1.5e-7 at 2.5'), whose double evaluation carries up to ~1e-4 relative
 * error in xi-(2.5') depending on the compiler's FMA contraction. */
#define OC_KERNEL_BODY(REAL, PI)                                                              \
  for (int i = i0; i < nt; i += stride) {                                                     \
    const double xmin_d = cos(exp(log(c->vtmin) + (i + 0.) * logdt));                       \
    const double xmax_d = cos(exp(log(c->vtmin) + (i + 1.) * logdt));                       \
    const REAL xmin = xmin_d, xmax = xmax_d;                                                 \
    REAL* Pmin = malloc(sizeof(REAL) * (L + 1));                                              \
    REAL* Pmax = malloc(sizeof(REAL) * (L + 1));                                              \
    REAL* dPmin = malloc(sizeof(REAL) * (L + 1));                                             \
    REAL* dPmax = malloc(sizeof(REAL) * (L + 1));                                             \
    for (int s = 0; s < 2; s++) {                                                             \
      REAL* P = s ? Pmax : Pmin;                                                              \
      REAL* dP = s ? dPmax : dPmin;                                                           \
      const REAL x = s ? xmax : xmin;                                                         \
      P[0] = 1; P[1] = x;                                                                     \
      REAL pm2 = 1, pm1 = x;                                                                  \
      for (int l = 2; l <= L; l++) {                                                          \
        const REAL p = (x * (2 * l - 1) * pm1 - (l - 1) * pm2) / l;                           \
        pm2 = pm1; pm1 = p; P[l] = p;                                                         \
      }                                                                                       \
      dP[0] = 0; dP[1] = 1;                                                                   \
      const REAL da = 1 + x, db = 1 - x;                                                      \
      for (int l = 2; l <= L; l++) dP[l] = -l * (x * P[l] - P[l - 1]) / (da * db);            \
    }                                                                                         \
    double* gp = Gp + (size_t)i * L;                                                          \
    double* gm = Gm + (size_t)i * L;                                                          \
    double* pg = Pgt + (size_t)i * L;                                                         \
    double* pw = Pw + (size_t)i * L;                                                          \
    gp[0] = gm[0] = pg[0] = pw[0] = 0.0;                                                      \
    for (int l = 1; l < L; l++) {                                                             \
      gp[l] = (double)((2. * l + 1) / (2. * PI * l * l * (l + 1) * (l + 1)) * (              \
          -l * (l - 1.) / 2 * (l + 2. / (2 * l + 1)) * (Pmin[l - 1] - Pmax[l - 1])           \
          - l * (l - 1.) * (2. - l) / 2 * (xmin * Pmin[l] - xmax * Pmax[l])                   \
          + l * (l - 1.) / (2. * l + 1) * (Pmin[l + 1] - Pmax[l + 1])                         \
          + (4 - l) * (dPmin[l] - dPmax[l])                                                   \
          + (l + 2) * (xmin * dPmin[l - 1] - xmax * dPmax[l - 1] - Pmin[l - 1] + Pmax[l - 1]) \
          + 2 * (l - 1) * (xmin * dPmin[l] - xmax * dPmax[l] - Pmin[l] + Pmax[l])             \
          - 2 * (l + 2) * (dPmin[l - 1] - dPmax[l - 1])) / (xmin - xmax));                    \
      gm[l] = (double)((2. * l + 1) / (2. * PI * l * l * (l + 1) * (l + 1)) * (              \
          -l * (l - 1.) / 2 * (l + 2. / (2 * l + 1)) * (Pmin[l - 1] - Pmax[l - 1])           \
          - l * (l - 1.) * (2. - l) / 2 * (xmin * Pmin[l] - xmax * Pmax[l])                   \
          + l * (l - 1.) / (2. * l + 1) * (Pmin[l + 1] - Pmax[l + 1])                         \
          + (4 - l) * (dPmin[l] - dPmax[l])                                                   \
          + (l + 2) * (xmin * dPmin[l - 1] - xmax * dPmax[l - 1] - Pmin[l - 1] + Pmax[l - 1]) \
          - 2 * (l - 1) * (xmin * dPmin[l] - xmax * dPmax[l] - Pmin[l] + Pmax[l])             \
          + 2 * (l + 2) * (dPmin[l - 1] - dPmax[l - 1])) / (xmin - xmax));                    \
      pg[l] = (double)((2. * l + 1) / (4. * PI * l * (l + 1) * (xmin - xmax)) *              \
              ((l + 2. / (2 * l + 1.)) * (Pmin[l - 1] - Pmax[l - 1]) +                        \
               (2 - l) * (xmin * Pmin[l] - xmax * Pmax[l]) -                                  \
               2. / (2 * l + 1.) * (Pmin[l + 1] - Pmax[l + 1])));                             \
      pw[l] = (double)((1.0 / (xmin - xmax)) * (1. / (4.0 * PI)) *                           \
                       (Pmin[l + 1] - Pmax[l + 1] - Pmin[l - 1] + Pmax[l - 1]));              \
    }                                                                                         \
    free(Pmin); free(Pmax); free(dPmin); free(dPmax);                                         \
  }

static int build_kernels(oc_ctx* c) {
  const int nt = c->ntheta, L = c->lmax;
  if (nt <= 0) { oc_set_err(c, "Ntable.Ntheta not initialized"); return -1; }
  free(c->K);
  c->K = (double*)calloc((size_t)4 * nt * L, sizeof(double));
  double* Gp = c->K;
  double* Gm = c->K + (size_t)nt * L;
  double* Pgt = c->K + (size_t)2 * nt * L;
  double* Pw = c->K + (size_t)3 * nt * L;
  const double logdt = (log(c->vtmax) - log(c->vtmin)) / nt;
  for (int i = 0; i < nt; i++) {
    const double xa = cos(exp(log(c->vtmin) + (i + 0.) * logdt));
    const double xb = cos(exp(log(c->vtmin) + (i + 1.) * logdt));
    if (fabs(xa) > 1 || fabs(xb) > 1) { oc_set_err(c, "Legendre argument out of range"); return -1; }
  }
  #pragma omp parallel
  {
#ifdef _OPENMP
    const int i0 = omp_get_thread_num(), stride = omp_get_num_threads();
#else
    const int i0 = 0, stride = 1;
#endif
    if (c->K_ld) {
      OC_KERNEL_BODY(long double, 3.14159265358979323846264338327950288L)
    } else {
      OC_KERNEL_BODY(double, M_PI)
    }
  }
  c->K_valid = 1;
  return 0;
}

int oc_set_kernel_precision(oc_ctx* c, int extended) {
  c->K_ld = extended ? 1 : 0;
  c->K_valid = 0;
  return 0;
}

int oc_kernels(oc_ctx* c, double* Gp, double* Gm, double* Pgt, double* Pw) {
  c->status = 0;
  if (!c->K_valid && build_kernels(c)) return -1;
  const size_t n = (size_t)c->ntheta * c->lmax;
  if (Gp) memcpy(Gp, c->K, n * sizeof(double));
  if (Gm) memcpy(Gm, c->K + n, n * sizeof(double));
  if (Pgt) memcpy(Pgt, c->K + 2 * n, n * sizeof(double));
  if (Pw) memcpy(Pw, c->K + 3 * n, n * sizeof(double));
  return 0;
}

/* plain dot over l = 1 .. LMAX-1 (cosmo2D.c:500-516 scalar branch) */
static void contract(const double* K, const double* Cl, int npair, int nt, int L, double* out) {
  #pragma omp parallel for collapse(2) schedule(static)
  for (int nz = 0; nz < npair; nz++)
    for (int i = 0; i < nt; i++) {
      const double* g = K + (size_t)i * L;
      const double* cl = Cl + (size_t)nz * L;
      double s = 0.0;
      for (int l = 1; l < L; l++) s += g[l] * cl[l];
      out[nz * nt + i] = s;
    }
}

int oc_xi_pm(oc_ctx* c, double* xip, double* xim) {
  c->status = 0;
  if (!c->K_valid && build_kernels(c)) return -1;
  const int np = oc_npairs_ss(c), L = c->lmax;
  double* cl = malloc(sizeof(double) * (size_t)np * L);
  if (oc_cl_full(c, cl, NULL, NULL)) { free(cl); return -1; }
  if (c->ia_model == 1) {
    /* xi+- from EE +- BB (cosmo2D.c:517-533); BB vanishes under NLA */
    double* bb = malloc(sizeof(double) * (size_t)np * L);
    gl_rule r64;
    gl_make(&r64, gl_order_ssgs(c->hdi));
    full_ss(c, &r64, bb, 0);
    double* cm = malloc(sizeof(double) * (size_t)np * L);
    for (size_t i = 0; i < (size_t)np * L; i++) {
      cm[i] = cl[i] - bb[i];
      cl[i] += bb[i];
    }
    contract(c->K, cl, np, c->ntheta, L, xip);
    contract(c->K + (size_t)c->ntheta * L, cm, np, c->ntheta, L, xim);
    free(bb); free(cm); free(cl);
    return c->status;
  }
  /* BB = 0 under NLA, so EE+BB = EE-BB = EE */
  contract(c->K, cl, np, c->ntheta, L, xip);
  contract(c->K + (size_t)c->ntheta * L, cl, np, c->ntheta, L, xim);
  free(cl);
  return c->status;
}

int oc_gammat(oc_ctx* c, double* out) {
  c->status = 0;
  if (!c->K_valid && build_kernels(c)) return -1;
  const int np = oc_npairs_gs(c), L = c->lmax;
  double* cl = malloc(sizeof(double) * (size_t)np * L);
  if (oc_cl_full(c, NULL, cl, NULL)) { free(cl); return -1; }
  contract(c->K + (size_t)2 * c->ntheta * L, cl, np, c->ntheta, L, out);
  free(cl);
  return c->status;
}

static int C_cl_tomo(oc_ctx* c, double* Cl, double tol);
#define OC_LMAX_NOLIMBER 150 /* structs.c:45 */

/* w_gg_tomo (cosmo2D.c:739-907): limber_gg = 1 fills every ell from the
   Limber tables (ell < LMIN_tab direct); 0 takes ell < LMAX_NOLIMBER from
   C_cl_tomo with tolerance 0.01 and the Limber fill above */
int oc_wtheta(oc_ctx* c, double* out) {
  c->status = 0;
  if (!c->K_valid && build_kernels(c)) return -1;
  const int L = c->lmax;
  double* cl = malloc(sizeof(double) * (size_t)c->nl * L);
  if (oc_cl_full(c, NULL, NULL, cl)) { free(cl); return -1; }
  if (!c->limber_gg) {
    double* nl = malloc(sizeof(double) * (size_t)c->nl * OC_LMAX_NOLIMBER);
    if (C_cl_tomo(c, nl, 0.01)) { free(nl); free(cl); return -1; }
    for (int i = 0; i < c->nl; i++)
      for (int l = 0; l < OC_LMAX_NOLIMBER && l < L; l++) cl[(size_t)i * L + l] = nl[(size_t)i * OC_LMAX_NOLIMBER + l];
    free(nl);
  }
  contract(c->K + (size_t)3 * c->ntheta * L, cl, c->nl, c->ntheta, L, out);
  free(cl);
  return c->status;
}

int oc_datavector_size(oc_ctx* c) {
  return c->ntheta * (2 * oc_npairs_ss(c) + oc_npairs_ggl(c) + c->nl);
}

/* compute_Mx2pt_N_masked<0,3> with every mask entry 1 and no point mass
 * (generic_interface.hpp:770-791 starts, 800-862 calibration, 968-1041 fill) */
/* bin-averaged theta [rad] (compute_binning_real_space, generic_interface.cpp:2037-2063) */
static void binning_theta(oc_ctx* c, double* theta) {
  const double logdt = (log(c->vtmax) - log(c->vtmin)) / c->ntheta;
  for (int i = 0; i < c->ntheta; i++) {
    const double t0 = exp(log(c->vtmin) + (i + 0.) * logdt);
    const double t1 = exp(log(c->vtmin) + (i + 1.) * logdt);
    theta[i] = 2. / 3. * (pow(t1, 3) - pow(t0, 3)) / (t1 * t1 - t0 * t0);
  }
}

/* PointMass::get_pm (generic_interface.cpp:2626-2641) */
static double point_mass(oc_ctx* c, int zl, int zs, double theta) {
  if (c->pm[zl] == 0.0) return 0.0;
  const double Goverc2 = 1.6e-23;
  const double a_lens = 1.0 / (1.0 + c->l_zmean[zl]);
  const double chi_lens = chi_of_a(c, a_lens);
  return 4 * M_PI * Goverc2 * c->pm[zl] * 1.e+13 * g_tomo(c, a_lens, zs) / (theta * theta) /
         (chi_lens * a_lens * a_lens * a_lens);
}

/* compute_data_vector_Mx2pt_N_starts (generic_interface.hpp:770-791):
 * stable_sort_index(ord, "ascend"), then start(i) = sum of sizes(indices(j))
 * over j < indices(i).  Returns -1 unless the blocks tile [0, n): a 3-cycle
 * ord such as (1,2,0) puts a block past the end of the data vector. */
static int block_starts(const int sizes[3], const int ord[3], int start[3]) {
  for (int i = 0; i < 3; i++)
    if (ord[i] < 0 || ord[i] > 2) return -1;
  if (ord[0] == ord[1] || ord[0] == ord[2] || ord[1] == ord[2]) return -1;
  int idx[3] = {0, 1, 2};
  for (int i = 1; i < 3; i++)
    for (int j = i; j > 0 && ord[idx[j - 1]] > ord[idx[j]]; j--) { int t = idx[j]; idx[j] = idx[j - 1]; idx[j - 1] = t; }
  for (int i = 0; i < 3; i++) {
    start[i] = 0;
    for (int j = 0; j < idx[i]; j++) start[i] += sizes[idx[j]];
  }
  const int n = sizes[0] + sizes[1] + sizes[2];
  for (int i = 0; i < 3; i++) {
    if (start[i] + sizes[i] > n) return -1;
    for (int j = 0; j < i; j++)
      if (sizes[i] && sizes[j] && start[i] < start[j] + sizes[j] && start[j] < start[i] + sizes[i]) return -1;
  }
  return 0;
}

int oc_datavector_3x2pt(oc_ctx* c, const int ord[3], double* dv) {
  c->status = 0;
  const int nt = c->ntheta, nss = oc_npairs_ss(c), ngs = oc_npairs_gs(c), ngg = c->nl;
  const int sizes[3] = {2 * nt * nss, nt * oc_npairs_ggl(c), nt * ngg};
  int start[3];
  if (block_starts(sizes, ord, start)) { oc_set_err(c, "ord: blocks do not tile the data vector"); return -1; }
  const int n = sizes[0] + sizes[1] + sizes[2];
  double* xp = malloc(sizeof(double) * nss * nt);
  double* xm = malloc(sizeof(double) * nss * nt);
  if (oc_xi_pm(c, xp, xm)) { free(xp); free(xm); return -1; }
  double* gt = NULL;
  double* w = NULL;
  if (c->nl > 0) {
    gt = malloc(sizeof(double) * ngs * nt);
    w = malloc(sizeof(double) * ngg * nt);
    if (oc_gammat(c, gt) || oc_wtheta(c, w)) { free(xp); free(xm); free(gt); free(w); return -1; }
  }
  for (int i = 0; i < n; i++) dv[i] = 0.0;
  for (int nz = 0; nz < nss; nz++) {
    int z1, z2;
    ss_pair(c->ns, nz, &z1, &z2);
    const double f = (1.0 + c->shear_m[z1]) * (1.0 + c->shear_m[z2]);
    for (int i = 0; i < nt; i++) {
      dv[start[0] + nt * nz + i] = xp[nz * nt + i] * f;
      dv[start[0] + nt * nz + i + nt * nss] = xm[nz * nt + i] * f;
    }
  }
  /* gamma_t: point mass, then (1 + m) (add_calib_and_set_mask_X_N,
     generic_interface.hpp:843-858; PointMass::get_pm, generic_interface.cpp:2626-2641) */
  double* theta = malloc(sizeof(double) * nt);
  binning_theta(c, theta);
  for (int nz = 0, r = 0; nz < ngs; nz++) {
    const int zs = nz % c->ns, zl = nz / c->ns;
    if (!oc_test_zoverlap(c, zl, zs)) continue;
    for (int i = 0; i < nt; i++)
      dv[start[1] + nt * r + i] = (gt[nz * nt + i] + point_mass(c, zl, zs, theta[i])) * (1.0 + c->shear_m[zs]);
    r++;
  }
  for (int nz = 0; nz < ngg; nz++)
    for (int i = 0; i < nt; i++) dv[start[2] + nt * nz + i] = w[nz * nt + i];
  /* masked entries are 0 */
  if (c->mask) {
    const int cs[3] = {0, 2 * nt * nss, 2 * nt * nss + sizes[1]};
    for (int k = 0; k < n; k++) {
      if (c->mask[k]) continue;
      const int b = k >= cs[2] ? 2 : (k >= cs[1] ? 1 : 0);
      dv[start[b] + (k - cs[b])] = 0.0;
    }
  }
  free(xp); free(xm); free(gt); free(w); free(theta);
  return c->status;
}

/* compute_Mx2pt_N_masked<1,3> (generic_interface.hpp:1137-1159): C_ss EE
 * (0 at ell >= lmax_shear, compute_X_N_masked<1,0>, :968-1004), C_gs, C_gg
 * at the given ells by the direct GL integrals, shear calibration
 * (add_calib_and_set_mask_X_N<1,M>, :800-862, no point mass in Fourier
 * space); mask (NULL = all ones) in the data-vector order of dv */
int oc_datavector_fourier(oc_ctx* c, const double* ell, int ncl, double lmax_shear, const int ord[3],
                          const int* mask, double* dv) {
  const int nss = oc_npairs_ss(c), ngs = oc_npairs_gs(c), ngg = c->nl;
  const int sizes[3] = {ncl * nss, ncl * oc_npairs_ggl(c), ncl * ngg};
  int start[3];
  c->status = 0;
  if (block_starts(sizes, ord, start)) { oc_set_err(c, "ord: blocks do not tile the data vector"); return -1; }
  double* ss = malloc(sizeof(double) * (nss * ncl + 1));
  double* gs = malloc(sizeof(double) * (ngs * ncl + 1));
  double* gg = malloc(sizeof(double) * (ngg * ncl + 1));
  int rc = oc_cl_ss_nointerp(c, ell, ncl, ss);
  if (!rc && ngg > 0) rc = oc_cl_gs_nointerp(c, ell, ncl, gs) || oc_cl_gg_nointerp(c, ell, ncl, gg);
  if (rc) { free(ss); free(gs); free(gg); return -1; }
  for (int k = 0; k < sizes[0] + sizes[1] + sizes[2]; k++) dv[k] = 0.0;
  for (int nz = 0; nz < nss; nz++) {
    int z1, z2;
    ss_pair(c->ns, nz, &z1, &z2);
    for (int i = 0; i < ncl; i++) {
      const int k = start[0] + ncl * nz + i;
      dv[k] = ell[i] < lmax_shear ? ss[nz * ncl + i] * (1.0 + c->shear_m[z1]) * (1.0 + c->shear_m[z2]) : 0.0;
    }
  }
  for (int nz = 0, r = 0; nz < ngs; nz++) {
    if (!oc_test_zoverlap(c, nz / c->ns, nz % c->ns)) continue;
    for (int i = 0; i < ncl; i++) dv[start[1] + ncl * r + i] = gs[nz * ncl + i] * (1.0 + c->shear_m[nz % c->ns]);
    r++;
  }
  for (int nz = 0; nz < ngg; nz++)
    for (int i = 0; i < ncl; i++) dv[start[2] + ncl * nz + i] = gg[nz * ncl + i];
  if (mask)
    for (int k = 0; k < sizes[0] + sizes[1] + sizes[2]; k++)
      if (!mask[k]) dv[k] = 0.0;
  free(ss); free(gs); free(gg);
  return 0;
}

/* chi^2 = sum_ij r_i C^-1_ij r_j over unmasked entries, r = dv - data
   (IP::get_chi2, generic_interface.cpp:2288-2325); mask and arrays in the
   data-vector order of dv */
double oc_chi2(const double* dv, const double* data, const double* inv_cov, const int* mask, int n) {
  double chi2 = 0.0;
  for (int i = 0; i < n; i++) {
    if (mask && !mask[i]) continue;
    const double x = dv[i] - data[i];
    double row = 0.0;
    for (int j = 0; j < n; j++) {
      if (mask && !mask[j]) continue;
      row += inv_cov[(size_t)i * n + j] * (dv[j] - data[j]);
    }
    chi2 += x * row;
  }
  return chi2;
}

/* init_IA(IA_MODEL = 1) plus set_nuisance_IA's A2 and b_TA
   (generic_interface.cpp:588-632, 1701-1774): mode 2 per-bin arrays, mode 3
   (A2, eta2) and one b_TA */
int oc_set_ia_tatt(oc_ctx* c, int model, const double* A2_bin, double A2, double eta2, const double* bta_bin,
                   double bta) {
  c->ia_model = model;
  c->ia2_A = A2;
  c->ia2_eta = eta2;
  c->bta_A = bta;
  for (int i = 0; i < OC_MAXB; i++) {
    c->ia2_bin[i] = A2_bin ? A2_bin[i < c->ns ? i : 0] : 0.0;
    c->bta_bin[i] = bta_bin ? bta_bin[i < c->ns ? i : 0] : 0.0;
  }
  return 0;
}

/* set_IA_PS (generic_interface.cpp:941-1001): PS[12][N], tables 0..9 in
   (Mpc/h)^3 -> divided by (c/H0)^3, k range in h/Mpc -> times c/H0 */
int oc_set_ia_ps(oc_ctx* c, const double* PS, double kmin, double kmax, double cutoff, int N) {
  (void)cutoff;
  if (N < 2) { oc_set_err(c, "set_IA_PS: N < 2"); return -1; }
  free(c->fpt);
  c->fpt = malloc(sizeof(double) * 10 * (size_t)N);
  const double h3 = OC_COVERH0 * OC_COVERH0 * OC_COVERH0;
  for (int t = 0; t < 10; t++)
    for (int j = 0; j < N; j++) c->fpt[(size_t)t * N + j] = PS[(size_t)t * N + j] / h3;
  c->fpt_n = N;
  c->fpt_kmin = kmin * OC_COVERH0;
  c->fpt_kmax = kmax * OC_COVERH0;
  return 0;
}

int oc_set_pointmass(oc_ctx* c, const double* pm, int n) {
  c->status = 0;
  if (n != c->nl) { oc_set_err(c, "set_pointmass: %d values for %d lens bins", n, c->nl); return -1; }
  for (int i = 0; i < n; i++) c->pm[i] = pm[i];
  return 0;
}

int oc_set_mask(oc_ctx* c, const int* mask, int n) {
  c->status = 0;
  free(c->mask);
  c->mask = NULL;
  c->mask_n = 0;
  if (!mask) return 0;
  c->mask = malloc(sizeof(int) * n);
  for (int i = 0; i < n; i++) c->mask[i] = mask[i];
  c->mask_n = n;
  return 0;
}

/* ---------------------------------------------------------------------- */
/* setters and probes                                                      */
/* ---------------------------------------------------------------------- */
int oc_set_ntable(oc_ctx* c, int n_ell, int n_a, int lmax, int high_def) {
  c->status = 0;
  if (n_ell < 2 || n_a < 2 || lmax < OC_LMIN_TAB + 1) { oc_set_err(c, "bad Ntable"); return -1; }
  c->n_ell = n_ell; c->n_a = n_a; c->lmax = lmax; c->hdi = high_def;
  c->g_valid = 0; c->K_valid = 0;
  if (c->nl > 0) compute_lens_zmean(c);
  return 0;
}

int oc_set_binning_real(oc_ctx* c, int ntheta, double tmin, double tmax) {
  c->status = 0;
  if (ntheta <= 0 || !(tmax > tmin) || !(tmin > 0)) { oc_set_err(c, "bad theta binning"); return -1; }
  c->ntheta = ntheta;
  c->vtmin = tmin * 2.90888208665721580e-4;
  c->vtmax = tmax * 2.90888208665721580e-4;
  c->K_valid = 0;
  return 0;
}

int oc_set_omega_m(oc_ctx* c, double om) { c->omega_m = om; c->g_valid = 0; return 0; }

static int any_nan(const double* x, int n) { for (int i = 0; i < n; i++) if (isnan(x[i])) return 1; return 0; }

int oc_set_distances(oc_ctx* c, const double* z, const double* chi, int n) {
  c->status = 0;
  if (n < 5 || any_nan(z, n) || any_nan(chi, n)) { oc_set_err(c, "bad distances"); return -1; }
  free(c->chi_z); free(c->chi_v);
  c->chi_z = dup(z, n); c->chi_v = dup(chi, n); c->chi_nz = n;
  c->g_valid = 0;
  return 0;
}

int oc_set_growth(oc_ctx* c, const double* z, const double* G, int n) {
  c->status = 0;
  if (n < 2 || any_nan(z, n) || any_nan(G, n)) { oc_set_err(c, "bad growth"); return -1; }
  free(c->G_z); free(c->G_v);
  c->G_z = dup(z, n); c->G_v = dup(G, n); c->G_nz = n;
  c->g_valid = 0;
  return 0;
}

int oc_set_pk(oc_ctx* c, const double* lk, int nk, const double* z, int nz, const double* lnP) {
  c->status = 0;
  if (nk < 2 || nz < 2 || any_nan(lk, nk) || any_nan(z, nz) || any_nan(lnP, nk * nz)) {
    oc_set_err(c, "bad P(k,z)"); return -1;
  }
  free(c->lk); free(c->pz); free(c->lnP);
  c->lk = dup(lk, nk); c->pz = dup(z, nz); c->lnP = dup(lnP, nk * nz);
  c->nk = nk; c->nzp = nz;
  c->g_valid = 0;
  return 0;
}

int oc_set_nuisance(oc_ctx* c, const double* src_dz, const double* lens_dz, const double* lens_st,
                    const double* b1, const double* bmag, const double* m) {
  c->status = 0;
  for (int i = 0; i < c->ns; i++) {
    if (src_dz) c->src_dz[i] = src_dz[i];
    if (m) c->shear_m[i] = m[i];
  }
  for (int i = 0; i < c->nl; i++) {
    if (lens_dz) c->lens_dz[i] = lens_dz[i];
    if (lens_st) c->lens_st[i] = lens_st[i];
    if (b1) c->b1[i] = b1[i];
    if (bmag) c->bmag[i] = bmag[i];
  }
  c->g_valid = 0;
  return 0;
}

int oc_set_ia(oc_ctx* c, int mode, const double* A_bin, double A, double eta) {
  c->status = 0;
  if (mode != 0 && mode != 2 && mode != 3) { oc_set_err(c, "IA mode not supported"); return -1; }
  c->ia_mode = mode;
  c->ia_A = A;
  c->ia_eta = eta;
  if (A_bin) for (int i = 0; i < c->ns; i++) c->ia_bin[i] = A_bin[i];
  return 0;
}

int oc_nz_source(oc_ctx* c, int bin, const double* z, int n, double* out) {
  if (bin < 0 || bin >= c->ns) return -1;
  for (int i = 0; i < n; i++) out[i] = nz_source(c, z[i], bin);
  return 0;
}
int oc_nz_lens(oc_ctx* c, int bin, const double* z, int n, double* out) {
  if (bin < 0 || bin >= c->nl) return -1;
  for (int i = 0; i < n; i++) out[i] = nz_lens(c, z[i], bin);
  return 0;
}
int oc_zmean_lens(oc_ctx* c, double* out) { for (int i = 0; i < c->nl; i++) out[i] = c->l_zmean[i]; return 0; }
int oc_g_tomo(oc_ctx* c, int bin, const double* a, int n, double* out) {
  c->status = 0;
  if (ensure_g(c)) return -1;
  for (int i = 0; i < n; i++) out[i] = g_tomo(c, a[i], bin);
  return c->status;
}
int oc_g_lens(oc_ctx* c, int bin, const double* a, int n, double* out) {
  c->status = 0;
  if (ensure_g(c)) return -1;
  for (int i = 0; i < n; i++) out[i] = g_lens(c, a[i], bin);
  return c->status;
}
int oc_chi_all(oc_ctx* c, const double* a, int n, double* chi, double* dchida) {
  c->status = 0;
  for (int i = 0; i < n; i++) chi_all(c, a[i], &chi[i], &dchida[i]);
  return c->status;
}
int oc_pdelta(oc_ctx* c, const double* k, const double* a, int n, double* out) {
  for (int i = 0; i < n; i++) out[i] = p_nonlin(c, k[i], a[i]);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Non-Limber clustering C_ell (C_cl_tomo, cosmo2D.c:4004-4769)            */
/* ---------------------------------------------------------------------- */
#define OC_NL_NCHI 500       /* Ntable.NL_Nchi, structs.c:272 */
#define OC_NL_BLOCK 16       /* cosmo2D.c:4672 */

/* next_fft_size (cosmo2D.c:4023-4034): the smallest m >= n whose prime
   factors are all <= 7 */
static long size_7smooth(long n) {
  static const int pr[4] = {2, 3, 5, 7};
  for (;; n++) {
    long m = n;
    for (int i = 0; i < 4; i++)
      while (m % pr[i] == 0) m /= pr[i];
    if (m == 1) return n;
  }
}

/* Unnormalised complex DFT out[k] = sum_j in[j stride] e^{sign 2 pi i jk/n}
   by recursive Cooley-Tukey over the smallest prime factor p of n (sub-DFTs
   of the p decimated sequences, then one p-point DFT per output group).  It
   stands in for FFTW's fftw_plan_dft_r2c_1d / c2r_1d (the reference's
   cfftlog_ells_p1/p2); pinned against numpy.fft in the tests. */
static void fft_rec(const double complex* in, double complex* out, int n, int stride, int sign) {
  if (n == 1) { out[0] = in[0]; return; }
  int p = 2;
  while (n % p) p++;
  const int m = n / p;
  for (int r = 0; r < p; r++) fft_rec(in + (size_t)r * stride, out + (size_t)r * m, m, stride * p, sign);
  double complex ybuf[16];
  double complex* y = p <= 16 ? ybuf : (double complex*)malloc(sizeof(double complex) * p);
  const double w0 = 2.0 * M_PI / n;
  for (int k = 0; k < m; k++) {
    for (int r = 0; r < p; r++) y[r] = out[(size_t)r * m + k];
    for (int q = 0; q < p; q++) {
      const long kk = k + (long)q * m;
      double complex s = 0.0;
      for (int r = 0; r < p; r++) {
        const long t = ((long)r * kk) % n;
        const double ang = sign * w0 * (double)t;
        s += y[r] * (cos(ang) + I * sin(ang));
      }
      out[kk] = s;
    }
  }
  if (y != ybuf) free(y);
}

int oc_fft(const double* re, const double* im, int n, int sign, double* out_re, double* out_im) {
  double complex* a = malloc(sizeof(double complex) * n);
  double complex* b = malloc(sizeof(double complex) * n);
  for (int i = 0; i < n; i++) a[i] = re[i] + I * (im ? im[i] : 0.0);
  fft_rec(a, b, n, 1, sign < 0 ? -1 : 1);
  for (int i = 0; i < n; i++) { out_re[i] = creal(b[i]); out_im[i] = cimag(b[i]); }
  free(a); free(b);
  return 0;
}

/* ln Gamma(a) for complex a: Lanczos approximation (g = 7, 9 terms, the
   coefficient table of cfftlog_ells_p2, cosmo2D.c:4160-4168, i.e. the
   published Godfrey set), with the reflection formula below Re a = 1/2
   (cosmo2D.c:4270-4297).  The imaginary part is defined modulo 2 pi. */
static const double lanczos_p[9] = {0.99999999999980993227684700473478, 676.520368121885098567009190444019,
                                    -1259.13921672240287047156078755283, 771.3234287776530788486528258894,
                                    -176.61502916214059906584551354, 12.507343278686904814458936853,
                                    -0.13857109526572011689554707, 9.984369578019570859563e-6,
                                    1.50563273514931155834e-7};
static double complex lanczos_lngamma1(double complex zz) { /* ln Gamma(zz + 1), Re zz > -1/2 */
  double complex s = lanczos_p[0];
  for (int w = 1; w < 9; w++) s += lanczos_p[w] / (zz + w);
  const double complex t = zz + 7.5;
  return 0.5 * log(2 * M_PI) + (zz + 0.5) * clog(t) - t + clog(s);
}
static double complex clngamma(double complex a) {
  if (creal(a) < 0.5) return clog(M_PI) - clog(csin(M_PI * a)) - lanczos_lngamma1(-a);
  return lanczos_lngamma1(a - 1.0);
}
int oc_clngamma(double re, double im, double* out) {
  const double complex v = clngamma(re + I * im);
  out[0] = creal(v);
  out[1] = cimag(v);
  return 0;
}

/* FFTLog configuration of C_cl_tomo (cosmo2D.c:4566-4580): density, RSD, magnification */
typedef struct { double nu, cww; int deriv, npad; } nl_cfg;
static const nl_cfg NL_CFG[3] = {{1.0, 0.25, 0, 200}, {1.01, 0.25, 2, 500}, {1.0, 0.25, 0, 500}};

/* g_l(z) of cfftlog_ells_p2 (cosmo2D.c:4253-4469): derivative 0:
   2^z Gamma((l+z)/2) / Gamma((3+l-z)/2); derivative 2:
   (z-1)(z-2) 2^(z-2) Gamma((l+z-2)/2) / Gamma((5+l-z)/2) */
static double complex nl_gl_direct(int deriv, int l, double complex z) {
  if (deriv == 0) return cexp(z * log(2.) + clngamma(0.5 * (l + z)) - clngamma(0.5 * (3 + l - z)));
  return (z - 1) * (z - 2) * cexp((z - 2) * log(2.) + clngamma(0.5 * (l + z - 2)) - clngamma(0.5 * (5 + l - z)));
}
/* the two-step recurrence the reference uses inside an ell block */
static double complex nl_gl_step(int deriv, int l, double complex z) {
  return deriv == 0 ? (l - 2 + z) / (l + 1 - z) : (l - 4 + z) / (l + 3 - z);
}

/* One FFTLog transform of fx on the log grid x (cfftlog_ells_p1 + p2 for one
   (bin, config)): Fy[l][q] at y = (l + 1) / x[nx-1-q] for l in [ls, le).
   Fy ~ int dx/x fx(x) j_l^(deriv)(x y). */
typedef struct {
  int N0, N2, nh;
  double dlnx, nu;
  double complex* T; /* windowed forward transform [nh] */
  double* eta;       /* [nh] */
} nl_fwd;

static void nl_forward(const double* x, const double* fx, int nx, const nl_cfg* cf, nl_fwd* F) {
  F->N0 = cf->npad;
  long raw = 2 * (long)cf->npad + nx;
  if (raw % 2) raw++;
  long N2 = size_7smooth(raw);
  if (N2 % 2) N2 = size_7smooth(N2 + 1);
  F->N2 = (int)N2;
  F->nh = F->N2 / 2 + 1;
  F->nu = cf->nu;
  F->dlnx = log(x[1] / x[0]);
  double complex* in = calloc(F->N2, sizeof(double complex));
  double complex* out = malloc(sizeof(double complex) * F->N2);
  for (int q = 0; q < nx; q++) in[F->N0 + q] = fx[q] / pow(x[q], cf->nu);
  fft_rec(in, out, F->N2, 1, -1);
  F->T = malloc(sizeof(double complex) * F->nh);
  for (int m = 0; m < F->nh; m++) F->T[m] = out[m];
  const int halfN = F->N2 / 2;
  const int kmax = (int)(halfN * cf->cww);
  for (int k = 0; k <= kmax; k++) { /* c_window_cfft (cosmo2D.c:4103-4116) */
    const double W = (double)k / kmax - sin(2. * M_PI * k / kmax) / (2. * M_PI);
    F->T[halfN - k] *= W;
  }
  F->eta = malloc(sizeof(double) * F->nh);
  for (int m = 0; m < F->nh; m++) F->eta[m] = (2.0 * M_PI / (F->dlnx * F->N2)) * m;
  free(in); free(out);
}

static void nl_backward(const double* x, int nx, const nl_cfg* cf, const nl_fwd* F, int ls, int le,
                        double* Fy /* [le-ls][nx] */) {
  const int N2 = F->N2, nh = F->nh;
  double complex* gl = malloc(sizeof(double complex) * (size_t)(le - ls) * nh);
  for (int m = 0; m < nh; m++) {
    const double complex z = F->nu + I * F->eta[m];
    for (int l = ls; l < le; l++)
      gl[(size_t)(l - ls) * nh + m] =
          (l < ls + 2) ? nl_gl_direct(cf->deriv, l, z) : gl[(size_t)(l - ls - 2) * nh + m] * nl_gl_step(cf->deriv, l, z);
  }
  const double base = x[0] / exp(2 * F->N0 * F->dlnx);
  double complex* spec = malloc(sizeof(double complex) * N2);
  double complex* out = malloc(sizeof(double complex) * N2);
  for (int l = ls; l < le; l++) {
    const double lnbase = log(base * (l + 1.) / x[nx - 1]);
    for (int m = 0; m < nh; m++) {
      const double ph = -F->eta[m] * lnbase;
      const double complex v = F->T[m] * (cos(ph) + I * sin(ph)) * gl[(size_t)(l - ls) * nh + m];
      spec[m] = conj(v);
    }
    /* c2r: Hermitian extension, imaginary parts of the DC and Nyquist terms dropped */
    spec[0] = creal(spec[0]);
    spec[N2 / 2] = creal(spec[N2 / 2]);
    for (int m = 1; m < N2 / 2; m++) spec[N2 - m] = conj(spec[m]);
    fft_rec(spec, out, N2, 1, +1);
    const double pre = sqrt(M_PI) / (4. * N2 * pow(l + 1., cf->nu));
    for (int q = 0; q < nx; q++) Fy[(size_t)(l - ls) * nx + q] = creal(out[F->N0 + q]) * pre * pow(x[nx - 1 - q], cf->nu);
  }
  free(gl); free(spec); free(out);
}

static void nl_free(nl_fwd* F) { free(F->T); free(F->eta); }

int oc_fftlog(const double* x, const double* fx, int nx, int cfg, int ls, int le, double* Fy) {
  if (cfg < 0 || cfg > 2 || nx < 2 || le <= ls || ls < 0) return -1;
  nl_fwd F;
  nl_forward(x, fx, nx, &NL_CFG[cfg], &F);
  nl_backward(x, nx, &NL_CFG[cfg], &F, ls, le, Fy);
  nl_free(&F);
  return 0;
}

/* C_cl_tomo (cosmo2D.c:4528-4769) with tol: Cl[nl][OC_LMAX_NOLIMBER] */
static int C_cl_tomo(oc_ctx* c, double* Cl, double tol) {
  if (!c->lnPL) { oc_set_err(c, "non-Limber w needs the linear P(k,z) (set_linear_power_spectrum)"); return -1; }
  const int nb = c->nl, nchi = OC_NL_NCHI, LN = OC_LMAX_NOLIMBER;
  const double rc = OC_COVERH0 / c->h0;
  const double chi_min = chi_of_a(c, 1. / (1.0 + 0.002)) * rc;
  const double chi_max = chi_of_a(c, 1. / (1.0 + 4.0)) * rc;
  const double dlnchi = log(chi_max / chi_min) / ((double)nchi - 1.0);
  const double dlnk = dlnchi;
  double* x = malloc(sizeof(double) * nchi);
  double* fx = calloc((size_t)nb * 3 * nchi, sizeof(double));
  for (int j = 0; j < nchi; j++) {
    x[j] = chi_min * exp(dlnchi * j);
    const double chi = x[j] / rc;
    const double a = a_chi(c, chi);
    const double z = 1. / a - 1.;
    const double hh = hoverh0(c, a);
    const double D = growfac(c, a), f = f_growth(c, z);
    for (int i = 0; i < nb; i++) {
      if (z < c->l_zmin[i] || z > c->l_zmax[i]) continue;
      const double pf = nz_lens(c, z, i);
      const double WM = W_mag(c, a, chi, i);
      fx[((size_t)i * 3 + 0) * nchi + j] = chi * pf * D * hh * c->b1[i];
      fx[((size_t)i * 3 + 1) * nchi + j] = -chi * pf * D * hh * f;
      fx[((size_t)i * 3 + 2) * nchi + j] = (WM / chi / (rc * rc)) * D;
    }
  }
  int bmag_zero = 1;
  for (int i = 0; i < nb; i++)
    if (fabs(c->bmag[i]) > 1.e-12) bmag_zero = 0;
  const int S2 = bmag_zero ? 2 : 3;
  nl_fwd* F = malloc(sizeof(nl_fwd) * (size_t)nb * 3);
  for (int i = 0; i < nb; i++)
    for (int jj = 0; jj < S2; jj++) nl_forward(x, fx + ((size_t)i * 3 + jj) * nchi, nchi, &NL_CFG[jj], &F[i * 3 + jj]);
  /* the gg Limber table (C_gg_tomo_limber) for the ells past convergence */
  double* tab = malloc(sizeof(double) * (size_t)nb * c->n_ell);
  table_gg(c, tab);
  double lim[3];
  gx_lim(c, lim);
  gl_rule r96;
  gl_make(&r96, gl_order_gg(c->hdi));
  int conv[OC_MAXB];
  for (int i = 0; i < nb; i++) { conv[i] = 0; c->nl_lmax[i] = LN; }
  double* Fy = malloc(sizeof(double) * 3 * OC_NL_BLOCK * nchi);
  for (int ks = 0; ks < LN;) {
    int all = 1;
    for (int i = 0; i < nb; i++) all &= conv[i];
    if (all) break;
    const int ke = ks + OC_NL_BLOCK < LN ? ks + OC_NL_BLOCK : LN;
    for (int i = 0; i < nb; i++) {
      if (conv[i]) continue;
      for (int jj = 0; jj < 3; jj++) {
        if (jj < S2) nl_backward(x, nchi, &NL_CFG[jj], &F[i * 3 + jj], ks, ke, Fy + (size_t)jj * OC_NL_BLOCK * nchi);
        else memset(Fy + (size_t)jj * OC_NL_BLOCK * nchi, 0, sizeof(double) * OC_NL_BLOCK * nchi);
      }
      const int kk = ke;
      for (int k = ks; k < kk; k++) {
        double tcl = 0.0;
        const double ell_prefactor = k * (k + 1.);
        for (int q = 0; q < nchi; q++) {
          const double ty = (k + 1.) / x[nchi - 1 - q];
          const double k1 = ty * rc;
          const size_t o = (size_t)(k - ks) * nchi + q;
          const double Fv = Fy[o] + Fy[OC_NL_BLOCK * nchi + o] +
                            c->bmag[i] * ell_prefactor * Fy[2 * OC_NL_BLOCK * nchi + o] / (ty * ty);
          tcl += Fv * Fv * (k1 * k1 * k1) * p_lin(c, k1, 1.0);
        }
        Cl[(size_t)i * LN + k] = tcl * dlnk * 2. / M_PI + C_gg_nointerp_ps(c, &r96, (double)k, i, 0) -
                                 C_gg_nointerp_ps(c, &r96, (double)k, i, 1);
      }
      const int L = kk - 1;
      const double denom = C_gg_nointerp(c, &r96, (double)L, i);
      if (fabs(denom) > 1e-300) {
        const double dev = Cl[(size_t)i * LN + L] / denom - 1.0;
        if (isfinite(dev) && fabs(dev) < tol) { conv[i] = 1; c->nl_lmax[i] = kk; }
      }
    }
    ks = ke;
  }
  for (int i = 0; i < nb; i++)
    for (int k = c->nl_lmax[i]; k < LN; k++)
      Cl[(size_t)i * LN + k] = (k > OC_LMIN_TAB) ? interpol1d(tab + (size_t)i * c->n_ell, c->n_ell, lim[0], lim[1], lim[2], log((double)k))
                                                 : C_gg_nointerp(c, &r96, (double)k, i);
  for (int i = 0; i < nb; i++)
    for (int jj = 0; jj < S2; jj++) nl_free(&F[i * 3 + jj]);
  free(F); free(Fy); free(tab); free(x); free(fx);
  return c->status;
}

int oc_cl_nonlimber(oc_ctx* c, double* cl /* [nl][150] */, int* lmax /* [nl] or NULL */) {
  c->status = 0;
  if (ensure_g(c)) return -1;
  if (C_cl_tomo(c, cl, 0.01)) return -1;
  if (lmax)
    for (int i = 0; i < c->nl; i++) lmax[i] = c->nl_lmax[i];
  return c->status;
}

int oc_set_h0(oc_ctx* c, double h0) { c->h0 = h0; return 0; }
int oc_set_limber_gg(oc_ctx* c, int limber) { c->limber_gg = limber != 0; return 0; }

/* set_linear_power_spectrum (generic_interface.cpp:1194-1270): lnPL k-major */
int oc_set_plin(oc_ctx* c, const double* lk, int nk, const double* z, int nz, const double* lnP) {
  c->status = 0;
  if (nk < 2 || nz < 2 || any_nan(lk, nk) || any_nan(z, nz) || any_nan(lnP, nk * nz)) {
    oc_set_err(c, "bad linear P(k,z)"); return -1;
  }
  free(c->lkL); free(c->pzL); free(c->lnPL);
  c->lkL = dup(lk, nk); c->pzL = dup(z, nz); c->lnPL = dup(lnP, nk * nz);
  c->nkL = nk; c->nzL = nz;
  return 0;
}

/* the FFTLog inputs of C_cl_tomo (cosmo2D.c:4586-4628): x[nchi] (Mpc) and
   fx[nl][3][nchi] (density, RSD, magnification), for the tests */
int oc_nl_inputs(oc_ctx* c, double* x_out, double* fx_out) {
  c->status = 0;
  if (ensure_g(c)) return -1;
  const int nb = c->nl, nchi = OC_NL_NCHI;
  const double rc = OC_COVERH0 / c->h0;
  const double chi_min = chi_of_a(c, 1. / (1.0 + 0.002)) * rc;
  const double chi_max = chi_of_a(c, 1. / (1.0 + 4.0)) * rc;
  const double dlnchi = log(chi_max / chi_min) / ((double)nchi - 1.0);
  for (size_t i = 0; i < (size_t)nb * 3 * nchi; i++) fx_out[i] = 0.0;
  for (int j = 0; j < nchi; j++) {
    const double x = chi_min * exp(dlnchi * j);
    x_out[j] = x;
    const double chi = x / rc;
    const double a = a_chi(c, chi);
    const double z = 1. / a - 1.;
    const double hh = hoverh0(c, a);
    const double D = growfac(c, a), f = f_growth(c, z);
    for (int i = 0; i < nb; i++) {
      if (z < c->l_zmin[i] || z > c->l_zmax[i]) continue;
      const double pf = nz_lens(c, z, i);
      fx_out[((size_t)i * 3 + 0) * nchi + j] = chi * pf * D * hh * c->b1[i];
      fx_out[((size_t)i * 3 + 1) * nchi + j] = -chi * pf * D * hh * f;
      fx_out[((size_t)i * 3 + 2) * nchi + j] = (W_mag(c, a, chi, i) / chi / (rc * rc)) * D;
    }
  }
  return c->status;
}

/* C_gg_tomo_limber_linpsopt_nointerp(l, i, i, use_linear_ps = 1) */
int oc_cl_gg_lin_nointerp(oc_ctx* c, const double* l, int nl, double* out) {
  c->status = 0;
  if (ensure_g(c)) return -1;
  if (!c->lnPL) { oc_set_err(c, "linear P(k,z) not set"); return -1; }
  gl_rule r;
  gl_make(&r, gl_order_gg(c->hdi));
  for (int k = 0; k < c->nl; k++)
    for (int i = 0; i < nl; i++) out[(size_t)k * nl + i] = C_gg_nointerp_ps(c, &r, l[i], k, 1);
  return c->status;
}

/* set_nuisance_nonlinear_bias / set_nuisance_nonlocal_bias
   (generic_interface.cpp:1533-1669): b2 per lens bin with
   bs2 = -4/7 (b1 - 1) unless b2 = 0 (then 0); b3 and bK per lens bin.
   NULL b3 / bk leave them unchanged. */
int oc_set_bias_nl(oc_ctx* c, const double* b1, const double* b2, const double* b3, const double* bk) {
  for (int i = 0; i < c->nl; i++) {
    if (b2) {
      c->b2[i] = b2[i];
      c->bs2[i] = (b2[i] == 0.0) ? 0.0 : (-4. / 7.) * (b1[i] - 1.0);
    }
    if (b3) c->b3[i] = b3[i];
    if (bk) c->bk[i] = bk[i];
  }
  return 0;
}

/* set_bias_PS (generic_interface.cpp:1008-1076): PS[12][N], rows 0..5 the
   FAST-PT bias spectra (d1d2, d2d2, d1s2, d2s2, s2s2, d1p3) in (Mpc/h)^3 ->
   / (c/H0)^3; k range in h/Mpc -> * c/H0; sigma4 in (Mpc/h)^3 -> / (c/H0)^3 */
int oc_set_bias_ps(oc_ctx* c, const double* PS, double kmin, double kmax, double cutoff, double sigma4, int N) {
  (void)cutoff;
  if (N < 2) { oc_set_err(c, "set_bias_PS: N < 2"); return -1; }
  free(c->fptb);
  c->fptb = malloc(sizeof(double) * 6 * (size_t)N);
  const double h3 = OC_COVERH0 * OC_COVERH0 * OC_COVERH0;
  for (int t = 0; t < 6; t++)
    for (int j = 0; j < N; j++) c->fptb[(size_t)t * N + j] = PS[(size_t)t * N + j] / h3;
  c->fptb_n = N;
  c->fptb_kmin = kmin * OC_COVERH0;
  c->fptb_kmax = kmax * OC_COVERH0;
  c->fptb_s4 = sigma4 / h3;
  return 0;
}

/* ---------------------------------------------------------------------- */
/* 6x2pt: CMB lensing convergence cross galaxies (gk), shear (ks), kk     */
/* ---------------------------------------------------------------------- */
/* g_cmb (redshift_spline.c:1111-1125) for a flat background (f_K(chi) = chi,
   Omega_v = 1 - Omega_m as set_cosmological_parameters sets it) */
static double g_cmb(oc_ctx* c, double a) {
  const double chi_cmb = chi_of_a(c, 1. / 1091.);
  return (chi_cmb - chi_of_a(c, a)) / chi_cmb;
}
/* W_k (radial_weights.c:107-113) */
static double W_k(oc_ctx* c, double a, double fK) { return (1.5 * c->omega_m * fK / a) * g_cmb(c, a); }

/* int_for_C_gk_tomo_limber (cosmo2D.c:2857-2954), RSD off (include_RSD_GK = 0) */
static double int_gk(oc_ctx* c, double a, int nl, double l) {
  const double ell = l + 0.5;
  double fK, dchida;
  chi_all(c, a, &fK, &dchida);
  const double hh = hoverh0v2(a, dchida);
  const double k = ell / fK;
  const double WK = W_k(c, a, fK);
  const double WGAL = W_gal(c, a, nl, hh), WMAG = W_mag(c, a, fK, nl);
  const double ell_prefactor = l * (l + 1.) / (ell * ell);
  const double PK = p_nonlin(c, k, a);
  const double res = WK * (WGAL * c->b1[nl] + WMAG * ell_prefactor * c->bmag[nl]) * PK;
  double oneloop = 0.0;
  if (nlb(c)) {
    if (!c->fptb) { oc_set_err(c, "one-loop bias needs the FAST-PT bias tables (set_bias_PS)"); return NAN; }
    const double g = growfac(c, a), g4 = g * g * g * g, lnk = log(k);
    oneloop = WK * WGAL *
              (g4 * (0.5 * c->b2[nl] * fptb_at(c, 0, lnk) + 0.5 * c->bs2[nl] * fptb_at(c, 2, lnk) +
                     0.5 * c->b3[nl] * fptb_at(c, 5, lnk)) +
               c->bk[nl] * k * k * PK);
  }
  return ((res + oneloop) * dchida / (fK * fK)) * ell_prefactor;
}

/* int_for_C_ks_tomo_limber (cosmo2D.c:3178-3211) */
static double int_ks(oc_ctx* c, double a, int ns, double l) {
  const double ell = l + 0.5;
  const double growfac_a = growfac(c, a);
  double fK, dchida;
  chi_all(c, a, &fK, &dchida);
  const double hh = hoverh0v2(a, dchida);
  const double k = ell / fK;
  const double PK = p_nonlin(c, k, a);
  const double WK1 = W_kappa(c, a, fK, ns), WK2 = W_k(c, a, fK);
  const double ell_prefactor1 = l * (l + 1.) / (ell * ell);
  const double tmp = (l - 1.) * l * (l + 1.) * (l + 2.);
  const double ell_prefactor2 = (tmp > 0) ? sqrt(tmp) / (ell * ell) : 0.0;
  double IA[2];
  IA_A1_Z1Z2(c, a, growfac_a, ns, ns, IA);
  const double WS1 = W_source(c, a, ns, hh) * IA[0];
  return ((WK1 - WS1) * WK2 * PK * dchida / (fK * fK)) * ell_prefactor1 * ell_prefactor2;
}

/* int_for_C_kk_limber (cosmo2D.c:3428-3446) */
static double int_kk(oc_ctx* c, double a, double l) {
  double fK, dchida;
  chi_all(c, a, &fK, &dchida);
  const double ell = l + 0.5;
  const double k = ell / fK;
  const double WK = W_k(c, a, fK);
  const double ell_prefactor = l * (l + 1.0) / (ell * ell);
  return WK * WK * p_nonlin(c, k, a) * (dchida / (fK * fK)) * ell_prefactor * ell_prefactor;
}

/* The Limber integrands at scale factors a[0..n) with the reference's
   parameter arrays (int_for_C_*_tomo_limber, cosmo2D.c:1521-1555,
   2079-2119, 2489-2606, 2857-2954, 3178-3211, 3428-3446): probe 0 ss
   {n1, n2, l, EE, deriv}, 1 gs {nl, ns, l, nonlinear_bias}, 2 gg {ni, nj, l,
   use_linear_ps, nonlinear_bias}, 3 gk {nl, l, nonlinear_bias}, 4 ks {ns, l},
   5 kk {l}.  ss with deriv = 1 returns the d C / d ln k integrand (dchi/da /
   f_K instead of / f_K^2). */
int oc_int_for(oc_ctx* c, int probe, const double* ar, const double* a, int n, double* out) {
  c->status = 0;
  if (ensure_g(c)) return -1;
  c->nlb_set = 1;
  c->nlb_on = probe == 1 ? (int)ar[3] : probe == 2 ? (int)ar[4] : probe == 3 ? (int)ar[2] : 0;
  for (int i = 0; i < n; i++) {
    double v;
    switch (probe) {
      case 0: {
        v = int_ss(c, a[i], (int)ar[0], (int)ar[1], ar[2], (int)ar[3]);
        if ((int)ar[4] == 1) v *= chi_of_a(c, a[i]);
        break;
      }
      case 1: v = int_gs(c, a[i], (int)ar[0], (int)ar[1], ar[2]); break;
      case 2: v = int_gg_ps(c, a[i], (int)ar[0], ar[2], (int)ar[3]); break;
      case 3: v = int_gk(c, a[i], (int)ar[0], ar[1]); break;
      case 4: v = int_ks(c, a[i], (int)ar[0], ar[1]); break;
      default: v = int_kk(c, a[i], ar[0]);
    }
    out[i] = v;
  }
  c->nlb_set = 0;
  return c->status;
}

/* C_gk_tomo_limber_nointerp (:2960-3015), C_ks_tomo_limber_nointerp
   (:3217-3264; amin_source / amax_source, redshift_spline.c:27-39),
   C_kk_limber_nointerp (:3449-3481): 96-point GL */
static double C_cmb_nointerp(oc_ctx* c, const gl_rule* r, int probe, double l, int bin) {
  double amin, amax;
  if (probe == 3) { amin = amin_lens(c, bin); amax = amax_lens(c, bin); }
  else if (probe == 4) { amin = 1. / (c->s_zmax_all + 1.); amax = 1. / (1. + fmax(c->s_zmin_all, 0.001)); }
  else { amin = OC_A_MIN * (1. + 1.e-5); amax = 0.99999; }
  double s = 0.0;
  for (int p = 0; p < r->n; p++) {
    double x, w;
    gl_point(r, amin, amax, p, &x, &w);
    s += w * (probe == 3 ? int_gk(c, x, bin, l) : probe == 4 ? int_ks(c, x, bin, l) : int_kk(c, x, l));
  }
  return s;
}

int oc_set_cmb(oc_ctx* c, double fwhm_arcmin, int lmin, int lmax, const double* win, int nwin) {
  c->cmb_fwhm = fwhm_arcmin * 2.90888208665721580e-4; /* init_cmb_cross_correlation, generic_interface.cpp:519 */
  c->cmb_lmin = lmin;
  c->cmb_lmax = lmax;
  free(c->cmb_win);
  c->cmb_win = NULL;
  c->cmb_nwin = 0;
  if (win && nwin > 0) {
    c->cmb_win = dup(win, nwin);
    c->cmb_nwin = nwin;
  }
  return 0;
}

/* beam_cmb x w_pixel (cosmo2D.c:268-279, 1003-1010) */
static double cmb_filter(oc_ctx* c, int l) {
  const double s = c->cmb_fwhm / sqrt(16.0 * log(2.0));
  double f = (l < c->cmb_lmin || l > c->cmb_lmax) ? 0.0 : exp(-l * (l + 1.0) * s * s);
  if (c->cmb_nwin > 0) f *= (l < c->cmb_nwin) ? c->cmb_win[l] : 0.0;
  return f;
}

int oc_cl_cmb_nointerp(oc_ctx* c, int probe, const double* l, int n, double* out) {
  c->status = 0;
  if (ensure_g(c)) return -1;
  if (probe < 3 || probe > 5) { oc_set_err(c, "probe %d", probe); return -1; }
  gl_rule r;
  gl_make(&r, gl_order_gg(c->hdi));
  const int nb = probe == 3 ? c->nl : probe == 4 ? c->ns : 1;
  #pragma omp parallel for collapse(2) schedule(static)
  for (int b = 0; b < nb; b++)
    for (int i = 0; i < n; i++) out[(size_t)b * n + i] = C_cmb_nointerp(c, &r, probe, l[i], b);
  return c->status;
}

/* C_gk / C_ks / C_kk tables at the gs/gg ell grid (cosmo2D.c:3020-3060,
   3270-3310, 3484-3516) */
static void table_cmb(oc_ctx* c, int probe, double* tab /* [nb][nell] */) {
  const int nell = c->n_ell;
  double lim[3];
  gx_lim(c, lim);
  gl_rule r;
  gl_make(&r, gl_order_gg(c->hdi));
  const int nb = probe == 3 ? c->nl : probe == 4 ? c->ns : 1;
  #pragma omp parallel for collapse(2) schedule(static)
  for (int b = 0; b < nb; b++)
    for (int i = 0; i < nell; i++) tab[(size_t)b * nell + i] = C_cmb_nointerp(c, &r, probe, exp(lim[0] + i * lim[2]), b);
}

/* w_gk_tomo / w_ks_tomo (cosmo2D.c:913-1296): C_ell direct below LMIN_tab,
   the table fill above, times the CMB filter, then the w (gk) or gamma_t
   (ks) bin-averaged kernel; out[nb][ntheta] */
static int w_cmb(oc_ctx* c, int probe, double* out) {
  if (!c->K_valid && build_kernels(c)) return -1;
  if (ensure_g(c)) return -1;
  const int L = c->lmax, nell = c->n_ell, nb = probe == 3 ? c->nl : c->ns;
  double* tab = malloc(sizeof(double) * (size_t)nb * nell);
  double* cl = malloc(sizeof(double) * (size_t)nb * L);
  table_cmb(c, probe, tab);
  double lim[3];
  gx_lim(c, lim);
  gl_rule r;
  gl_make(&r, gl_order_gg(c->hdi));
  #pragma omp parallel for schedule(dynamic)
  for (int b = 0; b < nb; b++) {
    double* o = cl + (size_t)b * L;
    o[0] = 0.0;
    for (int l = 1; l < OC_LMIN_TAB; l++) o[l] = C_cmb_nointerp(c, &r, probe, (double)l, b) * cmb_filter(c, l);
    fill(tab + (size_t)b * nell, lim, nell, OC_LMIN_TAB, L, o);
    for (int l = OC_LMIN_TAB; l < L; l++) o[l] *= cmb_filter(c, l);
  }
  contract(c->K + (size_t)(probe == 3 ? 3 : 2) * c->ntheta * L, cl, nb, c->ntheta, L, out);
  free(tab); free(cl);
  return c->status;
}
int oc_w_gk(oc_ctx* c, double* out) { c->status = 0; return w_cmb(c, 3, out); }
int oc_w_ks(oc_ctx* c, double* out) { c->status = 0; return w_cmb(c, 4, out); }

/* the kk block at like.ell (generic_interface.hpp:1085-1094): direct at
   ell <= LMIN_tab, C_kk_limber's table interpolation above */
int oc_cl_kk(oc_ctx* c, const double* ell, int n, double* out) {
  c->status = 0;
  if (ensure_g(c)) return -1;
  double* tab = malloc(sizeof(double) * (size_t)c->n_ell);
  table_cmb(c, 5, tab);
  double lim[3];
  gx_lim(c, lim);
  gl_rule r;
  gl_make(&r, gl_order_gg(c->hdi));
  for (int i = 0; i < n; i++)
    out[i] = (ell[i] <= OC_LMIN_TAB) ? C_cmb_nointerp(c, &r, 5, ell[i], 0)
                                    : interpol1d(tab, c->n_ell, lim[0], lim[1], lim[2], log(ell[i]));
  free(tab);
  return c->status;
}

int oc_set_kk_bandpower(oc_ctx* c, int nbp, int lmin, int lmax, const double* M, const double* off) {
  free(c->kk_M); free(c->kk_off);
  c->kk_M = c->kk_off = NULL;
  c->kk_nbp = 0;
  if (nbp <= 0) return 0;
  c->kk_nbp = nbp;
  c->kk_lmin = lmin;
  c->kk_lmax = lmax;
  c->kk_M = dup(M, nbp * (lmax - lmin + 1));
  c->kk_off = malloc(sizeof(double) * nbp);
  for (int j = 0; j < nbp; j++) c->kk_off[j] = off ? off[j] : 0.0;
  return 0;
}

/* the kk block as band powers (generic_interface.hpp:1095-1118):
   sum_L M[j][L - lmin] C_kk(L) - offset[j], C_kk direct at L <= LMIN_tab */
static int kk_bandpower(oc_ctx* c, double* out) {
  double* tab = malloc(sizeof(double) * (size_t)c->n_ell);
  table_cmb(c, 5, tab);
  double lim[3];
  gx_lim(c, lim);
  gl_rule r;
  gl_make(&r, gl_order_gg(c->hdi));
  const int nL = c->kk_lmax - c->kk_lmin + 1;
  for (int j = 0; j < c->kk_nbp; j++) out[j] = 0.0;
  for (int L = c->kk_lmin; L <= c->kk_lmax; L++) {
    const double ckk = (L <= OC_LMIN_TAB) ? C_cmb_nointerp(c, &r, 5, (double)L, 0)
                                          : interpol1d(tab, c->n_ell, lim[0], lim[1], lim[2], log((double)L));
    for (int j = 0; j < c->kk_nbp; j++) out[j] += ckk * c->kk_M[(size_t)j * nL + (L - c->kk_lmin)];
  }
  for (int j = 0; j < c->kk_nbp; j++) out[j] -= c->kk_off[j];
  free(tab);
  return c->status;
}

int oc_datavector_size_6x2pt(oc_ctx* c, int ncl) {
  return c->ntheta * (2 * oc_npairs_ss(c) + oc_npairs_ggl(c) + 2 * c->nl + c->ns) + (c->kk_nbp > 0 ? c->kk_nbp : ncl);
}

/* compute_Mx2pt_N_masked<0,6> (generic_interface.hpp:1137-1159) with every
   mask entry 1: blocks xi+- | gamma_t | w | w_gk | w_ks | kk placed by ord
   (compute_data_vector_Mx2pt_N_starts<0,6>, :770-791); point mass and
   (1+m) on gamma_t, (1+m_s) on w_ks (add_calib_and_set_mask_X_N, :800-920);
   kk at the ncl ells (like.ell) */
int oc_datavector_6x2pt(oc_ctx* c, const int ord[6], const double* ell, int ncl, double* dv) {
  c->status = 0;
  const int nt = c->ntheta, nss = oc_npairs_ss(c), ngs = oc_npairs_ggl(c), nl = c->nl, ns = c->ns;
  const int nkk = c->kk_nbp > 0 ? c->kk_nbp : ncl;
  const int sizes[6] = {2 * nt * nss, nt * ngs, nt * nl, nt * nl, nt * ns, nkk};
  for (int i = 0; i < 6; i++)
    if (ord[i] < 0 || ord[i] > 5) { oc_set_err(c, "ord"); return -1; }
  int idx[6] = {0, 1, 2, 3, 4, 5};
  for (int i = 1; i < 6; i++)
    for (int j = i; j > 0 && ord[idx[j - 1]] > ord[idx[j]]; j--) { int t = idx[j]; idx[j] = idx[j - 1]; idx[j - 1] = t; }
  int start[6];
  int n = 0;
  for (int i = 0; i < 6; i++) {
    n += sizes[i];
    start[i] = 0;
    for (int j = 0; j < idx[i]; j++) start[i] += sizes[idx[j]];
  }
  for (int i = 0; i < 6; i++) {
    if (start[i] + sizes[i] > n) { oc_set_err(c, "ord: blocks do not tile the data vector"); return -1; }
    for (int j = 0; j < i; j++)
      if (sizes[i] && sizes[j] && start[i] < start[j] + sizes[j] && start[j] < start[i] + sizes[i]) {
        oc_set_err(c, "ord: blocks do not tile the data vector");
        return -1;
      }
  }
  const int n3 = oc_datavector_size(c);
  double* d3 = malloc(sizeof(double) * n3);
  const int ord3[3] = {0, 1, 2};
  if (oc_datavector_3x2pt(c, ord3, d3)) { free(d3); return -1; }
  double* gk = malloc(sizeof(double) * (size_t)nl * nt);
  double* ks = malloc(sizeof(double) * (size_t)ns * nt);
  double* kk = malloc(sizeof(double) * (size_t)(nkk > 0 ? nkk : 1));
  if (oc_w_gk(c, gk) || oc_w_ks(c, ks) || (c->kk_nbp > 0 ? kk_bandpower(c, kk) : oc_cl_kk(c, ell, ncl, kk))) {
    free(d3); free(gk); free(ks); free(kk);
    return -1;
  }
  for (int i = 0; i < n; i++) dv[i] = 0.0;
  for (int k = 0; k < sizes[0]; k++) dv[start[0] + k] = d3[k];
  for (int k = 0; k < sizes[1]; k++) dv[start[1] + k] = d3[sizes[0] + k];
  for (int k = 0; k < sizes[2]; k++) dv[start[2] + k] = d3[sizes[0] + sizes[1] + k];
  for (int k = 0; k < sizes[3]; k++) dv[start[3] + k] = gk[k];
  for (int z = 0; z < ns; z++)
    for (int i = 0; i < nt; i++) dv[start[4] + nt * z + i] = ks[z * nt + i] * (1.0 + c->shear_m[z]);
  for (int k = 0; k < nkk; k++) dv[start[5] + k] = kk[k];
  free(d3); free(gk); free(ks); free(kk);
  return c->status;
}
