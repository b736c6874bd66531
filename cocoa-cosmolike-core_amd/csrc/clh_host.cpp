// Host runtime of the projection library: context, setters, cosmology-
// independent precomputation (n(z) splines, lens zmean, Gauss-Legendre rules,
// bin-averaged Legendre kernels and their fold), batch assembly and launches.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cosmolike_hip.h"
#include "clh_internal.hpp"
#include "clh_place.hpp"
#include "clh_fft.hpp"

hipError_t clh_launch_gtab(const DrawDesc* D, const int2* work, int n, hipStream_t s);
hipError_t clh_launch_nodes(const DrawDesc* D, const int2* work, int n, hipStream_t s);
hipError_t clh_launch_limber_sx(const DrawDesc* D, const int4* work, int n, int f32, int keep, hipStream_t s,
                                int ol = 0);
hipError_t clh_launch_limber_sx_tatt(const DrawDesc* D, const int4* work, int n, hipStream_t s);
hipError_t clh_launch_nl_forward(const DrawDesc* D, const int2* wprep, int nprep, const int4* wfwd, int nfwd,
                                 const int4* wgl, int ngl, const int4* wchk, int nchk, const int2* wbin, int nbin,
                                 const int4* widft, int nidft, hipStream_t s);
hipError_t clh_launch_nl_cl(const DrawDesc* D, const int2* work, int n, hipStream_t s);
hipError_t clh_launch_nl_cl_adj(const DrawDesc* D, const int2* wbin, int nbin, hipStream_t s);
hipError_t clh_launch_nl_tangent(const DrawDesc* D, const int2* wprep, int nprep, const int4* wfwd4, int nfwd4,
                                 const int4* widft, int nidft, hipStream_t s);
hipError_t clh_launch_cmb(const DrawDesc* D, const int4* work, int n, const int* wplace, int nplace, hipStream_t s);
hipError_t clh_launch_cmb_place_adj(const DrawDesc* D, const int* wplace, int nplace, hipStream_t s);
hipError_t clh_launch_limber_cmb_adj(const DrawDesc* D, const int4* work, int n, hipStream_t s);
hipError_t clh_launch_limber_gg(const DrawDesc* D, const int4* work, int n, int f32, int stage_doubles, int keep,
                                hipStream_t s, int ol = 0);
hipError_t clh_launch_legendre(const DrawDesc* D, const int4* work, int n, hipStream_t s);
hipError_t clh_launch_gtab_tangent(const DrawDesc* D, const int2* work, int n, int tables, hipStream_t s);
hipError_t clh_launch_fill(double* p, size_t n, double v, hipStream_t s);
hipError_t clh_launch_chi2_resid(int B, int n, const double* dv, const double* data, const double* mask,
                                 double* R, hipStream_t s);
hipError_t clh_launch_chi2_reduce(int B, int n, const double* R, const double* Y, const double* mask,
                                  double* chi2, double* grad, hipStream_t s);
hipError_t clh_launch_set_params(const DrawDesc* D, DrawParams* P, int B, int nknz, const double* lnP,
                                 const double* src, const double* lens, const double* b1,
                                 const double* iaA, const double* m, hipStream_t s);
hipError_t clh_launch_leg_adjoint(const DrawDesc* D, const int4* work, int n, hipStream_t s);
hipError_t clh_launch_m_grad(const DrawDesc* D, int B, hipStream_t s);
hipError_t clh_launch_limber_sx_adj(const DrawDesc* D, const int4* work, int n, int lds_doubles, int ns, int ck,
                                   int lim, int gs, int ol, hipStream_t s);
hipError_t clh_launch_limber_gg_adj(const DrawDesc* D, const int4* work, int n, int stage_doubles, int ck,
                                   int lim, int f32, int ol, hipStream_t s);
hipError_t clh_launch_nodes_tan(const DrawDesc* D, const int2* work, int n, hipStream_t s);
hipError_t clh_launch_limber_hdi_adj(const DrawDesc* D, const int4* work, int n, hipStream_t s);
hipError_t clh_launch_tatt_adj(const DrawDesc* D, const int4* work, int n, int lds_doubles, int ns, int gs,
                               hipStream_t s);
hipError_t clh_launch_limber_sx_tatt(const DrawDesc* D, const int4* work, int n, hipStream_t s);
hipError_t clh_launch_nodes_adj(const DrawDesc* D, const int2* work, int n, hipStream_t s);
hipError_t clh_launch_lim_grad(const DrawDesc* D, const DrawDesc* P, const DrawDesc* M, const int2* work, int n,
                               hipStream_t s);
hipError_t clh_launch_limber_hdi(const DrawDesc* D, const int4* work, int n, hipStream_t s);
hipError_t clh_launch_int_for(const DrawDesc* D, int probe, int i, int j, int f1, int f2, double l, const double* a,
                              int n, double* out, hipStream_t s);
hipError_t clh_launch_lnp_gather(const DrawDesc* D, const int2* work, int n, int ic, int nz, int nl,
                                 hipStream_t s);
hipError_t clh_launch_legendre_mx(const DrawDesc* D, const int4* work, int n, int mode, hipStream_t s);
hipError_t clh_launch_legendre_full(const DrawDesc* D, const int2* work, int n, int max_nrow,
                                    hipStream_t s);

namespace {

using dvec = std::vector<double>;

// ----------------------------------------------------------------- numerics
// n-point Gauss-Legendre rule on [-1,1], ascending (Newton on P_n); the rule
// gsl_integration_glfixed tabulates (cosmo2D.c:1583,1651,2226,2641).
void gauss_legendre(int n, double* x, double* w) {
  const int m = (n + 1) / 2;
  for (int i = 0; i < m; i++) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; it++) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; j++) {
        const double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      const double pp = n * (z * p1 - p2) / (z * z - 1.0);
      const double z1 = z;
      z = z1 - p1 / pp;
      if (std::fabs(z - z1) < 1e-16) break;
    }
    double p1 = 1.0, p2 = 0.0;
    for (int j = 1; j <= n; j++) {
      const double p3 = p2;
      p2 = p1;
      p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
    }
    const double pp = n * (z * p1 - p2) / (z * z - 1.0);
    x[i] = -z;
    x[n - 1 - i] = z;
    w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
  }
  if (n & 1) x[n / 2] = 0.0;
}

int bs_lo(const double* x, int n, double v) {
  int ilo = 0, ihi = n - 1;
  while (ihi > ilo + 1) {
    const int ll = (ihi + ilo) >> 1;
    if (x[ll] > v) ihi = ll; else ilo = ll;
  }
  return ilo;
}

// natural cubic spline (the gsl cspline the reference picks for
// photoz_interpolation_type = 0, basics.c:31-44): per-interval (y, b, c, d)
void natural_spline(int n, const double* x, const double* y, double* coef) {
  std::vector<double> c(n, 0.0);
  const int sys = n - 2;
  if (sys > 0) {
    std::vector<double> diag(sys), off(sys), g(sys);
    for (int i = 0; i < sys; i++) {
      const double hi = x[i + 1] - x[i], hip = x[i + 2] - x[i + 1];
      off[i] = hip;
      diag[i] = 2.0 * (hip + hi);
      g[i] = 3.0 * ((y[i + 2] - y[i + 1]) / hip - (y[i + 1] - y[i]) / hi);
    }
    for (int i = 1; i < sys; i++) {
      const double m = off[i - 1] / diag[i - 1];
      diag[i] -= m * off[i - 1];
      g[i] -= m * g[i - 1];
    }
    c[sys] = g[sys - 1] / diag[sys - 1];
    for (int i = sys - 2; i >= 0; i--) c[i + 1] = (g[i] - off[i] * c[i + 2]) / diag[i];
  }
  for (int i = 0; i < n - 1; i++) {
    const double h = x[i + 1] - x[i];
    coef[4 * i + 0] = y[i];
    coef[4 * i + 1] = (y[i + 1] - y[i]) / h - h * (c[i + 1] + 2.0 * c[i]) / 3.0;
    coef[4 * i + 2] = c[i];
    coef[4 * i + 3] = (c[i + 1] - c[i]) / (3.0 * h);
  }
}

double spline_eval(const dvec& zc, const double* coef, double x) {
  const int n = (int)zc.size();
  int i = bs_lo(zc.data(), n, x);
  if (i > n - 2) i = n - 2;
  const double* c = coef + 4 * i;
  const double dx = x - zc[i];
  return c[0] + dx * (c[1] + dx * (c[2] + dx * c[3]));
}

// -------------------------------------------------------------- structures
struct Sample {       // one n(z) sample as set_*_sample stores and splines it
  bool set = false;
  int nbin = 0, rows = 0;
  dvec zc, coef;      // spline knots, [bin][rows-1][4]
  double zmin_all = 0, zmax_all = 0;
  double zmin[CLH_MAXB] = {0}, zmax[CLH_MAXB] = {0};
};

struct Layout {
  Sample src, lens;
  double zmean[CLH_MAXB] = {0};
  int ntheta = 0;
  double vtmin = 0, vtmax = 0;
  int n_ell = 300, n_a = 225, lmax = 50000, hdi = 0;
  int ia_model = 0, ia_mode = 0;
  bool kern_valid = false;
  dvec kfold;  // [4][nrow][ntheta]
  dvec kfull;  // [4][lmax][ntheta] (FULL mode only)
  dvec ell_ss, ell_gx;
  dvec theta;               // [ntheta] bin-averaged theta [rad] (compute_binning_real_space)
  dvec ell_f;               // Fourier-space ell bins (init_binning_fourier), like.ell
  double lmax_shear = 0;    // like.lmax_shear
  std::vector<int> mask;    // likelihood mask in the (0,1,2) block order; empty = all ones
  bool nonlimber = false;   // !like.adopt_limber_gg: w from the non-Limber C_ell below ell 150
  dvec kwlow;               // [ntheta][CLH_NL_L] w kernel at integer ell (ell 0 = 0)
  // 6x2pt (init_probes("6x2pt")): CMB filter (init_cmb_cross_correlation) and
  // the folded w_gk / w_ks kernels with it
  bool sixpt = false;
  double cmb_fwhm = 0.0;    // rad
  int cmb_lmin = 0, cmb_lmax = 0;
  dvec cmb_win;             // HEALPix pixel window per ell (empty: none)
  dvec kf6;                 // [2][nrow][ntheta]
  // kk band powers (init_cmb_auto_bandpower); kk_nbp = 0: C_kk at like.ell
  int kk_nbp = 0, kk_lmin = 0, kk_lmax = 0;
  dvec kk_M;                // [kk_nbp][kk_lmax - kk_lmin + 1] binning matrix
  dvec kk_off;              // [kk_nbp] theory offset
  double kk_alpha = 1.0;    // Hartlap factor of the kkkk covariance block
  int kk_size() const { return kk_nbp > 0 ? kk_nbp : (int)ell_f.size(); }
  // the ells of the direct (untabulated) C_kk rows: like.ell, or the band
  // powers' integer L <= LMIN_tab
  dvec kk_direct_ells() const {
    if (kk_nbp <= 0) return ell_f;
    dvec e;
    for (int l = kk_lmin; l <= std::min(kk_lmax, CLH_LMIN_TAB); l++) e.push_back((double)l);
    return e;
  }
  // excluded (lens, source) pairs of gamma_t / C_gs (init_ggl_exclude,
  // generic_interface.cpp:874-896), flat (zl, zs) pairs.  Every pair keeps
  // its C_ell column (n_gs() = ns nl); the data vector holds the included
  // ones in N_ggl order (redshift_spline.c:199-223).
  std::vector<int> ggl_ex;
  bool zoverlap(int zl, int zs) const {  // test_zoverlap (redshift_spline.c:115-145)
    for (size_t k = 0; k + 1 < ggl_ex.size(); k += 2)
      if (ggl_ex[k] == zl && ggl_ex[k + 1] == zs) return false;
    return true;
  }
  // data-vector row of C_gs column zl ns + zs, -1 when excluded (N_ggl)
  std::vector<int> gs_rows() const {
    std::vector<int> r(n_gs());
    int n = 0;
    for (int zl = 0; zl < lens.nbin; zl++)
      for (int zs = 0; zs < src.nbin; zs++) r[zl * src.nbin + zs] = zoverlap(zl, zs) ? n++ : -1;
    return r;
  }
  int nrow() const { return CLH_NLOW + n_ell; }
  int n_ss() const { return src.nbin * (src.nbin + 1) / 2; }
  int n_gs() const { return src.nbin * lens.nbin; }
  int n_gs_dv() const {  // tomo.ggl_Npowerspectra
    int n = 0;
    for (int zl = 0; zl < lens.nbin; zl++)
      for (int zs = 0; zs < src.nbin; zs++) n += zoverlap(zl, zs) ? 1 : 0;
    return n;
  }
  int dv_size3() const { return ntheta * (2 * n_ss() + n_gs_dv() + lens.nbin); }
  int dv_size() const {
    return dv_size3() + (sixpt ? ntheta * (lens.nbin + src.nbin) + kk_size() : 0);
  }
};

struct Draw {  // one draw's cosmology + nuisance (host copy)
  double omega_m = 0.0;
  dvec chi_z, chi, G_z, G, lk, pz, lnP;
  double src_dz[CLH_MAXB] = {0}, lens_dz[CLH_MAXB] = {0}, lens_st[CLH_MAXB],
         b1[CLH_MAXB] = {0}, bmag[CLH_MAXB] = {0}, shear_m[CLH_MAXB] = {0};
  double ia_bin[CLH_MAXB] = {0};
  double pm[CLH_MAXB] = {0};
  double ia_A = 0.0, ia_eta = 0.0;
  // TATT: A2 and b_TA (per bin, or value / exponent), FAST-PT IA tables
  double ia2_bin[CLH_MAXB] = {0}, bta_bin[CLH_MAXB] = {0};
  double ia2_A = 0.0, ia2_eta = 0.0, bta_A = 0.0;
  dvec fpt;  // [10][fpt_n], already divided by (c/H0)^3
  int fpt_n = 0;
  double fpt_kmin = 0.0, fpt_kmax = 0.0;  // in H0/c
  double h0 = 0.0;           // H0 / 100 km/s/Mpc (set_cosmological_parameters)
  // one-loop galaxy bias: b2, bs2, b3, bK per lens bin and the FAST-PT bias tables
  double b2[CLH_MAXB] = {0}, bs2[CLH_MAXB] = {0}, b3[CLH_MAXB] = {0}, bk[CLH_MAXB] = {0};
  dvec fptb;                 // [6][fptb_n], already divided by (c/H0)^3
  int fptb_n = 0;
  double fptb_kmin = 0.0, fptb_kmax = 0.0, fptb_s4 = 0.0;  // k in H0/c; sigma4 / (c/H0)^3
  bool has_b2(int nl) const {
    for (int i = 0; i < nl; i++)
      if (b2[i] != 0.0) return true;
    return false;
  }
  dvec lkL, pzL, lnPL;       // linear P(k,z) (set_linear_power_spectrum)
  Draw() { for (int i = 0; i < CLH_MAXB; i++) lens_st[i] = 1.0; }
};

// While set (clh_compute_3x2pt_real_scratch), DBuf::ensure takes its memory
// from this bump allocator over caller-provided device memory instead of
// hipMalloc; such slices are never freed.
thread_local ClhArena* t_arena = nullptr;

template <class T>
struct DBuf {  // device buffer that grows
  T* p = nullptr;
  size_t cap = 0;
  bool ext = false;  // a slice of an arena: not owned
  hipError_t ensure(size_t n) {
    if (t_arena) {
      if (p && !ext) return hipErrorInvalidValue;  // an owned buffer is never traded for scratch
      const size_t bytes = (std::max<size_t>(n, 1) * sizeof(T) + 255) & ~size_t(255);
      if (t_arena->used + bytes > t_arena->cap) {
        t_arena->exhausted = true;
        return hipErrorOutOfMemory;
      }
      p = (T*)(t_arena->base + t_arena->used);
      t_arena->used += bytes;
      cap = n;
      ext = true;
      return hipSuccess;
    }
    if (ext) { p = nullptr; cap = 0; ext = false; }
    if (n <= cap) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    hipError_t e = hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T));
    if (e == hipSuccess) cap = n;
    return e;
  }
  void release() { if (p && !ext) (void)hipFree(p); p = nullptr; cap = 0; ext = false; }
};

struct DevLayout {  // device copy of one survey layout
  bool valid = false;
  DBuf<double> buf;
  DBuf<float> k32, ksc;  // folded kernels packed for the MFMA contraction
  DBuf<uint16_t> kbf;
  int ngrp = 0;
  DBuf<int> mask;       // likelihood mask (null when the layout has none)
  DBuf<int> gsrow;      // data-vector row per C_gs column (null without excluded pairs)
  size_t off_s_zc = 0, off_s_coef = 0, off_l_zc = 0, off_l_coef = 0, off_ell_ss = 0,
         off_ell_gx = 0, off_kfold = 0, off_kfoldT = 0, off_gl = 0, off_kfull = 0, off_fill = 0, off_theta = 0,
         off_kwlow = 0, off_nltw = 0, off_kf6 = 0, off_ellf = 0, off_kkM = 0, off_kkoff = 0;
  void release() {
    buf.release(); k32.release(); ksc.release(); kbf.release(); mask.release(); gsrow.release(); valid = false;
  }
};

struct LayoutCopy {  // a context's own copy of a captured layout, with its device copy
  Layout L;
  DevLayout D;
  uint64_t epoch = 0;
};

struct Batch {
  int B = 0;
  int max_nrow = 0;
  // ss / gs adjoint: the work list holds every ss item, then every gs item,
  // launched separately with their own LDS (doubles) so the small gs tiles
  // are not held to the occupancy of the ss ones
  int adj_lds_ss = 0, adj_lds_gs = 0, n_wl_ss = 0;
  bool ctx_layout = true;     // every draw uses the context's own layout
  uint64_t lay_epoch = 0;     // the context's layout_epoch when a ctx_layout batch was built
  bool uniform = true;        // every draw has the same layout
  std::vector<const Layout*> lays;
  std::vector<long long> dv_off;  // [B+1] data-vector offsets of the draws
  std::vector<DrawDesc> desc;
  std::vector<DrawParams> par;
  DBuf<DrawDesc> d_desc;
  DBuf<DrawParams> d_par;
  DBuf<double> d_tab, d_work, d_dv, d_ells, d_vjp, d_gscr, d_run;
  DBuf<int2> d_wg, d_wn, d_wl2, d_wlg;
  DBuf<int4> d_wa4;
  int n_wlg = 0, lnp_ic = 0;  // ln P gather: work items (draw, first k row), k rows per item
  DBuf<int4> d_wl, d_wsx, d_wgg, d_wga, d_wmx, d_wlf;
  int n_wg = 0, n_wn = 0, n_wl = 0, n_wl2 = 0, n_wa = 0, n_wsx = 0, n_wgg = 0, n_wga = 0, n_wmx = 0, n_wlf = 0;
  int n_wsx_std = 0;  // ss/gs tiles of non-TATT draws, first in d_wsx; the TATT draws' tiles follow
  // the VJP's lens-limit term: two views of each draw whose lens integration
  // limits take |dz| + lim_eps and |dz| - lim_eps (own node data and C_ell
  // rows in d_lim), the gs tiles of the non-TATT draws and the (draw, lens
  // bin) items
  std::vector<DrawDesc> desc_lim;
  DBuf<DrawDesc> d_desc_lim;
  DBuf<double> d_lim;
  DBuf<int4> d_wsx_gs, d_wsx_gs_t;  // the gs tiles of the non-TATT / TATT draws
  DBuf<int2> d_wlim;
  int n_wsx_gs = 0, n_wsx_gs_t = 0, n_wlim = 0;
  int nrow_custom = 0;  // rows of a custom-ell batch (0: the layouts' rows)
  int gg_stage = 0;     // LDS doubles for the gg kernels' staged tables (0: not staged)
  // the device state (lens-efficiency tables, node weights, ss/gs C_ell rows)
  // matches the draws' current parameters: set by a forward or a VJP, cleared
  // by uploads, parameter updates and precision changes.  The tables and
  // weights are FP64 in every mode; after a mixed-precision forward the ss
  // rows carry its FP32 node sums, which the VJP then uses only in the
  // shear-calibration sums T (relative effect on d/dm about 1e-6)
  bool state_valid = false;
  // checkpoint of the Limber integrands for the VJP (gg: t, P fac, dW_RSD;
  // ss / gs: P fac): written by forwards with keep_ckpt, current while
  // ckpt_fresh (cleared with state_valid and by forwards without it)
  bool keep_ckpt = false, ckpt_fresh = false;
  // with state_valid: the gs and gg C_ell rows are FP64 (the lens-limit term
  // of the VJP differences them against its views)
  bool rows64 = false;
  size_t ggk_per = 0, sxk_per = 0, dgk_per = 0;  // doubles per draw
  DBuf<double> d_ggk, d_sxk, d_dgk;  // dgk: lens-efficiency shift tangents
  int dv_size = 0;
  int nknz = 0;        // common ln P table size (0 if the draws differ)
  bool tatt = false;   // some draw uses TATT intrinsic alignments
  bool oneloop = false;  // some draw has one-loop galaxy bias
  bool cmb_any = false;  // some draw computes the 6x2pt CMB lensing blocks
  // draws at high_def_integration != 0: their Limber rows from k_limber_hdi
  // (draw, column, row0, rows) instead of the node kernels
  bool hdi_any = false;
  DBuf<int4> d_whd;
  int n_whd = 0;
  DBuf<int4> d_wcmb;
  DBuf<int> d_wcmbp;
  int n_wcmb = 0, n_wcmbp = 0;
  // non-Limber w: per such draw two views of it (gg Limber rows at ell =
  // 1..149 with P_nl and with P_lin) and the FFTLog work lists
  bool nl_any = false;
  std::vector<DrawDesc> desc_nl;
  DBuf<DrawDesc> d_desc_nl, d_desc_nlv;
  DBuf<double> d_nl_ells;
  DBuf<int2> d_wn_nl, d_wnl_prep, d_wnl_cl;
  DBuf<int4> d_wgg_nl, d_wnl_fwd, d_wnl_fwd4, d_wnl_gl, d_wnl_chk, d_wnl_idft;
  int n_wn_nl = 0, n_wgg_nl = 0, n_wnl_prep = 0, n_wnl_fwd = 0, n_wnl_fwd4 = 0, n_wnl_gl = 0, n_wnl_chk = 0,
      n_wnl_idft = 0, n_wnl_cl = 0;
  std::vector<DrawDesc> desc_nlv;  // the views' descriptors with the VJP's buffers
  int ord[3] = {0, 1, 2};
  void release() {
    d_desc.release(); d_par.release(); d_tab.release(); d_work.release(); d_dv.release(); d_ells.release();
    d_vjp.release(); d_gscr.release(); d_run.release(); d_wlg.release(); d_ggk.release(); d_sxk.release(); d_dgk.release();
    d_wg.release(); d_wn.release(); d_wl2.release(); d_wl.release(); d_wa4.release();
    d_wsx.release(); d_wgg.release(); d_wga.release(); d_wmx.release(); d_wlf.release();
    d_desc_nl.release(); d_desc_nlv.release(); d_nl_ells.release(); d_wn_nl.release(); d_wnl_prep.release(); d_wnl_cl.release();
    d_wgg_nl.release(); d_wnl_fwd.release(); d_wnl_fwd4.release(); d_wnl_gl.release(); d_wnl_chk.release();
    d_wnl_idft.release();
    d_wcmb.release(); d_wcmbp.release();
    d_desc_lim.release(); d_lim.release(); d_wsx_gs.release(); d_wsx_gs_t.release(); d_wlim.release();
    d_whd.release();
  }
};

}  // namespace

struct clh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;      // stream every launch goes to
  hipStream_t own_stream = nullptr;  // the context's own stream
  std::string err;
  Layout L;
  Draw cur;
  int legendre_mode = CLH_LEGENDRE_FOLDED;
  int precision = CLH_PREC_FP64;
  int tile_rows[4] = {16, 64, 64, 64};  // rows per tile: fwd ss/gs, fwd gg, adj ss/gs, adj gg
  const int* ord6 = nullptr;            // 6-block order of a clh_compute_6x2pt_real call in progress
  int gg_lds[2] = {0, 1};               // stage tables in LDS: fwd gg, adj gg
  int vjp_ckpt = 0;                     // batch forwards keep the gg checkpoint for a VJP
  int vjp_lim = 0;                      // lens-limit term: 0 nodes' motion, 1 finite-difference views, 2 none
  uint64_t version = 1, cached_version = 0;
  dvec cache_raw;   // raw (uncalibrated, ord 0,1,2) data vector of `cur`
  dvec cache_cl;    // cl rows of `cur`
  dvec cache_bb;    // B-mode C_ss rows of `cur` (zero under NLA)
  dvec cache_cl6;   // 6x2pt: C_gk [nl], C_ks [ns], C_kk columns of cache_nrow6 rows (empty otherwise)
  int cache_nrow6 = 0;
  dvec cache_nl;    // non-Limber C_ell [nl][150] then LMAX [nl] of `cur` (empty under Limber w)
  // w(theta) of `cur` through the path adopt_limber_gg does not select
  // (w_gg_tomo with the other `limber`), on a copy of L
  Layout alt_L;
  dvec alt_w;
  uint64_t alt_w_version = 0;
  int alt_w_limber = -1;
  // device copies of the layout
  DevLayout dl;                                        // device copy of L
  std::map<uint64_t, std::unique_ptr<LayoutCopy>> rl;  // layouts of a multi-layout batch
  uint64_t mode_epoch = 0;                             // bumped when device layouts go stale
  uint64_t batch_epoch = 0;
  // bumped by every change of the context's layout L; a batch uploaded on L
  // (ctx_layout) is refused once it differs from the batch's lay_epoch
  uint64_t layout_epoch = 0;
  Batch single, batch, scratch;
  Batch arena;  // clh_compute_3x2pt_real_scratch: every buffer a slice of the caller's scratch
  bool timing = false;
  bool use_graph = false;             // replay the batch forward from a hipGraph
  hipGraphExec_t gexec = nullptr;     // captured forward of `batch`
  hipStream_t gstream = nullptr;      // stream the graph was captured on
  hipEvent_t ev[14] = {nullptr};  // 0-5 forward stages, 6-10 and 13 VJP stages, 11-12 k_nodes_tan
  bool vjp_lim_tan = false;         // the last timed VJP took the lens-limit term from k_nodes_tan
  float last_ms[5] = {0, 0, 0, 0, 0};
  float last_vjp_ms[4] = {0, 0, 0, 0};
  DBuf<double> d_intf;  // scale factors and values of clh_int_for_limber
  bool vjp_timed = false;
  // likelihood (IP): data and inverse covariance in the block order like_ord
  int like_n = 0, like_ord[3] = {0, 1, 2};
  bool like_data_set = false, like_icov_set = false;
  dvec like_data_h;
  DBuf<double> like_data, like_icov, like_mask, like_R, like_Y;
  uint64_t like_mask_version = ~0ull;  // layout version the device mask copy was made for
  rocblas_handle blas = nullptr;
};

// the context's layout L changed: its device copy is stale and batches
// uploaded on it are refused
static void layout_changed(clh_ctx* c) {
  c->dl.valid = false;
  c->layout_epoch++;
}

static int fail(clh_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCHK(c, call)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(c, e_ == hipErrorOutOfMemory ? CLH_ERR_ALLOC : CLH_ERR_HIP, "%s: %s", \
                  #call, hipGetErrorString(e_));                                         \
  } while (0)

// first GPU use of a context: device check, stream and timing events
static int ensure_device(clh_ctx* c) {
  if (c->own_stream) return CLH_OK;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(c, CLH_ERR_HIP, "no HIP device");
  if (c->device >= n) return fail(c, CLH_ERR_ARG, "device %d of %d", c->device, n);
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking));
  c->stream = c->own_stream;
  for (auto& e : c->ev) HIPCHK(c, hipEventCreate(&e));
  return CLH_OK;
}

static bool has_nan(const double* x, int n) {
  for (int i = 0; i < n; i++) if (std::isnan(x[i])) return true;
  return false;
}

// ------------------------------------------------------------ n(z) samples
// set_source_sample / set_lens_sample (generic_interface.cpp:1802-1999) and
// the table build of nz_source_photoz / nz_lens_photoz (redshift_spline.c:446-512,
// 665-731)
static int build_sample(clh_ctx* c, Sample& S, const double* table, int nrows, int ntomo,
                        bool source) {
  if (!table || ntomo <= 0 || ntomo > CLH_MAXB || nrows < 5)
    return fail(c, CLH_ERR_ARG, "n(z) sample: ntomo=%d (1..%d), nrows=%d (>=5)", ntomo, CLH_MAXB, nrows);
  if (has_nan(table, nrows * (ntomo + 1))) return fail(c, CLH_ERR_NAN, "n(z) sample: NaN");
  dvec z(nrows);
  std::vector<dvec> col(ntomo, dvec(nrows));
  for (int i = 0; i < nrows; i++) {
    z[i] = table[(size_t)i * (ntomo + 1)];
    for (int k = 0; k < ntomo; k++) col[k][i] = table[(size_t)i * (ntomo + 1) + k + 1];
  }
  for (int i = 1; i < nrows; i++)
    if (!(z[i] > z[i - 1])) return fail(c, CLH_ERR_DOMAIN, "bad n(z) file (z vector not monotonic)");
  Sample T;
  T.nbin = ntomo;
  T.rows = nrows;
  T.zmin_all = std::fmax(z[0], 1.e-5);
  T.zmax_all = z[nrows - 1] + (z[nrows - 1] - z[0]) / ((double)nrows - 1.);
  for (int k = 0; k < ntomo; k++) {
    double mx = col[k][0];
    for (int i = 1; i < nrows; i++) mx = std::max(mx, col[k][i]);
    int first = -1, last = -1;
    for (int i = 0; i < nrows; i++)
      if (col[k][i] > 0.999e-8 * mx) { if (first < 0) first = i; last = i; }
    if (first < 0) return fail(c, CLH_ERR_DOMAIN, "zero/negative n(z) in bin %d", k);
    T.zmin[k] = source ? std::fmax(z[first], 1.001e-5) : z[first];
    T.zmax[k] = z[last];
  }
  if (source && (T.zmax_all < T.zmax[ntomo - 1] || T.zmin_all > T.zmin[0]))
    return fail(c, CLH_ERR_DOMAIN, "source n(z) range inconsistent");
  // histogram re-binned onto centres, normalised per bin
  auto histo = [&](double zz, int ni) -> double {
    if (zz >= T.zmin_all && zz < T.zmax_all) {
      const double dzh = (z[nrows - 1] - z[0]) / ((double)nrows - 1.);
      const int nj = (int)std::floor((zz - z[0]) / dzh);
      if (nj < 0 || nj > nrows - 1) return 0.0;
      return col[ni][nj];
    }
    return 0.0;
  };
  const double dz = (T.zmax_all - T.zmin_all) / ((double)nrows);
  T.zc.resize(nrows);
  for (int k = 0; k < nrows; k++) T.zc[k] = T.zmin_all + (k + 0.5) * dz;
  T.coef.assign((size_t)ntomo * 4 * (nrows - 1), 0.0);
  dvec y(nrows);
  for (int i = 0; i < ntomo; i++) {
    double NORM = 0.0;
    for (int k = 0; k < nrows; k++) NORM += histo(T.zc[k], i) * dz;
    if (!(NORM > 0.0)) return fail(c, CLH_ERR_DOMAIN, "zero/negative n(z) normalization for bin %d", i);
    for (int k = 0; k < nrows; k++) y[k] = histo(T.zc[k], i) / NORM;
    natural_spline(nrows, T.zc.data(), y.data(), T.coef.data() + (size_t)i * 4 * (nrows - 1));
  }
  T.set = true;
  S = std::move(T);
  return CLH_OK;
}

// GL orders of the Limber integrals by high_def_integration: ss and gs
// (cosmo2D.c:1577-1581, 2145-2149), gg (:2635-2639)
static int gl_order_ssgs(int hdi) {
  hdi = std::abs(hdi);
  return hdi == 0 ? 64 : hdi == 1 ? 96 : hdi == 2 ? 128 : hdi == 3 ? 256 : 512;
}
static int gl_order_gg(int hdi) {
  hdi = std::abs(hdi);
  return hdi == 0 ? 96 : hdi == 1 ? 128 : hdi == 2 ? 256 : hdi == 3 ? 512 : 1024;
}

// zmean (redshift_spline.c:794-844) as set_lens_sample evaluates it
// (generic_interface.cpp:1874-1876): zero shift, unit stretch, 256 / 512 /
// 1024-point GL by high_def_integration (:809-811)
static void lens_zmean(Layout& L) {
  const int hd = std::abs(L.hdi);
  const int n = hd == 0 ? 256 : hd == 1 ? 512 : 1024;
  std::vector<double> x(n), w(n);
  gauss_legendre(n, x.data(), w.data());
  const int rows = L.lens.rows;
  for (int b = 0; b < L.lens.nbin; b++) {
    const double lo = L.lens.zmin[b], hi = L.lens.zmax[b];
    const double A = 0.5 * (hi - lo), B = 0.5 * (hi + lo);
    double num = 0.0, den = 0.0;
    const double* coef = L.lens.coef.data() + (size_t)b * 4 * (rows - 1);
    for (int p = 0; p < n; p++) {
      const double zz = B + A * x[p];
      double v = 0.0;
      if (!(zz <= L.lens.zc[0] || zz >= L.lens.zc[rows - 1])) v = spline_eval(L.lens.zc, coef, zz);
      num += A * w[p] * zz * v;
      den += A * w[p] * v;
    }
    L.zmean[b] = num / den;
  }
}

// ------------------------------------------------------- Legendre kernels
// set_bin_average (basics.c:382-455: theta edges and P_l, P'_l at both edges;
// P'_l from (1-x^2) P'_l = l (P_{l-1} - x P_l)) and the bin-averaged kernels of
// xi_pm_tomo (cosmo2D.c:430-453), w_gammat_tomo (:631-639) and w_gg_tomo
// (:808-815); then the fold of the linear-in-ln(ell) fill (cosmo2D.c:1800-1807)
// into the kernels.
struct KernelKey {
  int nt, n_ell, lmax;
  double vtmin, vtmax;
  bool operator<(const KernelKey& o) const {
    if (nt != o.nt) return nt < o.nt;
    if (n_ell != o.n_ell) return n_ell < o.n_ell;
    if (lmax != o.lmax) return lmax < o.lmax;
    if (vtmin != o.vtmin) return vtmin < o.vtmin;
    return vtmax < o.vtmax;
  }
};
struct KernelSet {
  dvec kfold, ell_ss, ell_gx, kwlow;
};
// folded kernels depend only on the binning and Ntable: shared process-wide
static std::map<KernelKey, std::shared_ptr<const KernelSet>> g_kernel_cache;
static std::mutex g_kernel_mu;

static void build_kernels_uncached(Layout& L, bool full);

static void build_kernels(Layout& L, bool full) {
  const KernelKey key{L.ntheta, L.n_ell, L.lmax, L.vtmin, L.vtmax};
  if (L.sixpt) {  // the CMB filter enters the folded 6x2pt kernels: not shared
    build_kernels_uncached(L, full);
    return;
  }
  if (!full) {
    std::lock_guard<std::mutex> lk(g_kernel_mu);
    auto it = g_kernel_cache.find(key);
    if (it != g_kernel_cache.end()) {
      L.kfold = it->second->kfold;
      L.ell_ss = it->second->ell_ss;
      L.ell_gx = it->second->ell_gx;
      L.kwlow = it->second->kwlow;
      L.kfull.clear();
      L.kern_valid = true;
      return;
    }
  }
  build_kernels_uncached(L, full);
  std::lock_guard<std::mutex> lk(g_kernel_mu);
  if (g_kernel_cache.size() >= 64) g_kernel_cache.clear();
  auto ks = std::make_shared<KernelSet>();
  ks->kfold = L.kfold;
  ks->ell_ss = L.ell_ss;
  ks->ell_gx = L.ell_gx;
  ks->kwlow = L.kwlow;
  g_kernel_cache[key] = ks;
}

static void build_kernels_uncached(Layout& L, bool full) {
  const int nt = L.ntheta, Lm = L.lmax, nrow = L.nrow();
  L.kfold.assign((size_t)4 * nrow * nt, 0.0);
  L.kwlow.assign((size_t)nt * CLH_NL_L, 0.0);
  if (L.sixpt) L.kf6.assign((size_t)2 * nrow * nt, 0.0); else L.kf6.clear();
  // beam_cmb x w_pixel (cosmo2D.c:268-279, 1003-1010)
  const double cmb_s = L.cmb_fwhm / std::sqrt(16.0 * std::log(2.0));
  auto cmbf = [&](int l) {
    double f = (l < L.cmb_lmin || l > L.cmb_lmax) ? 0.0 : std::exp(-l * (l + 1.0) * cmb_s * cmb_s);
    if (!L.cmb_win.empty()) f *= (l < (int)L.cmb_win.size()) ? L.cmb_win[l] : 0.0;
    return f;
  };
  if (full) L.kfull.assign((size_t)4 * Lm * nt, 0.0); else L.kfull.clear();
  double lim_ss[3], lim_gx[3];
  lim_ss[0] = std::log(std::fmax(CLH_LMIN_TAB - 1., 1.0));
  lim_gx[0] = std::log(std::fmax((double)CLH_LMIN_TAB, 1.0));
  lim_ss[1] = lim_gx[1] = std::log((double)(Lm + 1));
  lim_ss[2] = (lim_ss[1] - lim_ss[0]) / ((double)L.n_ell - 1.);
  lim_gx[2] = (lim_gx[1] - lim_gx[0]) / ((double)L.n_ell - 1.);
  const double logdt = (std::log(L.vtmax) - std::log(L.vtmin)) / nt;
  // The bin-averaged G+- kernels are small differences of O(l^2) terms divided
  // by x_min - x_max (~1.5e-7 at 2.5'): evaluated in double they carry up to
  // ~1e-4 relative rounding error in xi-(2.5'), so the recurrences and the
  // kernel formulas run in 80-bit long double and are rounded to double once.
  using ld = long double;
  auto work = [&](int i) {
    const double xmin_d = std::cos(std::exp(std::log(L.vtmin) + (i + 0.) * logdt));
    const double xmax_d = std::cos(std::exp(std::log(L.vtmin) + (i + 1.) * logdt));
    const ld xmin = xmin_d, xmax = xmax_d;
    std::vector<ld> Pmin(Lm + 1), Pmax(Lm + 1), dPmin(Lm + 1), dPmax(Lm + 1);
    auto leg = [&](ld x, std::vector<ld>& P, std::vector<ld>& dP) {
      P[0] = 1.0L;
      P[1] = x;
      ld pm2 = 1.0L, pm1 = x;
      for (int l = 2; l <= Lm; l++) {
        const ld p = (x * (2 * l - 1) * pm1 - (l - 1) * pm2) / l;
        pm2 = pm1;
        pm1 = p;
        P[l] = p;
      }
      dP[0] = 0.0L;
      dP[1] = 1.0L;
      const ld da = 1.0L + x, db = 1.0L - x;
      for (int l = 2; l <= Lm; l++) dP[l] = -l * (x * P[l] - P[l - 1]) / (da * db);
    };
    leg(xmin, Pmin, dPmin);
    leg(xmax, Pmax, dPmax);
    const ld pi = 3.14159265358979323846264338327950288L;
    for (int l = 1; l < Lm; l++) {
      const ld dx = xmin - xmax;
      const ld lf = l;
      const ld common =
          -lf * (lf - 1) / 2 * (lf + 2.0L / (2 * lf + 1)) * (Pmin[l - 1] - Pmax[l - 1]) -
          lf * (lf - 1) * (2 - lf) / 2 * (xmin * Pmin[l] - xmax * Pmax[l]) +
          lf * (lf - 1) / (2 * lf + 1) * (Pmin[l + 1] - Pmax[l + 1]) + (4 - lf) * (dPmin[l] - dPmax[l]) +
          (lf + 2) * (xmin * dPmin[l - 1] - xmax * dPmax[l - 1] - Pmin[l - 1] + Pmax[l - 1]);
      const ld t1 = 2 * (lf - 1) * (xmin * dPmin[l] - xmax * dPmax[l] - Pmin[l] + Pmax[l]);
      const ld t2 = 2 * (lf + 2) * (dPmin[l - 1] - dPmax[l - 1]);
      const ld pre = (2 * lf + 1) / (2 * pi * lf * lf * (lf + 1) * (lf + 1));
      double g[4];
      g[0] = (double)(pre * (common + t1 - t2) / dx);
      g[1] = (double)(pre * (common - t1 + t2) / dx);
      g[2] = (double)((2 * lf + 1) / (4 * pi * lf * (lf + 1) * dx) *
                      ((lf + 2.0L / (2 * lf + 1)) * (Pmin[l - 1] - Pmax[l - 1]) +
                       (2 - lf) * (xmin * Pmin[l] - xmax * Pmax[l]) -
                       2.0L / (2 * lf + 1) * (Pmin[l + 1] - Pmax[l + 1])));
      g[3] = (double)((1.0L / dx) * (1.0L / (4 * pi)) * (Pmin[l + 1] - Pmax[l + 1] - Pmin[l - 1] + Pmax[l - 1]));
      if (l < CLH_NL_L) L.kwlow[(size_t)i * CLH_NL_L + l] = g[3];
      if (L.sixpt) {  // w_gk: w kernel, w_ks: gamma_t kernel, both times the CMB filter
        const double f = cmbf(l);
        for (int k6 = 0; k6 < 2; k6++) {
          const double gv = (k6 == 0 ? g[3] : g[2]) * f;
          double* kf = L.kf6.data() + (size_t)k6 * nrow * nt;
          if (l < CLH_LMIN_TAB) {
            kf[(size_t)(l - 1) * nt + i] += gv;
          } else {
            const double r = (std::log((double)l) - lim_gx[0]) * (1.0 / lim_gx[2]);
            const int ii = (int)std::floor(r);
            const int ic = ii < 0 ? 0 : (ii >= L.n_ell - 1 ? L.n_ell - 2 : ii);
            const double t = r - ic;
            kf[(size_t)(CLH_NLOW + ic) * nt + i] += gv * (1.0 - t);
            kf[(size_t)(CLH_NLOW + ic + 1) * nt + i] += gv * t;
          }
        }
      }
      for (int k = 0; k < 4; k++) {
        if (full) L.kfull[((size_t)k * Lm + l) * nt + i] = g[k];
        double* kf = L.kfold.data() + (size_t)k * nrow * nt;
        if (l < CLH_LMIN_TAB) {
          kf[(size_t)(l - 1) * nt + i] += g[k];
        } else {
          const double* lim = (k <= 1) ? lim_ss : lim_gx;
          const double r = (std::log((double)l) - lim[0]) * (1.0 / lim[2]);
          const int ii = (int)std::floor(r);
          const int ic = ii < 0 ? 0 : (ii >= L.n_ell - 1 ? L.n_ell - 2 : ii);
          const double t = r - ic;
          kf[(size_t)(CLH_NLOW + ic) * nt + i] += g[k] * (1.0 - t);
          kf[(size_t)(CLH_NLOW + ic + 1) * nt + i] += g[k] * t;
        }
      }
    }
  };
  const int nthr = std::max(1u, std::min<unsigned>(std::thread::hardware_concurrency(), 16u));
  std::vector<std::thread> th;
  for (int t = 0; t < nthr; t++)
    th.emplace_back([&, t]() { for (int i = t; i < nt; i += nthr) work(i); });
  for (auto& x : th) x.join();
  // the rows' own ell values
  L.ell_ss.resize(nrow);
  L.ell_gx.resize(nrow);
  for (int r = 0; r < CLH_NLOW; r++) L.ell_ss[r] = L.ell_gx[r] = (double)(r + 1);
  // the first table row of ss is exp(ln 19), which rounds a few ulp below
  // the direct row 19; the rows are kept non-decreasing (the adjoint's
  // per-node ln P runs rely on it), which moves that row by at most an ulp
  for (int i = 0; i < L.n_ell; i++) {
    L.ell_ss[CLH_NLOW + i] = std::max(std::exp(lim_ss[0] + i * lim_ss[2]), L.ell_ss[CLH_NLOW + i - 1]);
    L.ell_gx[CLH_NLOW + i] = std::max(std::exp(lim_gx[0] + i * lim_gx[2]), L.ell_gx[CLH_NLOW + i - 1]);
  }
  L.kern_valid = true;
}

// ------------------------------------------------------------- validation
static void host_chi_all(const Draw& d, double a, double& chi, double& dchida) {
  const double z = 1.0 / a - 1.0;
  const int n = (int)d.chi_z.size();
  int j = bs_lo(d.chi_z.data(), n, z);
  if (j > n - 3) j = n - 3;
  const double* zz = d.chi_z.data();
  const double* cc = d.chi.data();
  const double dy = (z - zz[j]) / (zz[j + 1] - zz[j]);
  const double out0 = cc[j] + dy * (cc[j + 1] - cc[j]);
  const double up = (cc[j + 2] - cc[j]) / (zz[j + 2] - zz[j]);
  const double down = (j > 0) ? (cc[j + 1] - cc[j - 1]) / (zz[j + 1] - zz[j - 1])
                              : (cc[j + 1] - cc[j]) / (zz[j + 1] - zz[j]);
  chi = out0 / CLH_COVERH0;
  dchida = (down + dy * (up - down)) / CLH_COVERH0 / (a * a);
}

static int check_draw(clh_ctx* c, const Layout& L, const Draw& d, int idx) {
  const int nc = (int)d.chi_z.size();
  if (nc < 5 || d.chi.size() != (size_t)nc) return fail(c, CLH_ERR_STATE, "draw %d: distances not set", idx);
  if (d.G_z.size() < 2 || d.G.size() != d.G_z.size()) return fail(c, CLH_ERR_STATE, "draw %d: growth not set", idx);
  if (d.lk.size() < 2 || d.pz.size() < 2 || d.lnP.size() != d.lk.size() * d.pz.size())
    return fail(c, CLH_ERR_STATE, "draw %d: P(k,z) not set", idx);
  if (!(d.omega_m > 0)) return fail(c, CLH_ERR_STATE, "draw %d: Omega_m not set", idx);
  if (L.ia_model == CLH_IA_TATT && (d.fpt_n < 2 || d.fpt.size() != (size_t)10 * d.fpt_n))
    return fail(c, CLH_ERR_STATE, "draw %d: TATT needs the FAST-PT IA tables (set_IA_PS)", idx);
  if (d.has_b2(L.lens.nbin)) {
    if (d.fptb_n < 2 || d.fptb.size() != (size_t)6 * d.fptb_n)
      return fail(c, CLH_ERR_STATE, "draw %d: one-loop bias needs the FAST-PT bias tables (set_bias_PS)", idx);
    if (L.ia_model == CLH_IA_TATT)
      return fail(c, CLH_ERR_ARG, "draw %d: one-loop bias together with TATT is not implemented", idx);
  }
  if (L.sixpt) {
    if (L.lens.nbin == 0) return fail(c, CLH_ERR_STATE, "6x2pt needs the lens sample");
    if (!(d.chi_z[nc - 2] > 1090.0))
      return fail(c, CLH_ERR_DOMAIN, "draw %d: 6x2pt needs chi(z) two nodes beyond z = 1090 (g_cmb)", idx);
  }
  if (L.nonlimber && L.lens.nbin > 0) {
    if (d.lkL.size() < 2 || d.pzL.size() < 2 || d.lnPL.size() != d.lkL.size() * d.pzL.size())
      return fail(c, CLH_ERR_STATE, "draw %d: non-Limber w needs the linear P(k,z) (set_linear_power_spectrum)", idx);
    if (!(d.h0 > 0)) return fail(c, CLH_ERR_STATE, "draw %d: non-Limber w needs H0 (set_cosmological_parameters)", idx);
    for (size_t i = 1; i < d.lkL.size(); i++)
      if (!(d.lkL[i] > d.lkL[i - 1])) return fail(c, CLH_ERR_DOMAIN, "draw %d: linear log10k not increasing", idx);
    for (size_t i = 1; i < d.pzL.size(); i++)
      if (!(d.pzL[i] > d.pzL[i - 1])) return fail(c, CLH_ERR_DOMAIN, "draw %d: linear P z not increasing", idx);
    if (!(d.chi_z[nc - 2] > 4.0)) return fail(c, CLH_ERR_DOMAIN, "draw %d: chi(z) must extend beyond z = 4", idx);
  }
  for (int i = 1; i < nc; i++)
    if (!(d.chi_z[i] > d.chi_z[i - 1]) || !(d.chi[i] > d.chi[i - 1]))
      return fail(c, CLH_ERR_DOMAIN, "draw %d: chi(z) table not strictly increasing", idx);
  for (size_t i = 1; i < d.G_z.size(); i++)
    if (!(d.G_z[i] > d.G_z[i - 1])) return fail(c, CLH_ERR_DOMAIN, "draw %d: growth z not increasing", idx);
  for (size_t i = 1; i < d.lk.size(); i++)
    if (!(d.lk[i] > d.lk[i - 1])) return fail(c, CLH_ERR_DOMAIN, "draw %d: log10k not increasing", idx);
  for (size_t i = 1; i < d.pz.size(); i++)
    if (!(d.pz[i] > d.pz[i - 1])) return fail(c, CLH_ERR_DOMAIN, "draw %d: P(k,z) z not increasing", idx);
  // chi_all reads chi[j+2] (cosmo3D.c:66): the largest queried z must lie below
  // the table's next-to-last node.  gg queries chi(a_min) at z = 40.
  double zq = L.src.zmax_all;
  if (L.lens.nbin > 0) zq = std::max(zq, 1.0 / CLH_A_MIN - 1.0);
  if (!(d.chi_z[nc - 2] > zq))
    return fail(c, CLH_ERR_DOMAIN, "draw %d: chi(z) table must extend two nodes beyond z=%g", idx, zq);
  for (int b = 0; b < L.lens.nbin; b++) {
    if (!(d.lens_st[b] != 0.0)) return fail(c, CLH_ERR_DOMAIN, "draw %d: zero lens stretch", idx);
    const double zmax = (L.lens.zmax[b] - L.zmean[b]) * d.lens_st[b] + L.zmean[b];
    const double amin = 1. / (1 + zmax + 2. * std::fabs(d.lens_dz[b]));
    const double zmin = (L.lens.zmin[b] - L.zmean[b]) * d.lens_st[b] + L.zmean[b];
    const double amax = (d.bmag[b] != 0) ? 1. / (1. + std::fmax(L.src.zmin_all, 0.001))
                                         : 1. / (1 + std::fmax(zmin - 2. * std::fabs(d.lens_dz[b]), 0.001));
    if (!(amin > 0) || !(amin < 1) || !(amax > 0) || !(amax < 1) || !(amin < amax))
      return fail(c, CLH_ERR_DOMAIN, "draw %d lens bin %d: 0 < amin < amax < 1 not true", idx, b);
  }
  return CLH_OK;
}

// ----------------------------------------------------------- device layout
// bfloat16 of a float, rounded to nearest even (finite inputs)
static uint16_t bf16_rne(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
static double bf16_value(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// FFTLog transform lengths of C_cl_tomo (cosmo2D.c:4632-4654): zero padding
// N_pad = 200, 500, 500 (density, RSD, magnification) on each side of the
// NL_Nchi nodes, rounded up to an even length without prime factors above 7
static long nl_fft_size(long n) {
  for (;; n++) {
    long m = n;
    for (int p : {2, 3, 5, 7})  // next_fft_size (cosmo2D.c:4023-4034)
      while (m % p == 0) m /= p;
    if (m == 1) return n;
  }
}
static void nl_sizes(int N[3], int N0[3]) {
  static const int pad[3] = {200, 500, 500};
  for (int j = 0; j < 3; j++) {
    N0[j] = pad[j];
    long raw = 2L * pad[j] + CLH_NL_NCHI;
    if (raw % 2) raw++;
    long n = nl_fft_size(raw);
    if (n % 2) n = nl_fft_size(n + 1);
    N[j] = (int)n;
  }
}

static int upload_layout(clh_ctx* c, const Layout& L, DevLayout& D) {
  const bool full = c->legendre_mode == CLH_LEGENDRE_FULL;
  if (!L.kern_valid || (full && L.kfull.empty()))
    return fail(c, CLH_ERR_STATE, "layout kernels not built for the Legendre mode");
  // GL rules: 64 and 96 points, then at high_def_integration != 0 the ss / gs
  // and gg rules of that level (k_limber_hdi)
  const int hss = gl_order_ssgs(L.hdi), hgg = gl_order_gg(L.hdi);
  std::vector<double> glv(2 * CLH_NP_SS + 2 * CLH_NP_GG + (L.hdi ? 2 * hss + 2 * hgg : 0));
  double* glbuf = glv.data();
  gauss_legendre(CLH_NP_SS, glbuf, glbuf + CLH_NP_SS);
  gauss_legendre(CLH_NP_GG, glbuf + 2 * CLH_NP_SS, glbuf + 2 * CLH_NP_SS + CLH_NP_GG);
  if (L.hdi) {
    double* h = glbuf + 2 * CLH_NP_SS + 2 * CLH_NP_GG;
    gauss_legendre(hss, h, h + hss);
    gauss_legendre(hgg, h + 2 * hss, h + 2 * hss + hgg);
  }
  size_t off = 0;
  auto take = [&](size_t n) { size_t o = off; off += (n + 31) & ~size_t(31); return o; };
  D.off_s_zc = take(L.src.zc.size());
  D.off_s_coef = take(L.src.coef.size());
  D.off_l_zc = take(L.lens.zc.size());
  D.off_l_coef = take(L.lens.coef.size());
  D.off_ell_ss = take(L.ell_ss.size());
  D.off_ell_gx = take(L.ell_gx.size());
  D.off_kfold = take(L.kfold.size());
  D.off_kfoldT = take(L.kfold.size());
  D.off_gl = take(glv.size());
  D.off_theta = take(L.theta.size());
  D.off_kfull = take(full ? L.kfull.size() : 0);
  // FULL mode: index and weight of the ln-ell fill at every integer ell
  // (C_*_tomo_limber_fill, cosmo2D.c:1800-1807)
  dvec fill;
  if (full) {
    fill.assign((size_t)4 * L.lmax, 0.0);
    for (int g = 0; g < 2; g++) {
      double lim0 = std::log(std::fmax(g == 0 ? CLH_LMIN_TAB - 1. : (double)CLH_LMIN_TAB, 1.0));
      const double lim1 = std::log((double)(L.lmax + 1));
      const double lim2 = (lim1 - lim0) / ((double)L.n_ell - 1.);
      const double inv_dx = 1.0 / lim2;
      for (int l = CLH_LMIN_TAB; l < L.lmax; l++) {
        const double r = (std::log((double)l) - lim0) * inv_dx;
        const int i = (int)std::floor(r);
        const int ic = i < 0 ? 0 : (i >= L.n_ell - 1 ? L.n_ell - 2 : i);
        fill[((size_t)g * L.lmax + l) * 2 + 0] = (double)ic;
        fill[((size_t)g * L.lmax + l) * 2 + 1] = r - ic;
      }
    }
  }
  D.off_fill = take(fill.size());
  // non-Limber w: the w kernel at integer ell < 150, and cos / sin of
  // 2 pi t / N_j for the three transform lengths
  int nlN[3], nlN0[3];
  nl_sizes(nlN, nlN0);
  // the device transforms (clh_fft.hpp) take N / 2 with factors 2, 3, 5 and the
  // LDS layout of k_nl_fft; the reference's paddings give N = 900, 1500, 1500
  for (int j = 0; j < 3; j++)
    if (!clh_fft_ok(nlN[j] / 2)) return fail(c, CLH_ERR_STATE, "non-Limber transform length %d has a factor 7", nlN[j]);
  if (nlN[1] != nlN[2] || nlN[0] + nlN[1] > 2400)
    return fail(c, CLH_ERR_STATE, "non-Limber transform lengths %d, %d, %d not supported", nlN[0], nlN[1], nlN[2]);
  dvec tw;
  for (int j = 0; j < 3; j++) {
    if (nlN[j] > 1500) return fail(c, CLH_ERR_STATE, "non-Limber transform length %d above 1500", nlN[j]);
    for (int t = 0; t < nlN[j]; t++) {
      const long double ang = 2.0L * 3.14159265358979323846264338327950288L * t / nlN[j];
      tw.push_back((double)std::cos(ang));
      tw.push_back((double)std::sin(ang));
    }
  }
  D.off_kwlow = take(L.kwlow.size());
  D.off_nltw = take(tw.size());
  D.off_kf6 = take(L.kf6.size());
  const dvec kk_ells = L.kk_direct_ells();
  D.off_ellf = take(kk_ells.size());
  D.off_kkM = take(L.kk_M.size());
  D.off_kkoff = take(L.kk_off.size());
  HIPCHK(c, D.buf.ensure(off));
  double* base = D.buf.p;
  auto put = [&](size_t o, const double* src, size_t n) -> hipError_t {
    if (n == 0) return hipSuccess;
    return hipMemcpyAsync(base + o, src, n * sizeof(double), hipMemcpyHostToDevice, c->stream);
  };
  HIPCHK(c, put(D.off_s_zc, L.src.zc.data(), L.src.zc.size()));
  HIPCHK(c, put(D.off_s_coef, L.src.coef.data(), L.src.coef.size()));
  HIPCHK(c, put(D.off_l_zc, L.lens.zc.data(), L.lens.zc.size()));
  HIPCHK(c, put(D.off_l_coef, L.lens.coef.data(), L.lens.coef.size()));
  HIPCHK(c, put(D.off_ell_ss, L.ell_ss.data(), L.ell_ss.size()));
  HIPCHK(c, put(D.off_ell_gx, L.ell_gx.data(), L.ell_gx.size()));
  HIPCHK(c, put(D.off_kfold, L.kfold.data(), L.kfold.size()));
  dvec kfoldT(L.kfold.size());  // [4][ntheta][nrow] for the adjoint contraction
  {
    const size_t nr = L.kfold.size() / (4 * (size_t)L.ntheta), nt = L.ntheta;
    for (size_t k = 0; k < 4; k++)
      for (size_t r = 0; r < nr; r++)
        for (size_t t = 0; t < nt; t++) kfoldT[(k * nt + t) * nr + r] = L.kfold[(k * nr + r) * nt + t];
  }
  HIPCHK(c, put(D.off_kfoldT, kfoldT.data(), kfoldT.size()));
  HIPCHK(c, put(D.off_gl, glbuf, glv.size()));
  HIPCHK(c, put(D.off_theta, L.theta.data(), L.theta.size()));
  HIPCHK(c, put(D.off_kwlow, L.kwlow.data(), L.kwlow.size()));
  HIPCHK(c, put(D.off_nltw, tw.data(), tw.size()));
  HIPCHK(c, put(D.off_kf6, L.kf6.data(), L.kf6.size()));
  HIPCHK(c, put(D.off_ellf, kk_ells.data(), kk_ells.size()));
  HIPCHK(c, put(D.off_kkM, L.kk_M.data(), L.kk_M.size()));
  HIPCHK(c, put(D.off_kkoff, L.kk_off.data(), L.kk_off.size()));
  if (!L.mask.empty()) {
    HIPCHK(c, D.mask.ensure(L.mask.size()));
    HIPCHK(c, hipMemcpyAsync(D.mask.p, L.mask.data(), L.mask.size() * sizeof(int), hipMemcpyHostToDevice,
                             c->stream));
  }
  if (!L.ggl_ex.empty() && L.n_gs() > 0) {
    const std::vector<int> rows = L.gs_rows();
    HIPCHK(c, D.gsrow.ensure(rows.size()));
    HIPCHK(c, hipMemcpyAsync(D.gsrow.p, rows.data(), rows.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
  }
  if (full) HIPCHK(c, put(D.off_kfull, L.kfull.data(), L.kfull.size()));
  if (full) HIPCHK(c, put(D.off_fill, fill.data(), fill.size()));
  // Folded kernels as MFMA A operands (k_legendre_mx): group g of 8 rows,
  // lane l = theta + 32 h, item t <-> row 8 g + 4 h + t.  The BF16 parts are
  // of the kernel scaled by 2^-e per 32-row band (e from the band's largest
  // |K|); the contraction scales that band's C_ell rows by 2^e.
  const int nrow = L.nrow(), nt = L.ntheta, ngrp = (nrow + 7) / 8;
  std::vector<float> k32((size_t)4 * ngrp * 64 * 4, 0.f), ksc((size_t)4 * ngrp, 1.f);
  std::vector<uint16_t> kbf((size_t)4 * ngrp * 64 * 12, 0);
  for (int k = 0; k < 4; k++) {
    const double* K = L.kfold.data() + (size_t)k * nrow * nt;
    for (int g0 = 0; g0 < ngrp; g0 += 4) {
      double mx = 0.0;
      for (int r = 8 * g0; r < std::min(nrow, 8 * (g0 + 4)); r++)
        for (int t = 0; t < nt; t++) mx = std::max(mx, std::fabs(K[(size_t)r * nt + t]));
      int e = 0;
      if (mx > 0.0) (void)std::frexp(mx, &e);
      for (int g = g0; g < std::min(ngrp, g0 + 4); g++) {
        ksc[(size_t)k * ngrp + g] = std::ldexp(1.0f, e);
        for (int l = 0; l < 64; l++)
          for (int t = 0; t < 4; t++) {
            const int th = l & 31, r = 8 * g + 4 * (l >> 5) + t;
            const double v = (th < nt && r < nrow) ? K[(size_t)r * nt + th] : 0.0;
            const size_t o = ((size_t)k * ngrp + g) * 64 + l;
            k32[o * 4 + t] = (float)v;
            const double vs = std::ldexp(v, -e);
            // three BF16 parts: hi, mid, lo (each of what the previous left)
            const uint16_t hi = bf16_rne((float)vs);
            const double r1 = vs - bf16_value(hi);
            const uint16_t mid = bf16_rne((float)r1);
            const uint16_t lo = bf16_rne((float)(r1 - bf16_value(mid)));
            kbf[o * 12 + t] = hi;
            kbf[o * 12 + 4 + t] = mid;
            kbf[o * 12 + 8 + t] = lo;
          }
      }
    }
  }
  HIPCHK(c, D.k32.ensure(k32.size()));
  HIPCHK(c, D.kbf.ensure(kbf.size()));
  HIPCHK(c, D.ksc.ensure(ksc.size()));
  HIPCHK(c, hipMemcpyAsync(D.k32.p, k32.data(), k32.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(D.kbf.p, kbf.data(), kbf.size() * sizeof(uint16_t), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(D.ksc.p, ksc.data(), ksc.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
  D.ngrp = ngrp;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  D.valid = true;
  return CLH_OK;
}

static int check_layout(clh_ctx* c, const Layout& L) {
  if (!L.src.set) return fail(c, CLH_ERR_STATE, "source sample not set");
  if (L.ntheta <= 0) return fail(c, CLH_ERR_STATE, "Ntable.Ntheta not initialized");
  if (L.ntheta > 32) return fail(c, CLH_ERR_ARG, "Ntheta=%d > 32 not supported", L.ntheta);
  if (L.n_a - 1 > 1024) return fail(c, CLH_ERR_ARG, "N_a=%d > 1025 not supported", L.n_a);
  return CLH_OK;
}

// node buffer of a draw: ss block, nl gs blocks, nl gg blocks, each with its
// search-result columns (node layout in clh_internal.hpp)
static size_t node_size(int ns, int nl) {
  return (size_t)CLH_NP_SS * (3 + ns + CLH_LUT_SX) +
         (size_t)nl * (CLH_NP_SS * (5 + ns + CLH_LUT_SX) + CLH_NP_GG * (5 + CLH_LUT_GG));
}

// the row tiles of one C_ell column: at most `cap` rows each, spread evenly
static void row_tiles(int nrow, int cap, std::vector<int2>& out) {
  out.clear();
  const int nt = (nrow + cap - 1) / cap;
  const int tr = (nrow + nt - 1) / nt;
  for (int r = 0; r < nrow; r += tr) out.push_back(make_int2(r, std::min(tr, nrow - r)));
}

// block starts of compute_data_vector_Mx2pt_N_starts (generic_interface.hpp:770-791)
static void block_starts(const int sizes[3], const int ord[3], int start[3]) {
  clh_block_starts(sizes, ord, 3, start);
}
static bool blocks_tile(const int sizes[3], const int ord[3]) {
  int start[3];
  return clh_blocks_tile(sizes, ord, 3, start);
}

static void dv_sizes(const Layout& L, int sizes[3]) {
  const int nt = L.ntheta;
  sizes[0] = 2 * nt * L.n_ss(); sizes[1] = nt * L.n_gs_dv(); sizes[2] = nt * L.lens.nbin;
}

// compute_data_vector_Mx2pt_N_starts<0,6> (generic_interface.hpp:770-791)
// for the six blocks of 6x2pt
static bool blocks_tile6(const int sizes[6], const int ord[6], int start[6]) {
  return clh_blocks_tile(sizes, ord, 6, start);
}
static void dv_sizes6(const Layout& L, int sizes[6]) {
  dv_sizes(L, sizes);
  sizes[3] = L.ntheta * L.lens.nbin;
  sizes[4] = L.ntheta * L.src.nbin;
  sizes[5] = L.kk_size();
}
// a 3-entry ord on a 6x2pt layout keeps gk, ks, kk after the 3x2pt blocks
static void ord6_of(const int ord[3], int o6[6]) {
  for (int i = 0; i < 3; i++) o6[i] = ord[i];
  o6[3] = 3; o6[4] = 4; o6[5] = 5;
}

// block starts of the 3x2pt blocks (start) and, for 6x2pt, of gk, ks, kk
// (start6); ord has 3 entries, or 6 when ord6 is set
static void dv_starts(const Layout& L, const int ord[3], int start[3], int start6[3] = nullptr,
                      const int* ord6 = nullptr) {
  if (!L.sixpt) {
    int sizes[3];
    dv_sizes(L, sizes);
    block_starts(sizes, ord, start);
    return;
  }
  int sizes[6], o6[6], st[6];
  dv_sizes6(L, sizes);
  if (ord6) for (int i = 0; i < 6; i++) o6[i] = ord6[i];
  else ord6_of(ord, o6);
  (void)blocks_tile6(sizes, o6, st);
  for (int i = 0; i < 3; i++) start[i] = st[i];
  if (start6)
    for (int i = 0; i < 3; i++) start6[i] = st[3 + i];
}

static bool ord_ok(const Layout& L, const int ord[3]) {
  if (L.sixpt) {
    int sizes[6], o6[6], st[6];
    dv_sizes6(L, sizes);
    ord6_of(ord, o6);
    return blocks_tile6(sizes, o6, st);
  }
  int sizes[3];
  dv_sizes(L, sizes);
  return blocks_tile(sizes, ord);
}

// Work lists of a batch (one entry per block; every stage is one launch for
// the whole batch).  rows: rows per tile of the forward ss/gs, forward gg,
// adjoint ss/gs and adjoint gg Limber stages (1..CLH_TILE_ROWS).
static int build_work(clh_ctx* c, Batch& bt, const int rows[4]) {
  std::vector<int2> wg, wn, wl2, wlim;
  std::vector<int4> wl, wl_gs, wsx, wsx_t, wsx_gs, wsx_gs_t, wgg, wga, wmx, wlf, wa, whd;
  int lds_ss = 0, lds_gs = 0;
  std::vector<int2> tiles, atiles_ss, atiles_gs;
  for (int i = 0; i < bt.B; i++) {
    const Layout& L = *bt.lays[i];
    const int ns = L.src.nbin, nl = L.lens.nbin;
    const int nrow = bt.nrow_custom ? bt.nrow_custom : L.nrow();
    const int n_ss = L.n_ss(), n_gs = L.n_gs();
    for (int b = 0; b < ns + nl; b++) wg.push_back(make_int2(i, b));
    const int ntask = 2 * n_ss + n_gs + nl;
    if (L.hdi != 0) {
      // every column (and the TATT B-mode ss columns) in tiles of CLH_HDI_ROWS
      // rows, for the forward and its adjoint (k_limber_hdi_adj); the
      // lens-limit items of the VJP's finite-difference views; no node items
      const int ncol = n_ss + n_gs + nl + (L.ia_model == CLH_IA_TATT ? n_ss : 0);
      for (int col = 0; col < ncol; col++)
        for (int r0 = 0; r0 < nrow; r0 += CLH_HDI_ROWS)
          whd.push_back(make_int4(i, col, r0, std::min(CLH_HDI_ROWS, nrow - r0)));
      for (int zl = 0; zl < nl; zl++) wlim.push_back(make_int2(i, zl));
    }
    if (L.hdi == 0) {
    for (int sg = 0; sg <= 2 * nl; sg++) wn.push_back(make_int2(i, sg));
    // forward Limber: ss, gs and gg tiles (draw, kind + 4 zl, row0, rows)
    // (TATT draws' tiles in their own list: they run in k_limber_sx_tatt)
    row_tiles(nrow, rows[0], tiles);
    std::vector<int4>& ws = L.ia_model == CLH_IA_TATT ? wsx_t : wsx;
    for (const int2& t : tiles) ws.push_back(make_int4(i, CLH_W_SS, t.x, t.y));
    for (int zl = 0; zl < nl; zl++)
      for (const int2& t : tiles) {
        ws.push_back(make_int4(i, CLH_W_GS + 4 * zl, t.x, t.y));
        (L.ia_model != CLH_IA_TATT ? wsx_gs : wsx_gs_t).push_back(make_int4(i, CLH_W_GS + 4 * zl, t.x, t.y));
      }
    for (int zl = 0; zl < nl; zl++) wlim.push_back(make_int2(i, zl));
    row_tiles(nrow, rows[1], tiles);
    for (int zl = 0; zl < nl; zl++)
      for (const int2& t : tiles) wgg.push_back(make_int4(i, zl, t.x, t.y));
    // adjoint ss/gs tiles: the tile's cotangent rows ([ns][ns][rows] for ss,
    // [ns][rows] for gs) fit in 64 KB of LDS
    // (TATT: the EE and BB cotangents, 2 ns^2 doubles per row, within 48 KB)
    const int cap_ss = std::max(1, std::min(rows[2], (L.ia_model == CLH_IA_TATT ? 3072 : 8192) / std::max(ns * ns, 1)));
    const int cap_gs = std::max(1, std::min(rows[2], 8192 / std::max(ns, 1)));
    row_tiles(nrow, cap_ss, atiles_ss);
    row_tiles(nrow, cap_gs, atiles_gs);
    for (const int2& t : atiles_ss) {
      wl.push_back(make_int4(i, CLH_W_SS, t.x, t.y));
      lds_ss = std::max(lds_ss, ns * ns * t.y);
    }
    for (int zl = 0; zl < nl; zl++)
      for (const int2& t : atiles_gs) {
        wl_gs.push_back(make_int4(i, CLH_W_GS + 4 * zl, t.x, t.y));
        lds_gs = std::max(lds_gs, ns * t.y);
      }
    row_tiles(nrow, rows[3], tiles);
    for (int zl = 0; zl < nl; zl++)
      for (const int2& t : tiles) wga.push_back(make_int4(i, zl, t.x, t.y));
    }  // L.hdi == 0
    for (int t = 0; t < ntask; t += 8) wl2.push_back(make_int2(i, t));
    // FP64 Legendre on MFMA: 16 columns of one probe kind per item, forward
    // (xi+, xi-, gamma_t, w kernels) and adjoint (ss with both xi kernels, gs, gg)
    const int nkc[4] = {n_ss, n_ss, n_gs, nl};
    for (int k = 0; k < 4; k++)
      for (int c0 = 0; c0 < nkc[k]; c0 += 16) wlf.push_back(make_int4(i, k, c0, 0));
    for (int k : {0, 2, 3})
      for (int c0 = 0; c0 < nkc[k]; c0 += 16) wa.push_back(make_int4(i, k, c0, 0));
    const int nck[4] = {n_ss, n_ss, n_gs, nl};
    for (int k = 0; k < 4; k++)
      for (int c0 = 0; c0 < nck[k]; c0 += 32) wmx.push_back(make_int4(i, k, c0, 0));
  }
  auto up2 = [&](DBuf<int2>& db, const std::vector<int2>& v, int& n) -> hipError_t {
    n = (int)v.size();
    hipError_t e = db.ensure(std::max<size_t>(v.size(), 1));
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpyAsync(db.p, v.data(), v.size() * sizeof(int2), hipMemcpyHostToDevice, c->stream);
  };
  auto up4 = [&](DBuf<int4>& db, const std::vector<int4>& v, int& n) -> hipError_t {
    n = (int)v.size();
    hipError_t e = db.ensure(std::max<size_t>(v.size(), 1));
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpyAsync(db.p, v.data(), v.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream);
  };
  HIPCHK(c, up2(bt.d_wg, wg, bt.n_wg));
  HIPCHK(c, up2(bt.d_wn, wn, bt.n_wn));
  HIPCHK(c, up2(bt.d_wl2, wl2, bt.n_wl2));

  bt.n_wl_ss = (int)wl.size();
  wl.insert(wl.end(), wl_gs.begin(), wl_gs.end());
  HIPCHK(c, up4(bt.d_wl, wl, bt.n_wl));
  bt.n_wsx_std = (int)wsx.size();
  wsx.insert(wsx.end(), wsx_t.begin(), wsx_t.end());
  HIPCHK(c, up4(bt.d_wsx, wsx, bt.n_wsx));
  HIPCHK(c, up4(bt.d_wgg, wgg, bt.n_wgg));
  HIPCHK(c, up4(bt.d_wsx_gs, wsx_gs, bt.n_wsx_gs));
  HIPCHK(c, up4(bt.d_wsx_gs_t, wsx_gs_t, bt.n_wsx_gs_t));
  HIPCHK(c, up2(bt.d_wlim, wlim, bt.n_wlim));
  HIPCHK(c, up4(bt.d_wga, wga, bt.n_wga));
  HIPCHK(c, up4(bt.d_wmx, wmx, bt.n_wmx));
  HIPCHK(c, up4(bt.d_wlf, wlf, bt.n_wlf));
  HIPCHK(c, up4(bt.d_wa4, wa, bt.n_wa));
  HIPCHK(c, up4(bt.d_whd, whd, bt.n_whd));
  bt.adj_lds_ss = lds_ss;
  bt.adj_lds_gs = lds_gs;
  // non-Limber w: the views' gg nodes and rows (ell = 1..149), then the FFTLog
  // stages per draw (node blocks; bin x configuration x m blocks;
  // configuration x m blocks; the ten block-end ells; every ell; bins)
  std::vector<int2> wn_nl, wnl_prep, wnl_cl;
  std::vector<int4> wgg_nl, wnl_fwd, wnl_fwd4, wnl_gl, wnl_chk, wnl_idft;
  int v = 0;
  for (int i = 0; i < bt.B; i++) {
    if (!bt.desc[i].nlw) continue;
    const int nl = bt.desc[i].nl;
    for (int u = v; u < v + 2; u++) {
      for (int sg = nl + 1; sg <= 2 * nl; sg++) wn_nl.push_back(make_int2(u, sg));
      row_tiles(CLH_NL_L - 1, rows[1], tiles);
      for (int zl = 0; zl < nl; zl++)
        for (const int2& t : tiles) wgg_nl.push_back(make_int4(u, zl, t.x, t.y));
    }
    v += 2;
    for (int b = 0; b < 2; b++) wnl_prep.push_back(make_int2(i, b));
    const int nmb = (CLH_NL_NH + 255) / 256;
    for (int zl = 0; zl < nl; zl++)
      for (int j = 0; j < 3; j++) wnl_fwd.push_back(make_int4(i, zl, j, 0));
    for (int zl = 0; zl < nl; zl++)  // the VJP's four tangent sets per bin
      for (int j = 0; j < 4; j++) wnl_fwd4.push_back(make_int4(i, zl, j, 0));
    for (int j = 0; j < 3; j++)
      for (int mb = 0; mb < nmb; mb++) wnl_gl.push_back(make_int4(i, j, mb, 0));
    for (int r = 0; r < 10; r++) wnl_chk.push_back(make_int4(i, r, 0, 0));  // the block-end ells
    for (int l = 0; l < CLH_NL_L; l++) wnl_idft.push_back(make_int4(i, l, 0, 0));
    for (int zl = 0; zl < nl; zl++) wnl_cl.push_back(make_int2(i, zl));
  }
  HIPCHK(c, up2(bt.d_wn_nl, wn_nl, bt.n_wn_nl));
  HIPCHK(c, up4(bt.d_wgg_nl, wgg_nl, bt.n_wgg_nl));
  HIPCHK(c, up2(bt.d_wnl_prep, wnl_prep, bt.n_wnl_prep));
  HIPCHK(c, up4(bt.d_wnl_fwd, wnl_fwd, bt.n_wnl_fwd));
  HIPCHK(c, up4(bt.d_wnl_fwd4, wnl_fwd4, bt.n_wnl_fwd4));
  HIPCHK(c, up4(bt.d_wnl_gl, wnl_gl, bt.n_wnl_gl));
  HIPCHK(c, up4(bt.d_wnl_chk, wnl_chk, bt.n_wnl_chk));
  HIPCHK(c, up4(bt.d_wnl_idft, wnl_idft, bt.n_wnl_idft));
  HIPCHK(c, up2(bt.d_wnl_cl, wnl_cl, bt.n_wnl_cl));
  // 6x2pt: gk per lens bin, ks per source bin, kk; tiles of 32 rows
  std::vector<int4> wcmb;
  std::vector<int> wcmbp;
  for (int i = 0; i < bt.B; i++) {
    const DrawDesc& q = bt.desc[i];
    if (!q.cl6) continue;
    if (!bt.nrow_custom) wcmbp.push_back(i);  // custom ells: C_ell columns only, no placement
    for (int r0 = 0; r0 < q.nrow6; r0 += 32) {
      const int nr = std::min(32, q.nrow6 - r0);
      for (int b = 0; b < q.nl; b++) wcmb.push_back(make_int4(i, 0 + 4 * b, r0, nr));
      for (int b = 0; b < q.ns; b++) wcmb.push_back(make_int4(i, 1 + 4 * b, r0, nr));
      wcmb.push_back(make_int4(i, 2, r0, nr));
    }
  }
  HIPCHK(c, up4(bt.d_wcmb, wcmb, bt.n_wcmb));
  bt.n_wcmbp = (int)wcmbp.size();
  HIPCHK(c, bt.d_wcmbp.ensure(std::max<size_t>(wcmbp.size(), 1)));
  if (!wcmbp.empty())
    HIPCHK(c, hipMemcpyAsync(bt.d_wcmbp.p, wcmbp.data(), wcmbp.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
  return CLH_OK;

}

// Builds the per-draw descriptors, device tables and work lists of a batch;
// draw i uses layout lays[i] with device copy devs[i].  `ells`: if non-null,
// nrow custom ell rows shared by all probes (the direct *_nointerp
// evaluation); otherwise the layout's rows.
static int build_batch(clh_ctx* c, Batch& bt, const std::vector<const Draw*>& draws,
                       const std::vector<const Layout*>& lays, const std::vector<const DevLayout*>& devs,
                       int apply_calib, const int ord[3], const dvec* ells, double* dv_out) {
  int rc;
  if ((rc = ensure_device(c))) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  const int B = (int)draws.size();
  for (int i = 0; i < B; i++) {
    if ((rc = check_layout(c, *lays[i]))) return rc;
    if ((rc = check_draw(c, *lays[i], *draws[i], i))) return rc;
    const Layout& L = *lays[i];
    if (L.hdi != 0 && (L.sixpt || (L.nonlimber && L.lens.nbin > 0 && !ells)))
      return fail(c, CLH_ERR_ARG, "high_def_integration=%d with %s is not implemented", L.hdi,
                  L.sixpt ? "the 6x2pt probes" : "the non-Limber w(theta)");
  }
  auto al = [](size_t n) { return (n + 31) & ~size_t(31); };
  // per-draw sizes and offsets (layouts may differ between draws)
  std::vector<size_t> woff(B + 1, 0), toff(B + 1, 0);
  std::vector<long long> dvoff(B + 1, 0);
  int max_nrow = 0;
  bool uniform = true;
  for (int i = 0; i < B; i++) {
    const Layout& L = *lays[i];
    const Draw& d = *draws[i];
    const int ns = L.src.nbin, nl = L.lens.nbin;
    const int nrow = ells ? (int)ells->size() : L.nrow();
    const size_t node_sz = node_size(ns, nl);
    const bool tatt = L.ia_model == CLH_IA_TATT;
    const bool nlw = L.nonlimber && nl > 0 && !ells;
    woff[i + 1] = woff[i] + al((size_t)(ns + nl) * L.n_a) + al(node_sz) +
                  al(8 * d.chi_z.size()) + al(3 * d.G_z.size()) + al((size_t)5 * nl * L.lens.rows) +
                  al((size_t)(L.n_ss() + L.n_gs() + nl) * nrow) +
                  (tatt ? al((size_t)(1 + nl) * (2 * ns + 1) * CLH_NP_SS) + al((size_t)L.n_ss() * nrow) : 0) +
                  (nlw ? clh_nl_off(nl).total + 2 * (al(node_sz) + al((size_t)(L.n_ss() + L.n_gs() + nl) *
                                                                       (CLH_NL_L - 1)))
                       : 0) +
                  (L.sixpt ? al((size_t)(nl + ns + 1) * (nrow + (ells ? 0 : L.kk_direct_ells().size()))) : 0);
    toff[i + 1] = toff[i] + 2 * d.chi_z.size() + 2 * d.G_z.size() + d.lk.size() + d.pz.size() + d.lnP.size() +
                  (tatt ? d.fpt.size() : 0) + d.lkL.size() + d.pzL.size() + d.lnPL.size() +
                  (d.has_b2(nl) ? d.fptb.size() : 0) + 32;
    dvoff[i + 1] = dvoff[i] + L.dv_size();
    max_nrow = std::max(max_nrow, nrow);
    if (lays[i] != lays[0]) uniform = false;
  }
  HIPCHK(c, bt.d_tab.ensure(toff[B]));
  HIPCHK(c, bt.d_work.ensure(woff[B]));
  if (!dv_out) HIPCHK(c, bt.d_dv.ensure((size_t)dvoff[B]));
  if (ells) {
    HIPCHK(c, bt.d_ells.ensure(ells->size()));
    HIPCHK(c, hipMemcpyAsync(bt.d_ells.p, ells->data(), ells->size() * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
  }
  std::vector<double> htab(toff[B]);
  bt.desc.assign(B, DrawDesc());
  bt.par.assign(B, DrawParams());
  std::vector<std::pair<int, double*>> nl_views;  // (draw, start of its two views' buffers)
  HIPCHK(c, bt.d_par.ensure(B));
  for (int i = 0; i < B; i++) {
    const Draw& d = *draws[i];
    const Layout& L = *lays[i];
    const DevLayout& DL = *devs[i];
    const double* lay = DL.buf.p;
    const int ns = L.src.nbin, nl = L.lens.nbin;
    const int nrow = ells ? (int)ells->size() : L.nrow();
    DrawDesc& q = bt.desc[i];
    std::memset(&q, 0, sizeof q);
    size_t o = toff[i];
    auto push = [&](const dvec& v) { const size_t oo = o; std::copy(v.begin(), v.end(), htab.begin() + o); o += v.size(); return bt.d_tab.p + oo; };
    q.chi_z = push(d.chi_z); q.chi_v = push(d.chi);
    q.G_z = push(d.G_z); q.G_v = push(d.G);
    q.lk = push(d.lk); q.pz = push(d.pz); q.lnP = push(d.lnP);
    const bool nlw = L.nonlimber && nl > 0 && !ells;
    if (!d.lkL.empty()) {  // linear P(k,z): the non-Limber w and int_for_C_gg_tomo_limber
      q.lkL = push(d.lkL); q.pzL = push(d.pzL); q.lnPL = push(d.lnPL);
      q.nkL = (int)d.lkL.size(); q.nzL = (int)d.pzL.size();
    }
    if (nlw) {
      q.h0 = d.h0;
      nl_sizes(q.nl_N, q.nl_N0);
    }
    if (d.has_b2(nl)) {
      q.oneloop = 1;
      q.fptb = push(d.fptb);
      q.fptb_n = d.fptb_n;
      q.fptb_l0 = std::log(d.fptb_kmin);
      q.fptb_l1 = std::log(d.fptb_kmax);
      q.fptb_dl = (q.fptb_l1 - q.fptb_l0) / d.fptb_n;
      q.fptb_s4 = d.fptb_s4;
    }
    if (L.ia_model == CLH_IA_TATT) {
      q.fpt = push(d.fpt);
      q.fpt_n = d.fpt_n;
      q.fpt_l0 = std::log(d.fpt_kmin);
      q.fpt_l1 = std::log(d.fpt_kmax);
      q.fpt_dl = (q.fpt_l1 - q.fpt_l0) / d.fpt_n;
    }
    q.chi_n = (int)d.chi_z.size(); q.G_n = (int)d.G_z.size();
    q.nk = (int)d.lk.size(); q.nzp = (int)d.pz.size();
    q.omega_m = d.omega_m;
    q.ns = ns; q.nl = nl; q.ntheta = L.ntheta; q.n_ell = L.n_ell; q.n_a = L.n_a; q.nrow = nrow;
    q.n_ss = L.n_ss(); q.n_gs = L.n_gs(); q.lmax = L.lmax;
    q.s_zc = lay + DL.off_s_zc; q.s_coef = lay + DL.off_s_coef; q.s_rows = L.src.rows;
    q.l_zc = lay + DL.off_l_zc; q.l_coef = lay + DL.off_l_coef; q.l_rows = L.lens.rows;
    q.s_zmin_all = L.src.zmin_all; q.s_zmax_all = L.src.zmax_all;
    q.l_zmin_all = L.lens.zmin_all; q.l_zmax_all = L.lens.zmax_all;
    DrawParams& pp = bt.par[i];
    for (int b = 0; b < CLH_MAXB; b++) {
      q.l_zmin[b] = L.lens.zmin[b]; q.l_zmax[b] = L.lens.zmax[b]; q.l_zmean[b] = L.zmean[b];
      pp.src_dz[b] = d.src_dz[b]; pp.lens_dz[b] = d.lens_dz[b]; pp.lens_st[b] = d.lens_st[b];
      pp.b1[b] = d.b1[b]; pp.bmag[b] = d.bmag[b]; pp.shear_m[b] = d.shear_m[b]; pp.ia_bin[b] = d.ia_bin[b];
      pp.pm[b] = d.pm[b];
    }
    pp.ia_A = d.ia_A; pp.ia_eta = d.ia_eta;
    for (int b = 0; b < CLH_MAXB; b++) { pp.ia2_bin[b] = d.ia2_bin[b]; pp.bta_bin[b] = d.bta_bin[b]; }
    pp.ia2_A = d.ia2_A; pp.ia2_eta = d.ia2_eta; pp.bta_A = d.bta_A;
    for (int b = 0; b < CLH_MAXB; b++) { pp.b2[b] = d.b2[b]; pp.bs2[b] = d.bs2[b]; pp.b3[b] = d.b3[b]; pp.bk[b] = d.bk[b]; }
    q.ia_model = L.ia_model;
    q.par = bt.d_par.p + i;
    q.ia_mode = L.ia_mode;
    q.ell_ss = ells ? bt.d_ells.p : lay + DL.off_ell_ss;
    q.ell_gx = ells ? bt.d_ells.p : lay + DL.off_ell_gx;
    q.kfold = lay + DL.off_kfold;
    q.kfoldT = lay + DL.off_kfoldT;
    q.kfull = (c->legendre_mode == CLH_LEGENDRE_FULL) ? lay + DL.off_kfull : nullptr;
    q.fill = (c->legendre_mode == CLH_LEGENDRE_FULL) ? lay + DL.off_fill : nullptr;
    q.gl = lay + DL.off_gl;
    q.gx = 25 * (1 + std::abs(L.hdi));
    q.np_ss = L.hdi ? gl_order_ssgs(L.hdi) : CLH_NP_SS;
    q.np_gg = L.hdi ? gl_order_gg(L.hdi) : CLH_NP_GG;
    q.gl_hd = L.hdi ? q.gl + 2 * CLH_NP_SS + 2 * CLH_NP_GG : nullptr;
    q.kmx32 = DL.k32.p; q.kmxbf = DL.kbf.p; q.kmxsc = DL.ksc.p;
    q.kmx_ngrp = DL.ngrp;
    q.theta = lay + DL.off_theta;
    q.mask = L.mask.empty() ? nullptr : DL.mask.p;
    q.gs_row = (L.ggl_ex.empty() || L.n_gs() == 0) ? nullptr : DL.gsrow.p;
    q.n_gs_dv = L.n_gs_dv();
    const size_t gtab_sz = (size_t)(ns + nl) * L.n_a;
    const size_t node_sz = node_size(ns, nl);
    double* wb = bt.d_work.p + woff[i];
    q.gtab = wb;
    q.nodes = wb + al(gtab_sz);
    q.cl = q.nodes + al(node_sz);
    double* wend = q.cl + al((size_t)(L.n_ss() + L.n_gs() + nl) * nrow);
    q.crec = wend;
    wend += al(8 * d.chi_z.size());
    q.grec = wend;
    wend += al(3 * d.G_z.size());
    q.zg_same = d.G_z == d.chi_z ? 1 : 0;
    q.nrec = wend;
    wend += al((size_t)5 * nl * L.lens.rows);
    if (L.ia_model == CLH_IA_TATT) {
      q.tatt = wend;
      q.cl_bb = q.tatt + al((size_t)(1 + nl) * (2 * ns + 1) * CLH_NP_SS);
      wend = q.cl_bb + al((size_t)L.n_ss() * nrow);
    }
    {  // ln-ell grid of the gs / gg tables (cosmo2D.c:2204-2207, 2703-2706)
      q.lim_gx[0] = std::log(std::fmax((double)CLH_LMIN_TAB, 1.0));
      q.lim_gx[1] = std::log((double)(L.lmax + 1));
      q.lim_gx[2] = (q.lim_gx[1] - q.lim_gx[0]) / ((double)L.n_ell - 1.0);
      q.lim_ss[0] = std::log(std::fmax(CLH_LMIN_TAB - 1., 1.0));
      q.lim_ss[1] = q.lim_gx[1];
      q.lim_ss[2] = (q.lim_ss[1] - q.lim_ss[0]) / ((double)L.n_ell - 1.0);
    }
    if (L.sixpt) {  // custom ells: the C_gk / C_ks / C_kk columns at those ells only
      q.n_f = ells ? 0 : (int)L.kk_direct_ells().size();
      q.nrow6 = nrow + q.n_f;
      if (!ells) {
        q.kf6 = lay + DL.off_kf6;
        q.ell_f = lay + DL.off_ellf;
        q.kk_nbp = L.kk_nbp;
        q.kk_lmin = L.kk_lmin;
        q.kk_nL = L.kk_lmax - L.kk_lmin + 1;
        q.kk_M = lay + DL.off_kkM;
        q.kk_off = lay + DL.off_kkoff;
      }
      double cc, dd;
      host_chi_all(d, 1.0 / 1091.0, cc, dd);
      q.chi_cmb = cc;
      q.cl6 = wend;
      wend += al((size_t)(nl + ns + 1) * q.nrow6);
    }
    if (nlw) {
      q.nlw = wend;
      wend += clh_nl_off(nl).total;
      q.nl_tw = lay + DL.off_nltw;
      q.nl_pw = lay + DL.off_kwlow;
      nl_views.push_back({i, wend});
    }
    q.dv = dv_out ? dv_out + dvoff[i] : bt.d_dv.p + dvoff[i];
    dv_starts(L, ord, q.dv_start, q.dv6_start, c->ord6);
    q.node_ss = 0;
    q.node_gs = CLH_NP_SS * (3 + ns + CLH_LUT_SX);
    q.node_gs_stride = CLH_NP_SS * (5 + ns + CLH_LUT_SX);
    q.node_gg_stride = CLH_NP_GG * (5 + CLH_LUT_GG);
    q.node_gg = q.node_gs + nl * q.node_gs_stride;
    double chi_amin, dd;
    host_chi_all(d, CLH_A_MIN, chi_amin, dd);
    q.chi_amin = chi_amin;
    {  // G(z=0) exactly as norm_growfac interpolates it (cosmo3D.c:151-167)
      const int j = bs_lo(d.G_z.data(), (int)d.G_z.size(), 0.0);
      const double dy = (0.0 - d.G_z[j]) / (d.G_z[j + 1] - d.G_z[j]);
      q.G0 = d.G[j] + dy * (d.G[j + 1] - d.G[j]);
    }
    q.apply_calib = apply_calib;
  }
  // non-Limber views: draw copies that integrate the gg Limber rows at
  // ell = 1..149, with P_nl and with P_lin, into their own nodes and rows
  bt.desc_nl.clear();
  if (!nl_views.empty()) {
    if (!bt.d_nl_ells.p) {
      dvec e(CLH_NL_L - 1);
      for (int l = 1; l < CLH_NL_L; l++) e[l - 1] = (double)l;
      HIPCHK(c, bt.d_nl_ells.ensure(e.size()));
      HIPCHK(c, hipMemcpy(bt.d_nl_ells.p, e.data(), e.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    for (auto& v : nl_views) {
      DrawDesc& q = bt.desc[v.first];
      const Layout& L = *lays[v.first];
      const size_t node_sz = node_size(L.src.nbin, L.lens.nbin);
      const size_t cl_sz = (size_t)(L.n_ss() + L.n_gs() + L.lens.nbin) * (CLH_NL_L - 1);
      double* p = v.second;
      for (int lin = 0; lin < 2; lin++) {
        DrawDesc u = q;
        u.ell_ss = u.ell_gx = bt.d_nl_ells.p;
        u.nrow = CLH_NL_L - 1;
        u.nodes = p; p += al(node_sz);
        u.cl = p; p += al(cl_sz);
        u.dv = nullptr;
        u.nlw = nullptr;
        u.tatt = nullptr;
        u.cl_bb = nullptr;
        u.kfull = nullptr;
        u.fill = nullptr;
        u.apply_calib = 0;
        if (lin) {  // C_gg_tomo_limber_linpsopt_nointerp(use_linear_ps = 1): no one-loop term
          u.lk = q.lkL; u.pz = q.pzL; u.lnP = const_cast<double*>(q.lnPL);
          u.nk = q.nkL; u.nzp = q.nzL;
          u.oneloop = 0;
        }
        (lin ? q.nl_llin : q.nl_lnl) = u.cl + (size_t)(L.n_ss() + L.n_gs()) * (CLH_NL_L - 1);
        bt.desc_nl.push_back(u);
      }
    }
    HIPCHK(c, bt.d_desc_nl.ensure(bt.desc_nl.size()));
    HIPCHK(c, hipMemcpyAsync(bt.d_desc_nl.p, bt.desc_nl.data(), bt.desc_nl.size() * sizeof(DrawDesc),
                             hipMemcpyHostToDevice, c->stream));
  }
  bt.nl_any = !nl_views.empty();
  HIPCHK(c, hipMemcpyAsync(bt.d_tab.p, htab.data(), toff[B] * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, bt.d_desc.ensure(B));
  HIPCHK(c, hipMemcpyAsync(bt.d_desc.p, bt.desc.data(), B * sizeof(DrawDesc), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(bt.d_par.p, bt.par.data(), B * sizeof(DrawParams), hipMemcpyHostToDevice, c->stream));
  bt.nknz = (int)draws[0]->lnP.size();
  for (int i = 1; i < B; i++)
    if ((int)draws[i]->lnP.size() != bt.nknz || draws[i]->pz.size() != draws[0]->pz.size()) bt.nknz = 0;
  for (int k = 0; k < 3; k++) bt.ord[k] = ord[k];
  bt.B = B;
  bt.max_nrow = max_nrow;
  bt.uniform = uniform;
  bt.lays = lays;
  bt.dv_off = dvoff;
  bt.dv_size = lays[0]->dv_size();
  bt.nrow_custom = ells ? (int)ells->size() : 0;
  bt.state_valid = false;
  bt.tatt = false;
  bt.oneloop = false;
  bt.cmb_any = false;
  bt.hdi_any = false;
  for (int i = 0; i < B; i++) {
    bt.hdi_any = bt.hdi_any || lays[i]->hdi != 0;
    bt.cmb_any = bt.cmb_any || bt.desc[i].cl6 != nullptr;
    bt.tatt = bt.tatt || lays[i]->ia_model == CLH_IA_TATT;
    bt.oneloop = bt.oneloop || bt.desc[i].oneloop;
  }
  {  // the gg kernels stage chi(z), G(z) and one lens n(z) spline in LDS when
     // the largest draw's tables fit in 44 KB (64 KB per block with the static LDS)
    size_t st = 0;
    for (int i = 0; i < B; i++) {
      const Layout& L = *lays[i];
      const size_t lr = L.lens.nbin > 0 ? (size_t)L.lens.rows : 0;
      st = std::max(st, 2 * draws[i]->chi_z.size() + 2 * draws[i]->G_z.size() + (lr ? lr + 4 * (lr - 1) : 0));
    }
    bt.gg_stage = st <= 5632 ? (int)st : 0;
  }
  if ((rc = build_work(c, bt, c->tile_rows))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CLH_OK;
}

// a batch whose draws all use the context's current layout
static int build_batch_ctx(clh_ctx* c, Batch& bt, const std::vector<const Draw*>& draws, int apply_calib,
                           const int ord[3], const dvec* ells, double* dv_out) {
  int rc = check_layout(c, c->L);
  if (rc) return rc;
  if ((rc = ensure_device(c))) return rc;
  const bool full = c->legendre_mode == CLH_LEGENDRE_FULL;
  if (!c->L.kern_valid || (full && c->L.kfull.empty())) build_kernels(c->L, full);
  if (!c->dl.valid && (rc = upload_layout(c, c->L, c->dl))) return rc;
  const std::vector<const Layout*> lays(draws.size(), &c->L);
  const std::vector<const DevLayout*> devs(draws.size(), &c->dl);
  rc = build_batch(c, bt, draws, lays, devs, apply_calib, ord, ells, dv_out);
  bt.ctx_layout = true;
  bt.lay_epoch = c->layout_epoch;
  return rc;
}

static int run_batch(clh_ctx* c, Batch& bt, bool legendre) {
  hipStream_t s = c->stream;
  if (c->timing) HIPCHK(c, hipEventRecord(c->ev[0], s));
  // with the checkpoint, the tables come with their photo-z shift tangents
  if (bt.keep_ckpt) HIPCHK(c, clh_launch_gtab_tangent(bt.d_desc.p, bt.d_wg.p, bt.n_wg, 1, s));
  else HIPCHK(c, clh_launch_gtab(bt.d_desc.p, bt.d_wg.p, bt.n_wg, s));
  if (c->timing) HIPCHK(c, hipEventRecord(c->ev[1], s));
  HIPCHK(c, clh_launch_nodes(bt.d_desc.p, bt.d_wn.p, bt.n_wn, s));
  if (c->timing) HIPCHK(c, hipEventRecord(c->ev[2], s));
  const bool mixed = c->precision != CLH_PREC_FP64;
  HIPCHK(c, clh_launch_limber_gg(bt.d_desc.p, bt.d_wgg.p, bt.n_wgg, mixed, c->gg_lds[0] ? bt.gg_stage : 0,
                                 bt.keep_ckpt && bt.ggk_per ? 1 : 0, s, bt.oneloop ? 1 : 0));
  if (c->timing) HIPCHK(c, hipEventRecord(c->ev[3], s));
  // ss/gs tiles: the other draws' in k_limber_sx, then the TATT draws' in
  // k_limber_sx_tatt (FP64 in every precision mode)
  HIPCHK(c, clh_launch_limber_sx(bt.d_desc.p, bt.d_wsx.p, bt.n_wsx_std, mixed, bt.keep_ckpt ? 1 : 0, s,
                                 bt.oneloop ? 1 : 0));
  if (bt.n_wsx > bt.n_wsx_std)
    HIPCHK(c, clh_launch_limber_sx_tatt(bt.d_desc.p, bt.d_wsx.p + bt.n_wsx_std, bt.n_wsx - bt.n_wsx_std, s));
  // draws at high_def_integration != 0: n-point GL rows of every column (FP64)
  if (bt.n_whd) HIPCHK(c, clh_launch_limber_hdi(bt.d_desc.p, bt.d_whd.p, bt.n_whd, s));
  if (c->timing) HIPCHK(c, hipEventRecord(c->ev[4], s));
  if (bt.cmb_any)  // 6x2pt CMB lensing columns and blocks: FP64 in every precision mode
    HIPCHK(c, clh_launch_cmb(bt.d_desc.p, bt.d_wcmb.p, bt.n_wcmb, bt.d_wcmbp.p, legendre ? bt.n_wcmbp : 0, s));
  if (legendre && bt.nl_any) {  // non-Limber w: FP64 in every precision mode
    HIPCHK(c, clh_launch_nodes(bt.d_desc_nl.p, bt.d_wn_nl.p, bt.n_wn_nl, s));
    HIPCHK(c, clh_launch_limber_gg(bt.d_desc_nl.p, bt.d_wgg_nl.p, bt.n_wgg_nl, 0, c->gg_lds[0] ? bt.gg_stage : 0, 0, s,
                                   bt.oneloop ? 1 : 0));
    HIPCHK(c, clh_launch_nl_forward(bt.d_desc.p, bt.d_wnl_prep.p, bt.n_wnl_prep, bt.d_wnl_fwd.p, bt.n_wnl_fwd,
                                    bt.d_wnl_gl.p, bt.n_wnl_gl, bt.d_wnl_chk.p, bt.n_wnl_chk, bt.d_wnl_cl.p,
                                    bt.n_wnl_cl, bt.d_wnl_idft.p, bt.n_wnl_idft, s));
  }
  if (legendre) {
    if (c->legendre_mode == CLH_LEGENDRE_FULL)
      HIPCHK(c, clh_launch_legendre_full(bt.d_desc.p, bt.d_wl2.p, bt.n_wl2, bt.max_nrow, s));
    else if (mixed)
      HIPCHK(c, clh_launch_legendre_mx(bt.d_desc.p, bt.d_wmx.p, bt.n_wmx, c->precision, s));
    else
      HIPCHK(c, clh_launch_legendre(bt.d_desc.p, bt.d_wlf.p, bt.n_wlf, s));
    if (bt.nl_any) HIPCHK(c, clh_launch_nl_cl(bt.d_desc.p, bt.d_wnl_cl.p, bt.n_wnl_cl, s));
  }
  if (c->timing) HIPCHK(c, hipEventRecord(c->ev[5], s));
  bt.state_valid = true;
  bt.ckpt_fresh = bt.keep_ckpt;
  bt.rows64 = !mixed;
  return CLH_OK;
}

// single-draw compute of the current state, cached by version
static int ensure_current(clh_ctx* c) {
  if (c->cached_version == c->version) return CLH_OK;
  const int ord[3] = {0, 1, 2};
  std::vector<const Draw*> v{&c->cur};
  int rc = build_batch_ctx(c, c->single, v, 0, ord, nullptr, nullptr);
  if (rc) return rc;
  if ((rc = run_batch(c, c->single, true))) return rc;
  c->cache_raw.resize(c->single.dv_size);
  const size_t ncl = (size_t)(c->L.n_ss() + c->L.n_gs() + c->L.lens.nbin) * c->L.nrow();
  c->cache_cl.resize(ncl);
  HIPCHK(c, hipMemcpyAsync(c->cache_raw.data(), c->single.desc[0].dv, c->cache_raw.size() * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->cache_cl.data(), c->single.desc[0].cl, ncl * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
  c->cache_nl.clear();
  if (c->single.desc[0].nlw) {
    const ClhNlOff off = clh_nl_off(c->L.lens.nbin);
    c->cache_nl.resize((size_t)c->L.lens.nbin * (CLH_NL_L + 1));
    HIPCHK(c, hipMemcpyAsync(c->cache_nl.data(), c->single.desc[0].nlw + off.used,
                             c->cache_nl.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  c->cache_cl6.clear();
  if (c->single.desc[0].cl6) {
    const DrawDesc& q = c->single.desc[0];
    c->cache_cl6.resize((size_t)(q.nl + q.ns + 1) * q.nrow6);
    c->cache_nrow6 = q.nrow6;
    HIPCHK(c, hipMemcpyAsync(c->cache_cl6.data(), q.cl6, c->cache_cl6.size() * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
  }
  c->cache_bb.assign((size_t)c->L.n_ss() * c->L.nrow(), 0.0);
  if (c->single.desc[0].cl_bb)
    HIPCHK(c, hipMemcpyAsync(c->cache_bb.data(), c->single.desc[0].cl_bb, c->cache_bb.size() * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->cached_version = c->version;
  return CLH_OK;
}

// the uploaded batch no longer matches the layouts it was built on: the
// Legendre mode changed, or (for a batch on the context's own layout) any
// layout setter ran since the upload.  The per-draw data-vector slices must
// still follow the layouts' sizes.
static bool batch_stale(const clh_ctx* c, const Batch& bt) {
  if (c->batch_epoch != c->mode_epoch) return true;
  if (bt.ctx_layout && (bt.lay_epoch != c->layout_epoch || !c->dl.valid)) return true;
  if ((int)bt.lays.size() != bt.B || (int)bt.dv_off.size() != bt.B + 1) return true;
  for (int i = 0; i < bt.B; i++)
    if (bt.dv_off[i + 1] - bt.dv_off[i] != (long long)bt.lays[i]->dv_size()) return true;
  return false;
}

// ============================================================== C ABI
extern "C" {

int clh_create(clh_ctx** out, int dev) {
  if (!out) return CLH_ERR_ARG;
  *out = nullptr;
  if (dev < 0) return CLH_ERR_ARG;
  clh_ctx* c = new (std::nothrow) clh_ctx();
  if (!c) return CLH_ERR_ALLOC;
  c->device = dev;  // the device, stream and events are set up on first use
  *out = c;
  return CLH_OK;
}

void clh_destroy(clh_ctx* c) {
  if (!c) return;
  if (c->own_stream) {
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
  }
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  c->single.release(); c->batch.release(); c->scratch.release(); c->arena.release();
  c->dl.release();
  c->rl.clear();
  c->like_data.release(); c->like_icov.release(); c->like_mask.release();
  c->like_R.release(); c->like_Y.release();
  c->d_intf.release();
  if (c->blas) (void)rocblas_destroy_handle(c->blas);
  for (auto& e : c->ev) if (e) (void)hipEventDestroy(e);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
}

const char* clh_last_error(const clh_ctx* c) { return c ? c->err.c_str() : "null context"; }

int clh_set_legendre_mode(clh_ctx* c, int mode) {
  if (!c) return CLH_ERR_ARG;
  if (mode != CLH_LEGENDRE_FOLDED && mode != CLH_LEGENDRE_FULL) return fail(c, CLH_ERR_ARG, "bad mode %d", mode);
  if (mode != c->legendre_mode) {
    c->legendre_mode = mode;
    layout_changed(c);
    c->mode_epoch++;
    c->version++;
  }
  return CLH_OK;
}

int clh_set_precision(clh_ctx* c, int mode) {
  if (!c) return CLH_ERR_ARG;
  if (mode != CLH_PREC_FP64 && mode != CLH_PREC_FP32 && mode != CLH_PREC_BF16X3)
    return fail(c, CLH_ERR_ARG, "bad precision mode %d", mode);
  if (mode != c->precision) {
    c->precision = mode;
    c->version++;
    c->batch.state_valid = false;
    if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
  }
  return CLH_OK;
}

int clh_init_ntable(clh_ctx* c, int n_ell, int n_a, int lmax, int hdi) {
  if (!c) return CLH_ERR_ARG;
  if (n_ell < 2 || n_a < 2 || n_a > 1025 || lmax < CLH_LMIN_TAB + 2)
    return fail(c, CLH_ERR_ARG, "bad Ntable (N_ell=%d, N_a=%d, LMAX=%d)", n_ell, n_a, lmax);
  c->L.n_ell = n_ell; c->L.n_a = n_a; c->L.lmax = lmax; c->L.hdi = hdi;
  if (c->L.lens.set) lens_zmean(c->L);  // its GL order follows high_def_integration
  c->L.kern_valid = false;
  layout_changed(c);
  c->version++;
  return CLH_OK;
}

int clh_init_binning_real_space(clh_ctx* c, int ntheta, double tmin, double tmax) {
  if (!c) return CLH_ERR_ARG;
  if (!(ntheta > 0) || ntheta > 32 || !(tmin > 0) || !(tmax > tmin))
    return fail(c, CLH_ERR_ARG, "bad theta binning (Ntheta=%d in 1..32, %g, %g)", ntheta, tmin, tmax);
  c->L.ntheta = ntheta;
  c->L.vtmin = tmin * 2.90888208665721580e-4;
  c->L.vtmax = tmax * 2.90888208665721580e-4;
  // bin-averaged theta of compute_binning_real_space (generic_interface.cpp:2037-2063)
  c->L.theta.resize(ntheta);
  {
    const double logdt = (std::log(c->L.vtmax) - std::log(c->L.vtmin)) / ntheta;
    for (int i = 0; i < ntheta; i++) {
      const double t0 = std::exp(std::log(c->L.vtmin) + (i + 0.) * logdt);
      const double t1 = std::exp(std::log(c->L.vtmin) + (i + 1.) * logdt);
      c->L.theta[i] = (2. / 3.) * (std::pow(t1, 3) - std::pow(t0, 3)) / (t1 * t1 - t0 * t0);
    }
  }
  if (!c->L.mask.empty() && (int)c->L.mask.size() != c->L.dv_size()) c->L.mask.clear();
  c->L.kern_valid = false;
  layout_changed(c);
  c->version++;
  return CLH_OK;
}

int clh_init_ia(clh_ctx* c, int model, int evol) {
  if (!c) return CLH_ERR_ARG;
  if (model != 0 && model != CLH_IA_TATT)
    return fail(c, CLH_ERR_ARG, "IA_MODEL=%d not implemented (NLA=0, TATT=1)", model);
  if (evol != 0 && evol != 2 && evol != 3) return fail(c, CLH_ERR_ARG, "IA redshift evolution %d not supported", evol);
  if (model != c->L.ia_model || evol != c->L.ia_mode) layout_changed(c);
  c->L.ia_model = model;
  c->L.ia_mode = evol;
  c->version++;
  return CLH_OK;
}

int clh_set_cosmological_parameters(clh_ctx* c, double om, double hubble) {
  if (!c) return CLH_ERR_ARG;
  if (std::isnan(om) || std::isnan(hubble)) return fail(c, CLH_ERR_NAN, "NaN cosmological parameter");
  c->cur.omega_m = om;
  if (hubble > 0) c->cur.h0 = hubble / 100.0;
  c->version++;
  return CLH_OK;
}

static int set_pair(clh_ctx* c, dvec& X, dvec& Y, const double* x, const double* y, int n, int nmin, const char* what) {
  if (!c) return CLH_ERR_ARG;
  if (!x || !y || n < nmin) return fail(c, CLH_ERR_ARG, "%s: size %d (>= %d)", what, n, nmin);
  if (has_nan(x, n) || has_nan(y, n)) return fail(c, CLH_ERR_NAN, "%s: NaN", what);
  X.assign(x, x + n);
  Y.assign(y, y + n);
  c->version++;
  return CLH_OK;
}

int clh_set_distances(clh_ctx* c, const double* z, const double* chi, int n) {
  return set_pair(c, c->cur.chi_z, c->cur.chi, z, chi, n, 5, "set_distances");
}
int clh_set_growth(clh_ctx* c, const double* z, const double* G, int n) {
  return set_pair(c, c->cur.G_z, c->cur.G, z, G, n, 2, "set_growth");
}

int clh_set_non_linear_power_spectrum(clh_ctx* c, const double* lk, int nk, const double* z, int nz,
                                      const double* lnP) {
  if (!c) return CLH_ERR_ARG;
  if (!lk || !z || !lnP || nk < 2 || nz < 2) return fail(c, CLH_ERR_ARG, "set_non_linear_power_spectrum: sizes");
  if (has_nan(lk, nk) || has_nan(z, nz) || has_nan(lnP, nk * nz)) return fail(c, CLH_ERR_NAN, "P(k,z): NaN");
  c->cur.lk.assign(lk, lk + nk);
  c->cur.pz.assign(z, z + nz);
  c->cur.lnP.assign(lnP, lnP + (size_t)nk * nz);
  c->version++;
  return CLH_OK;
}

// set_linear_power_spectrum (generic_interface.cpp:1194-1270): ln P_lin
// k-major, (Mpc/h)^3, k in h/Mpc; read by the non-Limber w only
int clh_set_linear_power_spectrum(clh_ctx* c, const double* lk, int nk, const double* z, int nz,
                                  const double* lnP) {
  if (!c) return CLH_ERR_ARG;
  if (!lk || !z || !lnP || nk < 2 || nz < 2) return fail(c, CLH_ERR_ARG, "set_linear_power_spectrum: sizes");
  if (has_nan(lk, nk) || has_nan(z, nz) || has_nan(lnP, nk * nz)) return fail(c, CLH_ERR_NAN, "P_lin(k,z): NaN");
  c->cur.lkL.assign(lk, lk + nk);
  c->cur.pzL.assign(z, z + nz);
  c->cur.lnPL.assign(lnP, lnP + (size_t)nk * nz);
  c->version++;
  return CLH_OK;
}

// set_nuisance_nonlinear_bias (generic_interface.cpp:1533-1568): b2 per lens
// bin, and bs2 = -4/7 (b1 - 1) unless b2 = 0 (then 0)
int clh_set_nuisance_nonlinear_bias(clh_ctx* c, const double* B1, const double* B2, int n) {
  if (!c) return CLH_ERR_ARG;
  const int nl = c->L.lens.nbin;
  if (nl == 0) return fail(c, CLH_ERR_STATE, "set_nuisance_nonlinear_bias: lens sample not set");
  if (!B1 || !B2 || n != nl) return fail(c, CLH_ERR_ARG, "set_nuisance_nonlinear_bias: need %d values", nl);
  if (has_nan(B1, n) || has_nan(B2, n)) return fail(c, CLH_ERR_NAN, "set_nuisance_nonlinear_bias: NaN");
  for (int i = 0; i < nl; i++) {
    c->cur.b2[i] = B2[i];
    c->cur.bs2[i] = B2[i] == 0.0 ? 0.0 : (-4. / 7.) * (B1[i] - 1.0);
  }
  c->version++;
  return CLH_OK;
}

// set_nuisance_nonlocal_bias (generic_interface.cpp:1610-1669): b3 and bK per lens bin
int clh_set_nuisance_nonlocal_bias(clh_ctx* c, const double* B3, const double* BK, int n) {
  if (!c) return CLH_ERR_ARG;
  const int nl = c->L.lens.nbin;
  if (nl == 0) return fail(c, CLH_ERR_STATE, "set_nuisance_nonlocal_bias: lens sample not set");
  if (!B3 || !BK || n != nl) return fail(c, CLH_ERR_ARG, "set_nuisance_nonlocal_bias: need %d values", nl);
  if (has_nan(B3, n) || has_nan(BK, n)) return fail(c, CLH_ERR_NAN, "set_nuisance_nonlocal_bias: NaN");
  for (int i = 0; i < nl; i++) {
    c->cur.b3[i] = B3[i];
    c->cur.bk[i] = BK[i];
  }
  c->version++;
  return CLH_OK;
}

// set_bias_PS (generic_interface.cpp:1008-1076): PS[12][N], rows 0..5 the
// FAST-PT bias spectra (d1d2, d2d2, d1s2, d2s2, s2s2, d1p3) in (Mpc/h)^3,
// k range and cutoff in h/Mpc, sigma4 in (Mpc/h)^3; stored as the reference
// stores them (/ (c/H0)^3, k * c/H0)
int clh_set_bias_PS(clh_ctx* c, const double* PS, double kmin, double kmax, double cutoff, double sigma4, int N) {
  if (!c) return CLH_ERR_ARG;
  (void)cutoff;
  if (!PS || N < 2) return fail(c, CLH_ERR_ARG, "set_bias_PS: N=%d", N);
  if (!(kmin > 0) || !(kmax > kmin)) return fail(c, CLH_ERR_ARG, "set_bias_PS: bad k range");
  if (has_nan(PS, 12 * N) || std::isnan(sigma4)) return fail(c, CLH_ERR_NAN, "set_bias_PS: NaN");
  const double h3 = CLH_COVERH0 * CLH_COVERH0 * CLH_COVERH0;
  c->cur.fptb.resize((size_t)6 * N);
  for (size_t i = 0; i < (size_t)6 * N; i++) c->cur.fptb[i] = PS[i] / h3;
  c->cur.fptb_n = N;
  c->cur.fptb_kmin = kmin * CLH_COVERH0;
  c->cur.fptb_kmax = kmax * CLH_COVERH0;
  c->cur.fptb_s4 = sigma4 / h3;
  c->version++;
  return CLH_OK;
}

// init_probes (generic_interface.cpp:640-701) for "3x2pt" and "6x2pt"
int clh_init_probes(clh_ctx* c, const char* probes) {
  if (!c || !probes) return CLH_ERR_ARG;
  const std::string p(probes);
  if (p != "3x2pt" && p != "6x2pt") return fail(c, CLH_ERR_ARG, "init_probes: '%s' not supported (3x2pt, 6x2pt)", probes);
  const bool six = p == "6x2pt";
  if (six != c->L.sixpt) {
    c->L.sixpt = six;
    c->L.kern_valid = false;
    layout_changed(c);
    c->L.mask.clear();
    c->version++;
  }
  return CLH_OK;
}

// init_cmb_cross_correlation (generic_interface.cpp:508-526): filter ell
// range, beam FWHM in arcmin and the HEALPix pixel window (column 1 of the
// reference's file, one value per ell; NULL / 0 = none)
int clh_init_cmb_cross_correlation(clh_ctx* c, int lmin, int lmax, double fwhm_arcmin, const double* win, int nwin) {
  if (!c) return CLH_ERR_ARG;
  if (std::isnan(fwhm_arcmin) || fwhm_arcmin < 0) return fail(c, CLH_ERR_ARG, "init_cmb_cross_correlation: fwhm");
  if (win && nwin > 0 && has_nan(win, nwin)) return fail(c, CLH_ERR_NAN, "init_cmb_cross_correlation: NaN window");
  c->L.cmb_fwhm = fwhm_arcmin * 2.90888208665721580e-4;
  c->L.cmb_lmin = lmin;
  c->L.cmb_lmax = lmax;
  if (win && nwin > 0) c->L.cmb_win.assign(win, win + nwin); else c->L.cmb_win.clear();
  c->L.kern_valid = false;
  layout_changed(c);
  c->version++;
  return CLH_OK;
}

// init_cmb_auto_bandpower (generic_interface.cpp:541-560, IPCMB::
// set_kk_binning_bandpower / set_kk_binning_mat / set_kk_theory_offset,
// :2512-2606): the kk block becomes nbins band powers
// sum_L M[j][L - lmin] C_kk(L) - offset[j] over L = lmin..lmax.  The matrix
// and the offset are the arrays the reference reads from its two files
// (offset NULL = zeros); alpha is the Hartlap factor the reference divides the
// kkkk covariance block by before inverting it: clh_like_set_inv_cov takes the
// inverse the caller formed, so alpha is kept for clh_cmb_kk_bandpower_info.
int clh_init_cmb_auto_bandpower(clh_ctx* c, int nbins, int lmin, int lmax, const double* binning_matrix,
                                const double* theory_offset, double alpha) {
  if (!c) return CLH_ERR_ARG;
  if (!(nbins > 0) || !(lmin > 0) || !(lmax >= lmin))
    return fail(c, CLH_ERR_ARG, "init_cmb_auto_bandpower: bad binning (nbins=%d, lmin=%d, lmax=%d)", nbins, lmin,
                lmax);
  if (!binning_matrix) return fail(c, CLH_ERR_ARG, "init_cmb_auto_bandpower: no binning matrix");
  if (!(alpha > 0)) return fail(c, CLH_ERR_ARG, "init_cmb_auto_bandpower: alpha must be > 0");
  const size_t nm = (size_t)nbins * (size_t)(lmax - lmin + 1);
  if (has_nan(binning_matrix, (int)nm) || (theory_offset && has_nan(theory_offset, nbins)))
    return fail(c, CLH_ERR_NAN, "init_cmb_auto_bandpower: NaN in the binning matrix or offset");
  Layout& L = c->L;
  L.kk_nbp = nbins;
  L.kk_lmin = lmin;
  L.kk_lmax = lmax;
  L.kk_M.assign(binning_matrix, binning_matrix + nm);
  if (theory_offset) L.kk_off.assign(theory_offset, theory_offset + nbins); else L.kk_off.assign(nbins, 0.0);
  L.kk_alpha = alpha;
  layout_changed(c);
  if (L.sixpt) L.mask.clear();  // the data-vector size changes
  c->version++;
  return CLH_OK;
}

int clh_cmb_kk_bandpower_info(clh_ctx* c, int* nbins, int* lmin, int* lmax, double* alpha) {
  if (!c) return CLH_ERR_ARG;
  if (nbins) *nbins = c->L.kk_nbp;
  if (lmin) *lmin = c->L.kk_lmin;
  if (lmax) *lmax = c->L.kk_lmax;
  if (alpha) *alpha = c->L.kk_alpha;
  return CLH_OK;
}

// like.adopt_limber_gg (generic_interface.cpp:269): 1 Limber w(theta) at every
// ell, 0 the non-Limber C_ell below ell 150 (C_cl_tomo)
int clh_set_adopt_limber_gg(clh_ctx* c, int limber) {
  if (!c) return CLH_ERR_ARG;
  if (limber != 0 && limber != 1) return fail(c, CLH_ERR_ARG, "adopt_limber_gg must be 0 or 1");
  if (c->L.nonlimber != (limber == 0)) {
    c->L.nonlimber = limber == 0;
    layout_changed(c);
    c->version++;
  }
  return CLH_OK;
}

int clh_set_source_sample(clh_ctx* c, const double* t, int nrows, int ntomo) {
  if (!c) return CLH_ERR_ARG;
  int rc = build_sample(c, c->L.src, t, nrows, ntomo, true);
  if (rc) return rc;
  layout_changed(c);
  c->version++;
  return CLH_OK;
}

int clh_set_lens_sample(clh_ctx* c, const double* t, int nrows, int ntomo) {
  if (!c) return CLH_ERR_ARG;
  int rc = build_sample(c, c->L.lens, t, nrows, ntomo, false);
  if (rc) return rc;
  lens_zmean(c->L);
  layout_changed(c);
  c->version++;
  return CLH_OK;
}

static int set_vec(clh_ctx* c, double* dst, const double* v, int n, int need, const char* what) {
  if (!c) return CLH_ERR_ARG;
  if (need <= 0) return fail(c, CLH_ERR_STATE, "%s: sample not set", what);
  if (!v || n != need) return fail(c, CLH_ERR_ARG, "%s: size %d (!= %d)", what, n, need);
  if (has_nan(v, n)) return fail(c, CLH_ERR_NAN, "%s: NaN", what);
  for (int i = 0; i < n; i++) dst[i] = v[i];
  c->version++;
  return CLH_OK;
}

int clh_set_nuisance_shear_calib(clh_ctx* c, const double* m, int n) {
  return set_vec(c, c->cur.shear_m, m, n, c ? c->L.src.nbin : 0, "set_nuisance_shear_calib");
}
int clh_set_nuisance_shear_photoz(clh_ctx* c, const double* v, int n) {
  return set_vec(c, c->cur.src_dz, v, n, c ? c->L.src.nbin : 0, "set_nuisance_shear_photoz");
}
int clh_set_nuisance_clustering_photoz(clh_ctx* c, const double* v, int n) {
  return set_vec(c, c->cur.lens_dz, v, n, c ? c->L.lens.nbin : 0, "set_nuisance_clustering_photoz");
}
int clh_set_nuisance_clustering_photoz_stretch(clh_ctx* c, const double* v, int n) {
  return set_vec(c, c->cur.lens_st, v, n, c ? c->L.lens.nbin : 0, "set_nuisance_clustering_photoz_stretch");
}
int clh_set_nuisance_linear_bias(clh_ctx* c, const double* v, int n) {
  return set_vec(c, c->cur.b1, v, n, c ? c->L.lens.nbin : 0, "set_nuisance_linear_bias");
}
int clh_set_nuisance_magnification_bias(clh_ctx* c, const double* v, int n) {
  return set_vec(c, c->cur.bmag, v, n, c ? c->L.lens.nbin : 0, "set_nuisance_magnification_bias");
}
int clh_set_nuisance_ia(clh_ctx* c, const double* A1, int n) {
  if (!c) return CLH_ERR_ARG;
  if (!A1 || n <= 0) return fail(c, CLH_ERR_ARG, "set_nuisance_IA: empty");
  if (has_nan(A1, n)) return fail(c, CLH_ERR_NAN, "set_nuisance_IA: NaN");
  if (c->L.ia_mode == 3) {
    if (n < 2) return fail(c, CLH_ERR_ARG, "set_nuisance_IA: need (A, eta)");
    c->cur.ia_A = A1[0];
    c->cur.ia_eta = A1[1];
  } else if (c->L.ia_mode == 2) {
    if (n < c->L.src.nbin) return fail(c, CLH_ERR_ARG, "set_nuisance_IA: need one A per source bin");
    for (int i = 0; i < c->L.src.nbin; i++) c->cur.ia_bin[i] = A1[i];
  }
  c->version++;
  return CLH_OK;
}

// set_nuisance_IA(A1, A2, BTA) (generic_interface.cpp:1701-1774): mode 2 one
// value per source bin in each array; mode 3 A1 = (A, eta), A2 = (A2, eta2),
// BTA[0] = b_TA
int clh_set_nuisance_ia_tatt(clh_ctx* c, const double* A1, const double* A2, const double* BTA, int n) {
  if (!c) return CLH_ERR_ARG;
  if (!A1 || !A2 || !BTA || n <= 0) return fail(c, CLH_ERR_ARG, "set_nuisance_IA: empty");
  if (has_nan(A1, n) || has_nan(A2, n) || has_nan(BTA, n)) return fail(c, CLH_ERR_NAN, "set_nuisance_IA: NaN");
  if (c->L.ia_mode == 3) {
    if (n < 2) return fail(c, CLH_ERR_ARG, "set_nuisance_IA: need (A, eta), (A2, eta2), b_TA");
    c->cur.ia_A = A1[0];
    c->cur.ia_eta = A1[1];
    c->cur.ia2_A = A2[0];
    c->cur.ia2_eta = A2[1];
    c->cur.bta_A = BTA[0];
  } else if (c->L.ia_mode == 2) {
    if (n < c->L.src.nbin) return fail(c, CLH_ERR_ARG, "set_nuisance_IA: need one value per source bin");
    for (int i = 0; i < c->L.src.nbin; i++) {
      c->cur.ia_bin[i] = A1[i];
      c->cur.ia2_bin[i] = A2[i];
      c->cur.bta_bin[i] = BTA[i];
    }
  }
  c->version++;
  return CLH_OK;
}

// set_IA_PS (generic_interface.cpp:941-1001): PS[12][N], rows 0..9 the
// FAST-PT IA spectra in (Mpc/h)^3, k range in h/Mpc; stored as the reference
// stores them (spectra / (c/H0)^3, k * c/H0)
int clh_set_IA_PS(clh_ctx* c, const double* PS, double kmin, double kmax, double cutoff, int N) {
  if (!c) return CLH_ERR_ARG;
  (void)cutoff;
  if (!PS || N < 2) return fail(c, CLH_ERR_ARG, "set_IA_PS: N=%d", N);
  if (!(kmin > 0) || !(kmax > kmin)) return fail(c, CLH_ERR_ARG, "set_IA_PS: bad k range");
  if (has_nan(PS, 12 * N)) return fail(c, CLH_ERR_NAN, "set_IA_PS: NaN");
  const double h3 = CLH_COVERH0 * CLH_COVERH0 * CLH_COVERH0;
  c->cur.fpt.resize((size_t)10 * N);
  for (size_t i = 0; i < (size_t)10 * N; i++) c->cur.fpt[i] = PS[i] / h3;
  c->cur.fpt_n = N;
  c->cur.fpt_kmin = kmin * CLH_COVERH0;
  c->cur.fpt_kmax = kmax * CLH_COVERH0;
  c->version++;
  return CLH_OK;
}

// PointMass::set_pm_vector (generic_interface.hpp:362-365)
int clh_set_nuisance_pointmass(clh_ctx* c, const double* pm, int n) {
  return set_vec(c, c->cur.pm, pm, n, c ? c->L.lens.nbin : 0, "set_pointmass");
}

// init_ggl_exclude (generic_interface.cpp:874-896): n/2 (lens, source) pairs,
// flat; n = 0 clears.  The pairs hold until the next call, whatever samples
// are set later (test_zoverlap reads them against the current bin counts).
int clh_init_ggl_exclude(clh_ctx* c, const int* pairs, int n) {
  if (!c) return CLH_ERR_ARG;
  if (n < 0 || (n > 1 && !pairs)) return fail(c, CLH_ERR_ARG, "init_ggl_exclude: %d entries", n);
  if (n / 2 > CLH_MAXB * CLH_MAXB) return fail(c, CLH_ERR_ARG, "init_ggl_exclude: %d pairs", n / 2);
  c->L.ggl_ex.assign(pairs ? pairs : nullptr, pairs ? pairs + 2 * (n / 2) : nullptr);
  if (!c->L.mask.empty() && (int)c->L.mask.size() != c->L.dv_size()) c->L.mask.clear();
  layout_changed(c);
  c->version++;
  return CLH_OK;
}

// test_zoverlap (redshift_spline.c:115-145): 1 if the pair is in the data
// vector, 0 if excluded, CLH_ERR_ARG for bins out of range
int clh_test_zoverlap(clh_ctx* c, int zl, int zs) {
  if (!c) return CLH_ERR_ARG;
  if (zl < 0 || zl >= c->L.lens.nbin || zs < 0 || zs >= c->L.src.nbin)
    return fail(c, CLH_ERR_ARG, "invalid bin input (ni, nj) = (%d, %d)", zl, zs);
  return c->L.zoverlap(zl, zs) ? 1 : 0;
}

// The Limber integrands at n scale factors (int_for_C_*_limber, cosmo2D.c),
// params as the reference's arrays; evaluated on the device for the current
// draw (its tables and lens efficiencies from the cached forward)
int clh_int_for_limber(clh_ctx* c, int probe, const double* ar, const double* a, int n, double* out) {
  if (!c) return CLH_ERR_ARG;
  if (!ar || !a || !out || n < 0 || probe < 0 || probe > 5) return fail(c, CLH_ERR_ARG, "int_for: bad arguments");
  if (n == 0) return CLH_OK;
  const Layout& L = c->L;
  const int np[6] = {5, 4, 5, 3, 2, 1};
  for (int k = 0; k < np[probe]; k++)
    if (std::isnan(ar[k])) return fail(c, CLH_ERR_NAN, "int_for: NaN parameter %d", k);
  for (int t = 0; t < n; t++)
    if (!(a[t] > 0) || !(a[t] < 1)) return fail(c, CLH_ERR_ARG, "a>0 and a<1 not true");
  const int ns = L.src.nbin, nl = L.lens.nbin;
  int i = 0, j = 0, f1 = 0, f2 = 0;
  double l = 0.0;
  auto bad = [&](int v, int hi) { return v < 0 || v >= hi; };
  switch (probe) {
    case 0:
      i = (int)ar[0]; j = (int)ar[1]; l = ar[2]; f1 = (int)ar[3]; f2 = (int)ar[4];
      if (bad(i, ns) || bad(j, ns)) return fail(c, CLH_ERR_ARG, "error in selecting bin number (ni,nj) = [%d,%d]", i, j);
      break;
    case 1:
      i = (int)ar[0]; j = (int)ar[1]; l = ar[2]; f1 = (int)ar[3];
      if (bad(i, nl) || bad(j, ns)) return fail(c, CLH_ERR_ARG, "error in selecting bin number (nl, ns) = [%d,%d]", i, j);
      break;
    case 2:
      i = (int)ar[0]; j = (int)ar[1]; l = ar[2]; f1 = (int)ar[3]; f2 = (int)ar[4];
      if (bad(i, nl) || bad(j, nl)) return fail(c, CLH_ERR_ARG, "error in selecting bin number (ni,nj) = [%d,%d]", i, j);
      if (i != j) return fail(c, CLH_ERR_ARG, "cross-tomography (ni,nj) = (%d,%d) bins not supported", i, j);
      break;
    case 3:
      i = (int)ar[0]; l = ar[1]; f1 = (int)ar[2];
      if (bad(i, nl)) return fail(c, CLH_ERR_ARG, "error in selecting bin number ni = %d", i);
      break;
    case 4:
      i = (int)ar[0]; l = ar[1];
      if (bad(i, ns)) return fail(c, CLH_ERR_ARG, "error in selecting bin number ni = %d", i);
      break;
    default:
      l = ar[0];
  }
  if (probe >= 3 && !L.sixpt) return fail(c, CLH_ERR_STATE, "CMB lensing probes need init_probes(\"6x2pt\")");
  int rc = ensure_current(c);
  if (rc) return rc;
  const DrawDesc& q = c->single.desc[0];
  // the one-loop term: FAST-PT bias tables of the draw (set_bias_PS), or no
  // bias terms for the bin, which make it zero
  const bool ol = (probe == 1 || probe == 3) ? f1 == 1 : (probe == 2 && f2 == 1 && f1 == 0);
  if (ol && !q.fptb) {
    const Draw& d = c->cur;
    if (d.b2[i] != 0 || d.bs2[i] != 0 || d.b3[i] != 0 || d.bk[i] != 0)
      return fail(c, CLH_ERR_STATE, "one-loop bias needs the FAST-PT bias tables (set_bias_PS)");
    if (probe == 2) f2 = 0; else f1 = 0;
  }
  if (probe == 2 && f1 == 1 && !q.lnPL) return fail(c, CLH_ERR_STATE, "use_linear_ps: linear P(k,z) not set");
  if (probe == 0 && L.ia_model == CLH_IA_TATT && !q.fpt) return fail(c, CLH_ERR_STATE, "TATT: FAST-PT IA tables not set");
  HIPCHK(c, c->d_intf.ensure(2 * (size_t)n));
  double* da = c->d_intf.p;
  HIPCHK(c, hipMemcpyAsync(da, a, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, clh_launch_int_for(c->single.d_desc.p, probe, i, j, f1, f2, l, da, n, da + n, c->stream));
  HIPCHK(c, hipMemcpyAsync(out, da + n, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CLH_OK;
}

// data-vector index in the block order `ord` of index k in the (0,1,2) order
static int dv_index_from_canonical(const Layout& L, const int ord[3], int k) {
  const int nt = L.ntheta;
  const int c3 = L.dv_size3();
  int start[3], start6[3];
  dv_starts(L, ord, start, start6);
  if (L.sixpt && k >= c3) {  // gk, ks, kk keep their canonical order after the 3x2pt blocks
    const int cs[3] = {c3, c3 + nt * L.lens.nbin, c3 + nt * (L.lens.nbin + L.src.nbin)};
    const int b = k >= cs[2] ? 2 : (k >= cs[1] ? 1 : 0);
    return start6[b] + (k - cs[b]);
  }
  const int canon_start[3] = {0, 2 * nt * L.n_ss(), nt * (2 * L.n_ss() + L.n_gs_dv())};
  const int b = k >= canon_start[2] ? 2 : (k >= canon_start[1] ? 1 : 0);
  return start[b] + (k - canon_start[b]);
}

// IP::set_mask (generic_interface.hpp:1234-1331): 0/1 per data-vector entry,
// in the block order ord; NULL clears the mask
int clh_like_set_mask(clh_ctx* c, const int* mask, int n, const int ord[3]) {
  if (!c) return CLH_ERR_ARG;
  if (!mask) {
    c->L.mask.clear();
  } else {
    if (!ord_ok(c->L, ord)) return fail(c, CLH_ERR_ARG, "set_mask: ord must be a permutation of 0,1,2 whose blocks tile the data vector");
    if (n != c->L.dv_size()) return fail(c, CLH_ERR_ARG, "set_mask: %d entries != data vector size %d", n, c->L.dv_size());
    std::vector<int> m(n);
    int kept = 0;
    for (int k = 0; k < n; k++) {
      const int v = mask[dv_index_from_canonical(c->L, ord, k)];
      if (v != 0 && v != 1) return fail(c, CLH_ERR_DOMAIN, "set_mask: inconsistent mask (entry %d)", k);
      m[k] = v;
      kept += v;
    }
    if (kept == 0) return fail(c, CLH_ERR_DOMAIN, "set_mask: the mask leaves no data points");
    c->L.mask = std::move(m);
  }
  layout_changed(c);
  c->version++;
  return CLH_OK;
}

// IP::set_data: the measured data vector, in the block order ord
int clh_like_set_data(clh_ctx* c, const double* data, int n, const int ord[3]) {
  if (!c || !data) return CLH_ERR_ARG;
  if (!ord_ok(c->L, ord)) return fail(c, CLH_ERR_ARG, "set_data: ord must be a permutation of 0,1,2");
  if (n != c->L.dv_size()) return fail(c, CLH_ERR_ARG, "set_data: %d entries != data vector size %d", n, c->L.dv_size());
  if (has_nan(data, n)) return fail(c, CLH_ERR_NAN, "set_data: NaN");
  if (c->like_icov_set && (c->like_n != n || std::memcmp(ord, c->like_ord, sizeof(int) * 3) != 0))
    c->like_icov_set = false;  // a different size or order than the inverse covariance
  c->like_data_h.assign(data, data + n);
  c->like_n = n;
  for (int k = 0; k < 3; k++) c->like_ord[k] = ord[k];
  c->like_data_set = true;
  return CLH_OK;
}

// IP::set_inv_cov after its inversion: the inverse of the masked covariance
// (with any Hartlap factor), n x n, rows and columns in the block order ord.
// Masked entries drop out of chi^2 through the residuals.
int clh_like_set_inv_cov(clh_ctx* c, const double* inv_cov, int n, const int ord[3]) {
  if (!c || !inv_cov) return CLH_ERR_ARG;
  if (!ord_ok(c->L, ord)) return fail(c, CLH_ERR_ARG, "set_inv_cov: ord must be a permutation of 0,1,2");
  if (n != c->L.dv_size()) return fail(c, CLH_ERR_ARG, "set_inv_cov: %d rows != data vector size %d", n, c->L.dv_size());
  if (has_nan(inv_cov, n * n)) return fail(c, CLH_ERR_NAN, "set_inv_cov: NaN");
  int rc = ensure_device(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, c->like_icov.ensure((size_t)n * n));
  HIPCHK(c, hipMemcpyAsync(c->like_icov.p, inv_cov, sizeof(double) * n * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->like_data_set && (c->like_n != n || std::memcmp(ord, c->like_ord, sizeof(int) * 3) != 0))
    c->like_data_set = false;
  c->like_n = n;
  for (int k = 0; k < 3; k++) c->like_ord[k] = ord[k];
  c->like_icov_set = true;
  return CLH_OK;
}

// chi^2 of B data vectors (device, [B][n] in the block order ord) against
// the data: chi2[b] (device), and if grad != NULL d chi2 / d dv [B][n]
static int chi2_run(clh_ctx* c, const Layout& L, int B, const int ord[3], const double* dv, double* chi2,
                       double* grad) {
  if (!c->like_data_set || !c->like_icov_set) return fail(c, CLH_ERR_STATE, "chi2: data or inverse covariance not set");
  if (std::memcmp(ord, c->like_ord, sizeof(int) * 3) != 0)
    return fail(c, CLH_ERR_ARG, "chi2: ord differs from the order of the data and inverse covariance");
  const int n = c->like_n;
  if (n != L.dv_size()) return fail(c, CLH_ERR_STATE, "chi2: data vector size %d != %d", L.dv_size(), n);
  HIPCHK(c, hipSetDevice(c->device));
  if (c->like_mask_version != c->version || !c->like_mask.p) {
    std::vector<double> m(n, 1.0);
    if (!L.mask.empty())
      for (int k = 0; k < n; k++) m[dv_index_from_canonical(L, ord, k)] = L.mask[k];
    HIPCHK(c, c->like_mask.ensure(n));
    HIPCHK(c, c->like_data.ensure(n));
    HIPCHK(c, hipMemcpyAsync(c->like_mask.p, m.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->like_data.p, c->like_data_h.data(), sizeof(double) * n, hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->like_mask_version = c->version;
  }
  if (!c->blas && rocblas_create_handle(&c->blas) != rocblas_status_success) {
    c->blas = nullptr;
    return fail(c, CLH_ERR_HIP, "rocblas_create_handle failed");
  }
  HIPCHK(c, c->like_R.ensure((size_t)B * n));
  HIPCHK(c, c->like_Y.ensure((size_t)B * n));
  HIPCHK(c, clh_launch_chi2_resid(B, n, dv, c->like_data.p, c->like_mask.p, c->like_R.p, c->stream));
  if (rocblas_set_stream(c->blas, c->stream) != rocblas_status_success)
    return fail(c, CLH_ERR_HIP, "rocblas_set_stream failed");
  const double one = 1.0, zero = 0.0;
  // column-major: Y (n x B) = C^-1 (n x n, symmetric) R (n x B)
  if (rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, n, B, n, &one, c->like_icov.p, n,
                    c->like_R.p, n, &zero, c->like_Y.p, n) != rocblas_status_success)
    return fail(c, CLH_ERR_HIP, "rocblas_dgemm failed");
  HIPCHK(c, clh_launch_chi2_reduce(B, n, c->like_R.p, c->like_Y.p, c->like_mask.p, chi2, grad, c->stream));
  return CLH_OK;
}

int clh_batch_chi2(clh_ctx* c, const int ord[3], const double* dv_device, double* chi2_device, double* grad_device) {
  if (!c || !ord || !dv_device || !chi2_device) return CLH_ERR_ARG;
  Batch& bt = c->batch;
  if (bt.B <= 0) return fail(c, CLH_ERR_STATE, "no batch uploaded");
  if (batch_stale(c, bt)) return fail(c, CLH_ERR_STATE, "layout changed since the batch upload");
  if (!bt.uniform) return fail(c, CLH_ERR_ARG, "chi2 needs a batch with one layout for every draw");
  return chi2_run(c, *bt.lays[0], bt.B, ord, dv_device, chi2_device, grad_device);
}

// IP::get_chi2 (generic_interface.cpp:2288-2325) of one host data vector
int clh_chi2(clh_ctx* c, const double* dv, int n, double* chi2) {
  if (!c || !dv || !chi2) return CLH_ERR_ARG;
  if (n != c->L.dv_size()) return fail(c, CLH_ERR_ARG, "chi2: %d entries != data vector size %d", n, c->L.dv_size());
  int rc = ensure_device(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  DBuf<double> d_dv, d_out;
  HIPCHK(c, d_dv.ensure(n));
  HIPCHK(c, d_out.ensure(1));
  HIPCHK(c, hipMemcpyAsync(d_dv.p, dv, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  if ((rc = chi2_run(c, c->L, 1, c->like_ord, d_dv.p, d_out.p, nullptr))) return rc;
  double v = 0.0;
  HIPCHK(c, hipMemcpyAsync(&v, d_out.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  d_dv.release();
  d_out.release();
  if (!(v >= 0.0)) return fail(c, CLH_ERR_DOMAIN, "chi2 = %g (invalid)", v);
  *chi2 = v;
  return CLH_OK;
}

int clh_datavector_size(clh_ctx* c) { return c ? c->L.dv_size() : 0; }

// init_binning_fourier (generic_interface.cpp:425-470): like.ell at the
// log-centres of nells bins in [lmin, lmax]
int clh_init_binning_fourier(clh_ctx* c, int nells, int lmin, int lmax, int lmax_shear) {
  if (!c) return CLH_ERR_ARG;
  if (!(nells > 0) || !(lmin > 0) || !(lmax > lmin))
    return fail(c, CLH_ERR_ARG, "bad Fourier binning (nells=%d, lmin=%d, lmax=%d)", nells, lmin, lmax);
  c->L.ell_f.resize(nells);
  const double logdl = (std::log((double)lmax) - std::log((double)lmin)) / (double)nells;
  for (int i = 0; i < nells; i++) c->L.ell_f[i] = std::exp(std::log((double)lmin) + (i + 0.5) * logdl);
  c->L.lmax_shear = lmax_shear;
  layout_changed(c);  // like.ell also feeds the 6x2pt kk block
  if (c->L.sixpt) c->L.mask.clear();
  c->version++;
  return CLH_OK;
}

int clh_datavector_size_fourier(clh_ctx* c) {
  if (!c) return 0;
  return (int)c->L.ell_f.size() * (c->L.n_ss() + c->L.n_gs_dv() + c->L.lens.nbin);
}

// compute_Mx2pt_N_masked<1,3> (generic_interface.hpp:1137-1159, 968-1041)
// C_ss (EE, zero at ell >= lmax_shear), C_gs, C_gg at like.ell by the direct
// GL integrals, shear calibration and the mask in data-vector order
// (add_calib_and_set_mask_X_N<1,M>, generic_interface.hpp:800-862; no point
// mass in Fourier space); mask NULL = all ones
int clh_compute_3x2pt_fourier(clh_ctx* c, const int ord[3], const int* mask, double* dv, int n) {
  if (!c || !ord || !dv) return CLH_ERR_ARG;
  const Layout& L = c->L;
  const int ncl = (int)L.ell_f.size();
  if (ncl == 0) return fail(c, CLH_ERR_STATE, "Fourier binning not initialized");
  if (n != clh_datavector_size_fourier(c)) return fail(c, CLH_ERR_ARG, "dv size %d != %d", n, clh_datavector_size_fourier(c));
  const int nss = L.n_ss(), ngs = L.n_gs(), nl = L.lens.nbin, ns = L.src.nbin;
  const int sizes[3] = {ncl * nss, ncl * L.n_gs_dv(), ncl * nl};
  if (!blocks_tile(sizes, ord))
    return fail(c, CLH_ERR_ARG, "ord must be a permutation of 0,1,2 whose blocks tile the data vector");
  int start[3];
  block_starts(sizes, ord, start);
  const int o3[3] = {0, 1, 2};
  std::vector<const Draw*> v{&c->cur};
  int rc = build_batch_ctx(c, c->scratch, v, 0, o3, &L.ell_f, nullptr);
  if (rc) return rc;
  if ((rc = run_batch(c, c->scratch, false))) return rc;
  dvec cl((size_t)(nss + ngs + nl) * ncl);
  HIPCHK(c, hipMemcpyAsync(cl.data(), c->scratch.desc[0].cl, cl.size() * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<double> out(n, 0.0);
  const std::vector<int> rows = L.gs_rows();
  if (clh_place_fourier(ns, nl, ncl, L.ell_f.data(), L.lmax_shear, c->cur.shear_m, rows.data(), start, cl.data(),
                        out.data(), n) != n)
    return fail(c, CLH_ERR_STATE, "Fourier placement does not fill the %d entries", n);
  if (mask)
    for (int k = 0; k < n; k++)
      if (!mask[k]) out[k] = 0.0;
  std::copy(out.begin(), out.end(), dv);
  return CLH_OK;
}

int clh_compute_3x2pt_real(clh_ctx* c, const int ord[3], double* dv, int n) {
  if (!c || !ord || !dv) return CLH_ERR_ARG;
  if (n != c->L.dv_size()) return fail(c, CLH_ERR_ARG, "dv size %d != %d", n, c->L.dv_size());
  if (!ord_ok(c->L, ord))
    return fail(c, CLH_ERR_ARG, "ord must be a permutation of 0,1,2 whose blocks tile the data vector");
  std::vector<const Draw*> v{&c->cur};
  int rc = build_batch_ctx(c, c->single, v, 1, ord, nullptr, nullptr);
  if (rc) return rc;
  if ((rc = run_batch(c, c->single, true))) return rc;
  HIPCHK(c, hipMemcpyAsync(dv, c->single.desc[0].dv, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->cached_version = 0;  // the single batch now holds a calibrated vector
  return CLH_OK;
}

}  // extern "C"

// clh_compute_3x2pt_real with every device buffer of the draw (tables,
// lens-efficiency tables, nodes, C_ell rows, descriptors, work lists, data
// vector) carved from `scratch` (bytes long, 256-byte aligned): no device
// allocation and no free.  The layout's device copy must already be current
// (an earlier clh_compute_3x2pt_real after the last layout setter made it).
int clh_compute_3x2pt_real_scratch(clh_ctx* c, const int ord[3], double* dv, int n, void* scratch, size_t bytes) {
  if (!c || !ord || !dv || !scratch) return CLH_ERR_ARG;
  if (n != c->L.dv_size()) return fail(c, CLH_ERR_ARG, "dv size %d != %d", n, c->L.dv_size());
  if (!ord_ok(c->L, ord))
    return fail(c, CLH_ERR_ARG, "ord must be a permutation of 0,1,2 whose blocks tile the data vector");
  int rc = check_layout(c, c->L);
  if (rc) return rc;
  if (c->L.sixpt) return fail(c, CLH_ERR_ARG, "allocation-free compute: 6x2pt is not supported");
  if (c->L.nonlimber && c->L.lens.nbin > 0)
    return fail(c, CLH_ERR_ARG, "allocation-free compute: the non-Limber w is not supported");
  const bool full = c->legendre_mode == CLH_LEGENDRE_FULL;
  if (!c->own_stream || !c->L.kern_valid || (full && c->L.kfull.empty()) || !c->dl.valid)
    return fail(c, CLH_ERR_STATE,
                "allocation-free compute: the layout's device copy is not current (call "
                "clh_compute_3x2pt_real once after the last layout setter)");
  ClhArena a{(char*)scratch, bytes, 0, false};
  const std::vector<const Draw*> v{&c->cur};
  const std::vector<const Layout*> lays{&c->L};
  const std::vector<const DevLayout*> devs{&c->dl};
  t_arena = &a;
  rc = build_batch(c, c->arena, v, lays, devs, 1, ord, nullptr, nullptr);
  t_arena = nullptr;
  if (rc) {
    if (a.exhausted)
      return fail(c, CLH_ERR_ARG,
                  "allocation-free compute: this layout and draw need more than the %zu bytes of scratch", bytes);
    return rc;
  }
  if ((rc = run_batch(c, c->arena, true))) return rc;
  HIPCHK(c, hipMemcpyAsync(dv, c->arena.desc[0].dv, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CLH_OK;
}

int clh_context_device(const clh_ctx* c) { return c ? c->device : -1; }

void clh_record_error(clh_ctx* c, int code, const char* msg) { (void)fail(c, code, "%s", msg); }

extern "C" {

// compute_Mx2pt_N_masked<0,6>(ord) (generic_interface.hpp:1137-1159): the
// 3x2pt blocks, w_gk, w_ks (x (1 + m)) and C_kk at like.ell, placed by ord
int clh_compute_6x2pt_real(clh_ctx* c, const int ord[6], double* dv, int n) {
  if (!c || !ord || !dv) return CLH_ERR_ARG;
  if (!c->L.sixpt) return fail(c, CLH_ERR_STATE, "6x2pt not selected (clh_init_probes(ctx, \"6x2pt\"))");
  if (n != c->L.dv_size()) return fail(c, CLH_ERR_ARG, "dv size %d != %d", n, c->L.dv_size());
  int sizes[6], st[6];
  dv_sizes6(c->L, sizes);
  if (!blocks_tile6(sizes, ord, st))
    return fail(c, CLH_ERR_ARG, "ord must be a permutation of 0..5 whose blocks tile the data vector");
  const int ord3[3] = {0, 1, 2};
  std::vector<const Draw*> v{&c->cur};
  c->ord6 = ord;
  int rc = build_batch_ctx(c, c->single, v, 1, ord3, nullptr, nullptr);
  c->ord6 = nullptr;
  if (rc) return rc;
  if ((rc = run_batch(c, c->single, true))) return rc;
  HIPCHK(c, hipMemcpyAsync(dv, c->single.desc[0].dv, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->cached_version = 0;
  return CLH_OK;
}

int clh_xi_pm(clh_ctx* c, double* xip, double* xim) {
  if (!c || !xip || !xim) return CLH_ERR_ARG;
  int rc = ensure_current(c);
  if (rc) return rc;
  const int m = c->L.n_ss() * c->L.ntheta;
  std::copy(c->cache_raw.begin(), c->cache_raw.begin() + m, xip);
  std::copy(c->cache_raw.begin() + m, c->cache_raw.begin() + 2 * m, xim);
  return CLH_OK;
}

int clh_gammat(clh_ctx* c, double* out) {
  if (!c || !out) return CLH_ERR_ARG;
  int rc = ensure_current(c);
  if (rc) return rc;
  const int m = c->L.n_ss() * c->L.ntheta;
  std::copy(c->cache_raw.begin() + 2 * m, c->cache_raw.begin() + 2 * m + c->L.n_gs_dv() * c->L.ntheta, out);
  return CLH_OK;
}

int clh_wtheta(clh_ctx* c, double* out) {
  if (!c || !out) return CLH_ERR_ARG;
  int rc = ensure_current(c);
  if (rc) return rc;
  const int m = c->L.ntheta * (2 * c->L.n_ss() + c->L.n_gs_dv());
  std::copy(c->cache_raw.begin() + m, c->cache_raw.begin() + m + c->L.lens.nbin * c->L.ntheta, out);
  return CLH_OK;
}

int clh_cl_tables(clh_ctx* c, double* ss, double* gs, double* gg) {
  if (!c) return CLH_ERR_ARG;
  int rc = ensure_current(c);
  if (rc) return rc;
  const int nrow = c->L.nrow(), ne = c->L.n_ell;
  const int nss = c->L.n_ss(), ngs = c->L.n_gs(), nl = c->L.lens.nbin;
  for (int q = 0; q < nss + ngs + nl; q++) {
    double* dst = q < nss ? (ss ? ss + (size_t)q * ne : nullptr)
                : q < nss + ngs ? (gs ? gs + (size_t)(q - nss) * ne : nullptr)
                : (gg ? gg + (size_t)(q - nss - ngs) * ne : nullptr);
    if (dst) std::copy(c->cache_cl.begin() + (size_t)q * nrow + CLH_NLOW, c->cache_cl.begin() + (size_t)(q + 1) * nrow, dst);
  }
  return CLH_OK;
}

// C_cl_tomo (cosmo2D.c:4528-4769) as w_gg_tomo(limber = 0) reads it
int clh_cl_nonlimber(clh_ctx* c, double* cl, int* lmax) {
  if (!c || !cl) return CLH_ERR_ARG;
  if (!c->L.nonlimber) return fail(c, CLH_ERR_STATE, "non-Limber w not selected (clh_set_adopt_limber_gg)");
  if (c->L.lens.nbin == 0) return fail(c, CLH_ERR_STATE, "lens sample not set");
  int rc = ensure_current(c);
  if (rc) return rc;
  const int nl = c->L.lens.nbin;
  std::copy(c->cache_nl.begin(), c->cache_nl.begin() + (size_t)nl * CLH_NL_L, cl);
  if (lmax)
    for (int i = 0; i < nl; i++) lmax[i] = (int)c->cache_nl[(size_t)nl * CLH_NL_L + i];
  return CLH_OK;
}

// C_ss_tomo_limber(l, ni, nj, EE = 0) tables (cosmo2D.c:1617): the B mode,
// zero under NLA
int clh_cl_tables_bb(clh_ctx* c, double* bb) {
  if (!c || !bb) return CLH_ERR_ARG;
  int rc = ensure_current(c);
  if (rc) return rc;
  const int nrow = c->L.nrow(), ne = c->L.n_ell;
  for (int q = 0; q < c->L.n_ss(); q++)
    std::copy(c->cache_bb.begin() + (size_t)q * nrow + CLH_NLOW, c->cache_bb.begin() + (size_t)(q + 1) * nrow,
              bb + (size_t)q * ne);
  return CLH_OK;
}

int clh_cl_nointerp(clh_ctx* c, int probe, const double* ell, int nl_, double* out) {
  if (!c || !ell || !out || nl_ <= 0) return CLH_ERR_ARG;
  if (probe < 0 || probe > 6) return fail(c, CLH_ERR_ARG, "probe %d", probe);
  if ((probe == 1 || probe == 2 || probe == 4) && c->L.lens.nbin == 0)
    return fail(c, CLH_ERR_STATE, "lens sample not set");
  if (probe >= 4 && !c->L.sixpt) return fail(c, CLH_ERR_STATE, "CMB lensing probes need init_probes(\"6x2pt\")");
  if (has_nan(ell, nl_)) return fail(c, CLH_ERR_NAN, "ell: NaN");
  dvec ells(ell, ell + nl_);
  const int ord[3] = {0, 1, 2};
  std::vector<const Draw*> v{&c->cur};
  int rc = build_batch_ctx(c, c->scratch, v, 0, ord, &ells, nullptr);
  if (rc) return rc;
  if ((rc = run_batch(c, c->scratch, false))) return rc;
  const int nss = c->L.n_ss(), ngs = c->L.n_gs(), nl = c->L.lens.nbin;
  if (probe >= 4) {  // gk [nl], ks [ns], kk [1] columns of nrow6 = nl_ rows
    const DrawDesc& q = c->scratch.desc[0];
    const int col0 = probe == 4 ? 0 : probe == 5 ? nl : nl + c->L.src.nbin;
    const int ncol = probe == 4 ? nl : probe == 5 ? c->L.src.nbin : 1;
    if (!q.cl6 || q.nrow6 != nl_) return fail(c, CLH_ERR_STATE, "6x2pt columns missing");
    HIPCHK(c, hipMemcpyAsync(out, q.cl6 + (size_t)col0 * nl_, (size_t)ncol * nl_ * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLH_OK;
  }
  if (probe == 3) {  // B mode (EE = 0)
    const double* bb = c->scratch.desc[0].cl_bb;
    if (bb)
      HIPCHK(c, hipMemcpyAsync(out, bb, (size_t)nss * nl_ * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    else
      std::fill(out, out + (size_t)nss * nl_, 0.0);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLH_OK;
  }
  const int col0 = probe == 0 ? 0 : probe == 1 ? nss : nss + ngs;
  const int ncol = probe == 0 ? nss : probe == 1 ? ngs : nl;
  HIPCHK(c, hipMemcpyAsync(out, c->scratch.desc[0].cl + (size_t)col0 * nl_, (size_t)ncol * nl_ * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CLH_OK;
}

// the table sizes, NaNs and required arrays of one uploaded draw
static int check_upload(clh_ctx* c, const clh_draw& s, const Layout& L, int i) {
  if (s.chi_n < 5 || s.G_n < 2 || s.nk < 2 || s.npz < 2) return fail(c, CLH_ERR_ARG, "draw %d: table sizes", i);
  if (has_nan(s.chi_z, s.chi_n) || has_nan(s.chi, s.chi_n) || has_nan(s.G_z, s.G_n) || has_nan(s.G, s.G_n) ||
      has_nan(s.log10k, s.nk) || has_nan(s.pk_z, s.npz) || has_nan(s.lnP, s.nk * s.npz))
    return fail(c, CLH_ERR_NAN, "draw %d: NaN in a table", i);
  if (L.lens.nbin > 0 && !s.b1) return fail(c, CLH_ERR_ARG, "draw %d: b1 missing", i);
  if (L.ia_model == CLH_IA_TATT) {
    if (!s.fpt || s.fpt_n < 2 || !(s.fpt_kmin > 0) || !(s.fpt_kmax > s.fpt_kmin))
      return fail(c, CLH_ERR_ARG, "draw %d: TATT needs the FAST-PT IA tables", i);
    if (has_nan(s.fpt, 10 * s.fpt_n)) return fail(c, CLH_ERR_NAN, "draw %d: NaN in the IA tables", i);
  }
  if (s.fptb) {
    if (s.fptb_n < 2 || !(s.fptb_kmin > 0) || !(s.fptb_kmax > s.fptb_kmin))
      return fail(c, CLH_ERR_ARG, "draw %d: bad FAST-PT bias tables", i);
    if (has_nan(s.fptb, 12 * s.fptb_n) || std::isnan(s.fptb_sigma4))
      return fail(c, CLH_ERR_NAN, "draw %d: NaN in the bias tables", i);
  }
  if (L.nonlimber && L.lens.nbin > 0) {
    if (!s.plin_log10k || !s.plin_z || !s.lnPL || s.plin_nk < 2 || s.plin_nz < 2 || !(s.h > 0))
      return fail(c, CLH_ERR_ARG, "draw %d: non-Limber w needs the linear P(k,z) and h", i);
    if (has_nan(s.plin_log10k, s.plin_nk) || has_nan(s.plin_z, s.plin_nz) || has_nan(s.lnPL, s.plin_nk * s.plin_nz))
      return fail(c, CLH_ERR_NAN, "draw %d: NaN in the linear P(k,z)", i);
  }
  return CLH_OK;
}

static void fill_draw(Draw& d, const clh_draw& s, const Layout& L) {
  d.omega_m = s.omega_m;
  if (s.chi_z && s.chi && s.chi_n > 0) { d.chi_z.assign(s.chi_z, s.chi_z + s.chi_n); d.chi.assign(s.chi, s.chi + s.chi_n); }
  if (s.G_z && s.G && s.G_n > 0) { d.G_z.assign(s.G_z, s.G_z + s.G_n); d.G.assign(s.G, s.G + s.G_n); }
  if (s.log10k && s.pk_z && s.lnP && s.nk > 0 && s.npz > 0) {
    d.lk.assign(s.log10k, s.log10k + s.nk);
    d.pz.assign(s.pk_z, s.pk_z + s.npz);
    d.lnP.assign(s.lnP, s.lnP + (size_t)s.nk * s.npz);
  }
  for (int i = 0; i < L.src.nbin; i++) {
    d.src_dz[i] = s.src_dz ? s.src_dz[i] : 0.0;
    d.shear_m[i] = s.shear_m ? s.shear_m[i] : 0.0;
    d.ia_bin[i] = s.ia_bin ? s.ia_bin[i] : 0.0;
  }
  for (int i = 0; i < L.lens.nbin; i++) {
    d.lens_dz[i] = s.lens_dz ? s.lens_dz[i] : 0.0;
    d.lens_st[i] = s.lens_stretch ? s.lens_stretch[i] : 1.0;
    d.b1[i] = s.b1 ? s.b1[i] : 0.0;
    d.bmag[i] = s.bmag ? s.bmag[i] : 0.0;
    d.pm[i] = s.pm ? s.pm[i] : 0.0;
  }
  d.ia_A = s.ia_A;
  d.ia_eta = s.ia_eta;
  for (int i = 0; i < L.lens.nbin; i++) {
    d.b2[i] = s.b2 ? s.b2[i] : 0.0;
    d.bs2[i] = s.bs2 ? s.bs2[i] : (d.b2[i] == 0.0 ? 0.0 : (-4. / 7.) * (d.b1[i] - 1.0));
    d.b3[i] = s.b3 ? s.b3[i] : 0.0;
    d.bk[i] = s.bk ? s.bk[i] : 0.0;
  }
  if (s.fptb && s.fptb_n >= 2) {
    const double h3 = CLH_COVERH0 * CLH_COVERH0 * CLH_COVERH0;
    d.fptb.resize((size_t)6 * s.fptb_n);
    for (size_t i = 0; i < d.fptb.size(); i++) d.fptb[i] = s.fptb[i] / h3;
    d.fptb_n = s.fptb_n;
    d.fptb_kmin = s.fptb_kmin * CLH_COVERH0;
    d.fptb_kmax = s.fptb_kmax * CLH_COVERH0;
    d.fptb_s4 = s.fptb_sigma4 / h3;
  }
  d.h0 = s.h;
  if (s.plin_log10k && s.plin_z && s.lnPL && s.plin_nk > 0 && s.plin_nz > 0) {
    d.lkL.assign(s.plin_log10k, s.plin_log10k + s.plin_nk);
    d.pzL.assign(s.plin_z, s.plin_z + s.plin_nz);
    d.lnPL.assign(s.lnPL, s.lnPL + (size_t)s.plin_nk * s.plin_nz);
  }
  if (L.ia_model == CLH_IA_TATT) {
    for (int i = 0; i < L.src.nbin; i++) {
      d.ia2_bin[i] = s.ia2_bin ? s.ia2_bin[i] : 0.0;
      d.bta_bin[i] = s.bta_bin ? s.bta_bin[i] : 0.0;
    }
    d.ia2_A = s.ia2_A;
    d.ia2_eta = s.ia2_eta;
    d.bta_A = s.bta_A;
    if (s.fpt && s.fpt_n >= 2) {
      const double h3 = CLH_COVERH0 * CLH_COVERH0 * CLH_COVERH0;
      d.fpt.resize((size_t)10 * s.fpt_n);
      for (size_t i = 0; i < d.fpt.size(); i++) d.fpt[i] = s.fpt[i] / h3;
      d.fpt_n = s.fpt_n;
      d.fpt_kmin = s.fpt_kmin * CLH_COVERH0;
      d.fpt_kmax = s.fpt_kmax * CLH_COVERH0;
    }
  }
}

int clh_batch_upload(clh_ctx* c, int B, const clh_draw* draws) {
  if (!c || B <= 0 || !draws) return CLH_ERR_ARG;
  std::vector<Draw> ds(B);
  std::vector<const Draw*> ptr(B);
  for (int i = 0; i < B; i++) {
    const clh_draw& s = draws[i];
    const int rc = check_upload(c, s, c->L, i);
    if (rc) return rc;
    fill_draw(ds[i], s, c->L);
    ptr[i] = &ds[i];
  }
  const int ord[3] = {0, 1, 2};
  if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
  c->batch.B = 0;
  c->batch_epoch = c->mode_epoch;
  const int rc = build_batch_ctx(c, c->batch, ptr, 1, ord, nullptr, nullptr);
  if (rc) c->batch.B = 0;
  c->rl.clear();
  return rc;
}

}  // extern "C"

// ======================================================= captured layouts
struct clh_layout {
  Layout L;
  uint64_t id = 0;
};
static std::atomic<uint64_t> g_layout_id{1};

extern "C" {

int clh_layout_capture(clh_ctx* c, clh_layout** out) {
  if (!c || !out) return CLH_ERR_ARG;
  *out = nullptr;
  int rc = check_layout(c, c->L);
  if (rc) return rc;
  const bool full = c->legendre_mode == CLH_LEGENDRE_FULL;
  if (!c->L.kern_valid || (full && c->L.kfull.empty())) build_kernels(c->L, full);
  clh_layout* l = new (std::nothrow) clh_layout();
  if (!l) return fail(c, CLH_ERR_ALLOC, "layout allocation");
  l->L = c->L;
  l->id = g_layout_id++;
  *out = l;
  return CLH_OK;
}

void clh_layout_destroy(clh_layout* l) { delete l; }

int clh_layout_datavector_size(const clh_layout* l) { return l ? l->L.dv_size() : 0; }

int clh_batch_upload_layouts(clh_ctx* c, int B, const clh_draw* draws, const clh_layout* const* layouts) {
  if (!c || B <= 0 || !draws || !layouts) return CLH_ERR_ARG;
  std::vector<Draw> ds(B);
  std::vector<const Draw*> ptr(B);
  for (int i = 0; i < B; i++) {
    const clh_draw& s = draws[i];
    if (!layouts[i]) return fail(c, CLH_ERR_ARG, "draw %d: no layout", i);
    const Layout& L = layouts[i]->L;
    const int rc = check_upload(c, s, L, i);
    if (rc) return rc;
    fill_draw(ds[i], s, L);
    ptr[i] = &ds[i];
  }
  int rc = ensure_device(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
  c->batch.B = 0;
  // the context's copies of the layouts, with device copies (kept from the
  // previous multi-layout batch where still valid)
  const bool full = c->legendre_mode == CLH_LEGENDRE_FULL;
  std::map<uint64_t, std::unique_ptr<LayoutCopy>> keep;
  std::vector<const Layout*> lays(B);
  std::vector<const DevLayout*> devs(B);
  for (int i = 0; i < B; i++) {
    const clh_layout* l = layouts[i];
    auto it = keep.find(l->id);
    if (it == keep.end()) {
      std::unique_ptr<LayoutCopy> e;
      auto old = c->rl.find(l->id);
      if (old != c->rl.end() && old->second->epoch == c->mode_epoch && old->second->D.valid) {
        e = std::move(old->second);
      } else {
        e.reset(new (std::nothrow) LayoutCopy());
        if (!e) return fail(c, CLH_ERR_ALLOC, "layout copy");
        e->L = l->L;
        if (!e->L.kern_valid || (full && e->L.kfull.empty())) build_kernels(e->L, full);
        if ((rc = check_layout(c, e->L))) return rc;
        if ((rc = upload_layout(c, e->L, e->D))) return rc;
        e->epoch = c->mode_epoch;
      }
      it = keep.emplace(l->id, std::move(e)).first;
    }
    lays[i] = &it->second->L;
    devs[i] = &it->second->D;
  }
  const int ord[3] = {0, 1, 2};
  rc = build_batch(c, c->batch, ptr, lays, devs, 1, ord, nullptr, nullptr);
  c->rl = std::move(keep);
  if (rc) { c->batch.B = 0; return rc; }
  c->batch.ctx_layout = false;
  c->batch_epoch = c->mode_epoch;
  return CLH_OK;
}

long long clh_batch_datavector_offset(clh_ctx* c, int i) {
  if (!c || i < 0 || i > c->batch.B) return -1;
  return c->batch.dv_off[i];
}

// Checkpoint buffers and descriptor pointers of the next forward: the forward
// stores its Limber integrands (and the lens-efficiency shift tangents) when
// the context asks for it and the batch has one layout.  `changed`: some
// descriptor changed on the host and must be uploaded.
static int prepare_checkpoint(clh_ctx* c, Batch& cb, bool& changed) {
  cb.keep_ckpt = c->vjp_ckpt && cb.uniform && !cb.nrow_custom && !cb.tatt && !cb.nl_any && !cb.oneloop &&
                 !cb.cmb_any && !cb.hdi_any;
  if (cb.keep_ckpt) {
    const size_t nl = cb.lays[0]->lens.nbin, nr = cb.lays[0]->nrow();
    cb.ggk_per = (nl * nr * 3 * CLH_NP_GG + 31) & ~size_t(31);
    cb.sxk_per = ((1 + nl) * nr * CLH_NP_SS + 31) & ~size_t(31);
    cb.dgk_per = (((size_t)cb.lays[0]->src.nbin + nl) * cb.lays[0]->n_a + 31) & ~size_t(31);
    if (cb.ggk_per) HIPCHK(c, cb.d_ggk.ensure(cb.ggk_per * cb.B));
    HIPCHK(c, cb.d_sxk.ensure(cb.sxk_per * cb.B));
    HIPCHK(c, cb.d_dgk.ensure(cb.dgk_per * cb.B));
  }
  for (int i = 0; i < cb.B; i++) {
    DrawDesc& q = cb.desc[i];
    double* gk = (cb.keep_ckpt && cb.ggk_per) ? cb.d_ggk.p + cb.ggk_per * i : nullptr;
    double* sk = cb.keep_ckpt ? cb.d_sxk.p + cb.sxk_per * i : nullptr;
    if (q.ggk != gk || q.sxk != sk) {
      q.ggk = gk;
      q.sxk = sk;
      changed = true;
    }
    if (cb.keep_ckpt && q.dgtab != cb.d_dgk.p + cb.dgk_per * i) {
      q.dgtab = cb.d_dgk.p + cb.dgk_per * i;
      changed = true;
    }
  }
  return CLH_OK;
}

int clh_batch_forward(clh_ctx* c, const int ord[3], double* dv_device) {
  if (!c || !ord || !dv_device) return CLH_ERR_ARG;
  if (c->batch.B <= 0) return fail(c, CLH_ERR_STATE, "no batch uploaded");
  if (batch_stale(c, c->batch))
    return fail(c, CLH_ERR_STATE, "layout changed since the batch upload");
  for (int i = 0; i < c->batch.B; i++)
    if ((i == 0 || c->batch.lays[i] != c->batch.lays[i - 1]) && !ord_ok(*c->batch.lays[i], ord))
      return fail(c, CLH_ERR_ARG, "ord must be a permutation of 0,1,2 whose blocks tile the data vector");
  // integrand checkpoint for a following VJP (batches with one layout)
  bool changed = false;
  int rc = prepare_checkpoint(c, c->batch, changed);
  if (rc) return rc;
  // per-call placement: output pointer and block order go into the descriptors
  for (int i = 0; i < c->batch.B; i++) {
    DrawDesc& q = c->batch.desc[i];
    int start[3];
    dv_starts(*c->batch.lays[i], ord, start);
    double* want = dv_device + c->batch.dv_off[i];
    if (q.dv != want || q.dv_start[0] != start[0] || q.dv_start[1] != start[1] || q.dv_start[2] != start[2]) {
      q.dv = want;
      for (int k = 0; k < 3; k++) q.dv_start[k] = start[k];
      changed = true;
    }
  }
  for (int k = 0; k < 3; k++) c->batch.ord[k] = ord[k];
  if (changed)
    HIPCHK(c, hipMemcpyAsync(c->batch.d_desc.p, c->batch.desc.data(), c->batch.B * sizeof(DrawDesc),
                             hipMemcpyHostToDevice, c->stream));
  if (c->use_graph && !c->timing) {
    // the kernels read placement and inputs from the descriptors in device
    // memory, so one captured graph serves every later call on this batch
    if (!c->gexec || c->gstream != c->stream) {
      if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
      hipGraph_t g = nullptr;
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int rc = run_batch(c, c->batch, true);
      hipError_t e = hipStreamEndCapture(c->stream, &g);
      if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
      if (e != hipSuccess) return fail(c, CLH_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
      e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) { c->gexec = nullptr; return fail(c, CLH_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e)); }
      c->gstream = c->stream;
    }
    HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
    c->batch.state_valid = true;
    c->batch.ckpt_fresh = c->batch.keep_ckpt;
    c->batch.rows64 = c->precision == CLH_PREC_FP64;
    return CLH_OK;
  }
  return run_batch(c, c->batch, true);
}

int clh_set_stream(clh_ctx* c, void* stream) {
  if (!c) return CLH_ERR_ARG;
  int rc = ensure_device(c);
  if (rc) return rc;
  if (c->stream != (hipStream_t)stream) HIPCHK(c, hipStreamSynchronize(c->stream));
  c->stream = (hipStream_t)stream;  // NULL: the HIP default (legacy) stream
  return CLH_OK;
}

int clh_synchronize(clh_ctx* c) {
  if (!c) return CLH_ERR_ARG;
  if (!c->own_stream) return CLH_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->timing) {
    for (int k = 0; k < 5; k++) HIPCHK(c, hipEventElapsedTime(&c->last_ms[k], c->ev[k], c->ev[k + 1]));
    if (c->vjp_timed) {
      // launch order sx_adj, gather, gg_adj, lens limits; kept as gg_adj,
      // launch order: ss adjoint (6-7), gather (7-8), gs adjoint (8-13), gg
      // adjoint (13-9), lens limits (9-10); kept as gg_adj, sx_adj (ss + gs),
      // gather, lens limits
      float t_ss = 0, t_gs = 0;
      HIPCHK(c, hipEventElapsedTime(&t_ss, c->ev[6], c->ev[7]));
      HIPCHK(c, hipEventElapsedTime(&c->last_vjp_ms[2], c->ev[7], c->ev[8]));
      HIPCHK(c, hipEventElapsedTime(&t_gs, c->ev[8], c->ev[13]));
      HIPCHK(c, hipEventElapsedTime(&c->last_vjp_ms[0], c->ev[13], c->ev[9]));
      HIPCHK(c, hipEventElapsedTime(&c->last_vjp_ms[3], c->ev[9], c->ev[10]));
      c->last_vjp_ms[1] = t_ss + t_gs;
      // lens limits from the nodes' motion: k_nodes_tan (the rest of the
      // term is inside the gs / gg adjoint kernels' times)
      if (c->vjp_lim_tan) HIPCHK(c, hipEventElapsedTime(&c->last_vjp_ms[3], c->ev[11], c->ev[12]));
    }
  }
  return CLH_OK;
}

int clh_batch_use_graph(clh_ctx* c, int on) {
  if (!c) return CLH_ERR_ARG;
  c->use_graph = on != 0;
  if (!c->use_graph && c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
  return CLH_OK;
}

int clh_set_vjp_checkpoint(clh_ctx* c, int on) {
  if (!c) return CLH_ERR_ARG;
  if ((on != 0) != (c->vjp_ckpt != 0)) {
    c->vjp_ckpt = on != 0;
    if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
  }
  return CLH_OK;
}

int clh_set_vjp_lens_limits(clh_ctx* c, int mode) {
  if (!c || mode < 0 || mode > 2) return CLH_ERR_ARG;
  if (mode != c->vjp_lim) {
    c->vjp_lim = mode;
    if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
  }
  return CLH_OK;
}

int clh_enable_timing(clh_ctx* c, int on) {
  if (!c) return CLH_ERR_ARG;
  c->timing = on != 0;
  return CLH_OK;
}

int clh_last_kernel_times(clh_ctx* c, const char** names, float* ms, int cap) {
  static const char* nm[9] = {"k_gtab", "k_nodes", "k_limber_gg", "k_limber_sx", "k_legendre",
                              "k_limber_gg_adj", "k_limber_sx_adj", "k_lnp_gather", "lens_limits"};
  static const char* nm_full = "k_legendre_full";
  static const char* nm32[3] = {"k_limber_gg_f32", "k_limber_sx_f32", "k_legendre_mx"};
  if (!c) return CLH_ERR_ARG;
  const bool mixed = c->precision != CLH_PREC_FP64;
  const int n = c->vjp_timed ? 9 : 5;
  for (int k = 0; k < n && k < cap; k++) {
    if (names) {
      names[k] = (mixed && k >= 2 && k < 5) ? nm32[k - 2] : nm[k];
      if (k == 4 && c->legendre_mode == CLH_LEGENDRE_FULL) names[k] = nm_full;
    }
    if (ms) ms[k] = k < 5 ? c->last_ms[k] : c->last_vjp_ms[k - 5];
  }
  return n;
}

int clh_set_tile_rows(clh_ctx* c, const int rows[4]) {
  if (!c || !rows) return CLH_ERR_ARG;
  for (int k = 0; k < 4; k++)
    if (rows[k] < 1 || rows[k] > CLH_TILE_ROWS) return fail(c, CLH_ERR_ARG, "tile rows %d not in 1..%d", rows[k], CLH_TILE_ROWS);
  for (int k = 0; k < 4; k++) c->tile_rows[k] = rows[k];
  if (c->batch.B > 0 && !batch_stale(c, c->batch)) {  // a stale batch is refused by every batch call
    int rc = build_work(c, c->batch, c->tile_rows);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
  }
  return CLH_OK;
}

int clh_get_tile_rows(clh_ctx* c, int rows[4]) {
  if (!c || !rows) return CLH_ERR_ARG;
  for (int k = 0; k < 4; k++) rows[k] = c->tile_rows[k];
  return CLH_OK;
}

int clh_autotune(clh_ctx* c, int rows_out[4], int lds_out[2]) {
  if (!c) return CLH_ERR_ARG;
  Batch& bt = c->batch;
  if (bt.B <= 0) return fail(c, CLH_ERR_STATE, "no batch uploaded");
  if (batch_stale(c, bt))
    return fail(c, CLH_ERR_STATE, "layout changed since the batch upload");
  HIPCHK(c, hipSetDevice(c->device));
  // the forward writes into the batch's own output buffer while tuning
  HIPCHK(c, bt.d_dv.ensure((size_t)bt.dv_off[bt.B]));
  bool ck_changed = false;
  int rc0 = prepare_checkpoint(c, bt, ck_changed);  // tune what the next forward will run
  if (rc0) return rc0;
  for (int i = 0; i < bt.B; i++) bt.desc[i].dv = bt.d_dv.p + bt.dv_off[i];
  HIPCHK(c, hipMemcpyAsync(bt.d_desc.p, bt.desc.data(), bt.B * sizeof(DrawDesc), hipMemcpyHostToDevice, c->stream));
  const bool adj = bt.uniform && bt.nknz > 0 && !(bt.tatt && bt.lays[0]->src.nbin > 10) && !bt.nl_any &&
                   !bt.hdi_any && !bt.nrow_custom;  // (no adjoint tiles to tune for these)
  // unit cotangent, ln P and lens-shift gradients of the adjoint stages (the
  // ln P runs and the lens-limit term are part of what is timed)
  DBuf<double> dvbar, glnp, glens;
  if (adj) {
    hipError_t e = dvbar.ensure((size_t)bt.dv_off[bt.B]);
    if (e == hipSuccess) e = clh_launch_fill(dvbar.p, (size_t)bt.dv_off[bt.B], 1.0, c->stream);
    if (e == hipSuccess) e = glnp.ensure((size_t)bt.nknz * bt.B);
    if (e == hipSuccess) e = glens.ensure((size_t)CLH_MAXB * bt.B);
    if (e != hipSuccess) {
      dvbar.release();
      glnp.release();
      glens.release();
      return fail(c, CLH_ERR_HIP, "autotune: %s", hipGetErrorString(e));
    }
  }
  const bool was_timing = c->timing;
  const int lds_was[2] = {c->gg_lds[0], c->gg_lds[1]};
  c->timing = true;
  static const int cand[3] = {16, 32, 64};
  float best[4] = {1e30f, 1e30f, 1e30f, 1e30f};
  int pick[4], pick_lds[2] = {c->gg_lds[0], c->gg_lds[1]};
  for (int k = 0; k < 4; k++) pick[k] = c->tile_rows[k];
  const int ord[3] = {0, 1, 2};
  const int nst = bt.gg_stage > 0 ? 2 : 1;  // LDS staging of the gg tables is an option when they fit
  int rc = CLH_OK;
  for (int k = 0; k < 3 && rc == CLH_OK; k++) {
    const int t[4] = {cand[k], cand[k], cand[k], cand[k]};
    if ((rc = build_work(c, bt, t))) break;
    for (int st = 0; st < nst && rc == CLH_OK; st++) {
      c->gg_lds[0] = c->gg_lds[1] = st;
      for (int rep = 0; rep < 3 && rc == CLH_OK; rep++) {
        if ((rc = run_batch(c, bt, false))) break;
        if (adj && (rc = clh_batch_vjp(c, ord, dvbar.p, glnp.p, nullptr, glens.p, nullptr, nullptr, nullptr)))
          break;
        if ((rc = clh_synchronize(c))) break;
        const float ms[4] = {c->last_ms[3], c->last_ms[2], adj ? c->last_vjp_ms[1] : 1e30f,
                             adj ? c->last_vjp_ms[0] : 1e30f};
        for (int s2 = 0; s2 < 4; s2++) {
          if (ms[s2] < best[s2]) {
            best[s2] = ms[s2];
            pick[s2] = cand[k];
            if (s2 == 1) pick_lds[0] = st;
            if (s2 == 3) pick_lds[1] = st;
          }
        }
      }
    }
  }
  // forward + VJP (checkpointing forward, tuned for pairs): the per-stage
  // minima need not make the fastest whole pass (a stage's tile and LDS
  // choice changes what the next stage finds in cache), so the forward gg
  // tile and both LDS choices are also checked jointly, timing whole passes
  if (adj && bt.keep_ckpt && rc == CLH_OK) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    float best_all = 1e30f;
    int jt[4] = {pick[0], pick[1], pick[2], pick[3]}, jl[2] = {pick_lds[0], pick_lds[1]};
    for (int g = 0; g < 2 && he == hipSuccess && rc == CLH_OK; g++) {
      const int t[4] = {pick[0], g == 0 ? pick[1] : (pick[1] == 64 ? 32 : 64), pick[2], pick[3]};
      if ((rc = build_work(c, bt, t))) break;
      for (int l0 = 0; l0 < nst && rc == CLH_OK; l0++)
        for (int l1 = 0; l1 < nst && rc == CLH_OK; l1++) {
          c->gg_lds[0] = l0;
          c->gg_lds[1] = l1;
          for (int rep = 0; rep < 3 && rc == CLH_OK; rep++) {
            if ((he = hipEventRecord(e0, c->stream)) != hipSuccess) break;
            if ((rc = run_batch(c, bt, false))) break;
            if ((rc = clh_batch_vjp(c, ord, dvbar.p, glnp.p, nullptr, glens.p, nullptr, nullptr, nullptr))) break;
            if ((he = hipEventRecord(e1, c->stream)) != hipSuccess) break;
            if ((he = hipEventSynchronize(e1)) != hipSuccess) break;
            float ms = 0.0f;
            if ((he = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) break;
            if (ms < best_all) {
              best_all = ms;
              for (int k = 0; k < 4; k++) jt[k] = t[k];
              jl[0] = l0;
              jl[1] = l1;
            }
          }
          if (he != hipSuccess) break;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc == CLH_OK && he != hipSuccess) rc = fail(c, CLH_ERR_HIP, "autotune: %s", hipGetErrorString(he));
    if (rc == CLH_OK) {
      for (int k = 0; k < 4; k++) pick[k] = jt[k];
      pick_lds[0] = jl[0];
      pick_lds[1] = jl[1];
    }
  }
  c->timing = was_timing;
  (void)hipStreamSynchronize(c->stream);
  dvbar.release();  // local buffers (hipFree)
  glnp.release();
  glens.release();
  if (rc) {
    c->gg_lds[0] = lds_was[0];
    c->gg_lds[1] = lds_was[1];
    (void)build_work(c, bt, c->tile_rows);
    return rc;
  }
  for (int k = 0; k < 4; k++) c->tile_rows[k] = pick[k];
  c->gg_lds[0] = pick_lds[0];
  c->gg_lds[1] = pick_lds[1];
  if ((rc = build_work(c, bt, c->tile_rows))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
  if (rows_out)
    for (int k = 0; k < 4; k++) rows_out[k] = pick[k];
  if (lds_out)
    for (int k = 0; k < 2; k++) lds_out[k] = pick_lds[k];
  return CLH_OK;
}

}  // extern "C"

// ============================================================ batch VJP
extern "C" {

int clh_batch_set_params(clh_ctx* c, const double* lnP, const double* src_dz, const double* lens_dz,
                         const double* b1, const double* ia_A, const double* shear_m) {
  if (!c) return CLH_ERR_ARG;
  Batch& bt = c->batch;
  if (bt.B <= 0) return fail(c, CLH_ERR_STATE, "no batch uploaded");
  if (batch_stale(c, bt)) return fail(c, CLH_ERR_STATE, "layout changed since the batch upload");
  if (lnP && bt.nknz <= 0) return fail(c, CLH_ERR_ARG, "draws of the batch have different P(k,z) shapes");
  if (!bt.uniform) return fail(c, CLH_ERR_ARG, "per-draw parameter arrays need a batch with one layout");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, clh_launch_set_params(bt.d_desc.p, bt.d_par.p, bt.B, bt.nknz, lnP, src_dz, lens_dz, b1, ia_A,
                                  shear_m, c->stream));
  bt.state_valid = false;
  return CLH_OK;
}

int clh_batch_vjp(clh_ctx* c, const int ord[3], const double* dv_bar, double* g_lnP, double* g_src_dz,
                  double* g_lens_dz, double* g_b1, double* g_ia_A, double* g_m) {
  if (!c || !dv_bar || !ord) return CLH_ERR_ARG;
  Batch& bt = c->batch;
  if (bt.B <= 0) return fail(c, CLH_ERR_STATE, "no batch uploaded");
  if (bt.nknz <= 0) return fail(c, CLH_ERR_ARG, "draws of the batch have different P(k,z) shapes");
  if (batch_stale(c, bt))
    return fail(c, CLH_ERR_STATE, "layout changed since the batch upload");
  if (!bt.uniform) return fail(c, CLH_ERR_ARG, "VJP needs a batch with one layout for every draw");
  if (bt.tatt && bt.lays[0]->src.nbin > 10)
    return fail(c, CLH_ERR_ARG, "VJP under TATT takes at most 10 source bins (%d)", bt.lays[0]->src.nbin);
  if (bt.nl_any && (bt.oneloop || bt.tatt || bt.lays[0]->sixpt || c->vjp_lim == 1))
    return fail(c, CLH_ERR_ARG,
                "VJP of the non-Limber w(theta) together with one-loop bias, TATT or 6x2pt, or with the "
                "finite-difference lens limits, is not implemented");
  if (bt.nrow_custom) return fail(c, CLH_ERR_ARG, "VJP of a custom-ell batch (no data vector)");
  if (bt.hdi_any && (bt.tatt || bt.oneloop))
    return fail(c, CLH_ERR_ARG, "VJP at high_def_integration != 0 with TATT or one-loop bias is not implemented");
  const Layout& L = *bt.lays[0];
  if (!ord_ok(L, ord))
    return fail(c, CLH_ERR_ARG, "ord must be a permutation of 0,1,2 whose blocks tile the data vector");
  const int B = bt.B, ns = L.src.nbin, nl = L.lens.nbin, nrow = L.nrow();
  const int ncol = L.n_ss() + L.n_gs() + nl;
  HIPCHK(c, hipSetDevice(c->device));
  // per-draw adjoint buffers
  auto al = [](size_t n) { return (n + 31) & ~size_t(31); };
  const size_t n_cbar = al((size_t)ncol * nrow), n_raw = al(bt.dv_size), n_abar = al((size_t)ns * CLH_NP_SS),
               n_bbar = al((size_t)nl * ns * CLH_NP_SS),
               n_wbar = al((size_t)nl * 2 * CLH_NP_SS + (size_t)nl * 2 * CLH_NP_GG),
               n_dg = al((size_t)(ns + nl) * L.n_a),
               n_tan = al((size_t)nl * (6 + ns) * CLH_NP_SS + (size_t)nl * 8 * CLH_NP_GG),
               n_cbb = bt.tatt ? al((size_t)L.n_ss() * nrow) : 0,
               n_tb = bt.tatt ? al((size_t)(1 + nl) * 2 * ns * CLH_NP_SS) : 0;
  const bool six = L.sixpt;
  const int nrow6 = six ? bt.desc[0].nrow6 : 0;
  const size_t n_c6 = six ? al((size_t)(nl + ns + 1) * nrow6) : 0;
  // non-Limber w: the tangent area, the two views' row cotangents and (with
  // the node tangents) the views' own node tangents
  const bool nlv = bt.nl_any;
  const size_t n_nlt = nlv ? clh_nl_tan(nl).total : 0, n_nlc = nlv ? al((size_t)2 * ncol * (CLH_NL_L - 1)) : 0;
  const size_t per0 = n_cbar + n_raw + n_abar + n_bbar + n_wbar + n_dg + n_tan + n_cbb + n_tb + n_c6;
  const size_t per = per0 + n_nlt + n_nlc + (nlv ? 2 * n_tan : 0);
  // the lens-limit part of the lens-shift gradient from the nodes' motion
  // (k_nodes_tan and the LIM adjoint instantiations), unless the
  // finite-difference views are asked for
  // (one-loop, TATT and 6x2pt batches take the term from the
  // finite-difference views: the node-motion tangent covers neither the
  // one-loop and TATT terms nor the gk integral's nodes)
  // (and so do batches at high_def_integration != 0, whose nodes k_limber_hdi
  // places itself)
  const bool hdi = bt.hdi_any;
  const bool lim_tan = g_lens_dz && nl > 0 && c->vjp_lim == 0 && !bt.oneloop && !bt.tatt && !six && !hdi;
  const bool lim_fd =
      g_lens_dz && nl > 0 && (c->vjp_lim == 1 || (c->vjp_lim == 0 && (bt.oneloop || bt.tatt || six || hdi)));
  const int ol = bt.oneloop ? 1 : 0;
  HIPCHK(c, bt.d_vjp.ensure(per * B));
  // per-node ln P cotangent runs of the ss / gs nodes and the gather work
  // list (k rows per item: the LDS tile of k rows x nz columns plus the
  // nodes' (j, dy) fits 64 KB); the gg nodes fold theirs in LDS
  const int nk = bt.desc[0].nk, nzp = bt.desc[0].nzp;
  const size_t n_run = al((size_t)CLH_NP_SS * nk);  // the ss nodes' rows
  if (g_lnP) {
    const size_t nq = (size_t)CLH_NP_SS;
    auto lds = [&](int ic) { return sizeof(double) * ((size_t)ic * nzp + nq) + sizeof(int) * nq; };
    int ic = 16;
    while (ic > 2 && lds(ic) > 65536) ic /= 2;
    if (lds(ic) > 65536)
      return fail(c, CLH_ERR_ARG, "ln P gradient: %d z values of P(k,z) with %d lens bins exceed the LDS tile",
                  nzp, nl);
    // the per-node runs need rows in non-decreasing ell (as the layout builds them)
    for (const dvec* e : {&L.ell_ss, &L.ell_gx})
      for (size_t r = 1; r < e->size(); r++)
        if ((*e)[r] < (*e)[r - 1]) return fail(c, CLH_ERR_STATE, "ell rows not sorted");
    HIPCHK(c, bt.d_run.ensure(n_run * B));
    if (bt.lnp_ic != ic || bt.n_wlg != B * ((nk + ic - 1) / ic)) {
      std::vector<int2> wlg;
      for (int i = 0; i < B; i++)
        for (int i0 = 0; i0 < nk; i0 += ic) wlg.push_back(make_int2(i, i0));
      HIPCHK(c, bt.d_wlg.ensure(wlg.size()));
      HIPCHK(c, hipMemcpyAsync(bt.d_wlg.p, wlg.data(), wlg.size() * sizeof(int2), hipMemcpyHostToDevice, c->stream));
      bt.n_wlg = (int)wlg.size();
      bt.lnp_ic = ic;
    }
  }
  // scratch for gradients the caller does not want
  const size_t gper = al(bt.nknz) + 4 * al(CLH_MAXB);
  HIPCHK(c, bt.d_gscr.ensure(gper * B));
  {  // block starts of dv_bar (generic_interface.hpp:770-791)
    int start[3];
    dv_starts(L, ord, start);
    for (int i = 0; i < B; i++)
      for (int k = 0; k < 3; k++) bt.desc[i].dv_start[k] = start[k];
  }
  for (int i = 0; i < B; i++) {
    DrawDesc& q = bt.desc[i];
    double* v = bt.d_vjp.p + per * i;
    q.cbar = v; v += n_cbar;
    q.raw = v; v += n_raw;
    q.abar = v; v += n_abar;
    q.bbar = v; v += n_bbar;
    q.wbar = v; v += n_wbar;
    double* g = bt.d_gscr.p + gper * i;
    q.dvbar = dv_bar + (size_t)bt.dv_size * i;
    q.g_lnP = g_lnP ? g_lnP + (size_t)bt.nknz * i : g;
    q.run = g_lnP ? bt.d_run.p + n_run * i : nullptr;
    // the forward's checkpoints, if they still match the parameters
    const bool ck = bt.state_valid && bt.ckpt_fresh;
    q.ggk = (ck && bt.ggk_per) ? bt.d_ggk.p + bt.ggk_per * i : nullptr;
    q.sxk = ck ? bt.d_sxk.p + bt.sxk_per * i : nullptr;
    q.dgtab = ck ? bt.d_dgk.p + bt.dgk_per * i : v;
    q.ntan = lim_tan ? v + n_dg : nullptr;
    q.cbar_bb = bt.tatt ? v + n_dg + n_tan : nullptr;
    q.tbar = bt.tatt ? v + n_dg + n_tan + n_cbb : nullptr;
    q.cbar6 = six ? v + n_dg + n_tan + n_cbb + n_tb : nullptr;
    double* vn = bt.d_vjp.p + per * i + per0;  // (non-Limber buffers after the others)
    q.nlt = nlv ? vn : nullptr;
    q.nlc_bar = nlv ? vn + n_nlt : nullptr;
    q.g_src = g_src_dz ? g_src_dz + (size_t)ns * i : g + al(bt.nknz);
    q.g_lens = g_lens_dz ? g_lens_dz + (size_t)nl * i : g + al(bt.nknz) + al(CLH_MAXB);
    q.g_b1 = g_b1 ? g_b1 + (size_t)nl * i : g + al(bt.nknz) + 2 * al(CLH_MAXB);
    q.g_A = g_ia_A ? g_ia_A + i : g + al(bt.nknz) + 3 * al(CLH_MAXB);
    q.g_m = g_m ? g_m + (size_t)ns * i : g + al(bt.nknz) + 3 * al(CLH_MAXB) + 1;
  }
  hipStream_t s = c->stream;
  HIPCHK(c, hipMemcpyAsync(bt.d_desc.p, bt.desc.data(), B * sizeof(DrawDesc), hipMemcpyHostToDevice, s));
  if (nlv) {
    // the non-Limber views (P_nl, then P_lin, per draw) with the VJP's
    // buffers: their gg rows at ell = 1..149 have their own cotangents; their
    // gg nodes sit where the draw's do, so the node cotangents, b1, lens
    // shift and (P_nl view) ln P go to the draw's accumulators
    if (bt.desc_nl.size() != (size_t)2 * B) return fail(c, CLH_ERR_STATE, "non-Limber views missing");
    bt.desc_nlv.assign(bt.desc_nl.begin(), bt.desc_nl.end());
    for (int i = 0; i < B; i++)
      for (int lin = 0; lin < 2; lin++) {
        const DrawDesc& q = bt.desc[i];
        DrawDesc& u = bt.desc_nlv[2 * i + lin];
        u.cbar = q.nlc_bar + (size_t)lin * ncol * (CLH_NL_L - 1);
        u.wbar = q.wbar;
        u.abar = q.abar;
        u.bbar = q.bbar;
        u.run = lin ? nullptr : q.run;  // (the P_lin rows carry no ln P cotangent)
        u.g_lnP = q.g_lnP;
        u.g_src = q.g_src; u.g_lens = q.g_lens; u.g_b1 = q.g_b1; u.g_A = q.g_A; u.g_m = q.g_m;
        u.dgtab = q.dgtab;
        u.ntan = lim_tan ? q.nlc_bar + n_nlc + (size_t)lin * n_tan : nullptr;  // (their own z weights)
        u.dvbar = nullptr; u.raw = nullptr; u.ggk = nullptr; u.sxk = nullptr;
        u.cbar_bb = nullptr; u.tbar = nullptr; u.cbar6 = nullptr; u.nlt = nullptr; u.nlc_bar = nullptr;
      }
    HIPCHK(c, bt.d_desc_nlv.ensure(2 * B));
    HIPCHK(c, hipMemcpyAsync(bt.d_desc_nlv.p, bt.desc_nlv.data(), 2 * B * sizeof(DrawDesc), hipMemcpyHostToDevice,
                             s));
  }
  // zero the accumulators (fill kernels, not memsets: safe under graph capture)
  HIPCHK(c, clh_launch_fill(bt.d_vjp.p, per * B, 0.0, s));
  HIPCHK(c, clh_launch_fill(bt.d_gscr.p, gper * B, 0.0, s));
  if (g_lnP) HIPCHK(c, clh_launch_fill(bt.d_run.p, n_run * B, 0.0, s));  // g_lnP: every cell stored by the gather
  if (g_src_dz) HIPCHK(c, clh_launch_fill(g_src_dz, (size_t)ns * B, 0.0, s));
  if (g_lens_dz && nl) HIPCHK(c, clh_launch_fill(g_lens_dz, (size_t)nl * B, 0.0, s));
  if (g_b1 && nl) HIPCHK(c, clh_launch_fill(g_b1, (size_t)nl * B, 0.0, s));
  if (g_ia_A) HIPCHK(c, clh_launch_fill(g_ia_A, (size_t)B, 0.0, s));
  // forward recompute: lens-efficiency tables and their shift tangents, node
  // weights, and the ss / gs C_ell rows (the shear-calibration gradient needs
  // them; the gg rows carry no calibration and are not recomputed)
  const bool ck_all = bt.state_valid && bt.ckpt_fresh;
  if (!ck_all)  // else the checkpointing forward left tables and shift tangents
    HIPCHK(c, clh_launch_gtab_tangent(bt.d_desc.p, bt.d_wg.p, bt.n_wg, bt.state_valid ? 0 : 1, s));
  if (!bt.state_valid) {  // else the last forward / VJP left them for these parameters
    bt.ckpt_fresh = false;  // a checkpoint predates these parameters
    bt.rows64 = false;      // the gg rows predate them too
    HIPCHK(c, clh_launch_nodes(bt.d_desc.p, bt.d_wn.p, bt.n_wn, s));
    HIPCHK(c, clh_launch_limber_sx(bt.d_desc.p, bt.d_wsx.p, bt.n_wsx_std, 0, 0, s, ol));
    if (bt.n_wsx > bt.n_wsx_std)  // the TATT draws' tiles (EE and BB rows)
      HIPCHK(c, clh_launch_limber_sx_tatt(bt.d_desc.p, bt.d_wsx.p + bt.n_wsx_std, bt.n_wsx - bt.n_wsx_std, s));
    if (six)  // the C_gk / C_ks / C_kk columns (d/dm reads C_ks; the views' base rows)
      HIPCHK(c, clh_launch_cmb(bt.d_desc.p, bt.d_wcmb.p, bt.n_wcmb, nullptr, 0, s));
    if (hdi)  // the n-point GL rows of every column
      HIPCHK(c, clh_launch_limber_hdi(bt.d_desc.p, bt.d_whd.p, bt.n_whd, s));
    if (nlv) {  // the non-Limber views' nodes and rows, the FFTLog stages and the LMAX switch
      HIPCHK(c, clh_launch_nodes(bt.d_desc_nl.p, bt.d_wn_nl.p, bt.n_wn_nl, s));
      HIPCHK(c, clh_launch_limber_gg(bt.d_desc_nl.p, bt.d_wgg_nl.p, bt.n_wgg_nl, 0, 0, 0, s, 0));
      HIPCHK(c, clh_launch_nl_forward(bt.d_desc.p, bt.d_wnl_prep.p, bt.n_wnl_prep, bt.d_wnl_fwd.p, bt.n_wnl_fwd,
                                      bt.d_wnl_gl.p, bt.n_wnl_gl, bt.d_wnl_chk.p, bt.n_wnl_chk, bt.d_wnl_cl.p,
                                      bt.n_wnl_cl, bt.d_wnl_idft.p, bt.n_wnl_idft, s));
    }
  }
  // reverse sweep
  HIPCHK(c, clh_launch_leg_adjoint(bt.d_desc.p, bt.d_wa4.p, bt.n_wa, s));
  if (six) HIPCHK(c, clh_launch_cmb_place_adj(bt.d_desc.p, bt.d_wcmbp.p, bt.n_wcmbp, s));
  // non-Limber w: its cotangent onto the gg rows, the views' rows and the FFTLog sums
  if (nlv) HIPCHK(c, clh_launch_nl_cl_adj(bt.d_desc.p, bt.d_wnl_cl.p, bt.n_wnl_cl, s));
  HIPCHK(c, clh_launch_m_grad(bt.d_desc.p, B, s));
  if (lim_tan) {
    if (c->timing) HIPCHK(c, hipEventRecord(c->ev[11], s));
    HIPCHK(c, clh_launch_nodes_tan(bt.d_desc.p, bt.d_wn.p, bt.n_wn, s));
    if (nlv) HIPCHK(c, clh_launch_nodes_tan(bt.d_desc_nlv.p, bt.d_wn_nl.p, bt.n_wn_nl, s));
    if (c->timing) HIPCHK(c, hipEventRecord(c->ev[12], s));
  }
  c->vjp_lim_tan = lim_tan;
  // ss adjoint, the gather that stores its part of g_lnP (every cell), then
  // the gs and gg adjoints adding theirs from their LDS bands on top (events
  // 6, 7, 8, 13, 9 in this order)
  if (c->timing) HIPCHK(c, hipEventRecord(c->ev[6], s));
  if (bt.tatt)  // TATT: the EE and BB cotangent tiles
    HIPCHK(c, clh_launch_tatt_adj(bt.d_desc.p, bt.d_wl.p, bt.n_wl_ss, 2 * bt.adj_lds_ss, ns, 0, s));
  else
    HIPCHK(c, clh_launch_limber_sx_adj(bt.d_desc.p, bt.d_wl.p, bt.n_wl_ss, bt.adj_lds_ss, ns, ck_all ? 1 : 0, 0, 0,
                                       0, s));
  if (c->timing) HIPCHK(c, hipEventRecord(c->ev[7], s));
  if (g_lnP) {
    if (hdi)  // no ss nodes (k_limber_hdi places its own): g_lnP starts at zero
      HIPCHK(c, clh_launch_fill(g_lnP, (size_t)bt.nknz * B, 0.0, s));
    else
      HIPCHK(c, clh_launch_lnp_gather(bt.d_desc.p, bt.d_wlg.p, bt.n_wlg, bt.lnp_ic, nzp, nl, s));
  }
  if (c->timing) HIPCHK(c, hipEventRecord(c->ev[8], s));
  // (with the lens-limit term the gs and gg adjoints evaluate their
  // integrands again: the term needs the P(k,z) cells and the a_1 walks)
  if (bt.tatt)
    HIPCHK(c, clh_launch_tatt_adj(bt.d_desc.p, bt.d_wl.p + bt.n_wl_ss, bt.n_wl - bt.n_wl_ss, bt.adj_lds_gs, ns, 1, s));
  else
    HIPCHK(c, clh_launch_limber_sx_adj(bt.d_desc.p, bt.d_wl.p + bt.n_wl_ss, bt.n_wl - bt.n_wl_ss, bt.adj_lds_gs, ns,
                                       (ck_all && !lim_tan) ? 1 : 0, lim_tan ? 1 : 0, 1, ol, s));
  if (c->timing) HIPCHK(c, hipEventRecord(c->ev[13], s));
  const bool use_ggk = bt.state_valid && bt.ckpt_fresh && bt.ggk_per && !lim_tan;
  HIPCHK(c, clh_launch_limber_gg_adj(bt.d_desc.p, bt.d_wga.p, bt.n_wga,
                                     (c->gg_lds[1] && !use_ggk) ? bt.gg_stage : 0, use_ggk ? 1 : 0,
                                     lim_tan ? 1 : 0, c->precision != CLH_PREC_FP64 ? 1 : 0, ol, s));
  // 6x2pt: the gk / ks / kk Limber adjoint (its ln P runs added after the gather)
  if (six) HIPCHK(c, clh_launch_limber_cmb_adj(bt.d_desc.p, bt.d_wcmb.p, bt.n_wcmb, s));
  // high_def_integration != 0: the adjoint of every column's n-point rows (the
  // node-based adjoints above have no items; the gather stored zeros)
  if (hdi) HIPCHK(c, clh_launch_limber_hdi_adj(bt.d_desc.p, bt.d_whd.p, bt.n_whd, s));
  if (nlv) {  // non-Limber w: the views' gg rows (FP64), then the FFTLog sums' tangents
    HIPCHK(c, clh_launch_limber_gg_adj(bt.d_desc_nlv.p, bt.d_wgg_nl.p, bt.n_wgg_nl, 0, 0, lim_tan ? 1 : 0, 0, 0, s));
    HIPCHK(c, clh_launch_nl_tangent(bt.d_desc.p, bt.d_wnl_prep.p, bt.n_wnl_prep, bt.d_wnl_fwd4.p, bt.n_wnl_fwd4,
                                    bt.d_wnl_idft.p, bt.n_wnl_idft, s));
  }
  if (c->timing) HIPCHK(c, hipEventRecord(c->ev[9], s));
  c->vjp_timed = c->timing;
  if (lim_fd) {
    // The lens integration limits move with |dz| (redshift_spline.c:50-79),
    // and the GL nodes of the bin's gs and gg integrals with them; the
    // adjoint kernels above hold the nodes fixed.  That part comes from the
    // gs / gg rows of two views whose limits take |dz| +- eps, against the
    // FP64 rows at these parameters, contracted with their cotangent
    // (k_lim_grad).  eps: the rounding of the rows' differences costs about
    // 1e-3 of the term (itself below 1% of the lens-shift gradient), and about
    // one row in 1e4 has a node crossing a table knot within eps.
    constexpr double eps = 1e-8;
    const int gg_st = c->gg_lds[0] ? bt.gg_stage : 0;
    if (!bt.rows64) {  // base rows: gs (unless recomputed above in FP64) and gg
      if (bt.state_valid) HIPCHK(c, clh_launch_limber_sx(bt.d_desc.p, bt.d_wsx_gs.p, bt.n_wsx_gs, 0, 0, s, ol));
      HIPCHK(c, clh_launch_limber_gg(bt.d_desc.p, bt.d_wgg.p, bt.n_wgg, 0, gg_st, 0, s, ol));
      bt.rows64 = true;
    }
    // (TATT: the views' own B_b, D_b node vectors for their gs tiles)
    const size_t nsz = al(node_size(ns, nl)), csz = al((size_t)ncol * std::max(nrow, bt.desc[0].nrow)),
                 tsz = bt.tatt ? al((size_t)(1 + nl) * (2 * ns + 1) * CLH_NP_SS) : 0,
                 vsz = nsz + csz + tsz + n_c6;  // (6x2pt: the views' own cl6 columns)
    HIPCHK(c, bt.d_lim.ensure(vsz * 2 * B));
    HIPCHK(c, bt.d_desc_lim.ensure(2 * B));
    bt.desc_lim.resize(2 * B);
    for (int i = 0; i < 2 * B; i++) {
      DrawDesc& q = bt.desc_lim[i];
      q = bt.desc[i % B];
      q.lim_eps = i < B ? eps : -eps;
      q.nodes = bt.d_lim.p + vsz * i;
      q.cl = q.nodes + nsz;
      q.dv = nullptr; q.cl_bb = nullptr; q.nlw = nullptr;
      q.tatt = bt.tatt ? q.cl + csz : nullptr;
      q.cl6 = six ? q.cl + csz + tsz : nullptr;
      q.ggk = nullptr; q.sxk = nullptr;
    }
    HIPCHK(c, hipMemcpyAsync(bt.d_desc_lim.p, bt.desc_lim.data(), 2 * B * sizeof(DrawDesc), hipMemcpyHostToDevice,
                             s));
    for (int h = 0; h < 2; h++) {
      const DrawDesc* V = bt.d_desc_lim.p + h * B;
      HIPCHK(c, clh_launch_nodes(V, bt.d_wn.p, bt.n_wn, s));
      HIPCHK(c, clh_launch_limber_gg(V, bt.d_wgg.p, bt.n_wgg, 0, gg_st, 0, s, ol));
      HIPCHK(c, clh_launch_limber_sx(V, bt.d_wsx_gs.p, bt.n_wsx_gs, 0, 0, s, ol));
      if (bt.n_wsx_gs_t) HIPCHK(c, clh_launch_limber_sx_tatt(V, bt.d_wsx_gs_t.p, bt.n_wsx_gs_t, s));
      if (six) HIPCHK(c, clh_launch_cmb(V, bt.d_wcmb.p, bt.n_wcmb, nullptr, 0, s));
      if (hdi) HIPCHK(c, clh_launch_limber_hdi(V, bt.d_whd.p, bt.n_whd, s));
    }
    HIPCHK(c, clh_launch_lim_grad(bt.d_desc.p, bt.d_desc_lim.p, bt.d_desc_lim.p + B, bt.d_wlim.p, bt.n_wlim, s));
  }
  if (c->timing) HIPCHK(c, hipEventRecord(c->ev[10], s));
  bt.state_valid = true;
  HIPCHK(c, clh_launch_nodes_adj(bt.d_desc.p, bt.d_wn.p, bt.n_wn, s));
  return CLH_OK;
}

}  // extern "C"

// ================================================= reference-named entries
// Scalar accessors with the signatures of cosmolike/cosmo2D.h:30-96, served
// from a process-wide context.  Errors return NaN and record the message
// (the reference logs and exits instead).
#include <limits>
#include <mutex>

static clh_ctx* g_default = nullptr;
static std::mutex g_default_mu;

static double nanv() { return std::numeric_limits<double>::quiet_NaN(); }

static int n_shear_idx(int ns, int ni, int nj) {  // N_shear: redshift_spline.c:277-303
  const int i = std::min(ni, nj), j = std::max(ni, nj);
  int q = 0;
  for (int k = 0; k < i; k++) q += ns - k;
  return q + (j - i);
}

// interpol1d over one table column (basics.c:457-480)
static double table_lookup(const clh_ctx* c, int col, bool ss, double l, bool bb = false, bool cmb = false) {
  const Layout& L = c->L;
  const double lim0 = std::log(std::fmax(ss ? CLH_LMIN_TAB - 1. : (double)CLH_LMIN_TAB, 1.0));
  const double lim1 = std::log((double)(L.lmax + 1));
  const double dx = (lim1 - lim0) / ((double)L.n_ell - 1.);
  const double* f = cmb ? c->cache_cl6.data() + (size_t)col * c->cache_nrow6 + CLH_NLOW
                        : (bb ? c->cache_bb.data() : c->cache_cl.data()) + (size_t)col * L.nrow() + CLH_NLOW;
  const double x = std::log(l);
  if (x < lim0) return f[0];
  const double r = (x - lim0) / dx;
  const int i = (int)std::floor(r);
  if (i + 1 >= L.n_ell) return f[L.n_ell - 1];
  return (r - i) * (f[i + 1] - f[i]) + f[i];
}

// w(theta) of the current draw with the Limber (limber = 1) or non-Limber
// (0) C_gg below ell 150, [nl][ntheta]: the cached data vector when the
// context's adopt_limber_gg selects that path, else one evaluation of the
// other variant on a copy of the layout (cached by the context version)
static const double* w_block(clh_ctx* c, int limber) {
  if (limber == (c->L.nonlimber ? 0 : 1)) {
    if (ensure_current(c)) return nullptr;
    return c->cache_raw.data() + (size_t)(2 * c->L.n_ss() + c->L.n_gs_dv()) * c->L.ntheta;
  }
  if (c->alt_w_version == c->version && c->alt_w_limber == limber) return c->alt_w.data();
  if (check_layout(c, c->L) || ensure_device(c)) return nullptr;
  const bool full = c->legendre_mode == CLH_LEGENDRE_FULL;
  if (!c->L.kern_valid || (full && c->L.kfull.empty())) build_kernels(c->L, full);
  if (!c->dl.valid && upload_layout(c, c->L, c->dl)) return nullptr;
  c->alt_L = c->L;  // the device copy of L serves both variants
  c->alt_L.nonlimber = limber == 0;
  const int ord[3] = {0, 1, 2};
  std::vector<const Draw*> v{&c->cur};
  const std::vector<const Layout*> lays{&c->alt_L};
  const std::vector<const DevLayout*> devs{&c->dl};
  if (build_batch(c, c->scratch, v, lays, devs, 0, ord, nullptr, nullptr)) return nullptr;
  if (run_batch(c, c->scratch, true)) return nullptr;
  const int nw = c->L.lens.nbin * c->L.ntheta;
  c->alt_w.resize(nw);
  const double* src = c->scratch.desc[0].dv + (size_t)(2 * c->L.n_ss() + c->L.n_gs_dv()) * c->L.ntheta;
  if (hipMemcpyAsync(c->alt_w.data(), src, nw * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess) {
    fail(c, CLH_ERR_HIP, "w_gg_tomo: copy of w(theta) failed");
    return nullptr;
  }
  c->alt_w_version = c->version;
  c->alt_w_limber = limber;
  return c->alt_w.data();
}

extern "C" {

clh_ctx* clh_default(void) {
  std::lock_guard<std::mutex> g(g_default_mu);
  if (!g_default) clh_create(&g_default, 0);
  return g_default;
}

double xi_pm_tomo(const int pm, const int nt, const int ni, const int nj, const int limber) {
  clh_ctx* c = clh_default();
  if (!c) return nanv();
  if (limber != 1) { fail(c, CLH_ERR_ARG, "NonLimber not implemented"); return nanv(); }
  const int ns = c->L.src.nbin;
  if (nt < 0 || nt > c->L.ntheta - 1) { fail(c, CLH_ERR_ARG, "error in selecting bin number nt = %d", nt); return nanv(); }
  if (ni < 0 || ni > ns - 1 || nj < 0 || nj > ns - 1) { fail(c, CLH_ERR_ARG, "error in selecting bin number (ni,nj) = [%d,%d]", ni, nj); return nanv(); }
  if (ensure_current(c)) return nanv();
  const int q = n_shear_idx(ns, ni, nj) * c->L.ntheta + nt;
  const int m = c->L.n_ss() * c->L.ntheta;
  return (pm > 0) ? c->cache_raw[q] : c->cache_raw[m + q];
}

double w_gammat_tomo(const int nt, const int ni, const int nj, const int limber) {
  clh_ctx* c = clh_default();
  if (!c) return nanv();
  if (limber != 1) { fail(c, CLH_ERR_ARG, "NonLimber not implemented"); return nanv(); }
  if (nt < 0 || nt > c->L.ntheta - 1) { fail(c, CLH_ERR_ARG, "error in selecting bin number nt = %d", nt); return nanv(); }
  if (ni < 0 || ni > c->L.lens.nbin - 1 || nj < 0 || nj > c->L.src.nbin - 1) {
    fail(c, CLH_ERR_ARG, "error in selecting bin number (ni, nj) = [%d,%d]", ni, nj);
    return nanv();
  }
  if (!c->L.zoverlap(ni, nj)) return 0.0;  // cosmo2D.c:722-732
  if (ensure_current(c)) return nanv();
  const int q = c->L.gs_rows()[ni * c->L.src.nbin + nj] * c->L.ntheta + nt;
  return c->cache_raw[2 * c->L.n_ss() * c->L.ntheta + q];
}

double w_gg_tomo(const int nt, const int ni, const int nj, const int limber) {
  clh_ctx* c = clh_default();
  if (!c) return nanv();
  if (limber != 0 && limber != 1) { fail(c, CLH_ERR_ARG, "limber must be 0 or 1"); return nanv(); }
  if (nt < 0 || nt > c->L.ntheta - 1) { fail(c, CLH_ERR_ARG, "error in selecting bin number nt = %d", nt); return nanv(); }
  if (ni < 0 || ni > c->L.lens.nbin - 1 || nj < 0 || nj > c->L.lens.nbin - 1) {
    fail(c, CLH_ERR_ARG, "error in selecting bin number (ni,nj) = [%d,%d]", ni, nj);
    return nanv();
  }
  if (ni != nj) { fail(c, CLH_ERR_ARG, "ni != nj tomography not supported"); return nanv(); }
  // `limber` picks the path of this evaluation only (cosmo2D.c:845-857); the
  // context's adopt_limber_gg, which the data vectors follow, is left as set
  const double* w = w_block(c, limber);
  return w ? w[ni * c->L.ntheta + nt] : nanv();
}

double C_ss_tomo_limber(const double l, const int ni, const int nj, const int EE) {
  clh_ctx* c = clh_default();
  if (!c) return nanv();
  const int ns = c->L.src.nbin;
  if (ni < 0 || ni > ns - 1 || nj < 0 || nj > ns - 1) { fail(c, CLH_ERR_ARG, "error in selecting bin number (ni,nj) = [%d,%d]", ni, nj); return nanv(); }
  if (ensure_current(c)) return nanv();
  return table_lookup(c, n_shear_idx(ns, ni, nj), true, l, EE != 1);  // EE = 0: the B mode
}

double C_gs_tomo_limber(const double l, const int ni, const int nj) {
  clh_ctx* c = clh_default();
  if (!c) return nanv();
  if (ni < 0 || ni > c->L.lens.nbin - 1 || nj < 0 || nj > c->L.src.nbin - 1) {
    fail(c, CLH_ERR_ARG, "error in selecting bin number (ni, nj) = [%d,%d]", ni, nj);
    return nanv();
  }
  if (!c->L.zoverlap(ni, nj)) return 0.0;  // cosmo2D.c:2376-2391
  if (ensure_current(c)) return nanv();
  return table_lookup(c, c->L.n_ss() + ni * c->L.src.nbin + nj, false, l);
}

double C_gg_tomo_limber(const double l, const int ni, const int nj) {
  clh_ctx* c = clh_default();
  if (!c) return nanv();
  if (ni < 0 || ni > c->L.lens.nbin - 1 || nj < 0 || nj > c->L.lens.nbin - 1) {
    fail(c, CLH_ERR_ARG, "error in selecting bin number (ni,nj) = [%d,%d]", ni, nj);
    return nanv();
  }
  if (ni != nj) { fail(c, CLH_ERR_ARG, "cross-tomography not supported"); return nanv(); }
  if (ensure_current(c)) return nanv();
  return table_lookup(c, c->L.n_ss() + c->L.n_gs() + ni, false, l);
}

// init = 1 only primes the reference's static caches; its value is discarded
// by every caller, so it returns 0 here.
double C_ss_tomo_limber_nointerp(const double l, const int ni, const int nj, const int EE, const int init) {
  clh_ctx* c = clh_default();
  if (!c) return nanv();
  const int ns = c->L.src.nbin;
  if (ni < 0 || ni > ns - 1 || nj < 0 || nj > ns - 1) { fail(c, CLH_ERR_ARG, "invalid bin input (ni, nj) = (%d, %d)", ni, nj); return nanv(); }
  if (init == 1) return 0.0;
  // EE = 0: the B mode, zero under NLA (cosmo2D.c:1488-1504) and integrated
  // under TATT (:1562-1601)
  if (EE != 1 && c->L.ia_model != CLH_IA_TATT) return 0.0;
  std::vector<double> out(c->L.n_ss());
  if (clh_cl_nointerp(c, EE != 1 ? 3 : 0, &l, 1, out.data())) return nanv();
  return out[n_shear_idx(ns, ni, nj)];
}

double C_gs_tomo_limber_nointerp(const double l, const int nl, const int ns, const int init) {
  clh_ctx* c = clh_default();
  if (!c) return nanv();
  if (nl < 0 || nl > c->L.lens.nbin - 1 || ns < 0 || ns > c->L.src.nbin - 1) {
    fail(c, CLH_ERR_ARG, "invalid bin input (ni, nj) = (%d, %d)", nl, ns);
    return nanv();
  }
  if (init == 1) return 0.0;
  std::vector<double> out(c->L.n_gs());
  if (clh_cl_nointerp(c, 1, &l, 1, out.data())) return nanv();
  return out[nl * c->L.src.nbin + ns];
}

double C_gg_tomo_limber_nointerp(const double l, const int ni, const int nj, const int init) {
  clh_ctx* c = clh_default();
  if (!c) return nanv();
  if (ni < 0 || ni > c->L.lens.nbin - 1 || nj < 0 || nj > c->L.lens.nbin - 1) {
    fail(c, CLH_ERR_ARG, "error in selecting bin number (ni, nj) = [%d,%d]", ni, nj);
    return nanv();
  }
  if (ni != nj) { fail(c, CLH_ERR_ARG, "cross-tomography (ni,nj) = (%d,%d) bins not supported", ni, nj); return nanv(); }
  if (init == 1) return 0.0;
  std::vector<double> out(c->L.lens.nbin);
  if (clh_cl_nointerp(c, 2, &l, 1, out.data())) return nanv();
  return out[ni];
}

// 6x2pt CMB lensing scalars on the default context: w_gk_tomo (cosmo2D.c:913),
// w_ks_tomo (:1107) from the cached raw data vector (the gk and ks blocks
// after the 3x2pt ones, without calibration), C_gk / C_ks / C_kk tables at the
// gs grid of ln ell (:3020, 3270, 3484) and the direct integrals at any ell
// (:2960, 3217, 3449; init = 1 returns 0, the reference's cache set-up call)
static clh_ctx* sixpt_default() {
  clh_ctx* c = clh_default();
  if (!c) return nullptr;
  if (!c->L.sixpt) {
    fail(c, CLH_ERR_STATE, "CMB lensing probes need init_probes(\"6x2pt\")");
    return nullptr;
  }
  return c;
}
static double cmb_w(int nt, int ni, int limber, bool ks) {
  clh_ctx* c = sixpt_default();
  if (!c) return nanv();
  const Layout& L = c->L;
  if (limber != 1) { fail(c, CLH_ERR_ARG, "NonLimber not implemented"); return nanv(); }
  const int nb = ks ? L.src.nbin : L.lens.nbin;
  if (nt < 0 || nt > L.ntheta - 1) { fail(c, CLH_ERR_ARG, "error in selecting bin number nt = %d", nt); return nanv(); }
  if (ni < 0 || ni > nb - 1) { fail(c, CLH_ERR_ARG, "error in selecting bin number ni = %d", ni); return nanv(); }
  if (ensure_current(c)) return nanv();
  const size_t o = (size_t)L.dv_size3() + (size_t)(ks ? L.lens.nbin : 0) * L.ntheta + (size_t)ni * L.ntheta + nt;
  return c->cache_raw[o];
}
double w_gk_tomo(const int nt, const int ni, const int limber) { return cmb_w(nt, ni, limber, false); }
double w_ks_tomo(const int nt, const int ni, const int limber) { return cmb_w(nt, ni, limber, true); }

static double cmb_table(int col, double l) {
  clh_ctx* c = sixpt_default();
  if (!c || ensure_current(c)) return nanv();
  if (c->cache_cl6.empty()) { fail(c, CLH_ERR_STATE, "6x2pt columns missing"); return nanv(); }
  return table_lookup(c, col, false, l, false, true);
}
double C_gk_tomo_limber(const double l, const int ni) {
  clh_ctx* c = sixpt_default();
  if (!c) return nanv();
  if (ni < 0 || ni > c->L.lens.nbin - 1) { fail(c, CLH_ERR_ARG, "error in selecting bin number ni = %d", ni); return nanv(); }
  return cmb_table(ni, l);
}
double C_ks_tomo_limber(const double l, const int ni) {
  clh_ctx* c = sixpt_default();
  if (!c) return nanv();
  if (ni < 0 || ni > c->L.src.nbin - 1) { fail(c, CLH_ERR_ARG, "error in selecting bin number ni = %d", ni); return nanv(); }
  return cmb_table(c->L.lens.nbin + ni, l);
}
double C_kk_limber(const double l) {
  clh_ctx* c = sixpt_default();
  if (!c) return nanv();
  return cmb_table(c->L.lens.nbin + c->L.src.nbin, l);
}
static double cmb_direct(int probe, double l, int bin, int init) {
  clh_ctx* c = sixpt_default();
  if (!c) return nanv();
  const int nb = probe == 4 ? c->L.lens.nbin : probe == 5 ? c->L.src.nbin : 1;
  if (bin < 0 || bin > nb - 1) { fail(c, CLH_ERR_ARG, "error in selecting bin number ni = %d", bin); return nanv(); }
  if (init == 1) return 0.0;
  std::vector<double> out(nb);
  if (clh_cl_nointerp(c, probe, &l, 1, out.data())) return nanv();
  return out[bin];
}
double C_gk_tomo_limber_nointerp(const double l, const int nl, const int init) { return cmb_direct(4, l, nl, init); }
double C_ks_tomo_limber_nointerp(const double l, const int ns, const int init) { return cmb_direct(5, l, ns, init); }
double C_kk_limber_nointerp(const double l, const int init) { return cmb_direct(6, l, 0, init); }

// int_for_C_*_limber (cosmo2D.h:116-130) at one scale factor on the default
// context; NaN (message recorded) where the reference exits
static double int_for_default(int probe, double a, void* params) {
  clh_ctx* c = clh_default();
  if (!c) return nanv();
  double out = nanv();
  if (clh_int_for_limber(c, probe, (const double*)params, &a, 1, &out)) return nanv();
  return out;
}
double int_for_C_ss_tomo_limber(double a, void* params) { return int_for_default(0, a, params); }
double int_for_C_gs_tomo_limber(double a, void* params) { return int_for_default(1, a, params); }
double int_for_C_gg_tomo_limber(double a, void* params) { return int_for_default(2, a, params); }
double int_for_C_gk_tomo_limber(double a, void* params) { return int_for_default(3, a, params); }
double int_for_C_ks_tomo_limber(double a, void* params) { return int_for_default(4, a, params); }
double int_for_C_kk_limber(double a, void* params) { return int_for_default(5, a, params); }

// the gs pair maps of redshift_spline.c:115-223 on the default context
int test_zoverlap(int ni, int nj) {
  clh_ctx* c = clh_default();
  if (!c) return -1;
  const int r = clh_test_zoverlap(c, ni, nj);
  return r < 0 ? -1 : r;
}

int N_ggl(int ni, int nj) {
  clh_ctx* c = clh_default();
  if (!c) return -1;
  if (ni < 0 || ni >= c->L.lens.nbin || nj < 0 || nj >= c->L.src.nbin) {
    fail(c, CLH_ERR_ARG, "invalid bin input (ni, nj) = (%d, %d)", ni, nj);
    return -1;
  }
  return c->L.gs_rows()[ni * c->L.src.nbin + nj];
}

static int ggl_pair_bin(int nz, bool lens) {
  clh_ctx* c = clh_default();
  if (!c) return -1;
  const std::vector<int> rows = c->L.gs_rows();
  for (size_t q = 0; q < rows.size(); q++)
    if (rows[q] == nz && nz >= 0) return lens ? (int)q / c->L.src.nbin : (int)q % c->L.src.nbin;
  fail(c, CLH_ERR_ARG, "invalid bin input %d (max %d)", nz, c->L.n_gs_dv());
  return -1;
}
int ZL(int ni) { return ggl_pair_bin(ni, true); }
int ZS(int nj) { return ggl_pair_bin(nj, false); }

}  // extern "C"
