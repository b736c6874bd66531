// Shared host/device definitions of the projection library.
#pragma once
#include <stdint.h>

#define CLH_MAXB 20          // MAX_SIZE_ARRAYS (cosmolike/structs.h:11)
#define CLH_LMIN_TAB 20      // limits.LMIN_tab (cosmolike/structs.c:44)
#define CLH_NLOW 19          // directly integrated ell = 1 .. LMIN_tab-1
#define CLH_COVERH0 2997.92458
#define CLH_A_MIN (1.0 / (1.0 + 40.0))  // limits.a_min (structs.c:41)
#define CLH_NP_SS 64         // GL order for ss/gs at high_def_integration = 0
#define CLH_NP_GG 96         // GL order for gg at high_def_integration = 0
#define CLH_HDI_ROWS 16      // rows per work item of k_limber_hdi (high_def_integration != 0)

// Work-item kinds of the Limber kernel
#define CLH_W_SS 0
#define CLH_W_GS 1
#define CLH_W_GG 2
#define CLH_IA_TATT 1        // nuisance.IA_MODEL (cosmolike/IA.h:8)

// Precision modes (CLH_PRECISION_* of the C ABI)
#define CLH_PREC_FP64 0
#define CLH_PREC_FP32 1
#define CLH_PREC_BF16X3 2

// Rows of a Limber table column: CLH_NLOW direct rows (ell = r+1), then the
// N_ell log-spaced rows.  A Limber work tile covers at most CLH_TILE_ROWS
// consecutive rows (the host spreads a column's rows evenly over its tiles);
// the forward kernels process a tile in sub-tiles of CLH_SUB_ROWS rows.
#define CLH_TILE_ROWS 64
#define CLH_SUB_ROWS 16
#define CLH_LDS_LK 512       // log10 k tables up to this length are staged in LDS

// Per-node search results written by k_nodes after each node block's weight
// columns (shared by every row tile of the segment):
//  ss / gs nodes: CLH_LUT_SX columns  z index of P(k,z), its weight dy,
//                 log10(chi c/H0)
//  gg nodes:      CLH_LUT_GG columns  the same three, then the a_0 side of
//                 W_RSD (radial_weights.c:98-99): chi(z) index at chi,
//                 z index of the dchi/dz stencil, growth index, lens n(z)
//                 knot index, n_lens h f and d n_lens/d dz h f at a_0
#define CLH_LUT_SX 3
#define CLH_LUT_GG 9

// Non-Limber clustering (C_cl_tomo, cosmolike/cosmo2D.c:4528-4769): NL_Nchi
// log-spaced chi nodes (structs.c:272), ell < LMAX_NOLIMBER (structs.c:45) in
// blocks of 16 (cosmo2D.c:4672), three FFTLog configurations (density, RSD,
// magnification) whose transform lengths N_j and zero padding N0_j follow
// from NL_Nchi; a draw's work area holds, at the offsets of ClhNlOff:
//   x[512]                 chi nodes (Mpc)
//   fb[nl][3][512]         FFTLog inputs f(x) / x^nu
//   T[nl][3][NH][2]        windowed forward transforms (complex)
//   G[3][NH][LT][2]        g_l(nu + i eta_m) e^{-i eta_m ln base_l}, ell fastest
//   part[nl][LT][QT]       partial q sums of the C_ell integrand
//   partc[nl][16][QT]      the same at the block ends ell = 15, 31, ..., 149
//   used[nl][L]            the C_ell the w sum reads below LMAX_NOLIMBER
//   lmx[nl]                the ell where each bin switched to Limber (LMAX)
//   scal[8]                chi_min, dln chi, c/H0 in Mpc
#define CLH_NL_NCHI 500
#define CLH_NL_L 150
#define CLH_NL_LT 160          // ell rows padded to 16
#define CLH_NL_NH 752          // N/2 + 1 padded to even, N <= 1500
struct ClhNlOff {
  size_t x, xnu, fb, T, G, part, partc, used, lmx, scal, total;
};
#ifdef __HIPCC__
__host__ __device__
#endif
inline ClhNlOff clh_nl_off(int nl) {
  ClhNlOff o;
  size_t p = 0;
  o.x = p; p += 512;
  o.xnu = p; p += 512;  // x^1.01 (the RSD configuration's nu)
  o.fb = p; p += (size_t)nl * 3 * 512;
  o.T = p; p += (size_t)nl * 3 * CLH_NL_NH * 2;
  o.G = p; p += (size_t)3 * CLH_NL_LT * CLH_NL_NH * 2;
  o.part = p; p += (size_t)nl * CLH_NL_LT;   // sum_q F^2 k^3 P_lin per (bin, ell)
  o.partc = p; p += (size_t)nl * 16;         // the same at the block-end ells
  o.used = p; p += (size_t)nl * CLH_NL_L;
  o.lmx = p; p += (size_t)nl;
  o.scal = p; p += 8;
  o.total = (p + 31) & ~(size_t)31;
  return o;
}

// VJP tangent area of a non-Limber draw (DrawDesc::nlt): FFTLog inputs of
// the two tangent directions per lens bin (sets 0-2: d/d dz of the density,
// RSD and magnification inputs; set 3: d/d b1 of the density input), their
// transforms, and the cotangent of sum_q F^2 k^3 P_lin per (bin, ell)
struct ClhNlTan {
  size_t fb, T, sbar, total;
};
#ifdef __HIPCC__
__host__ __device__
#endif
inline ClhNlTan clh_nl_tan(int nl) {
  ClhNlTan o;
  size_t p = 0;
  o.fb = p; p += (size_t)nl * 4 * 512;
  o.T = p; p += (size_t)nl * 4 * CLH_NL_NH * 2;
  o.sbar = p; p += (size_t)nl * CLH_NL_LT;
  o.total = (p + 31) & ~(size_t)31;
  return o;
}

// Bump allocator over device memory the caller provides (the allocation-free
// entry point's scratch array)
struct ClhArena {
  char* base;
  size_t cap, used;
  bool exhausted;
};
struct clh_ctx;
// clh_compute_3x2pt_real with every device buffer carved from `scratch`
int clh_compute_3x2pt_real_scratch(clh_ctx* c, const int ord[3], double* dv, int n, void* scratch, size_t bytes);
int clh_context_device(const clh_ctx* c);
void clh_record_error(clh_ctx* c, int code, const char* msg);

// Everything one draw needs on the device.  Plain data; lives in an array in
// device memory and is read with uniform (scalar) loads.
struct DrawParams;

struct DrawDesc {
  // cosmology tables (cosmopara, cosmolike/structs.h:128-191)
  const double* chi_z;
  const double* chi_v;
  const double* G_z;
  const double* G_v;
  const double* lk;   // log10 k [h/Mpc]
  const double* pz;   // z of the P(k,z) table
  double* lnP;        // [nk][nzp] (writable: device-side parameter updates)
  int chi_n, G_n, nk, nzp;
  double omega_m;

  // survey layout
  int ns, nl, ntheta, n_ell, n_a, nrow, n_ss, n_gs;
  const double* s_zc;    // [s_rows] n(z) spline knots (histogram centres)
  const double* s_coef;  // [ns][s_rows-1][4] (y, b, c, d) per interval
  const double* l_zc;
  const double* l_coef;
  int s_rows, l_rows;
  double s_zmin_all, s_zmax_all, l_zmin_all, l_zmax_all;
  double l_zmin[CLH_MAXB], l_zmax[CLH_MAXB], l_zmean[CLH_MAXB];
  const double* ell_ss;   // [nrow] ell of every ss row
  const double* ell_gx;   // [nrow] ell of every gs/gg row
  const double* kfold;    // [4][nrow][ntheta] (xi+, xi-, gamma_t, w), row-major in r
  const double* kfoldT;   // the same kernels as [4][ntheta][nrow]
  const double* kfull;    // [4][lmax][ntheta] full-length kernels (FULL mode) or null
  const double* fill;     // [2][lmax][2] (index, weight) of the ln-ell fill, ss then gs/gg
  double lim_ss[3], lim_gx[3];
  int lmax;
  // folded kernels packed as MFMA A operands, rows in groups of 8
  // (group g, lane l = theta + 32 h, item t <-> row 8 g + 4 h + t):
  const float* kmx32;        // [4][ngrp][64][4]   FP32
  const uint16_t* kmxbf;     // [4][ngrp][64][12]  BF16 hi[4], mid[4], lo[4] of the band-scaled kernel
  const float* kmxsc;        // [4][ngrp] power-of-two factor applied to C_ell of the band
  int kmx_ngrp;

  // nuisance values of the draw (nuisanceparams, cosmolike/structs.h:203-298)
  const struct DrawParams* par;
  int ia_mode;

  // per-draw work buffers
  double* gtab;    // [ns + nl][n_a] lens-efficiency tables (g_tomo, then g_lens)
  // per-interval coefficients of the chi(z), dchi/dz and G(z) tables read by
  // the gg integrand's a_1 walks (written by k_gtab, clh_cell_records):
  // crec [chi_n][8] = chi, z, zA, zS, D0, D1, G0, slG of interval j; grec
  // [G_n][3] = z, G0, slG of the G(z) intervals when G_z differs from chi_z
  double* crec;
  double* grec;
  int zg_same;     // G(z) is tabulated on the z grid of chi(z)
  // nrec [nl][l_rows][5] = knot z and the spline coefficients / lens stretch
  // of each lens n(z) interval (k_gtab, lens blocks)
  double* nrec;
  // the lens integration limits take |dz| + lim_eps for dz != 0 (0 in a
  // draw; a small step in the views that give the VJP the limits' motion)
  double lim_eps;
  double* nodes;   // node data, offsets below
  double* cl;      // [n_ss + n_gs + nl][nrow]
  double* dv;      // data vector
  int dv_start[3];
  int node_ss, node_gs, node_gg;  // offsets into nodes
  int node_gs_stride, node_gg_stride;
  double chi_amin;  // chi(a_min) for the gg RSD cut
  // Ntable.high_def_integration: gx = 25 (1 + |hdi|) fine-grid points per
  // lens-efficiency interval (redshift_spline.c:865, 939, 1027); for hdi != 0
  // the Limber rows come from k_limber_hdi with np_ss (ss, gs) and np_gg
  // (gg) GL nodes, gl_hd = x, w of the ss / gs rule then x, w of the gg rule
  int gx, np_ss, np_gg;
  const double* gl_hd;
  double G0;        // G(z=0) by the growth table's own interpolation
  const double* gl; // GL abscissae/weights on [-1,1]: x64 w64 x96 w96
  int apply_calib;  // 1: shear calibration, point mass and mask applied to the data vector
  const double* theta;  // [ntheta] bin-averaged theta [rad] (compute_binning_real_space)
  const int* mask;      // [dv size] 0/1 in the (0,1,2) block order, or null (all ones)
  // data-vector row of C_gs column zl ns + zs, -1 for an excluded pair
  // (init_ggl_exclude), or null (every pair, row = column); n_gs_dv rows
  const int* gs_row;
  int n_gs_dv;

  // adjoint (VJP) buffers of the draw
  const double* dvbar;  // cotangent of the data vector
  double* raw;          // [ncol] dvbar . uncalibrated data vector per C_ell column
  double* cbar;         // [ncol][nrow] cotangent of the C_ell rows
  double* abar;         // [ns][64] cotangent of A_b at the ss nodes
  double* bbar;         // [nl][ns][64] cotangent of B_zs at the gs nodes
  double* wbar;         // [nl][2][64] gs nodes (W_gal, W_mag), then [nl][2][96] gg nodes
  double* dgtab;        // [ns+nl][n_a] d g / d(photo-z shift of its own bin)
  double* g_lnP;        // [nk][nzp]
  double* run;          // [64 (1 + nl)][nk] per-GL-node ln P cotangent runs of the ss / gs nodes, or null
  double* sxk;          // [1 + nl][nrow][64] ss / gs checkpoint of the forward (S = P fac per
                        // row and node) for the VJP, or null
  // TATT VJP: cotangents of the B-mode ss rows [n_ss][nrow] and of the node
  // vectors B_b, D_b at the ss and gs nodes [(1 + nl)][2][ns][64] (null
  // otherwise)
  double* cbar_bb;
  double* tbar;
  // 6x2pt VJP: cotangent of the cl6 columns [nl + ns + 1][nrow6] (null otherwise)
  double* cbar6;
  // non-Limber VJP (null otherwise): the tangent area (ClhNlTan) and the
  // cotangents of the two views' C_ell rows at ell = 1..149, [2][ncol][149]
  // (P_nl view, then P_lin view; only the gg columns are used)
  double* nlt;
  double* nlc_bar;
  // the motion of the gs / gg nodes with the lens photo-z shift (the lens
  // integration limits move with |dz|), per node of lens bin zl, for the VJP
  // (k_nodes_tan), or null: gs [nl][6 + ns][64] = da/d dz, d ln(GL weight x
  // fac)/d dz, then per unit a: d log10 k, d dy (P(k,z) z weight), W_gal,
  // W_mag, B_zs[ns]; then gg [nl][8][96] = the first six, t0 (the a_0 half of
  // W_RSD) and chi
  double* ntan;
  // TATT intrinsic alignments (ia_model == CLH_IA_TATT)
  int ia_model;
  double* cl_bb;        // [n_ss][nrow] B-mode C_ss rows, or null (NLA: B mode is zero)
  double* tatt;         // [1 + nl][2 ns + 1][64] per ss / gs node: B_b, D_b, D(a)^4
  const double* fpt;    // [10][fpt_n] FAST-PT IA tables (set_IA_PS units: / (c/H0)^3)
  int fpt_n;
  double fpt_l0, fpt_l1, fpt_dl;  // ln k range (k in H0/c) and step (l1 - l0) / fpt_n
  double* ggk;          // [nl][nrow][3][96] gg integrand checkpoint of the forward (bracket t,
                        // P fac, dW_RSD/d dz per row and node) for the VJP, or null
  // one-loop galaxy bias (some b2 != 0, has_b2_galaxies): FAST-PT bias tables
  int oneloop;
  const double* fptb;         // [6][fptb_n] d1d2, d2d2, d1s2, d2s2, s2s2, d1p3 (/ (c/H0)^3)
  int fptb_n;
  double fptb_l0, fptb_l1, fptb_dl, fptb_s4;  // ln k range (k in H0/c), step, sigma4 (/ (c/H0)^3)
  // 6x2pt CMB lensing probes (cl6 null otherwise): C_gk [nl], C_ks [ns], C_kk
  // columns of nrow6 = nrow + n_f rows (the gs/gg ell rows, then the direct
  // C_kk ells)
  double* cl6;
  int nrow6;
  const double* kf6;          // [2][nrow][ntheta] folded w (gk) and gamma_t (ks) kernels x CMB filter
  const double* ell_f;        // [n_f] direct C_kk ells: like.ell, or band-power L <= LMIN_tab
  int n_f;
  int kk_nbp, kk_lmin, kk_nL; // kk band powers (kk_nbp 0: C_kk at like.ell), L = kk_lmin + [0, kk_nL)
  const double* kk_M;         // [kk_nbp][kk_nL] binning matrix
  const double* kk_off;       // [kk_nbp] theory offset
  double chi_cmb;             // chi(z = 1090) in c/H0 (g_cmb, redshift_spline.c:1111-1125)
  int dv6_start[3];           // data-vector starts of the gk, ks, kk blocks
  // non-Limber w (like.adopt_limber_gg = 0), nlw null otherwise
  double* nlw;                // work area, ClhNlOff offsets
  const double* nl_tw;        // [3][2 N_j] cos, sin of 2 pi t / N_j (layout)
  const double* nl_pw;        // [ntheta][CLH_NL_L] w kernel at integer ell (layout)
  const double* nl_lnl;       // [nl][CLH_NL_L - 1] gg Limber rows at ell = 1..149 with P_nl
  const double* nl_llin;      // the same with the linear P(k,z)
  const double* lkL;          // linear P(k,z): log10 k, z, ln P k-major
  const double* pzL;
  const double* lnPL;
  int nkL, nzL;
  int nl_N[3], nl_N0[3];      // transform lengths and zero padding per configuration
  double h0;                  // H0 / 100 km/s/Mpc
  double* g_src;        // [ns]
  double* g_lens;       // [nl]
  double* g_b1;         // [nl]
  double* g_A;          // [1]
  double* g_m;          // [ns]
};

struct DrawParams {
  double src_dz[CLH_MAXB], lens_dz[CLH_MAXB], lens_st[CLH_MAXB];
  double b1[CLH_MAXB], bmag[CLH_MAXB], shear_m[CLH_MAXB];
  double ia_bin[CLH_MAXB];
  double pm[CLH_MAXB];  // point-mass amplitudes per lens bin (PointMass::pm_)
  double ia_A, ia_eta;
  // TATT: nuisance.ia[1][*] (A2: per bin, or value and exponent) and ia[2][*] (b_TA)
  double ia2_bin[CLH_MAXB], bta_bin[CLH_MAXB];
  double ia2_A, ia2_eta, bta_A;
  // one-loop galaxy bias: nuisance.gb[1], gb[2], gb[3], gb[5] per lens bin
  double b2[CLH_MAXB], bs2[CLH_MAXB], b3[CLH_MAXB], bk[CLH_MAXB];
};

// Node buffer layout (doubles):
//  ss block  : a[64] fK[64] fac[64] A[ns][64] LUT[3][64]      fac = wt dchi/da / fK^2
//              A_b = W_kappa_b - C1 W_source_b
//  gs block zl: a fK fac WGAL WMAG [64 each], B[ns][64] LUT[3][64]
//              B_zs = W_kappa_zs - C1_zs W_source_zs
//  gg block zl: a fK fac WGAL WMAG [96 each] LUT[9][96]
// Sizes: ss 64 (6 + ns), gs stride 64 (8 + ns), gg stride 96 (5 + 9).
