/*
 * cl_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of CosmoLike's tomographic two-point projection path
 * (Limber C_ell for ss/gs/gg + bin-averaged Legendre transform to xi+-,
 * gamma_t, w).  It exists to check the HIP library and to time a CPU baseline;
 * the shipped library never links, loads or calls it.
 *
 * Every function follows a reference routine cited at its definition
 * (cosmolike/cosmo2D.c, cosmo3D.c, redshift_spline.c, radial_weights.c,
 * basics.c, IA.c, bias.c, generic_interface.{cpp,hpp}).  State lives in an
 * explicit context instead of the reference's global structs.
 */
#ifndef CL_ORACLE_H
#define CL_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct oc_ctx oc_ctx;

oc_ctx* oc_create(void);
void oc_free(oc_ctx* c);

/* Ntable (structs.c:262-279): N_ell, N_a, LMAX, high_def_integration */
int oc_set_ntable(oc_ctx* c, int n_ell, int n_a, int lmax, int high_def);
/* init_binning_real_space (generic_interface.cpp:479-505), arcmin */
int oc_set_binning_real(oc_ctx* c, int ntheta, double tmin_arcmin, double tmax_arcmin);
int oc_set_omega_m(oc_ctx* c, double omega_m);
int oc_set_distances(oc_ctx* c, const double* z, const double* chi, int n);
int oc_set_growth(oc_ctx* c, const double* z, const double* G, int n);
/* lnP is k-major: lnP[i*nz + j] (generic_interface.cpp:1343) */
int oc_set_pk(oc_ctx* c, const double* log10k, int nk, const double* z, int nz,
              const double* lnP);
/* table is row-major (nrows, ntomo+1): column 0 = z, columns 1.. = n_i(z) */
int oc_set_source_sample(oc_ctx* c, const double* table, int nrows, int ntomo);
int oc_set_lens_sample(oc_ctx* c, const double* table, int nrows, int ntomo);
int oc_set_nuisance(oc_ctx* c, const double* src_dz, const double* lens_dz,
                    const double* lens_stretch, const double* b1, const double* bmag,
                    const double* shear_m);
/* IA: model 0 = NLA only; mode 0=none, 2=per-bin (A[i]), 3=A*(a^-1/1.62)^eta */
int oc_set_ia(oc_ctx* c, int ia_mode, const double* A_per_bin, double A, double eta);

/* ------------- pieces (for pinning tests) ------------------------------ */
int oc_nz_source(oc_ctx* c, int bin, const double* z, int n, double* out);
int oc_nz_lens(oc_ctx* c, int bin, const double* z, int n, double* out);
int oc_zmean_lens(oc_ctx* c, double* out);
int oc_g_tomo(oc_ctx* c, int bin, const double* a, int n, double* out);
int oc_g_lens(oc_ctx* c, int bin, const double* a, int n, double* out);
int oc_chi_all(oc_ctx* c, const double* a, int n, double* chi, double* dchida);
int oc_pdelta(oc_ctx* c, const double* k, const double* a, int n, double* out);
int oc_gl_nodes(int n, double* x, double* w);
int oc_legendre_pl_dpl(int lmax, double x, double* P, double* dP);
/* 1 (default): kernels evaluated in long double; 0: in double, as the reference */
int oc_set_kernel_precision(oc_ctx* c, int extended);
/* G kernels [ntheta][LMAX] for xi+, xi-, gamma_t, w (cosmo2D.c:430-453,631-639,808-815) */
int oc_kernels(oc_ctx* c, double* Gp, double* Gm, double* Pgt, double* Pw);

/* Limber tables at the N_ell log-spaced ell (cosmo2D.c:1617,2195,2693):
 * ss[npairs][N_ell] (EE), gs[n_ggl][N_ell], gg[nlens][N_ell] */
int oc_cl_tables(oc_ctx* c, double* ss, double* gs, double* gg);
/* direct GL integrals at arbitrary ell (the *_nointerp routines):
 * out[pair][nl] */
int oc_cl_ss_nointerp(oc_ctx* c, const double* l, int nl, double* out);
int oc_cl_gs_nointerp(oc_ctx* c, const double* l, int nl, double* out);
int oc_cl_gg_nointerp(oc_ctx* c, const double* l, int nl, double* out);
/* B-mode C_ss (nonzero under TATT): table at the table ells, direct GL */
int oc_cl_tables_bb(oc_ctx* c, double* ss_bb);
int oc_cl_ss_bb_nointerp(oc_ctx* c, const double* l, int nl, double* out);
/* TATT: IA model (0 NLA, 1 TATT), A2 and b_TA (per bin for mode 2; value and
 * exponent / one value for mode 3), and the FAST-PT IA tables of set_IA_PS */
int oc_set_ia_tatt(oc_ctx* c, int model, const double* A2_bin, double A2, double eta2, const double* bta_bin,
                   double bta);
int oc_set_ia_ps(oc_ctx* c, const double* PS, double kmin, double kmax, double cutoff, int N);
/* one FAST-PT IA table value at k (c/H0 units) times g4, as the integrands read it */
double oc_fpt_at(oc_ctx* c, int t, double k, double g4);
/* Fourier-space 3x2pt data vector at the given ells (compute_Mx2pt_N_masked<1,3>) */
int oc_datavector_fourier(oc_ctx* c, const double* ell, int ncl, double lmax_shear, const int ord[3],
                          const int* mask, double* dv);
/* C_ell filled at integer ell 0..LMAX-1 as the real-space routines use them:
 * ss[npairs][LMAX], gs[n_ggl][LMAX], gg[nlens][LMAX] (NULL to skip) */
int oc_cl_full(oc_ctx* c, double* ss, double* gs, double* gg);

/* xi_pm_tomo (cosmo2D.c:353-556): xip/xim[npairs*ntheta] (q = nz*Ntheta+i) */
int oc_xi_pm(oc_ctx* c, double* xip, double* xim);
int oc_gammat(oc_ctx* c, double* out);   /* w_gammat_tomo, [n_ggl*ntheta] */
int oc_wtheta(oc_ctx* c, double* out);   /* w_gg_tomo limber=1, [nlens*ntheta] */
/* compute_Mx2pt_N_masked<0,3>: shear calibration, point mass on gamma_t,
 * masked entries 0 (all-ones mask unless oc_set_mask was called) */
int oc_datavector_3x2pt(oc_ctx* c, const int ord[3], double* dv);
/* PointMass::set_pm_vector: amplitudes per lens bin (default 0) */
int oc_set_pointmass(oc_ctx* c, const double* pm, int n);
/* likelihood mask in the (0,1,2) block order (NULL clears) */
int oc_set_mask(oc_ctx* c, const int* mask, int n);
/* IP::get_chi2: chi^2 of dv against data with the inverse covariance, over
 * the entries where mask != 0 (mask NULL = all); arrays in dv's order */
double oc_chi2(const double* dv, const double* data, const double* inv_cov, const int* mask, int n);
int oc_datavector_size(oc_ctx* c);

int oc_npairs_ss(oc_ctx* c);
int oc_npairs_gs(oc_ctx* c);
int oc_test_zoverlap(oc_ctx* c, int zl, int zs);
int oc_npairs_ggl(oc_ctx* c);
int oc_set_ggl_exclude(oc_ctx* c, const int* pairs, int n);
int oc_int_for(oc_ctx* c, int probe, const double* ar, const double* a, int n, double* out);
/* invalidates the cosmology-dependent caches (a new draw) */
void oc_touch(oc_ctx* c);
const char* oc_last_error(oc_ctx* c);

/* ---- non-Limber clustering (C_cl_tomo, cosmo2D.c:4004-4769) ---- */
/* h0 = H0 / 100 (set_cosmological_parameters); default 0.7 */
int oc_set_h0(oc_ctx* c, double h0);
/* set_linear_power_spectrum (generic_interface.cpp:1194-1270), lnP k-major */
int oc_set_plin(oc_ctx* c, const double* log10k, int nk, const double* z, int nz, const double* lnP);
/* like.adopt_limber_gg: 1 (default here) Limber w, 0 non-Limber below ell 150 */
int oc_set_limber_gg(oc_ctx* c, int limber);
/* C_cl_tomo with tolerance 0.01: cl[nl][150] (ell 0..149) and the ell where
   each bin switched to Limber (LMAX[i]) */
int oc_cl_nonlimber(oc_ctx* c, double* cl, int* lmax);
/* one FFTLog transform (configuration cfg 0 density, 1 RSD, 2 magnification
   of C_cl_tomo) of fx on the log grid x: Fy[l - ls][q] at y = (l+1)/x[nx-1-q] */
int oc_fftlog(const double* x, const double* fx, int nx, int cfg, int ls, int le, double* Fy);
/* the DFT behind it (sign -1 forward, +1 backward, unnormalised) */
int oc_fft(const double* re, const double* im, int n, int sign, double* out_re, double* out_im);
/* complex ln Gamma (Lanczos g = 7, 9 terms; imaginary part modulo 2 pi) */
int oc_clngamma(double re, double im, double* out);
/* FFTLog inputs of C_cl_tomo: x[500] (Mpc), fx[nl][3][500] */
int oc_nl_inputs(oc_ctx* c, double* x, double* fx);
/* Limber C_gg with the linear P(k,z) at arbitrary ell: out[nl][n] */
int oc_cl_gg_lin_nointerp(oc_ctx* c, const double* l, int n, double* out);

/* ---- one-loop galaxy bias (cosmo2D.c:1970-2057, 2560-2603) ---- */
/* b2 (with bs2 = -4/7 (b1 - 1) unless b2 = 0), b3, bK per lens bin; NULL = unchanged */
int oc_set_bias_nl(oc_ctx* c, const double* b1, const double* b2, const double* b3, const double* bk);
/* set_bias_PS: PS[12][N] FAST-PT bias spectra (rows 0..5 used), k range h/Mpc */
int oc_set_bias_ps(oc_ctx* c, const double* PS, double kmin, double kmax, double cutoff, double sigma4, int N);

/* ---- 6x2pt CMB lensing probes (cosmo2D.c:913-1296, 2857-3516) ---- */
/* init_cmb_cross_correlation: beam FWHM in arcmin, filter ell range, HEALPix
   pixel window per ell (NULL / 0: none) */
int oc_set_cmb(oc_ctx* c, double fwhm_arcmin, int lmin, int lmax, const double* win, int nwin);
/* direct GL C_ell: probe 3 gk [nl][n], 4 ks [ns][n], 5 kk [1][n] */
int oc_cl_cmb_nointerp(oc_ctx* c, int probe, const double* l, int n, double* out);
int oc_w_gk(oc_ctx* c, double* out);  /* [nl][ntheta] */
int oc_w_ks(oc_ctx* c, double* out);  /* [ns][ntheta] */
int oc_cl_kk(oc_ctx* c, const double* ell, int n, double* out);
/* init_cmb_auto_bandpower: the kk block as nbp band powers of C_kk(L),
   L = lmin..lmax (M [nbp][lmax - lmin + 1], offset [nbp] or NULL = 0); nbp = 0 clears */
int oc_set_kk_bandpower(oc_ctx* c, int nbp, int lmin, int lmax, const double* M, const double* off);
int oc_datavector_size_6x2pt(oc_ctx* c, int ncl);
int oc_datavector_6x2pt(oc_ctx* c, const int ord[6], const double* ell, int ncl, double* dv);

#ifdef __cplusplus
}
#endif
#endif
