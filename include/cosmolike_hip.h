/*
 * cosmolike_hip.h -- C ABI of the MI355X (gfx950) tomographic two-point
 * projection library: Limber C_ell for shear-shear (ss), galaxy-shear (gs)
 * and galaxy-galaxy (gg), and their bin-averaged Legendre transforms
 * xi+-(theta), gamma_t(theta), w(theta).
 *
 * Plain pointers and sizes only.  Every entry point that can fail returns an
 * int status (CLH_OK = 0, negative on error); it never exits the process and
 * writes nothing to output buffers on error.  clh_last_error() gives the text.
 *
 * Three layers:
 *  1. Context setters mirroring the cosmolike_interface setters that feed the
 *     reference's global structs (each cites the routine it replaces).
 *  2. Bulk compute calls: the real-space 3x2pt data vector in the reference's
 *     order, per-probe arrays, C_ell tables, and a device-resident batch.
 *  3. The reference's own scalar entry points (xi_pm_tomo, w_gammat_tomo,
 *     w_gg_tomo, C_*_tomo_limber[_nointerp]) with identical signatures,
 *     served from a process-wide default context; see clh_default().
 */
#ifndef COSMOLIKE_HIP_H
#define COSMOLIKE_HIP_H
#ifdef __cplusplus
extern "C" {
#endif

#define CLH_OK 0
#define CLH_ERR_ARG (-1)     /* bad size, index or enum                        */
#define CLH_ERR_NAN (-2)     /* NaN in an input table                          */
#define CLH_ERR_STATE (-3)   /* an input the call needs has not been set       */
#define CLH_ERR_DOMAIN (-4)  /* input outside the domain the algorithm needs   */
#define CLH_ERR_HIP (-5)     /* HIP runtime failure                            */
#define CLH_ERR_ALLOC (-6)   /* device or host allocation failed               */

/* Legendre contraction variants (all FP64). FOLDED contracts the Limber
 * table rows with cosmology-independent kernels into which the reference's
 * linear-in-ln(ell) fill has been folded; FULL generates C_ell at every
 * integer ell < LMAX on the fly and contracts over all of them, as the
 * reference does. */
#define CLH_LEGENDRE_FOLDED 0
#define CLH_LEGENDRE_FULL 1

/* Arithmetic of the Limber and folded-Legendre stages.  FP64 is the plain
 * path (the reference's arithmetic, default).  FP32: the ss and gg Limber
 * integrands take P(k,z) as the FP32 exp of the FP64 bilinear ln P; the node
 * sums stay FP64 (ss pair sums on FP64 MFMA), and the gs rows, table
 * searches, chi(z)/growth tables, lens efficiencies and Legendre kernels stay
 * FP64; Legendre contraction on FP32 MFMA.  BF16X3: as FP32, with the
 * Legendre contraction on BF16 MFMA: each operand split into three BF16
 * parts a0 + a1 + a2 (each the nearest BF16 of what the previous left), the
 * six products a2b0 + a1b1 + a0b2 + a1b0 + a0b1 + a0b0 summed smallest first
 * in FP32, with power-of-two scaling of 32-row ell bands.  FULL Legendre mode
 * always contracts in FP64. */
#define CLH_PRECISION_FP64 0
#define CLH_PRECISION_FP32 1
#define CLH_PRECISION_BF16X3 2

typedef struct clh_ctx clh_ctx;

/* ---------------------------------------------------------------- context */
int clh_create(clh_ctx** out, int hip_device);
void clh_destroy(clh_ctx* ctx);
const char* clh_last_error(const clh_ctx* ctx);
/* process-wide context behind the reference-named scalar entry points */
clh_ctx* clh_default(void);
int clh_set_legendre_mode(clh_ctx* ctx, int mode);
/* one of CLH_PRECISION_*; no reference counterpart (the reference is FP64 only) */
int clh_set_precision(clh_ctx* ctx, int mode);

/* ---------------------------------------------------------------- setters */
/* init_accuracy_boost / init_ntable_lmax (generic_interface.cpp:284-342):
 * N_ell log-spaced Limber ell, N_a lens-efficiency table size, LMAX,
 * high_def_integration (Ntab.high_def_integration, structs.h:52-126): the
 * Limber GL orders 64 / 96 / 128 / 256 / 512 (ss, gs) and 96 / 128 / 256 /
 * 512 / 1024 (gg) for |hdi| = 0 .. 4+ (cosmo2D.c:1577-1581, 2635-2639), the
 * lens-efficiency fine grid 25 (1 + |hdi|) (redshift_spline.c:865) and the
 * lens zmean GL order 256 / 512 / 1024 (:809-811).  At hdi != 0 the Limber
 * rows are direct n-point sums of the integrands (the node kernels serve
 * hdi = 0); the 6x2pt probes and the non-Limber w refuse it, and so does
 * the VJP under TATT or one-loop bias. */
int clh_init_ntable(clh_ctx* ctx, int n_ell, int n_a, int lmax, int high_def);
/* init_binning_real_space (generic_interface.cpp:479-505), arcmin */
int clh_init_binning_real_space(clh_ctx* ctx, int ntheta, double theta_min_arcmin,
                                double theta_max_arcmin);
/* init_IA (generic_interface.cpp:628-634): ia_model 0 (NLA) or 1 (TATT);
 * ia_redshift_evol 0 (none), 2 (per-bin A) or 3 (A ((1+z)/1.62)^eta).
 * TATT computes the E- and B-mode C_ss (xi+- = EE +- BB transforms) and the
 * TATT gamma_t in FP64 in every precision mode; clh_batch_vjp takes TATT
 * batches of at most 10 source bins. */
int clh_init_ia(clh_ctx* ctx, int ia_model, int ia_redshift_evol);
/* set_cosmological_parameters (generic_interface.cpp:912-939); H0 in
 * km/s/Mpc sets c/H0 in Mpc for the non-Limber w */
int clh_set_cosmological_parameters(clh_ctx* ctx, double omega_matter, double hubble);
/* set_distances (generic_interface.cpp:1085-1143): z, chi [Mpc/h], n >= 5 */
int clh_set_distances(clh_ctx* ctx, const double* z, const double* chi, int n);
/* set_growth (generic_interface.cpp:1149-1188): G = D/a */
int clh_set_growth(clh_ctx* ctx, const double* z, const double* G, int n);
/* set_non_linear_power_spectrum (generic_interface.cpp:1276-1352):
 * lnP k-major, lnP[i*nz + j], P in (Mpc/h)^3, k in h/Mpc */
int clh_set_non_linear_power_spectrum(clh_ctx* ctx, const double* log10k, int nk,
                                      const double* z, int nz, const double* lnP);
/* set_linear_power_spectrum (generic_interface.cpp:1194-1270), the same
 * layout; read only by the non-Limber w (p_lin, cosmo3D.c:306-358) */
int clh_set_linear_power_spectrum(clh_ctx* ctx, const double* log10k, int nk,
                                  const double* z, int nz, const double* lnP);
/* like.adopt_limber_gg (generic_interface.cpp:269): 1 (this library's
 * default) Limber w(theta) at every ell; 0 (the reference's default) the
 * non-Limber C_gg of C_cl_tomo (cosmo2D.c:4528-4769) below ell 150, FFTLog
 * with tolerance 0.01, Limber above.  Needs the linear P(k,z) and H0;
 * clh_batch_vjp includes it (not with one-loop bias, TATT or 6x2pt). */
int clh_set_adopt_limber_gg(clh_ctx* ctx, int limber);
/* set_source_sample / set_lens_sample (generic_interface.cpp:1907-1999,
 * 1802-1883): row-major (nrows, ntomo+1) table, column 0 = z */
int clh_set_source_sample(clh_ctx* ctx, const double* table, int nrows, int ntomo);
int clh_set_lens_sample(clh_ctx* ctx, const double* table, int nrows, int ntomo);
/* set_nuisance_* (generic_interface.cpp:1358-1610, 1701-1775) */
int clh_set_nuisance_shear_calib(clh_ctx* ctx, const double* m, int n);
int clh_set_nuisance_shear_photoz(clh_ctx* ctx, const double* dz, int n);
int clh_set_nuisance_clustering_photoz(clh_ctx* ctx, const double* dz, int n);
int clh_set_nuisance_clustering_photoz_stretch(clh_ctx* ctx, const double* s, int n);
int clh_set_nuisance_linear_bias(clh_ctx* ctx, const double* b1, int n);
int clh_set_nuisance_magnification_bias(clh_ctx* ctx, const double* bmag, int n);
/* set_nuisance_IA: mode 2 uses A1[0..ns-1]; mode 3 uses A1[0]=A, A1[1]=eta */
int clh_set_nuisance_ia(clh_ctx* ctx, const double* A1, int n);
/* set_nuisance_IA with A2 and b_TA (generic_interface.cpp:1701-1774), for
 * TATT: mode 2 one value per source bin in each of A1, A2, BTA; mode 3
 * A1 = (A, eta), A2 = (A2, eta2), BTA[0] = b_TA */
int clh_set_nuisance_ia_tatt(clh_ctx* ctx, const double* A1, const double* A2, const double* BTA, int n);
/* set_nuisance_nonlinear_bias (generic_interface.cpp:1533-1568): b2 per lens
 * bin; bs2 = -4/7 (b1 - 1) where b2 != 0, else 0.  Any b2 != 0 switches on
 * the one-loop bias terms of the gg and gs integrands (cosmo2D.c:1970-2057,
 * 2560-2603; FP64 gs, gg in the precision mode); with TATT they are
 * refused.  The VJP differentiates them with bs2 held at its value (d/d b1
 * at fixed bs2, as after a b1-only update; the reference recomputes bs2 only
 * when b2 changes), and takes their lens-limit part from finite-difference
 * views. */
int clh_set_nuisance_nonlinear_bias(clh_ctx* ctx, const double* B1, const double* B2, int n);
/* set_nuisance_nonlocal_bias (generic_interface.cpp:1610-1669): b3, bK per lens bin */
int clh_set_nuisance_nonlocal_bias(clh_ctx* ctx, const double* B3, const double* BK, int n);
/* set_bias_PS (generic_interface.cpp:1008-1076): PS row-major [12][N], rows
 * 0..5 the FAST-PT bias spectra Pd1d2, Pd2d2, Pd1s2, Pd2s2, Ps2s2, Pd1p3
 * (Mpc/h)^3 on N log-spaced k in [kmin, kmax) h/Mpc; sigma4 (Mpc/h)^3;
 * cutoff accepted and ignored as the integrands ignore it */
int clh_set_bias_PS(clh_ctx* ctx, const double* PS, double kmin, double kmax, double cutoff, double sigma4, int N);
/* set_IA_PS (generic_interface.cpp:941-1001): PS row-major [12][N], rows
 * 0..9 the FAST-PT IA spectra (Mpc/h)^3 on N log-spaced k in [kmin, kmax)
 * h/Mpc (rows 10, 11 unused by the Limber integrands); cutoff is accepted and
 * ignored as the reference's lookup ignores it */
int clh_set_IA_PS(clh_ctx* ctx, const double* PS, double kmin, double kmax, double cutoff, int N);

/* ----------------------------------------------------------- bulk compute */
/* PointMass::set_pm_vector (generic_interface.hpp:362-365): point-mass
 * amplitudes per lens bin; gamma_t of the data vectors gets
 * 4 pi (G/c^2) pm 1e13 g_tomo(a_l, zs) / (theta^2 chi(a_l) a_l^3) with
 * a_l = 1/(1+zmean(zl)) before shear calibration (PointMass::get_pm,
 * generic_interface.cpp:2626-2641; add_calib_and_set_mask_X_N,
 * generic_interface.hpp:843-858). Default 0. */
int clh_set_nuisance_pointmass(clh_ctx* ctx, const double* pm, int n);
/* init_ggl_exclude (generic_interface.cpp:874-896): n/2 (lens, source) bin
 * pairs, flat [zl0, zs0, zl1, zs1, ...], left out of gamma_t / C_gs; n = 0
 * clears.  The gamma_t block of the data vectors then holds the included
 * pairs in N_ggl order (lens major; tomo.ggl_Npowerspectra of them,
 * generic_interface.cpp:775-802), clh_gammat returns those rows, and
 * w_gammat_tomo / C_gs_tomo_limber return 0 for an excluded pair
 * (cosmo2D.c:722-732, 2376-2391).  A mask of the old size is dropped. */
int clh_init_ggl_exclude(clh_ctx* ctx, const int* pairs, int n);
/* test_zoverlap (redshift_spline.c:115-145): 1 included, 0 excluded,
 * CLH_ERR_ARG for bins out of range */
int clh_test_zoverlap(clh_ctx* ctx, int zl, int zs);
/* The Limber integrands of cosmo2D.c as functions of the scale factor, at
 * the n points a[] (host arrays, 0 < a < 1) for the current draw, evaluated
 * on the device; ar is the reference's parameter array: probe 0 ss {n1, n2,
 * l, EE, deriv} (int_for_C_ss_tomo_limber, cosmo2D.c:1521-1555), 1 gs {nl,
 * ns, l, nonlinear_bias} (:2079-2119), 2 gg {ni, nj, l, use_linear_ps,
 * nonlinear_bias} (:2489-2606), 3 gk {nl, l, nonlinear_bias} (:2857-2954),
 * 4 ks {ns, l} (:3178-3211), 5 kk {l} (:3428-3446).  RSD enters gg only
 * (include_RSD_GS / _GK = 0, as the data vectors use them); deriv = 1 gives
 * the d C / d ln k integrand; nonlinear_bias = 1 needs the FAST-PT bias
 * tables (set_bias_PS) unless the bin has no bias terms; gk, ks, kk need the
 * 6x2pt probes. */
int clh_int_for_limber(clh_ctx* ctx, int probe, const double* ar, const double* a, int n, double* out);

/* ---------------------------------------------------- Fourier-space 3x2pt */
/* init_binning_fourier (generic_interface.cpp:425-470): like.ell at the
 * log-centres of nells bins in [lmin, lmax]; cosmic shear is kept below
 * lmax_shear */
int clh_init_binning_fourier(clh_ctx* ctx, int nells, int lmin, int lmax, int lmax_shear);
int clh_datavector_size_fourier(clh_ctx* ctx);
/* compute_Mx2pt_N_masked<1,3> (generic_interface.hpp:1137-1159): C_ss (EE,
 * 0 at ell >= lmax_shear), C_gs, C_gg at like.ell by the direct GL integrals
 * (C_*_tomo_limber_nointerp), shear calibration, blocks placed by ord as
 * compute_data_vector_Mx2pt_N_starts<1,3>; mask (NULL = all ones) holds n
 * entries in the same order as dv, masked entries are 0 */
int clh_compute_3x2pt_fourier(clh_ctx* ctx, const int ord[3], const int* mask, double* dv, int n);

/* ------------------------------------------------------------------ 6x2pt */
/* init_probes (generic_interface.cpp:640-701): "3x2pt" (default) or
 * "6x2pt".  6x2pt adds w_gk(theta) per lens bin and w_ks(theta) per source
 * bin in the real-space theta bins (w_gk_tomo, w_ks_tomo, cosmo2D.c:913-1296)
 * and C_kk at like.ell (the clh_init_binning_fourier bins;
 * generic_interface.hpp:1085-1094) to the real-space data vector, whose size
 * clh_datavector_size then reports; the 3-entry ord of the other entry
 * points keeps gk, ks, kk after the 3x2pt blocks.  Needs chi(z) beyond
 * z = 1090 (g_cmb) and clh_init_cmb_cross_correlation; clh_batch_vjp
 * includes these blocks.  Real space only, as in the reference (whose Fourier gk / ks
 * are not implemented).  With clh_init_cmb_auto_bandpower the kk block holds
 * band powers instead of C_kk at like.ell. */
int clh_init_probes(clh_ctx* ctx, const char* probes);
/* init_cmb_cross_correlation (generic_interface.cpp:508-526): the CMB filter
 * of w_gk / w_ks, exp(-l(l+1) s^2) with s = FWHM / sqrt(16 ln 2) on
 * [lmin, lmax] (beam_cmb, cosmo2D.c:268-271), times the HEALPix pixel window
 * win[l] (l < nwin, 0 above; NULL = none) */
int clh_init_cmb_cross_correlation(clh_ctx* ctx, int lmin, int lmax, double fwhm_arcmin, const double* win,
                                   int nwin);
/* init_cmb_auto_bandpower (generic_interface.cpp:541-560; IPCMB::
 * set_kk_binning_bandpower, set_kk_binning_mat, set_kk_theory_offset,
 * generic_interface.cpp:2512-2606): the kk block becomes nbins band powers
 * sum_L M[j][L - lmin] C_kk(L) - offset[j], L = lmin..lmax, C_kk direct at
 * L <= LMIN_tab and from the table above (generic_interface.hpp:1095-1118).
 * binning_matrix is [nbins][lmax - lmin + 1] row-major and theory_offset
 * [nbins] (NULL = zeros): the tables the reference reads from its two files.
 * alpha, the Hartlap factor the reference divides the kkkk covariance block
 * by, is stored only (clh_like_set_inv_cov takes an inverse the caller
 * formed).  nbins, lmin > 0, lmax >= lmin, alpha > 0. */
int clh_init_cmb_auto_bandpower(clh_ctx* ctx, int nbins, int lmin, int lmax, const double* binning_matrix,
                                const double* theory_offset, double alpha);
/* the band-power binning in use (nbins 0: C_kk at like.ell); any pointer may be NULL */
int clh_cmb_kk_bandpower_info(clh_ctx* ctx, int* nbins, int* lmin, int* lmax, double* alpha);
/* compute_Mx2pt_N_masked<0,6>(ord) (generic_interface.hpp:1137-1159): the
 * six blocks xi+- | gamma_t | w | w_gk | w_ks | kk placed by ord (6 entries),
 * calibrations and mask as compute_Mx2pt_N_masked applies them */
int clh_compute_6x2pt_real(clh_ctx* ctx, const int ord[6], double* dv, int n);

/* ------------------------------------------------------------ likelihood */
/* IP::set_mask (generic_interface.hpp:1234-1331): 0/1 per data-vector entry,
 * in the block order ord (NULL clears). Masked entries of every computed
 * data vector are 0 and carry no gradient. */
int clh_like_set_mask(clh_ctx* ctx, const int* mask, int n, const int ord[3]);
/* IP::set_data: the measured data vector (n entries, block order ord) */
int clh_like_set_data(clh_ctx* ctx, const double* data, int n, const int ord[3]);
/* IP::set_inv_cov after its inversion (generic_interface.cpp:2144-2278): the
 * inverse of the masked covariance (Hartlap factor included), n x n row-major
 * (symmetric), rows in the block order ord */
int clh_like_set_inv_cov(clh_ctx* ctx, const double* inv_cov, int n, const int ord[3]);
/* IP::get_chi2 (generic_interface.cpp:2288-2325) of one host data vector */
int clh_chi2(clh_ctx* ctx, const double* dv, int n, double* chi2);
/* chi^2 of every data vector of the uploaded batch (device dv [B][n] in the
 * block order ord, which must be the data's): chi2 [B] (device) and, if
 * grad != NULL, d chi2 / d dv [B][n] (device; zero on masked entries), the
 * cotangent to chain into clh_batch_vjp. */
int clh_batch_chi2(clh_ctx* ctx, const int ord[3], const double* dv_device, double* chi2_device,
                   double* grad_device);

int clh_datavector_size(clh_ctx* ctx);
/* compute_Mx2pt_N_masked<0,3>(ord) (generic_interface.hpp:1137-1159) with an
 * all-ones mask: xi+ | xi- | gamma_t | w blocks placed by ord, shear
 * calibration applied; dv has clh_datavector_size() entries (host memory) */
int clh_compute_3x2pt_real(clh_ctx* ctx, const int ord[3], double* dv, int n);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): clh_compute_3x2pt_real with every device buffer of the
 * draw (its tables, lens-efficiency tables, nodes, C_ell rows, descriptors,
 * work lists and the data vector) in a fixed 16 MiB array per GPU inside the
 * library; it calls no hipMalloc or hipFree.  The layout's device copy must
 * be current: call clh_compute_3x2pt_real once after the last layout setter
 * (CLH_ERR_STATE otherwise).  A layout and draw that need more than 16 MiB
 * are refused with CLH_ERR_ARG; 6x2pt and the non-Limber w are refused.
 * Calls that overlap on one GPU (several threads or contexts) race on the
 * shared array. */
int clh_compute_3x2pt_real_noalloc(clh_ctx* ctx, const int ord[3], double* dv, int n);
/* per-probe arrays as the reference's static caches hold them (q = pair*Ntheta + i),
 * without calibration: xi_pm_tomo (cosmo2D.c:353), w_gammat_tomo (:562; the
 * included pairs, clh_init_ggl_exclude), w_gg_tomo (:739; Limber or not as
 * clh_set_adopt_limber_gg selects) */
int clh_xi_pm(clh_ctx* ctx, double* xip, double* xim);
int clh_gammat(clh_ctx* ctx, double* out);
int clh_wtheta(clh_ctx* ctx, double* out);
/* Limber tables at the N_ell log-spaced ell (C_ss/gs/gg_tomo_limber tables,
 * cosmo2D.c:1617, 2195, 2693): ss[n_ss][N_ell], gs[n_gs][N_ell], gg[nl][N_ell] */
int clh_cl_tables(clh_ctx* ctx, double* ss, double* gs, double* gg);
/* C_cl_tomo's C_gg as w_gg_tomo(limber = 0) uses it: cl[nl][150] at ell
 * 0..149 (non-Limber below each bin's LMAX, Limber from there) and LMAX[nl]
 * (NULL = skip) */
int clh_cl_nonlimber(clh_ctx* ctx, double* cl, int* lmax);
/* the B-mode C_ss table, C_ss_tomo_limber(l, ni, nj, EE = 0) (cosmo2D.c:1617):
 * bb[n_ss][N_ell]; zero under NLA */
int clh_cl_tables_bb(clh_ctx* ctx, double* bb);
/* Direct GL C_ell at arbitrary ell (C_*_tomo_limber_nointerp,
 * cosmo2D.c:1562, 2125, 2672, 2960, 3217, 3449): probe 0=ss(EE), 1=gs,
 * 2=gg, 3=ss(BB), and with the 6x2pt probes 4=gk, 5=ks, 6=kk;
 * out[pair][nl] */
int clh_cl_nointerp(clh_ctx* ctx, int probe, const double* ell, int nl, double* out);

/* ------------------------------------------------ device-resident batches */
/* One draw's cosmology and nuisance (host pointers; NULL nuisance arrays
 * mean zero shift / unit stretch / zero m / zero b_mag; b1 must be given when
 * the layout has lens bins). The survey layout (n(z), binning, Ntable) is the
 * context's (clh_batch_upload) or a captured one (clh_batch_upload_layouts). */
typedef struct {
  double omega_m;
  const double* chi_z; const double* chi; int chi_n;
  const double* G_z; const double* G; int G_n;
  const double* log10k; const double* pk_z; const double* lnP; int nk; int npz;
  const double* src_dz; const double* lens_dz; const double* lens_stretch;
  const double* b1; const double* bmag; const double* shear_m;
  double ia_A; double ia_eta; const double* ia_bin;
  const double* pm;  /* point-mass amplitudes per lens bin (NULL = 0) */
  /* TATT only (ignored for NLA): A2 / b_TA as ia_A / ia_bin, and the
   * set_IA_PS tables PS [12][fpt_n] over [fpt_kmin, fpt_kmax) h/Mpc */
  double ia2_A; double ia2_eta; const double* ia2_bin;
  double bta_A; const double* bta_bin;
  const double* fpt; int fpt_n; double fpt_kmin; double fpt_kmax;
  /* non-Limber w only (clh_set_adopt_limber_gg(ctx, 0)): h = H0/100 and the
   * linear P(k,z) as clh_set_linear_power_spectrum takes it */
  double h;
  const double* plin_log10k; const double* plin_z; const double* lnPL; int plin_nk; int plin_nz;
  /* one-loop galaxy bias per lens bin (NULL = 0; bs2 NULL = -4/7 (b1 - 1)
   * where b2 != 0, as set_nuisance_nonlinear_bias sets it) and the
   * set_bias_PS tables PS [12][fptb_n] over [fptb_kmin, fptb_kmax) h/Mpc */
  const double* b2; const double* bs2; const double* b3; const double* bk;
  const double* fptb; int fptb_n; double fptb_kmin; double fptb_kmax; double fptb_sigma4;
} clh_draw;

/* uploads B draws into device memory owned by the context */
int clh_batch_upload(clh_ctx* ctx, int B, const clh_draw* draws);
/* Captured survey layouts, for batches whose draws differ in layout (number
 * of source / lens bins, n(z), theta binning).  clh_layout_capture snapshots
 * the context's current layout (the state set by clh_init_ntable,
 * clh_init_binning_real_space, clh_init_ia, clh_set_source_sample,
 * clh_set_lens_sample); later setter calls do not change the snapshot.  No
 * counterpart in the reference, whose layout is one set of global structs. */
typedef struct clh_layout clh_layout;
int clh_layout_capture(clh_ctx* ctx, clh_layout** out);
void clh_layout_destroy(clh_layout* layout);
int clh_layout_datavector_size(const clh_layout* layout);
/* uploads B draws, draw i with layout layouts[i] (the context keeps its own
 * copies; the caller may destroy the layouts afterwards).  Forward only:
 * clh_batch_set_params and clh_batch_vjp need every draw on one layout. */
int clh_batch_upload_layouts(clh_ctx* ctx, int B, const clh_draw* draws,
                             const clh_layout* const* layouts);
/* offset (doubles) of draw i's data vector in the batch output; i = B gives
 * the total length; -1 on a bad index */
long long clh_batch_datavector_offset(clh_ctx* ctx, int i);
/* runs the whole path for every uploaded draw on the context's stream;
 * dv_device: device pointer to clh_batch_datavector_offset(ctx, B) doubles
 * (B * clh_datavector_size() for clh_batch_upload), draw i at its offset */
int clh_batch_forward(clh_ctx* ctx, const int ord[3], double* dv_device);
int clh_synchronize(clh_ctx* ctx);
/* on != 0: clh_batch_forward captures its launches into a hipGraph on first
 * use and replays it afterwards (re-captured after clh_batch_upload) */
int clh_batch_use_graph(clh_ctx* ctx, int on);
/* launch every subsequent call on `stream` (a hipStream_t, e.g. the caller's
 * current stream; NULL = the HIP default stream) instead of the context's own
 * non-blocking stream */
int clh_set_stream(clh_ctx* ctx, void* stream);

/* Device-side update of the uploaded batch's differentiable inputs (device
 * pointers, [B][...] row-major, NULL = keep): lnP [B][nk*nz] (all draws must
 * share the table shape), src_dz [B][ns], lens_dz [B][nl], b1 [B][nl],
 * ia_A [B], shear_m [B][ns]. No counterpart in the reference, which has no
 * gradients; it serves the autograd layer. */
int clh_batch_set_params(clh_ctx* ctx, const double* lnP, const double* src_dz,
                         const double* lens_dz, const double* b1, const double* ia_A,
                         const double* shear_m);
/* Vector-Jacobian product of the batch's data vectors (blocks placed by ord,
 * as clh_batch_forward) with the cotangent dv_bar [B][n] (device). Writes, for
 * every non-NULL device pointer, d<dv_bar, dv>/d(input): g_lnP [B][nk*nz],
 * g_src_dz [B][ns], g_lens_dz [B][nl], g_b1 [B][nl], g_ia_A [B], g_m [B][ns].
 * Differentiates the discrete algorithm (no reference counterpart).  Refused
 * (CLH_ERR_ARG) for batches with TATT with more than 10 source bins, more
 * than one layout, custom ells, high_def_integration != 0 under TATT or
 * one-loop bias, and for the non-Limber w together with one-loop bias, TATT,
 * 6x2pt or
 * clh_set_vjp_lens_limits(1).  With 6x2pt the w_gk, w_ks and C_kk blocks
 * are included; with the non-Limber w its correction below ell 150.
 * Under TATT the gradient runs through A1 (ia_A, redshift evolution mode 3)
 * with A2, eta2 and b_TA held at their uploaded values. */
int clh_batch_vjp(clh_ctx* ctx, const int ord[3], const double* dv_bar, double* g_lnP,
                  double* g_src_dz, double* g_lens_dz, double* g_b1, double* g_ia_A, double* g_m);
/* Rows per work tile of the four Limber stages: rows[0] forward ss/gs,
 * rows[1] forward gg, rows[2] adjoint ss/gs, rows[3] adjoint gg (each
 * 1..64; the adjoint ss/gs tile is further capped so its cotangents fit in
 * 64 KB of LDS).  clh_autotune times 16, 32 and 64 rows for every stage, and
 * for the two gg stages with and without their chi(z) / G(z) / lens n(z)
 * tables staged in LDS, on the uploaded batch (HIP events, best of 3; the
 * adjoint stages only for a batch with one layout, with a unit cotangent),
 * keeps the fastest and returns the choice in rows_out and lds_out
 * ([forward gg, adjoint gg], 1 = staged; either may be NULL).  It overwrites
 * the batch's own output buffer and invalidates a captured graph.  No
 * reference counterpart. */
int clh_set_tile_rows(clh_ctx* ctx, const int rows[4]);
int clh_get_tile_rows(clh_ctx* ctx, int rows[4]);
int clh_autotune(clh_ctx* ctx, int rows_out[4], int lds_out[2]);
/* per-kernel timing of the last clh_batch_forward (and, after a
 * clh_batch_vjp, its two Limber adjoint stages, the ln P gather and the
 * lens-limit term of the lens-shift gradient: the finite-difference views, or
 * k_nodes_tan when the term comes from the nodes' motion, whose remainder is
 * then part of the adjoint stages) with
 * HIP events on the context stream: names[i], milliseconds[i]; returns the
 * kernel count */
int clh_last_kernel_times(clh_ctx* ctx, const char** names, float* ms, int cap);
int clh_enable_timing(clh_ctx* ctx, int on);
/* on != 0: batch forwards (one layout per batch) also store the gg integrand
 * of every (ell row, GL node) -- bracket t, P fac, dW_RSD/d dz, 3 doubles --
 * so a following clh_batch_vjp with unchanged parameters reads it instead of
 * evaluating the integrand again (nl x nrow x 288 doubles per draw of device
 * memory).  After a mixed-precision forward the checkpoint holds its FP32
 * P(k,z) products.  Default off.  No reference counterpart. */
int clh_set_vjp_checkpoint(clh_ctx* ctx, int on);
/* How clh_batch_vjp obtains the part of the lens-shift gradient that comes
 * from the lens integration limits, which move with |dz| (amin_lens /
 * amax_lens, redshift_spline.c:50-79), and the gs / gg GL nodes with them.
 * mode 0 (default): the derivative of the integrands along each node's
 * motion, inside the gs and gg adjoint kernels (exact for the discrete
 * forward away from the tables' knots).  mode 1: finite differences of two
 * views whose limits take |dz| +- 1e-8, a cross-check of mode 0 at about
 * twice the VJP's cost.  mode 2: left out (the gradient at fixed limits).
 * With mode 0 and the lens-shift gradient requested,
 * the gs and gg adjoints evaluate their integrands again even when a
 * checkpoint is present.  No reference counterpart. */
int clh_set_vjp_lens_limits(clh_ctx* ctx, int mode);

/* ---------------------------------------- reference-named scalar entries */
/* Identical signatures to cosmolike/cosmo2D.h:30-96, served from clh_default().
 * On an error they return NaN and record the message (clh_last_error on
 * clh_default()) instead of exiting.  xi_pm_tomo and w_gammat_tomo take
 * limber = 1 only; w_gg_tomo's limber selects the Limber (1) or non-Limber
 * (0) w for the default context, as clh_set_adopt_limber_gg does. */
double xi_pm_tomo(const int pm, const int nt, const int ni, const int nj, const int limber);
double w_gammat_tomo(const int nt, const int ni, const int nj, const int limber);
double w_gg_tomo(const int nt, const int ni, const int nj, const int limber);
double C_ss_tomo_limber(const double l, const int ni, const int nj, const int EE);
double C_gs_tomo_limber(const double l, const int ni, const int nj);
double C_gg_tomo_limber(const double l, const int ni, const int nj);
double C_ss_tomo_limber_nointerp(const double l, const int ni, const int nj, const int EE,
                                 const int init);
double C_gs_tomo_limber_nointerp(const double l, const int nl, const int ns, const int init);
double C_gg_tomo_limber_nointerp(const double l, const int ni, const int nj, const int init);
/* gs pair maps, redshift_spline.h (redshift_spline.c:115-223): test_zoverlap,
 * N_ggl (data-vector pair index of (lens, source), -1 if excluded), ZL / ZS
 * (lens / source bin of pair index n); -1 on a bin out of range */
int test_zoverlap(int ni, int nj);
int N_ggl(int ni, int nj);
int ZL(int ni);
int ZS(int nj);
/* 6x2pt CMB lensing (cosmo2D.h:43-45, 66-70, 98-102), after
 * init_probes("6x2pt"): w_gk_tomo / w_ks_tomo without calibration (limber =
 * 1 only), the C_gk / C_ks / C_kk tables and the direct integrals (init = 1
 * returns 0) */
double w_gk_tomo(const int nt, const int ni, const int limber);
double w_ks_tomo(const int nt, const int ni, const int limber);
double C_gk_tomo_limber(const double l, const int ni);
double C_ks_tomo_limber(const double l, const int ni);
double C_kk_limber(const double l);
double C_gk_tomo_limber_nointerp(const double l, const int nl, const int init);
double C_ks_tomo_limber_nointerp(const double l, const int ns, const int init);
double C_kk_limber_nointerp(const double l, const int init);
/* the integrands, cosmo2D.h:116-130 (clh_int_for_limber at one point) */
double int_for_C_ss_tomo_limber(double a, void* params);
double int_for_C_gs_tomo_limber(double a, void* params);
double int_for_C_gg_tomo_limber(double a, void* params);
double int_for_C_gk_tomo_limber(double a, void* params);
double int_for_C_ks_tomo_limber(double a, void* params);
double int_for_C_kk_limber(double a, void* params);

#ifdef __cplusplus
}
#endif
#endif
