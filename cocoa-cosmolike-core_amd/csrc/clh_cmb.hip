// 6x2pt CMB lensing probes: galaxy x convergence (gk), shear x convergence
// (ks) and convergence auto (kk) Limber C_ell (int_for_C_gk/ks/kk_tomo_limber,
// cosmolike/cosmo2D.c:2857-3481), and their data-vector blocks: w_gk(theta)
// and w_ks(theta) with the CMB beam / pixel filter (w_gk_tomo, w_ks_tomo,
// cosmo2D.c:913-1296) and C_kk at like.ell (compute_X_N_masked<0,5>,
// generic_interface.hpp:1085-1094).  FP64 in every precision mode; these
// are 2 nl + ns + 1 columns against the 3x2pt's few hundred.
#include "clh_device.hpp"

// ----------------------------------------------------------------------------
// k_limber_cmb: one block of 128 threads per work item (draw, probe + 4 bin,
// row0, rows), probe 0 gk (lens bin), 1 ks (source bin), 2 kk.  Thread p < 96
// holds GL node p of the probe's a range (96 points, high_def_integration 0):
// gk amin_lens..amax_lens, ks amin_source..amax_source (redshift_spline.c:
// 27-39), kk a_min (1 + 1e-5)..0.99999; for every row the node values are
// summed over the block.
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(128) void k_limber_cmb(const DrawDesc* __restrict__ D, const int4* __restrict__ work) {
  __shared__ double part[2];
  const int4 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const int probe = w.y & 3, bin = w.y >> 2, row0 = w.z, nr = w.w;
  const int p = threadIdx.x;
  constexpr int np = CLH_NP_GG;
  const bool node = p < np;
  double amin, amax;
  if (probe == 0) {
    amin = clh_amin_lens(d, bin);
    amax = clh_amax_lens(d, bin);
  } else if (probe == 1) {
    amin = 1. / (d.s_zmax_all + 1.);
    amax = 1. / (1. + fmax(d.s_zmin_all, 0.001));
  } else {
    amin = CLH_A_MIN * (1. + 1.e-5);
    amax = 0.99999;
  }
  // node constants: weight x dchi/da / fK^2 and the radial weights
  double a = 0.5, fK = 1.0, cw = 0.0, wA = 0.0, wB = 0.0, jz = 0.0, dz = 0.0;
  if (node) {
    const double* glx = d.gl + 2 * CLH_NP_SS;
    const double* glw = glx + np;
    const double A = 0.5 * (amax - amin), B = 0.5 * (amin + amax);
    a = B + A * glx[p];
    double dchida;
    clh_chi_all(d, a, fK, dchida);
    cw = A * glw[p] * dchida / (fK * fK);
    const double z = 1.0 / a - 1.0;
    const double WK = (1.5 * d.omega_m * fK / a) * ((d.chi_cmb - fK) / d.chi_cmb);
    const double hh = 1.0 / ((a * a) * dchida);
    if (probe == 0) {
      wA = WK * clh_nz_lens(d, z, bin) * hh;                        // W_k W_gal
      wB = WK * (1.5 * d.omega_m * fK / a) * clh_g_lens(d, a, bin);  // W_k W_mag
    } else if (probe == 1) {
      const double gf = clh_growfac(d, a);
      wA = WK * ((1.5 * d.omega_m * fK / a) * clh_g_tomo(d, a, bin) - clh_nz_src(d, z, bin) * hh * clh_ia_c1(d, a, gf, bin));
    } else {
      wA = WK * WK;
    }
    const int j = clh_bs(d.pz, d.nzp, z);
    jz = j;
    dz = (z - d.pz[j]) / (d.pz[j + 1] - d.pz[j]);
  }
  double ol_g4 = 0.0;
  if (node && probe == 0 && d.oneloop) {
    const double g = clh_growfac(d, a);
    ol_g4 = g * g * g * g;
  }
  const int ncol = probe == 0 ? bin : (probe == 1 ? d.nl + bin : d.nl + d.ns);
  double* out = d.cl6 + (size_t)ncol * d.nrow6;
  for (int r = 0; r < nr; r++) {
    const int row = row0 + r;
    const double l = row < d.nrow ? d.ell_gx[row] : d.ell_f[row - d.nrow];
    const double ell = l + 0.5;
    double v = 0.0;
    if (node) {
      const double k = ell / fK;
      const double log10k = log10(k / CLH_COVERH0);
      const int i = clh_bs(d.lk, d.nk, log10k);
      const int j = (int)jz;
      const double dx = (log10k - d.lk[i]) / (d.lk[i + 1] - d.lk[i]);
      const double* P = d.lnP + (size_t)i * d.nzp + j;
      const double lnp = (1 - dx) * (1 - dz) * P[0] + (1 - dx) * dz * P[1] + dx * (1 - dz) * P[d.nzp] +
                         dx * dz * P[d.nzp + 1];
      const double pk = exp(lnp) / (CLH_COVERH0 * CLH_COVERH0 * CLH_COVERH0);
      const double pref = l * (l + 1.) / (ell * ell);
      if (probe == 0) {
        double t = (wA * d.par->b1[bin] + wB * pref * d.par->bmag[bin]) * pk;
        if (d.oneloop) {  // cosmo2D.c:2925-2953
          const double lnk = log(k);
          t += wA * (0.5 * ol_g4 *
                         (d.par->b2[bin] * clh_fptb(d, 0, lnk) + d.par->bs2[bin] * clh_fptb(d, 2, lnk) +
                          d.par->b3[bin] * clh_fptb(d, 5, lnk)) +
                     d.par->bk[bin] * k * k * pk);
        }
        v = t * cw * pref;
      } else if (probe == 1) {
        const double tmp = (l - 1.) * l * (l + 1.) * (l + 2.);
        const double pref2 = (tmp > 0) ? sqrt(tmp) / (ell * ell) : 0.0;
        v = wA * pk * cw * pref * pref2;
      } else {
        v = wA * pk * cw * pref * pref;
      }
    }
    v = clh_wave_sum(v);
    if ((p & 63) == 0) part[p >> 6] = v;
    __syncthreads();
    if (p == 0) out[row] = part[0] + part[1];
    __syncthreads();
  }
}

// C_kk at ell from the kk column C: the direct row `direct` (ell <= LMIN_tab)
// or the table's linear interpolation in ln ell (C_kk_limber,
// cosmo2D.c:3484-3516)
__device__ __forceinline__ double clh_ckk(const DrawDesc& d, const double* C, double l, int direct) {
  if (l <= CLH_LMIN_TAB) return C[d.nrow + direct];
  const double* f = C + CLH_NLOW;
  const double x = log(l);
  const int n = d.n_ell;
  if (x < d.lim_gx[0]) return f[0];
  const double r = (x - d.lim_gx[0]) / d.lim_gx[2];
  const int i0 = (int)floor(r);
  return (i0 + 1 >= n) ? f[n - 1] : (r - i0) * (f[i0 + 1] - f[i0]) + f[i0];
}

// ----------------------------------------------------------------------------
// k_cmb_place: the gk, ks and kk blocks of one draw.  w_gk / w_ks: the folded
// kernels (w and gamma_t kernels times the CMB filter, ln-ell fill folded in)
// against the C_gk / C_ks rows; (1 + m_s) on ks; the kk block either C_kk at
// like.ell or, with band powers (init_cmb_auto_bandpower), the block sums
// sum_L M[j][L - lmin] C_kk(L) - offset[j] over integer L
// (generic_interface.hpp:1095-1118).  Masked entries are 0.
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_cmb_place(const DrawDesc* __restrict__ D, const int* __restrict__ work) {
  const DrawDesc& d = D[work[blockIdx.x]];
  const int nt = d.ntheta, nrow = d.nrow;
  const int base = nt * (2 * d.n_ss + d.n_gs_dv + d.nl);  // canonical start of the gk block
  const int nw = (d.nl + d.ns) * nt;
  const double* Ckk = d.cl6 + (size_t)(d.nl + d.ns) * d.nrow6;
  const int nf = d.kk_nbp > 0 ? 0 : d.n_f;
  for (int t = threadIdx.x; t < nw + nf; t += 256) {
    double v;
    int idx, canon;
    if (t < nw) {
      const int col = t / nt, th = t % nt;
      const bool ks = col >= d.nl;
      const double* K = d.kf6 + (size_t)(ks ? 1 : 0) * nrow * nt + th;
      const double* C = d.cl6 + (size_t)col * d.nrow6;
      v = 0.0;
      for (int r = 0; r < nrow; r++) v += K[(size_t)r * nt] * C[r];
      if (ks) {
        const int zs = col - d.nl;
        if (d.apply_calib) v *= 1.0 + d.par->shear_m[zs];
        idx = d.dv6_start[1] + nt * zs + th;
      } else {
        idx = d.dv6_start[0] + nt * col + th;
      }
      canon = base + t;
    } else {
      const int i = t - nw;
      v = clh_ckk(d, Ckk, d.ell_f[i], i);
      idx = d.dv6_start[2] + i;
      canon = base + nw + i;
    }
    d.dv[idx] = clh_masked(d, canon) ? 0.0 : v;
  }
  if (d.kk_nbp > 0) {  // band powers: one block-wide sum over L per band
    __shared__ double part[4];
    for (int j = 0; j < d.kk_nbp; j++) {
      const double* M = d.kk_M + (size_t)j * d.kk_nL;
      double v = 0.0;
      for (int i = threadIdx.x; i < d.kk_nL; i += 256) {
        const int L = d.kk_lmin + i;
        v += clh_ckk(d, Ckk, (double)L, i) * M[i];  // direct row i holds L = lmin + i
      }
      v = clh_wave_sum(v);
      if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
      __syncthreads();
      if (threadIdx.x == 0) {
        const double bp = (part[0] + part[1]) + (part[2] + part[3]) - d.kk_off[j];
        d.dv[d.dv6_start[2] + j] = clh_masked(d, base + nw + j) ? 0.0 : bp;
      }
      __syncthreads();
    }
  }
}

hipError_t clh_launch_cmb(const DrawDesc* D, const int4* work, int n, const int* wplace, int nplace, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_limber_cmb, dim3(n), dim3(128), 0, s, D, work);
  if (nplace > 0) hipLaunchKernelGGL(k_cmb_place, dim3(nplace), dim3(256), 0, s, D, wplace);
  return hipGetLastError();
}
