// 6x2pt CMB lensing probes: galaxy x convergence (gk), shear x convergence
// (ks) and convergence auto (kk) Limber C_ell (int_for_C_gk/ks/kk_tomo_limber,
// cosmolike/cosmo2D.c:2857-3481), and their data-vector blocks: w_gk(theta)
// and w_ks(theta) with the CMB beam / pixel filter (w_gk_tomo, w_ks_tomo,
// cosmo2D.c:913-1296) and C_kk at like.ell (compute_X_N_masked<0,5>,
// generic_interface.hpp:1085-1094).  FP64 in every precision mode; these
// are 2 nl + ns + 1 columns against the 3x2pt's few hundred.
#include "clh_device.hpp"

// ----------------------------------------------------------------------------
// One GL node of a CMB-lensing Limber integral (probe 0 gk of lens bin `bin`,
// 1 ks of source bin `bin`, 2 kk): 96 points of the probe's a range
// (high_def_integration 0): gk amin_lens..amax_lens, ks amin_source..
// amax_source (redshift_spline.c:27-39), kk a_min (1 + 1e-5)..0.99999.
// cw = GL weight x dchi/da / fK^2; wA, wB the radial weights (gk: W_k W_gal,
// W_k W_mag; ks: W_k (W_K - W_S C1); kk: W_k^2); jz, dz the z cell of P(k,z).
// ----------------------------------------------------------------------------
struct ClhCmbNode {
  double a, fK, cw, wA, wB, dz, ol_g4;
  double z, hh, WK, wkp, gf;  // (for the adjoint: z, H factor, W_k, 1.5 Omega_m fK / a, D(a))
  int jz;
};

__device__ __forceinline__ ClhCmbNode clh_cmb_node(const DrawDesc& d, int probe, int bin, int p) {
  constexpr int np = CLH_NP_GG;
  ClhCmbNode n;
  n.a = 0.5; n.fK = 1.0; n.cw = 0.0; n.wA = 0.0; n.wB = 0.0; n.dz = 0.0; n.ol_g4 = 0.0;
  n.z = 1.0; n.hh = 0.0; n.WK = 0.0; n.wkp = 0.0; n.gf = 1.0; n.jz = 0;
  if (p >= np) return n;
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
  const double* glx = d.gl + 2 * CLH_NP_SS;
  const double* glw = glx + np;
  const double A = 0.5 * (amax - amin), B = 0.5 * (amin + amax);
  const double a = B + A * glx[p];
  double fK, dchida;
  clh_chi_all(d, a, fK, dchida);
  n.a = a;
  n.fK = fK;
  n.cw = A * glw[p] * dchida / (fK * fK);
  const double z = 1.0 / a - 1.0;
  const double WK = (1.5 * d.omega_m * fK / a) * ((d.chi_cmb - fK) / d.chi_cmb);
  const double hh = 1.0 / ((a * a) * dchida);
  n.z = z; n.hh = hh; n.WK = WK; n.wkp = 1.5 * d.omega_m * fK / a;
  if (probe == 0) {
    n.wA = WK * clh_nz_lens(d, z, bin) * hh;                        // W_k W_gal
    n.wB = WK * (1.5 * d.omega_m * fK / a) * clh_g_lens(d, a, bin);  // W_k W_mag
    if (d.oneloop) {
      const double g = clh_growfac(d, a);
      n.ol_g4 = g * g * g * g;
    }
  } else if (probe == 1) {
    const double gf = clh_growfac(d, a);
    n.gf = gf;
    n.wA = WK * ((1.5 * d.omega_m * fK / a) * clh_g_tomo(d, a, bin) - clh_nz_src(d, z, bin) * hh * clh_ia_c1(d, a, gf, bin));
  } else {
    n.wA = WK * WK;
  }
  const int j = clh_bs(d.pz, d.nzp, z);
  n.jz = j;
  n.dz = (z - d.pz[j]) / (d.pz[j + 1] - d.pz[j]);
  return n;
}

// The (row, node) term v of the integrand at multipole l, and for the
// adjoint (ADJ): g = dv/d ln P at the node's bilinear cell (k cell i, weight
// dx; z cell jz, dz), tA / tB = dv/d wA / dv/d wB and tb1 = dv/d b1 (gk; the
// one-loop terms at the uploaded bs2, cosmo2D.c:2925-2953).
struct ClhCmbTerm {
  double v, g, tA, tB, tb1, dx;
  int i;
};

template <bool ADJ>
__device__ __forceinline__ ClhCmbTerm clh_cmb_term(const DrawDesc& d, const ClhCmbNode& n, int probe, int bin,
                                                   double l) {
  ClhCmbTerm o;
  const double ell = l + 0.5;
  const double k = ell / n.fK;
  const double log10k = log10(k / CLH_COVERH0);
  const int i = clh_bs(d.lk, d.nk, log10k);
  const int j = n.jz;
  const double dx = (log10k - d.lk[i]) / (d.lk[i + 1] - d.lk[i]);
  const double* P = d.lnP + (size_t)i * d.nzp + j;
  const double dz = n.dz;
  const double lnp = (1 - dx) * (1 - dz) * P[0] + (1 - dx) * dz * P[1] + dx * (1 - dz) * P[d.nzp] +
                     dx * dz * P[d.nzp + 1];
  const double pk = exp(lnp) / (CLH_COVERH0 * CLH_COVERH0 * CLH_COVERH0);
  const double pref = l * (l + 1.) / (ell * ell);
  o.i = i;
  o.dx = dx;
  o.g = o.tA = o.tB = o.tb1 = 0.0;
  if (probe == 0) {
    const double b1 = d.par->b1[bin], bm = d.par->bmag[bin];
    double t = (n.wA * b1 + n.wB * pref * bm) * pk;
    double tp = t, ta = b1 * pk;  // d t / d ln P, d t / d wA
    if (d.oneloop) {  // cosmo2D.c:2925-2953
      const double lnk = log(k);
      const double q = 0.5 * n.ol_g4 *
                       (d.par->b2[bin] * clh_fptb(d, 0, lnk) + d.par->bs2[bin] * clh_fptb(d, 2, lnk) +
                        d.par->b3[bin] * clh_fptb(d, 5, lnk));
      const double kq = d.par->bk[bin] * k * k * pk;
      t += n.wA * (q + kq);
      if (ADJ) {
        tp += n.wA * kq;
        ta += q + kq;
      }
    }
    o.v = t * n.cw * pref;
    if (ADJ) {
      const double c = n.cw * pref;
      o.g = tp * c;
      o.tA = ta * c;
      o.tB = pref * bm * pk * c;
      o.tb1 = n.wA * pk * c;
    }
  } else if (probe == 1) {
    const double tmp = (l - 1.) * l * (l + 1.) * (l + 2.);
    const double pref2 = (tmp > 0) ? sqrt(tmp) / (ell * ell) : 0.0;
    o.v = n.wA * pk * n.cw * pref * pref2;
    if (ADJ) {
      o.g = o.v;
      o.tA = pk * n.cw * pref * pref2;
    }
  } else {
    o.v = n.wA * pk * n.cw * pref * pref;
    if (ADJ) o.g = o.v;
  }
  return o;
}

// column of the probe's C_ell in cl6: C_gk [nl], C_ks [ns], C_kk
__device__ __forceinline__ int clh_cmb_col(const DrawDesc& d, int probe, int bin) {
  return probe == 0 ? bin : (probe == 1 ? d.nl + bin : d.nl + d.ns);
}

// ----------------------------------------------------------------------------
// k_limber_cmb: one block of 128 threads per work item (draw, probe + 4 bin,
// row0, rows); thread p < 96 holds GL node p (clh_cmb_node); for every row
// the node terms are summed over the block.
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(128) void k_limber_cmb(const DrawDesc* __restrict__ D, const int4* __restrict__ work) {
  __shared__ double part[2];
  const int4 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const int probe = w.y & 3, bin = w.y >> 2, row0 = w.z, nr = w.w;
  const int p = threadIdx.x;
  const bool node = p < CLH_NP_GG;
  const ClhCmbNode n = clh_cmb_node(d, probe, bin, p);
  double* out = d.cl6 + (size_t)clh_cmb_col(d, probe, bin) * d.nrow6;
  for (int r = 0; r < nr; r++) {
    const int row = row0 + r;
    const double l = row < d.nrow ? d.ell_gx[row] : d.ell_f[row - d.nrow];
    double v = node ? clh_cmb_term<false>(d, n, probe, bin, l).v : 0.0;
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

// ----------------------------------------------------------------------------
// Adjoint (VJP) of the 6x2pt blocks.
// clh_ckk_adj: the cotangent u of C_kk(l) (clh_ckk) into its direct row or
// its two table rows of the kk column's cotangent Cb (atomics: band powers
// and ells share table rows).
// ----------------------------------------------------------------------------
__device__ __forceinline__ void clh_ckk_adj(const DrawDesc& d, double* Cb, double l, int direct, double u) {
  if (u == 0.0) return;
  if (l <= CLH_LMIN_TAB) {
    atomicAdd(Cb + d.nrow + direct, u);
    return;
  }
  double* f = Cb + CLH_NLOW;
  const double x = log(l);
  const int n = d.n_ell;
  if (x < d.lim_gx[0]) {
    atomicAdd(f, u);
    return;
  }
  const double r = (x - d.lim_gx[0]) / d.lim_gx[2];
  const int i0 = (int)floor(r);
  if (i0 + 1 >= n) {
    atomicAdd(f + n - 1, u);
    return;
  }
  const double t = r - i0;
  atomicAdd(f + i0, (1.0 - t) * u);
  atomicAdd(f + i0 + 1, t * u);
}

// ----------------------------------------------------------------------------
// k_cmb_place_adj: adjoint of k_cmb_place, one block per draw.  gk / ks row
// cotangents cbar6[col][r] = f sum_theta K[r][theta] u[col][theta] (u the
// masked data-vector cotangent, f = 1 + m_s on ks), the ks columns'
// cotangent-weighted uncalibrated sums sum_r (cbar6 / f) C into raw after
// the 3x2pt columns (d/dm, k_m_grad), and the kk cotangent through C_kk(l)
// at like.ell or, with band powers, u(L) = sum_j u_j M[j][L - lmin].
// cbar6 starts at zero.
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_cmb_place_adj(const DrawDesc* __restrict__ D, const int* __restrict__ work) {
  __shared__ double u[32];
  __shared__ double red[4];
  const DrawDesc& d = D[work[blockIdx.x]];
  const int nt = d.ntheta, nrow = d.nrow, nl = d.nl, ns = d.ns, tid = threadIdx.x;
  const int base = nt * (2 * d.n_ss + d.n_gs_dv + nl);  // canonical start of the gk block
  const int nw = (nl + ns) * nt;
  for (int col = 0; col < nl + ns; col++) {
    const bool ks = col >= nl;
    const int zs = col - nl;
    const int s0 = ks ? d.dv6_start[1] + nt * zs : d.dv6_start[0] + nt * col;
    if (tid < nt) u[tid] = clh_masked(d, base + col * nt + tid) ? 0.0 : d.dvbar[s0 + tid];
    __syncthreads();
    const double* K = d.kf6 + (size_t)(ks ? 1 : 0) * nrow * nt;
    const double* C = d.cl6 + (size_t)col * d.nrow6;
    double* cb = d.cbar6 + (size_t)col * d.nrow6;
    const double f = (ks && d.apply_calib) ? 1.0 + d.par->shear_m[zs] : 1.0;
    double tr = 0.0;
    for (int r = tid; r < nrow; r += 256) {
      double v = 0.0;
      for (int th = 0; th < nt; th++) v += K[(size_t)r * nt + th] * u[th];
      cb[r] = f * v;
      tr += v * C[r];
    }
    if (ks) {
      const double T = clh_block_sum(tr, red);
      if (tid == 0) d.raw[d.n_ss + d.n_gs + nl + zs] = T;
    }
    __syncthreads();
  }
  double* Cb = d.cbar6 + (size_t)(nl + ns) * d.nrow6;
  if (d.kk_nbp > 0) {
    for (int i = tid; i < d.kk_nL; i += 256) {
      double ub = 0.0;
      for (int j = 0; j < d.kk_nbp; j++)
        if (!clh_masked(d, base + nw + j)) ub += d.dvbar[d.dv6_start[2] + j] * d.kk_M[(size_t)j * d.kk_nL + i];
      clh_ckk_adj(d, Cb, (double)(d.kk_lmin + i), i, ub);
    }
  } else {
    for (int i = tid; i < d.n_f; i += 256)
      clh_ckk_adj(d, Cb, d.ell_f[i], i, clh_masked(d, base + nw + i) ? 0.0 : d.dvbar[d.dv6_start[2] + i]);
  }
}

// sum over a block of 128 threads (two waves)
__device__ __forceinline__ double clh_sum128(double v, double* lds) {
  v = clh_wave_sum(v);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  const double s = lds[0] + lds[1];
  __syncthreads();
  return s;
}

// ----------------------------------------------------------------------------
// k_limber_cmb_adj: adjoint of k_limber_cmb over the same work items.  Per
// (row, node) with c = cbar6 of the row: the ln P cotangent c dv/d ln P into
// the node's bilinear cells (a run per k cell, added to g_lnP when the cell
// changes; after the gather), and per node the sums of c dv/d wA, c dv/d wB,
// c dv/d b1; then gk: lens shift through W_gal = n_lens hh and W_mag =
// 1.5 Omega_m fK / a g_lens (at fixed nodes: the motion of the gk limits
// with |dz| comes from the finite-difference views), b1; ks: source shift
// through g_tomo and n_src, A_IA through C1.
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(128) void k_limber_cmb_adj(const DrawDesc* __restrict__ D,
                                                        const int4* __restrict__ work) {
  __shared__ double red[2];
  const int4 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const int probe = w.y & 3, bin = w.y >> 2, row0 = w.z, nr = w.w;
  const int p = threadIdx.x;
  const bool node = p < CLH_NP_GG;
  const ClhCmbNode n = clh_cmb_node(d, probe, bin, p);
  const double* cb = d.cbar6 + (size_t)clh_cmb_col(d, probe, bin) * d.nrow6;
  const bool lnp = d.run != nullptr;  // the ln P cotangent is wanted
  double sA = 0.0, sB = 0.0, sb1 = 0.0;
  int ci = -1;
  double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;  // the run in k cell ci: (ci, jz), (ci, jz+1), (ci+1, jz), (ci+1, jz+1)
  const size_t nzp = d.nzp;
  double* gl = d.g_lnP + n.jz;
  auto flush = [&]() {
    if (ci < 0) return;
    double* g = gl + (size_t)ci * nzp;
    if (q0 != 0.0) atomicAdd(g, q0);
    if (q1 != 0.0) atomicAdd(g + 1, q1);
    if (q2 != 0.0) atomicAdd(g + nzp, q2);
    if (q3 != 0.0) atomicAdd(g + nzp + 1, q3);
  };
  if (node) {
    for (int r = 0; r < nr; r++) {
      const int row = row0 + r;
      const double c = cb[row];
      if (c == 0.0) continue;
      const double l = row < d.nrow ? d.ell_gx[row] : d.ell_f[row - d.nrow];
      const ClhCmbTerm t = clh_cmb_term<true>(d, n, probe, bin, l);
      sA += c * t.tA;
      sB += c * t.tB;
      sb1 += c * t.tb1;
      if (lnp) {
        if (t.i != ci) {
          flush();
          ci = t.i;
          q0 = q1 = q2 = q3 = 0.0;
        }
        const double g = c * t.g;
        q0 += g * (1 - t.dx) * (1 - n.dz);
        q1 += g * (1 - t.dx) * n.dz;
        q2 += g * t.dx * (1 - n.dz);
        q3 += g * t.dx * n.dz;
      }
    }
    if (lnp) flush();
  }
  if (probe == 2) return;  // kk: ln P only (uniform per block)
  double gp = 0.0, gA = 0.0;
  if (node) {
    if (probe == 0) {
      gp = n.WK * (sA * clh_dnz_lens(d, n.z, bin) * n.hh + sB * n.wkp * clh_dg_lens(d, n.a, bin));
    } else {
      const double dWK = n.wkp * clh_dg_tomo(d, n.a, bin), WS = clh_nz_src(d, n.z, bin) * n.hh;
      const double dWS = clh_dnz_src(d, n.z, bin) * n.hh;
      gp = n.WK * sA * (dWK - clh_ia_c1(d, n.a, n.gf, bin) * dWS);
      gA = -n.WK * sA * WS * clh_ia_c1_unit(d, n.a, n.gf);
    }
  }
  gp = clh_sum128(gp, red);
  if (probe == 0) {
    const double b = clh_sum128(sb1, red);
    if (p == 0) {
      if (gp != 0.0) atomicAdd(d.g_lens + bin, gp);
      if (b != 0.0) atomicAdd(d.g_b1 + bin, b);
    }
  } else {
    gA = clh_sum128(gA, red);
    if (p == 0) {
      if (gp != 0.0) atomicAdd(d.g_src + bin, gp);
      if (gA != 0.0) atomicAdd(d.g_A, gA);
    }
  }
}

hipError_t clh_launch_cmb(const DrawDesc* D, const int4* work, int n, const int* wplace, int nplace, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_limber_cmb, dim3(n), dim3(128), 0, s, D, work);
  if (nplace > 0) hipLaunchKernelGGL(k_cmb_place, dim3(nplace), dim3(256), 0, s, D, wplace);
  return hipGetLastError();
}
hipError_t clh_launch_cmb_place_adj(const DrawDesc* D, const int* wplace, int nplace, hipStream_t s) {
  if (nplace <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_cmb_place_adj, dim3(nplace), dim3(256), 0, s, D, wplace);
  return hipGetLastError();
}
hipError_t clh_launch_limber_cmb_adj(const DrawDesc* D, const int4* work, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_limber_cmb_adj, dim3(n), dim3(128), 0, s, D, work);
  return hipGetLastError();
}
