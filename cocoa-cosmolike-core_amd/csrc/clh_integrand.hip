// The Limber integrands as functions of the scale factor (int_for_C_*_limber,
// cosmolike/cosmo2D.c): the same per-point physics the projection kernels
// sum over their GL nodes, here at caller-given a for one (probe, bins, ell),
// one thread per point.  Parameters follow the reference's arrays:
//   probe 0 ss {n1, n2, l, EE, deriv}          cosmo2D.c:1521-1555
//         1 gs {nl, ns, l, nonlinear_bias}     :2079-2119
//         2 gg {ni, nj, l, use_linear_ps, nonlinear_bias}  :2489-2606
//         3 gk {nl, l, nonlinear_bias}         :2857-2954
//         4 ks {ns, l}                         :3178-3211
//         5 kk {l}                             :3428-3446
#include "clh_device.hpp"

struct ClhIntArgs {
  int probe, i, j, f1, f2;
  double l;
};

// p_lin: cosmo3D.c:306-358 (bilinear in log10 k, z of ln P_lin)
__device__ __forceinline__ double clh_p_lin(const DrawDesc& d, double k, double a) {
  const double log10k = log10(k / CLH_COVERH0);
  const double z = 1.0 / a - 1.0;
  const int nz = d.nzL;
  const int i = clh_bs(d.lkL, d.nkL, log10k);
  const int j = clh_bs(d.pzL, nz, z);
  const double dx = (log10k - d.lkL[i]) / (d.lkL[i + 1] - d.lkL[i]);
  const double dy = (z - d.pzL[j]) / (d.pzL[j + 1] - d.pzL[j]);
  const double* P = d.lnPL + (size_t)i * nz + j;
  const double out = (1 - dx) * (1 - dy) * P[0] + (1 - dx) * dy * P[1] + dx * (1 - dy) * P[nz] + dx * dy * P[nz + 1];
  return exp(out) / (CLH_COVERH0 * CLH_COVERH0 * CLH_COVERH0);
}

__device__ double clh_int_for(const DrawDesc& d, const ClhIntArgs& q, double a) {
  double fK, dchida;
  clh_chi_all(d, a, fK, dchida);
  const double hh = clh_hoverh0v2(a, dchida);
  const double l = q.l, ell = l + 0.5, k = ell / fK, z = 1.0 / a - 1.0;
  const double wpre = 1.5 * d.omega_m * fK / a;  // W_kappa / g
  const double fac = dchida / (fK * fK);
  const double pref = l * (l + 1.) / (ell * ell);  // 1812.05995 eqs 74-79
  const double tmp = (l - 1.) * l * (l + 1.) * (l + 2.);
  const double pref2 = (tmp > 0) ? sqrt(tmp) / (ell * ell) : 0.0;
  const bool tatt = d.ia_model == CLH_IA_TATT;
  switch (q.probe) {
    case 0: {  // int_for_C_ss_tomo_limber_core (cosmo2D.c:1370-1515)
      const int n1 = q.i, n2 = q.j, EE = q.f1;
      const double g = clh_growfac(d, a);
      const double PK = clh_p_nonlin(d, k, a);
      const double ell4 = ell * ell * ell * ell;
      const double p4 = l * (l - 1.) * (l + 1.) * (l + 2.) / ell4;
      const double WK1 = wpre * clh_g_tomo(d, a, n1), WK2 = wpre * clh_g_tomo(d, a, n2);
      const double WS1 = clh_nz_src(d, z, n1) * hh, WS2 = clh_nz_src(d, z, n2) * hh;
      const double C11 = clh_ia_c1(d, a, g, n1), C12 = clh_ia_c1(d, a, g, n2);
      double ans;
      if (tatt) {
        const double g4 = g * g * g * g, lnk = log(k);
        const double C21 = clh_ia_c2(d, a, g, n1), C22 = clh_ia_c2(d, a, g, n2);
        const double bta1 = clh_ia_bta(d, n1), bta2 = clh_ia_bta(d, n2);
        if (EE == 1) {
          const double tt = g4 * clh_fpt(d, 0, lnk), ta_dE1 = g4 * clh_fpt(d, 2, lnk),
                       ta_dE2 = g4 * clh_fpt(d, 3, lnk), ta = g4 * clh_fpt(d, 4, lnk),
                       mixA = g4 * clh_fpt(d, 6, lnk), mixB = g4 * clh_fpt(d, 7, lnk),
                       mixEE = g4 * clh_fpt(d, 8, lnk);
          ans = WK1 * WK2 * PK - WS1 * WK2 * (C11 * PK + C11 * bta1 * (ta_dE1 + ta_dE2) - 5 * C21 * (mixA + mixB)) -
                WS2 * WK1 * (C12 * PK + C12 * bta2 * (ta_dE1 + ta_dE2) - 5 * C22 * (mixA + mixB)) +
                WS1 * WS2 *
                    (C11 * C12 * PK + C11 * C12 * (bta1 * bta2 * ta + (bta1 + bta2) * (ta_dE1 + ta_dE2)) -
                     5. * (C11 * C22 + C12 * C21) * (mixA + mixB) - 5. * (C11 * bta1 * C22 + C12 * bta2 * C21) * mixEE +
                     25. * C21 * C22 * tt);
        } else {
          const double tt = g4 * clh_fpt(d, 1, lnk), ta = g4 * clh_fpt(d, 5, lnk), mix = g4 * clh_fpt(d, 9, lnk);
          ans = WS1 * WS2 *
                (C11 * C12 * bta1 * bta2 * ta - 5. * (C11 * bta1 * C22 + C12 * bta2 * C21) * mix + 25. * C21 * C22 * tt);
        }
      } else {
        ans = EE == 1 ? WK1 * WK2 * PK - WS1 * WK2 * C11 * PK - WS2 * WK1 * C12 * PK + WS1 * WS2 * C11 * C12 * PK : 0.0;
      }
      return q.f2 == 1 ? ans * (dchida / fK) * p4 : ans * fac * p4;
    }
    case 1: {  // int_for_C_gs_tomo_limber_core (cosmo2D.c:1890-2073), no RSD (include_RSD_GS = 0)
      const int zl = q.i, zs = q.j;
      const double g = clh_growfac(d, a);
      const double PK = clh_p_nonlin(d, k, a);
      const double WK = wpre * clh_g_tomo(d, a, zs), WS = clh_nz_src(d, z, zs) * hh;
      const double WGAL = clh_nz_lens(d, z, zl) * hh, WMAG = wpre * clh_g_lens(d, a, zl);
      const double gal = WGAL * d.par->b1[zl] + WMAG * pref * d.par->bmag[zl];
      const double g4 = g * g * g * g;
      const double ol = q.f1 == 1 ? clh_oneloop_gs(d, zl, log(k), k, g4, PK) : 0.0;
      const double C1 = clh_ia_c1(d, a, g, zs);
      double ans;
      if (tatt) {
        const double lnk = log(k);
        const double mixA = g4 * clh_fpt(d, 6, lnk), mixB = g4 * clh_fpt(d, 7, lnk),
                     ta_dE1 = g4 * clh_fpt(d, 2, lnk), ta_dE2 = g4 * clh_fpt(d, 3, lnk);
        const double bta = clh_ia_bta(d, zs), C2 = clh_ia_c2(d, a, g, zs);
        ans = WK * (gal * PK + WGAL * ol) - WS * gal * (C1 * PK + C1 * bta * (ta_dE1 + ta_dE2) - 5 * C2 * (mixA + mixB));
      } else {
        ans = (WK - WS * C1) * (gal * PK + WGAL * ol);
      }
      return ans * fac * pref2;
    }
    case 2: {  // int_for_C_gg_tomo_limber (cosmo2D.c:2489-2606), RSD on, no HOD
      const int zl = q.i;
      if ((ell + 1.0) / k > d.chi_amin) return 0.0;  // chi_1 beyond chi(a_min)
      const double a0 = clh_a_chi(d, ell / k), a1 = clh_a_chi(d, (ell + 1.0) / k);
      const double WGAL = clh_nz_lens(d, z, zl) * hh, WMAG = wpre * clh_g_lens(d, a, zl);
      const double t = WGAL * d.par->b1[zl] + WMAG * pref * d.par->bmag[zl] + clh_w_rsd(d, ell, a0, a1, zl);
      const double PK = q.f1 == 1 ? clh_p_lin(d, k, a) : clh_p_nonlin(d, k, a);
      double ol = 0.0;
      if (q.f2 == 1 && q.f1 == 0) {
        const double g = clh_growfac(d, a);
        ol = WGAL * WGAL * clh_oneloop_gg(d, zl, log(k), k, g * g * g * g, PK);
      }
      return (t * t * PK + ol) * fac;
    }
    case 3: {  // int_for_C_gk_tomo_limber (cosmo2D.c:2857-2954), RSD off
      const int zl = q.i;
      const double WK = wpre * ((d.chi_cmb - fK) / d.chi_cmb);
      const double WGAL = clh_nz_lens(d, z, zl) * hh, WMAG = wpre * clh_g_lens(d, a, zl);
      const double PK = clh_p_nonlin(d, k, a);
      double r = WK * (WGAL * d.par->b1[zl] + WMAG * pref * d.par->bmag[zl]) * PK;
      if (q.f1 == 1) {
        const double g = clh_growfac(d, a);
        r += WK * WGAL * clh_oneloop_gs(d, zl, log(k), k, g * g * g * g, PK);
      }
      return r * fac * pref;
    }
    case 4: {  // int_for_C_ks_tomo_limber (cosmo2D.c:3178-3211)
      const int zs = q.i;
      const double g = clh_growfac(d, a);
      const double WK1 = wpre * clh_g_tomo(d, a, zs), WK2 = wpre * ((d.chi_cmb - fK) / d.chi_cmb);
      const double WS1 = clh_nz_src(d, z, zs) * hh * clh_ia_c1(d, a, g, zs);
      return ((WK1 - WS1) * WK2 * clh_p_nonlin(d, k, a) * fac) * pref * pref2;
    }
    default: {  // int_for_C_kk_limber (cosmo2D.c:3428-3446)
      const double WK = wpre * ((d.chi_cmb - fK) / d.chi_cmb);
      return WK * WK * clh_p_nonlin(d, k, a) * fac * pref * pref;
    }
  }
}

__global__ __launch_bounds__(256) void k_int_for(const DrawDesc* __restrict__ D, ClhIntArgs q,
                                                 const double* __restrict__ a, int n, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = clh_int_for(*D, q, a[t]);
}

// ----------------------------------------------------------------------------
// k_limber_hdi: the Limber C_ell rows of a draw at high_def_integration != 0
// (Ntable.high_def_integration): n-point GL of the integrands above over the
// reference's ranges, n = 96 / 128 / 256 / 512 for ss and gs and 128 / 256 /
// 512 / 1024 for gg (cosmo2D.c:1577-1581, 2145-2149, 2635-2639); ss on
// [1/(1 + z_max,src), 1/(1 + max(z_min,src, 1e-6))] (:1675-1676), gs and gg on
// [amin_lens, amax_lens] (:2159-2160, 2647-2648).  Work item (draw, column,
// row0, rows): column < n_ss + n_gs + nl is an ss (EE), gs or gg column of
// d.cl, column - (n_ss + n_gs + nl) a B-mode ss column of d.cl_bb (TATT).
// Rows go two at a time to 128 threads each (n <= 128) or one at a time to
// all 256; a thread sums its nodes in a fixed order and the row's sum is a
// fixed-order wave and LDS reduction, so repeated launches agree bit for bit.
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_limber_hdi(const DrawDesc* __restrict__ D, const int4* __restrict__ work) {
  __shared__ double part[4];
  const int4 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const int ncol = d.n_ss + d.n_gs + d.nl;
  const bool bb = w.y >= ncol;
  const int col = bb ? w.y - ncol : w.y, row0 = w.z, nr = w.w;
  ClhIntArgs q{0, 0, 0, 0, 0, 0.0};
  double amin, amax;
  int np;
  const double* gx;
  if (col < d.n_ss) {
    int i, j;
    clh_ss_pair(d.ns, col, i, j);
    q = ClhIntArgs{0, i, j, bb ? 0 : 1, 0, 0.0};
    amin = 1. / (d.s_zmax_all + 1.);
    amax = 1. / (1. + fmax(d.s_zmin_all, 1e-6));
    np = d.np_ss;
    gx = d.gl_hd;
  } else if (col < d.n_ss + d.n_gs) {
    const int zl = (col - d.n_ss) / d.ns, zs = (col - d.n_ss) % d.ns;
    q = ClhIntArgs{1, zl, zs, d.oneloop, 0, 0.0};
    amin = clh_amin_lens(d, zl);
    amax = clh_amax_lens(d, zl);
    np = d.np_ss;
    gx = d.gl_hd;
  } else {
    const int zl = col - d.n_ss - d.n_gs;
    q = ClhIntArgs{2, zl, zl, 0, d.oneloop, 0.0};
    amin = clh_amin_lens(d, zl);
    amax = clh_amax_lens(d, zl);
    np = d.np_gg;
    gx = d.gl_hd + 2 * d.np_ss;
  }
  const double A = 0.5 * (amax - amin), B = 0.5 * (amax + amin);
  const double* ells = col < d.n_ss ? d.ell_ss : d.ell_gx;
  double* out = bb ? d.cl_bb + (size_t)col * d.nrow : d.cl + (size_t)col * d.nrow;
  const int per = np <= 128 ? 128 : 256, rpb = 256 / per;  // threads per row, rows in flight
  const int lr = threadIdx.x % per, sub = threadIdx.x / per;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, wpr = per / 64;
  for (int r0 = 0; r0 < nr; r0 += rpb) {
    const int r = r0 + sub;
    double v = 0.0;
    if (r < nr) {
      ClhIntArgs qr = q;
      qr.l = ells[row0 + r];
      for (int n = lr; n < np; n += per) v += A * gx[np + n] * clh_int_for(d, qr, B + A * gx[n]);
    }
    v = clh_wave_sum(v);
    if (lane == 0) part[wv] = v;
    __syncthreads();
    if (threadIdx.x < rpb && r0 + (int)threadIdx.x < nr) {
      double t = 0.0;
      for (int k = 0; k < wpr; k++) t += part[threadIdx.x * wpr + k];
      out[row0 + r0 + threadIdx.x] = t;
    }
    __syncthreads();
  }
}

// ----------------------------------------------------------------------------
// k_limber_hdi_adj: adjoint of k_limber_hdi for NLA draws without one-loop
// terms (the VJP refuses the others at high_def_integration != 0).  Work
// item (draw, column, row0, rows) as the forward; thread = GL node n (n =
// tid, tid + 256, ...) walking the tile's rows.  Per (row, node) with c the
// row's cotangent: c dv/d ln P into the node's bilinear P(k,z) cell (a run
// per cell, added to g_lnP when the k cell changes; after the gather), and
// row sums that carry the node's radial weights to the parameters:
//   ss: v = A_1 A_2 P fac p4 w, A_b = W_K,b - C1_b W_S,b: source shifts
//       through dA_b = wpre dg_b - C1_b dW_S,b, A_IA through -c1 W_S,b;
//   gs: v = (W_K - C1 W_S)(W_gal b1 + W_mag pref b_mag) P fac pref2 w: the
//       source shift, A_IA, b1, and the lens shift through W_gal and W_mag;
//   gg: v = t^2 P fac w, t = W_gal b1 + W_mag pref b_mag + W_RSD: b1 and the
//       lens shift through W_gal, W_mag and the two n(z) of W_RSD.
// The motion of the gs / gg limits with |dz| comes from the
// finite-difference views.
// ----------------------------------------------------------------------------
// d W_RSD / d dz of the lens bin (clh_w_rsd with the n(z) shift tangents)
__device__ __forceinline__ double clh_w_rsd_ddz(const DrawDesc& d, double l, double a0, double a1, int ni) {
  double c0, dc0, c1, dc1;
  clh_chi_all(d, a0, c0, dc0);
  clh_chi_all(d, a1, c1, dc1);
  const double h0 = 1.0 / ((a0 * a0) * dc0);
  const double h1 = 1.0 / ((a1 * a1) * dc1);
  double w = (1 + 8. * l) / ((2. * l + 1.) * (2. * l + 1.)) * clh_dnz_lens(d, 1. / a0 - 1., ni) * h0 *
             clh_f_growth(d, 1.0 / a0 - 1.0);
  w -= 4. / (2 * l + 3.) * sqrt((2 * l + 1.) / (2 * l + 3.)) * clh_dnz_lens(d, 1. / a1 - 1., ni) * h1 *
       clh_f_growth(d, 1.0 / a1 - 1.0);
  return w;
}

__global__ __launch_bounds__(256) void k_limber_hdi_adj(const DrawDesc* __restrict__ D,
                                                        const int4* __restrict__ work) {
  __shared__ double red[4];
  __shared__ double cr[CLH_HDI_ROWS];
  const int4 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const int ncol = d.n_ss + d.n_gs + d.nl;
  const int col = w.y, row0 = w.z, nr = w.w;
  if (col >= ncol) return;  // (B-mode columns: TATT)
  int kind, bi, bj;  // ss: source bins bi <= bj; gs: lens bi, source bj; gg: lens bi
  double amin, amax;
  int np;
  const double* gx;
  if (col < d.n_ss) {
    kind = 0;
    clh_ss_pair(d.ns, col, bi, bj);
    amin = 1. / (d.s_zmax_all + 1.);
    amax = 1. / (1. + fmax(d.s_zmin_all, 1e-6));
    np = d.np_ss;
    gx = d.gl_hd;
  } else if (col < d.n_ss + d.n_gs) {
    kind = 1;
    bi = (col - d.n_ss) / d.ns;
    bj = (col - d.n_ss) % d.ns;
    amin = clh_amin_lens(d, bi);
    amax = clh_amax_lens(d, bi);
    np = d.np_ss;
    gx = d.gl_hd;
  } else {
    kind = 2;
    bi = bj = col - d.n_ss - d.n_gs;
    amin = clh_amin_lens(d, bi);
    amax = clh_amax_lens(d, bi);
    np = d.np_gg;
    gx = d.gl_hd + 2 * d.np_ss;
  }
  for (int t = threadIdx.x; t < nr; t += 256) cr[t] = d.cbar[(size_t)col * d.nrow + row0 + t];
  __syncthreads();
  const double* ells = kind == 0 ? d.ell_ss : d.ell_gx;
  const bool lnp = d.run != nullptr;  // the ln P cotangent is wanted
  const double A = 0.5 * (amax - amin), B = 0.5 * (amax + amin);
  const double b1 = kind ? d.par->b1[bi] : 0.0, bmag = kind ? d.par->bmag[bi] : 0.0;
  const int nz = d.nzp;
  double gi = 0.0, gj = 0.0, gl = 0.0, gb = 0.0, gA = 0.0;
  for (int n = threadIdx.x; n < np; n += 256) {
    const double a = B + A * gx[n], wn = A * gx[np + n];
    double fK, dchida;
    clh_chi_all(d, a, fK, dchida);
    const double hh = clh_hoverh0v2(a, dchida), z = 1.0 / a - 1.0;
    const double wpre = 1.5 * d.omega_m * fK / a, fac = dchida / (fK * fK);
    const double g = clh_growfac(d, a), c1u = clh_ia_c1_unit(d, a, g);
    // node constants: ss A_i, A_j (and their shift / A_IA tangents); gs, gg
    // the source and lens weights
    double X0 = 0, X1 = 0, dX0 = 0, dX1 = 0, aX0 = 0, aX1 = 0, WG = 0, WM = 0, dWG = 0, dWM = 0;
    if (kind == 0) {
      const double C1i = clh_ia_c1(d, a, g, bi), C1j = clh_ia_c1(d, a, g, bj);
      const double WSi = clh_nz_src(d, z, bi) * hh, WSj = clh_nz_src(d, z, bj) * hh;
      X0 = wpre * clh_g_tomo(d, a, bi) - C1i * WSi;
      X1 = wpre * clh_g_tomo(d, a, bj) - C1j * WSj;
      dX0 = wpre * clh_dg_tomo(d, a, bi) - C1i * clh_dnz_src(d, z, bi) * hh;
      dX1 = wpre * clh_dg_tomo(d, a, bj) - C1j * clh_dnz_src(d, z, bj) * hh;
      aX0 = -c1u * WSi;
      aX1 = -c1u * WSj;
    } else {
      WG = clh_nz_lens(d, z, bi) * hh;
      WM = wpre * clh_g_lens(d, a, bi);
      dWG = clh_dnz_lens(d, z, bi) * hh;
      dWM = wpre * clh_dg_lens(d, a, bi);
      if (kind == 1) {
        const double C1 = clh_ia_c1(d, a, g, bj), WS = clh_nz_src(d, z, bj) * hh;
        X0 = wpre * clh_g_tomo(d, a, bj) - WS * C1;
        dX0 = wpre * clh_dg_tomo(d, a, bj) - C1 * clh_dnz_src(d, z, bj) * hh;
        aX0 = -c1u * WS;
      }
    }
    double S0 = 0.0, S1 = 0.0;
    int ci = -1;
    double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;
    auto flush = [&]() {
      if (ci < 0) return;
      double* gp = d.g_lnP + ci;
      if (q0 != 0.0) atomicAdd(gp, q0);
      if (q1 != 0.0) atomicAdd(gp + 1, q1);
      if (q2 != 0.0) atomicAdd(gp + nz, q2);
      if (q3 != 0.0) atomicAdd(gp + nz + 1, q3);
    };
    for (int r = 0; r < nr; r++) {
      const double c = cr[r];
      if (c == 0.0) continue;
      const double l = ells[row0 + r], ell = l + 0.5, k = ell / fK;
      double gv;  // c dv / d ln P
      if (kind == 2 && (ell + 1.0) / k > d.chi_amin) continue;  // chi_1 beyond chi(a_min): zero
      int cell;
      double wc[4];
      const double PK = clh_p_nonlin_cells(d, k, a, cell, wc);
      const double pref = l * (l + 1.) / (ell * ell);
      if (kind == 0) {
        const double p4 = l * (l - 1.) * (l + 1.) * (l + 2.) / (ell * ell * ell * ell);
        const double s = c * PK * fac * p4 * wn;
        S0 += s;
        gv = s * X0 * X1;
      } else if (kind == 1) {
        const double tmp = (l - 1.) * l * (l + 1.) * (l + 2.);
        const double pref2 = (tmp > 0) ? sqrt(tmp) / (ell * ell) : 0.0;
        const double s = c * PK * fac * pref2 * wn;
        S0 += s;
        S1 += s * pref;
        gv = s * X0 * (WG * b1 + WM * pref * bmag);
      } else {
        const double a0 = clh_a_chi(d, ell / k), a1 = clh_a_chi(d, (ell + 1.0) / k);
        const double t = WG * b1 + WM * pref * bmag + clh_w_rsd(d, ell, a0, a1, bi);
        const double s = 2.0 * c * t * PK * fac * wn;
        gv = 0.5 * s * t;
        gb += s * WG;
        gl += s * (dWG * b1 + dWM * pref * bmag + clh_w_rsd_ddz(d, ell, a0, a1, bi));
      }
      if (lnp) {
        if (cell != ci) {
          flush();
          ci = cell;
          q0 = q1 = q2 = q3 = 0.0;
        }
        q0 += gv * wc[0];
        q1 += gv * wc[1];
        q2 += gv * wc[2];
        q3 += gv * wc[3];
      }
    }
    if (lnp) flush();
    if (kind == 0) {
      gi += S0 * X1 * dX0;
      gj += S0 * X0 * dX1;
      gA += S0 * (X1 * aX0 + X0 * aX1);
    } else if (kind == 1) {
      const double Gs = b1 * WG * S0 + bmag * WM * S1;
      gj += dX0 * Gs;
      gA += aX0 * Gs;
      gb += X0 * WG * S0;
      gl += X0 * (b1 * dWG * S0 + bmag * dWM * S1);
    }
  }
  gi = clh_block_sum(gi, red);
  gj = clh_block_sum(gj, red);
  gl = clh_block_sum(gl, red);
  gb = clh_block_sum(gb, red);
  gA = clh_block_sum(gA, red);
  if (threadIdx.x == 0) {
    if (kind == 0) {
      if (gi != 0.0) atomicAdd(d.g_src + bi, gi);
      if (gj != 0.0) atomicAdd(d.g_src + bj, gj);
    } else {
      if (kind == 1 && gj != 0.0) atomicAdd(d.g_src + bj, gj);
      if (gl != 0.0) atomicAdd(d.g_lens + bi, gl);
      if (gb != 0.0) atomicAdd(d.g_b1 + bi, gb);
    }
    if (gA != 0.0) atomicAdd(d.g_A, gA);
  }
}

hipError_t clh_launch_limber_hdi_adj(const DrawDesc* D, const int4* work, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_limber_hdi_adj, dim3(n), dim3(256), 0, s, D, work);
  return hipGetLastError();
}

hipError_t clh_launch_limber_hdi(const DrawDesc* D, const int4* work, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_limber_hdi, dim3(n), dim3(256), 0, s, D, work);
  return hipGetLastError();
}

hipError_t clh_launch_int_for(const DrawDesc* D, int probe, int i, int j, int f1, int f2, double l, const double* a,
                              int n, double* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  ClhIntArgs q{probe, i, j, f1, f2, l};
  hipLaunchKernelGGL(k_int_for, dim3((n + 255) / 256), dim3(256), 0, s, D, q, a, n, out);
  return hipGetLastError();
}
