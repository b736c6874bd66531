// Adjoint (vector-Jacobian product) kernels of the FP64 projection path.
//
// The reference has no gradient code; these differentiate the discrete
// algorithm the forward kernels implement (GL node sums, bilinear ln P
// lookups, trapezoid prefix sums of the lens efficiencies, natural splines of
// n(z), the folded Legendre contraction, shear calibration) with respect to
// the ln P(k,z) table, the source and lens photo-z shifts, the linear galaxy
// bias, the NLA amplitude A_IA and the shear calibration m.  Quantities at
// the Limber nodes are recomputed, not stored.
//
//  k_set_params   device-side update of per-draw parameters (autograd inputs)
//  k_leg_adjoint  C_ell-row cotangents from the data-vector cotangent
//  k_m_grad       d/dm from the uncalibrated data vector
//  k_limber_adj   ln P scatter + cotangents of the node weights
//  k_nodes_adj    node weights -> photo-z shifts (n(z) and g tangents), A_IA
#include <hip/hip_runtime.h>

#include <algorithm>

#include "clh_device.hpp"
#include "clh_internal.hpp"

#define CLH_BLOCK 256

__global__ void k_fill(double* __restrict__ p, size_t n, double v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// one block per draw; arrays are [B][...] with the batch's uniform shapes
__global__ __launch_bounds__(CLH_BLOCK) void k_set_params(const DrawDesc* __restrict__ D,
                                                          DrawParams* __restrict__ P, int nknz,
                                                          const double* __restrict__ lnP,
                                                          const double* __restrict__ src_dz,
                                                          const double* __restrict__ lens_dz,
                                                          const double* __restrict__ b1,
                                                          const double* __restrict__ ia_A,
                                                          const double* __restrict__ shear_m) {
  const int b = blockIdx.x;
  const DrawDesc& d = D[b];
  DrawParams& p = P[b];
  if (lnP)
    for (int i = threadIdx.x; i < nknz; i += CLH_BLOCK) d.lnP[i] = lnP[(size_t)b * nknz + i];
  const int t = threadIdx.x;
  if (t < d.ns) {
    if (src_dz) p.src_dz[t] = src_dz[(size_t)b * d.ns + t];
    if (shear_m) p.shear_m[t] = shear_m[(size_t)b * d.ns + t];
  }
  if (t < d.nl) {
    if (lens_dz) p.lens_dz[t] = lens_dz[(size_t)b * d.nl + t];
    if (b1) p.b1[t] = b1[(size_t)b * d.nl + t];
  }
  if (t == 0 && ia_A) p.ia_A = ia_A[b];
}

// ----------------------------------------------------------------------------
// k_leg_adjoint: work item (draw, first column); 8 columns per block.
// cbar[col][r] = sum_tasks sum_theta Kfold[kind][r][theta] * dvbar * calibration
// (the FULL-mode forward has the same derivative with respect to the rows).
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(CLH_BLOCK) void k_leg_adjoint(const DrawDesc* __restrict__ D,
                                                           const int2* __restrict__ work) {
  __shared__ double vb[8][2][32];
  const int2 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const int nt = d.ntheta, nrow = d.nrow, nss = d.n_ss, ngs = d.n_gs;
  const int ncol = nss + ngs + d.nl;
  for (int i = threadIdx.x; i < 8 * 2 * 32; i += CLH_BLOCK) {
    const int c = i >> 6, s = (i >> 5) & 1, th = i & 31;
    const int col = w.y + c;
    double v = 0.0;
    if (col < ncol && th < nt) {
      if (col < nss) {
        int a, b;
        clh_ss_pair(d.ns, col, a, b);
        const double f = d.apply_calib ? (1.0 + d.par->shear_m[a]) * (1.0 + d.par->shear_m[b]) : 1.0;
        v = d.dvbar[d.dv_start[0] + nt * col + th + (s ? nt * nss : 0)] * f;
      } else if (s == 0 && col < nss + ngs) {
        const int q = col - nss;
        const double f = d.apply_calib ? 1.0 + d.par->shear_m[q % d.ns] : 1.0;
        v = d.dvbar[d.dv_start[1] + nt * q + th] * f;
      } else if (s == 0) {
        v = d.dvbar[d.dv_start[2] + nt * (col - nss - ngs) + th];
      }
    }
    vb[c][s][th] = v;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 8 * nrow; i += CLH_BLOCK) {
    const int c = i / nrow, r = i % nrow;
    const int col = w.y + c;
    if (col >= ncol) continue;
    const int kind0 = col < nss ? 0 : (col < nss + ngs ? 2 : 3);
    const double* K0 = d.kfold + ((size_t)kind0 * nrow + r) * nt;
    double acc = 0.0;
    for (int th = 0; th < nt; th++) acc += K0[th] * vb[c][0][th];
    if (col < nss) {
      const double* K1 = d.kfold + ((size_t)nrow + r) * nt;
      for (int th = 0; th < nt; th++) acc += K1[th] * vb[c][1][th];
    }
    d.cbar[(size_t)col * nrow + r] = acc;
  }
}

// ----------------------------------------------------------------------------
// k_m_grad: one block per draw, one thread per source bin.
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_m_grad(const DrawDesc* __restrict__ D) {
  const DrawDesc& d = D[blockIdx.x];
  const int i = threadIdx.x;
  if (i >= d.ns || !d.apply_calib) {
    if (i < d.ns) d.g_m[i] = 0.0;
    return;
  }
  const int nt = d.ntheta, nss = d.n_ss;
  const double* m = d.par->shear_m;
  double g = 0.0;
  for (int q = 0; q < nss; q++) {
    int a, b;
    clh_ss_pair(d.ns, q, a, b);
    if (a != i && b != i) continue;
    const double der = (a == b) ? 2.0 * (1.0 + m[i]) : 1.0 + m[a == i ? b : a];
    double s = 0.0;
    for (int th = 0; th < nt; th++) {
      const int ip = d.dv_start[0] + nt * q + th;
      const int im = ip + nt * nss;
      s += d.dvbar[ip] * d.raw[ip] + d.dvbar[im] * d.raw[im];
    }
    g += der * s;
  }
  for (int zl = 0; zl < d.nl; zl++) {
    const int q = zl * d.ns + i;
    for (int th = 0; th < nt; th++) {
      const int idx = d.dv_start[1] + nt * q + th;
      g += d.dvbar[idx] * d.raw[idx];
    }
  }
  d.g_m[i] = g;
}

// ----------------------------------------------------------------------------
// k_limber_adj: same work items as k_limber (draw, kind, zl, row0).
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(CLH_BLOCK) void k_limber_adj(const DrawDesc* __restrict__ D,
                                                          const int4* __restrict__ work) {
  __shared__ double S[CLH_GG_ROWS * CLH_NP_GG];
  __shared__ double W[(CLH_MAXB + 2) * CLH_NP_SS];
  __shared__ double Cb[CLH_MAXB * (CLH_MAXB + 1) / 2 * CLH_SS_ROWS];
  __shared__ double pr[3][CLH_GG_ROWS];
  __shared__ double red[8];
  const int4 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const int kind = w.y, zl = w.z, row0 = w.w;
  const int tid = threadIdx.x;
  const int nrow = d.nrow;

  if (kind == CLH_W_SS || kind == CLH_W_GS) {
    const int np = CLH_NP_SS;
    const double* o = (kind == CLH_W_SS) ? d.nodes + d.node_ss
                                         : d.nodes + d.node_gs + (size_t)zl * d.node_gs_stride;
    const double* ells = (kind == CLH_W_SS) ? d.ell_ss : d.ell_gx;
    const int rr_own = tid >> 6, p_own = tid & 63;
    const int r_own = row0 + rr_own;
    double pk = 0.0, wts[4] = {0, 0, 0, 0};
    int cell = 0;
    if (r_own < nrow) {
      const double ell = ells[r_own] + 0.5;
      pk = clh_p_nonlin_cells(d, ell / o[np + p_own], o[p_own], cell, wts);
    }
    S[rr_own * np + p_own] = pk * o[2 * np + p_own];
    const int nw = (kind == CLH_W_SS) ? d.ns : d.ns + 2;
    for (int i = tid; i < nw * np; i += CLH_BLOCK) W[i] = o[3 * np + i];
    const int npair = (kind == CLH_W_SS) ? d.n_ss : d.ns;
    const int col0 = (kind == CLH_W_SS) ? 0 : d.n_ss + zl * d.ns;
    for (int i = tid; i < npair * CLH_SS_ROWS; i += CLH_BLOCK) {
      const int q = i / CLH_SS_ROWS, rr = i % CLH_SS_ROWS;
      const int r = row0 + rr;
      Cb[i] = (r < nrow) ? d.cbar[(size_t)(col0 + q) * nrow + r] : 0.0;
    }
    if (tid < CLH_SS_ROWS) {
      const int r = row0 + tid;
      double p0 = 0, p1 = 0, p2 = 0;
      if (r < nrow) {
        const double l = ells[r], ell = l + 0.5, ell4 = ell * ell * ell * ell;
        p0 = l * (l - 1.) * (l + 1.) * (l + 2.) / ell4;
        p1 = l * (l + 1.) / (ell * ell);
        const double tmp = (l - 1.) * l * (l + 1.) * (l + 2.);
        p2 = (tmp > 0) ? sqrt(tmp) / (ell * ell) : 0.0;
      }
      pr[0][tid] = p0; pr[1][tid] = p1; pr[2][tid] = p2;
    }
    __syncthreads();
    if (kind == CLH_W_SS) {
      // ln P scatter of the own (row, node)
      if (r_own < nrow) {
        double sb = 0.0;
        for (int q = 0; q < d.n_ss; q++) {
          int i, j;
          clh_ss_pair(d.ns, q, i, j);
          sb += Cb[q * CLH_SS_ROWS + rr_own] * W[i * np + p_own] * W[j * np + p_own];
        }
        sb *= pr[0][rr_own] * o[2 * np + p_own];
        clh_scatter_lnP(d, cell, wts, sb * pk);
      }
      // A_b cotangents
      for (int it = tid; it < d.ns * np; it += CLH_BLOCK) {
        const int b = it / np, p = it % np;
        double acc = 0.0;
        for (int rr = 0; rr < CLH_SS_ROWS; rr++) {
          double s = 0.0;
          for (int j = 0; j < d.ns; j++) {
            const double f = (j == b) ? 2.0 : 1.0;
            s += f * Cb[clh_ss_index(d.ns, b, j) * CLH_SS_ROWS + rr] * W[j * np + p];
          }
          acc += pr[0][rr] * S[rr * np + p] * s;
        }
        if (acc != 0.0) atomicAdd(d.abar + (size_t)b * np + p, acc);
      }
    } else {
      const double b1 = d.par->b1[zl], bmag = d.par->bmag[zl];
      const double* WG = W;
      const double* WM = W + np;
      if (r_own < nrow) {
        const double g = WG[p_own] * b1 + WM[p_own] * pr[1][rr_own] * bmag;
        double sb = 0.0;
        for (int zs = 0; zs < d.ns; zs++) sb += Cb[zs * CLH_SS_ROWS + rr_own] * W[(2 + zs) * np + p_own];
        sb *= pr[2][rr_own] * g * o[2 * np + p_own];
        clh_scatter_lnP(d, cell, wts, sb * pk);
      }
      // B_zs cotangents
      for (int it = tid; it < d.ns * np; it += CLH_BLOCK) {
        const int zs = it / np, p = it % np;
        double acc = 0.0;
        for (int rr = 0; rr < CLH_SS_ROWS; rr++)
          acc += pr[2][rr] * Cb[zs * CLH_SS_ROWS + rr] * (WG[p] * b1 + WM[p] * pr[1][rr] * bmag) * S[rr * np + p];
        if (acc != 0.0) atomicAdd(d.bbar + ((size_t)zl * d.ns + zs) * np + p, acc);
      }
      // W_gal, W_mag cotangents and b1
      double b1p = 0.0;
      if (tid < np) {
        const int p = tid;
        double t = 0.0, tm = 0.0;
        for (int rr = 0; rr < CLH_SS_ROWS; rr++) {
          double s = 0.0;
          for (int zs = 0; zs < d.ns; zs++) s += Cb[zs * CLH_SS_ROWS + rr] * W[(2 + zs) * np + p];
          s *= pr[2][rr] * S[rr * np + p];
          t += s;
          tm += s * pr[1][rr];
        }
        atomicAdd(d.wbar + ((size_t)zl * 2 + 0) * np + p, t * b1);
        atomicAdd(d.wbar + ((size_t)zl * 2 + 1) * np + p, tm * bmag);
        b1p = t * WG[p];
      }
      if (tid < 64) {
        b1p = clh_wave_sum(b1p);
        if (tid == 0 && b1p != 0.0) atomicAdd(d.g_b1 + zl, b1p);
      }
    }
  } else {
    const int np = CLH_NP_GG;
    const double* o = d.nodes + d.node_gg + (size_t)zl * d.node_gg_stride;
    const double b1 = d.par->b1[zl], bmag = d.par->bmag[zl];
    const int colg = d.n_ss + d.n_gs + zl;
    double b1p = 0.0, lzp = 0.0;
    for (int it = tid; it < CLH_GG_ROWS * np; it += CLH_BLOCK) {
      const int rr = it / np, p = it % np;
      const int r = row0 + rr;
      double tb = 0.0;
      if (r < nrow) {
        const double cb = d.cbar[(size_t)colg * nrow + r];
        const double l = d.ell_gx[r], ell = l + 0.5;
        const double a = o[p], fK = o[np + p], fac = o[2 * np + p];
        const double k = ell / fK;
        const double chi_0 = ell / k;
        const double chi_1 = (ell + 1.0) / k;
        if (!(chi_1 > d.chi_amin) && cb != 0.0) {
          const double a_0 = clh_a_chi(d, chi_0);
          const double a_1 = clh_a_chi(d, chi_1);
          double c0, dc0, c1, dc1;
          clh_chi_all(d, a_0, c0, dc0);
          clh_chi_all(d, a_1, c1, dc1);
          const double h0 = 1.0 / ((a_0 * a_0) * dc0);
          const double h1 = 1.0 / ((a_1 * a_1) * dc1);
          const double al = (1 + 8. * ell) / ((2. * ell + 1.) * (2. * ell + 1.)) * h0 * clh_f_growth(d, 1.0 / a_0 - 1.0);
          const double be = 4. / (2 * ell + 3.) * sqrt((2 * ell + 1.) / (2 * ell + 3.)) * h1 *
                            clh_f_growth(d, 1.0 / a_1 - 1.0);
          const double wr = al * clh_nz_lens(d, 1. / a_0 - 1., zl) - be * clh_nz_lens(d, 1. / a_1 - 1., zl);
          const double dwr = al * clh_dnz_lens(d, 1. / a_0 - 1., zl) - be * clh_dnz_lens(d, 1. / a_1 - 1., zl);
          const double pref = l * (l + 1.) / (ell * ell);
          const double t = o[3 * np + p] * b1 + o[4 * np + p] * pref * bmag + wr;
          int cell;
          double wts[4];
          const double pk = clh_p_nonlin_cells(d, k, a, cell, wts);
          clh_scatter_lnP(d, cell, wts, cb * t * t * fac * pk);
          tb = cb * 2.0 * t * pk * fac;
          b1p += tb * o[3 * np + p];
          lzp += tb * dwr;
        }
      }
      S[it] = tb;
    }
    // mag-bias weight needs pref per row: keep the row factor alongside
    if (tid < CLH_GG_ROWS) {
      const int r = row0 + tid;
      double pf = 0.0;
      if (r < nrow) { const double l = d.ell_gx[r], ell = l + 0.5; pf = l * (l + 1.) / (ell * ell); }
      pr[1][tid] = pf;
    }
    __syncthreads();
    if (tid < np) {
      double tg = 0.0, tm = 0.0;
      for (int rr = 0; rr < CLH_GG_ROWS; rr++) {
        tg += S[rr * np + tid];
        tm += S[rr * np + tid] * pr[1][rr];
      }
      const size_t base = (size_t)2 * d.nl * CLH_NP_SS + (size_t)zl * 2 * np;
      if (tg != 0.0) atomicAdd(d.wbar + base + tid, tg * b1);
      if (tm != 0.0 && bmag != 0.0) atomicAdd(d.wbar + base + np + tid, tm * bmag);
    }
    // block reductions of b1 and lens-shift partials
    const int lane = tid & 63, wid = tid >> 6;
    b1p = clh_wave_sum(b1p);
    lzp = clh_wave_sum(lzp);
    if (lane == 0) { red[wid] = b1p; red[4 + wid] = lzp; }
    __syncthreads();
    if (tid == 0) {
      const double sb = red[0] + red[1] + red[2] + red[3];
      const double sl = red[4] + red[5] + red[6] + red[7];
      if (sb != 0.0) atomicAdd(d.g_b1 + zl, sb);
      if (sl != 0.0) atomicAdd(d.g_lens + zl, sl);
    }
  }
}

// ----------------------------------------------------------------------------
// k_nodes_adj: (draw, segment) like k_nodes; one thread per node.
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(128) void k_nodes_adj(const DrawDesc* __restrict__ D,
                                                   const int2* __restrict__ work) {
  const int2 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const int seg = w.y;
  const int p = threadIdx.x;
  const int np = (seg <= d.nl) ? CLH_NP_SS : CLH_NP_GG;
  const bool act = p < np;
  const int zl = seg == 0 ? -1 : (seg <= d.nl ? seg - 1 : seg - 1 - d.nl);
  double a = 0.5, fK = 1.0, dchida = 1.0;
  if (act) {
    const double* o = seg == 0 ? d.nodes + d.node_ss
                    : seg <= d.nl ? d.nodes + d.node_gs + (size_t)zl * d.node_gs_stride
                                  : d.nodes + d.node_gg + (size_t)zl * d.node_gg_stride;
    a = o[p];
    clh_chi_all(d, a, fK, dchida);
  }
  const double gf = act ? clh_growfac(d, a) : 1.0;
  const double hh = clh_hoverh0v2(a, dchida);
  const double z = 1.0 / a - 1.0;
  const double wk_pre = 1.5 * d.omega_m * fK / a;
  const double c1u = clh_ia_c1_unit(d, a, gf);
  double gA = 0.0, gl = 0.0;
  // per-source-bin shift partials, reduced per wave below
  for (int b = 0; b < d.ns; b++) {
    double gs = 0.0;
    if (act && seg <= d.nl) {
      const double bar = (seg == 0) ? d.abar[(size_t)b * CLH_NP_SS + p]
                                    : d.bbar[((size_t)zl * d.ns + b) * CLH_NP_SS + p];
      if (bar != 0.0) {
        const double C1 = clh_ia_c1(d, a, gf, b);
        const double WS = clh_nz_src(d, z, b) * hh;
        const double dWS = clh_dnz_src(d, z, b) * hh;
        const double dWK = wk_pre * clh_dg_tomo(d, a, b);
        gs = bar * (dWK - C1 * dWS);
        gA += bar * (-c1u * WS);
      }
    }
    gs = clh_wave_sum(gs);
    if ((p & 63) == 0 && gs != 0.0) atomicAdd(d.g_src + b, gs);
  }
  if (act && seg > 0) {
    const double* wb = (seg <= d.nl) ? d.wbar + (size_t)zl * 2 * CLH_NP_SS
                                     : d.wbar + (size_t)2 * d.nl * CLH_NP_SS + (size_t)zl * 2 * CLH_NP_GG;
    gl = wb[p] * clh_dnz_lens(d, z, zl) * hh + wb[np + p] * wk_pre * clh_dg_lens(d, a, zl);
  }
  gA = clh_wave_sum(gA);
  gl = clh_wave_sum(gl);
  if ((p & 63) == 0) {
    if (gA != 0.0) atomicAdd(d.g_A, gA);
    if (zl >= 0 && gl != 0.0) atomicAdd(d.g_lens + zl, gl);
  }
}

// ----------------------------------------------------------------------------
hipError_t clh_launch_fill(double* p, size_t n, double v, hipStream_t s) {
  if (!p || n == 0) return hipSuccess;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}
hipError_t clh_launch_set_params(const DrawDesc* D, DrawParams* P, int B, int nknz, const double* lnP,
                                 const double* src, const double* lens, const double* b1,
                                 const double* iaA, const double* m, hipStream_t s) {
  hipLaunchKernelGGL(k_set_params, dim3(B), dim3(CLH_BLOCK), 0, s, D, P, nknz, lnP, src, lens, b1, iaA, m);
  return hipGetLastError();
}
hipError_t clh_launch_leg_adjoint(const DrawDesc* D, const int2* work, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_leg_adjoint, dim3(n), dim3(CLH_BLOCK), 0, s, D, work);
  return hipGetLastError();
}
hipError_t clh_launch_m_grad(const DrawDesc* D, int B, hipStream_t s) {
  hipLaunchKernelGGL(k_m_grad, dim3(B), dim3(64), 0, s, D);
  return hipGetLastError();
}
hipError_t clh_launch_limber_adj(const DrawDesc* D, const int4* work, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_limber_adj, dim3(n), dim3(CLH_BLOCK), 0, s, D, work);
  return hipGetLastError();
}
hipError_t clh_launch_nodes_adj(const DrawDesc* D, const int2* work, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_nodes_adj, dim3(n), dim3(128), 0, s, D, work);
  return hipGetLastError();
}
