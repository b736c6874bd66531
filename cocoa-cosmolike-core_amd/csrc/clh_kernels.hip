// HIP kernels of the FP64 projection path (gfx950).
//
//  k_gtab     lens-efficiency tables g_tomo / g_lens: fine-grid integrand,
//             per-interval trapezoid sums, block prefix scan
//             (cosmolike/redshift_spline.c:852-923, 1014-1109)
//  k_nodes    Gauss-Legendre nodes in a and every weight the Limber sums need
//             (create_cosmo_nodes cosmo2D.c:1331-1356; radial_weights.c)
//  k_limber_sx  Limber node sums for ss / gs tiles (P(k,z) and weights staged
//               in LDS; cosmo2D.c:1655-1747, 2229-2348) -> C_ell rows
//  k_limber_gg  gg integrand with RSD (cosmo2D.c:2490-2666) -> C_ell rows
//  k_legendre contraction of C_ell rows with the bin-averaged Legendre
//             kernels into xi+, xi-, gamma_t, w and data-vector placement with
//             shear calibration (cosmo2D.c:487-534; generic_interface.hpp:800-862)
#include <hip/hip_runtime.h>

#include "clh_device.hpp"
#include "clh_internal.hpp"

#define CLH_BLOCK 256

// ----------------------------------------------------------------------------
// block-wide inclusive scan of one double per thread (256 threads, wave64)
// ----------------------------------------------------------------------------
__device__ __forceinline__ double clh_block_scan(double v, double* lds /* >= 4 */) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  if (lane == 63) lds[wid] = v;
  __syncthreads();
  double base = 0.0;
  for (int w = 0; w < wid; w++) base += lds[w];
  __syncthreads();
  return v + base;
}

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

// ----------------------------------------------------------------------------
// k_gtab: one block per (draw, bin).  bin < ns -> g_tomo of source bin,
// otherwise g_lens of lens bin (bin - ns).
// ----------------------------------------------------------------------------
// TANGENT: the same prefix sums over d(integrand)/d(photo-z shift of the bin),
// i.e. the derivative of the table with respect to that shift.
template <bool TANGENT>
__global__ __launch_bounds__(CLH_BLOCK) void k_gtab(const DrawDesc* __restrict__ D,
                                                    const int2* __restrict__ work) {
  __shared__ double sP[1024], sQ[1024];
  __shared__ double red[8];
  const int2 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const bool src = w.y < d.ns;
  const int b = src ? w.y : w.y - d.ns;
  const int x = 25;  // 25 (1 + |high_def_integration|), high_def_integration = 0
  const int nint = d.n_a - 1;
  const int Na = x * nint + 1;
  const double amax = 0.999999;
  const double amin = 1.0 / ((src ? d.s_zmax_all : d.l_zmax_all) + 1.0);
  const double da = (amax - amin) / ((double)Na - 1.0);

  // g_lens pre-integral from a_min of the shear sample (redshift_spline.c:1052-1071)
  double P0 = 0.0, Q0 = 0.0;
  if (!src) {
    const double amin_shear = 1.0 / (d.s_zmax_all + 1.0);
    const int Na0 = (amin_shear < amin) ? (int)ceil((amin - amin_shear) / da) + 1 : 1;
    const double da0 = (Na0 > 1) ? (amin - amin_shear) / ((double)Na0 - 1.) : 0.0;
    double p = 0.0, q = 0.0;
    for (int i = 1 + threadIdx.x; i < Na0; i += CLH_BLOCK) {
      const double apl = amin_shear + (i - 1) * da0;
      const double aph = amin_shear + i * da0;
      const double fl = (TANGENT ? clh_dnz_lens(d, 1.0 / apl - 1.0, b)
                                 : clh_nz_lens(d, 1.0 / apl - 1.0, b)) / (apl * apl);
      const double fh = (TANGENT ? clh_dnz_lens(d, 1.0 / aph - 1.0, b)
                                 : clh_nz_lens(d, 1.0 / aph - 1.0, b)) / (aph * aph);
      p += 0.5 * da0 * (fl + fh);
      q += 0.5 * da0 * (fl / clh_chi(d, apl) + fh / clh_chi(d, aph));
    }
    P0 = clh_block_sum(p, red);
    Q0 = clh_block_sum(q, red);
  }

  // per-interval trapezoid sums of n/a^2 and n/(chi a^2)
  for (int t = threadIdx.x; t < nint; t += CLH_BLOCK) {
    double p = 0.0, q = 0.0;
    double fprev = 0.0, gprev = 0.0;
    for (int s = 0; s <= x; s++) {
      const int i = t * x + s;
      const double a = amin + i * da;
      double n;
      if (TANGENT) n = src ? clh_dnz_src(d, 1. / a - 1., b) : clh_dnz_lens(d, 1.0 / a - 1.0, b);
      else n = src ? clh_nz_src(d, 1. / a - 1., b) : clh_nz_lens(d, 1.0 / a - 1.0, b);
      const double f = n / (a * a);
      const double g = f / clh_chi(d, amin + i * da);
      if (s > 0) {
        p += 0.5 * da * (fprev + f);
        q += 0.5 * da * (gprev + g);
      }
      fprev = f;
      gprev = g;
    }
    sP[t] = p;
    sQ[t] = q;
  }
  __syncthreads();

  // chunked prefix scan over intervals
  const int cpt = (nint + CLH_BLOCK - 1) / CLH_BLOCK;
  const int t0 = threadIdx.x * cpt;
  const int t1 = min(t0 + cpt, nint);
  double cp = 0.0, cq = 0.0;
  for (int t = t0; t < t1; t++) { cp += sP[t]; cq += sQ[t]; }
  const double ip = clh_block_scan(cp, red);
  const double iq = clh_block_scan(cq, red);
  double P = P0 + ip - cp, Q = Q0 + iq - cq;
  double* tab = (TANGENT ? d.dgtab : d.gtab) + (size_t)w.y * d.n_a;
  for (int t = t0; t < t1; t++) {
    P += sP[t];
    Q += sQ[t];
    const int i = (t + 1) * x;
    tab[t + 1] = P - clh_chi(d, amin + i * da) * Q;
  }
  if (threadIdx.x == 0) tab[0] = src ? 0.0 : P0 - clh_chi(d, amin) * Q0;
}

// ----------------------------------------------------------------------------
// k_nodes: one block per (draw, segment); segment 0 = ss, 1..nl = gs lens
// bin, nl+1..2nl = gg lens bin.  One thread per GL node.
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(128) void k_nodes(const DrawDesc* __restrict__ D,
                                               const int2* __restrict__ work) {
  const int2 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const int seg = w.y;
  const int p = threadIdx.x;
  const int np = (seg <= d.nl) ? CLH_NP_SS : CLH_NP_GG;
  if (p >= np) return;
  const double* glx = d.gl + (np == CLH_NP_SS ? 0 : 2 * CLH_NP_SS);
  const double* glw = glx + np;
  double amin, amax;
  int zl = -1;
  if (seg == 0) {
    amin = 1. / (d.s_zmax_all + 1.);
    amax = 1. / (1. + fmax(d.s_zmin_all, 1e-6));
  } else {
    zl = (seg <= d.nl) ? seg - 1 : seg - 1 - d.nl;
    amin = clh_amin_lens(d, zl);
    amax = clh_amax_lens(d, zl);
  }
  const double A = 0.5 * (amax - amin), B = 0.5 * (amin + amax);
  const double a = B + A * glx[p];
  const double wt = A * glw[p];
  double fK, dchida;
  clh_chi_all(d, a, fK, dchida);
  const double gf = clh_growfac(d, a);
  const double hh = clh_hoverh0v2(a, dchida);
  const double fac = wt * (dchida / (fK * fK));
  const double z = 1.0 / a - 1.0;
  const double wk_pre = 1.5 * d.omega_m * fK / a;
  if (seg == 0) {
    double* o = d.nodes + d.node_ss;
    o[p] = a;
    o[np + p] = fK;
    o[2 * np + p] = fac;
    for (int bb = 0; bb < d.ns; bb++) {
      const double WK = wk_pre * clh_g_tomo(d, a, bb);
      const double WS = clh_nz_src(d, z, bb) * hh;
      o[(3 + bb) * np + p] = WK - clh_ia_c1(d, a, gf, bb) * WS;
    }
  } else if (seg <= d.nl) {
    double* o = d.nodes + d.node_gs + (size_t)zl * d.node_gs_stride;
    o[p] = a;
    o[np + p] = fK;
    o[2 * np + p] = fac;
    o[3 * np + p] = clh_nz_lens(d, z, zl) * hh;
    o[4 * np + p] = wk_pre * clh_g_lens(d, a, zl);
    for (int zs = 0; zs < d.ns; zs++) {
      const double WK = wk_pre * clh_g_tomo(d, a, zs);
      const double WS = clh_nz_src(d, z, zs) * hh;
      o[(5 + zs) * np + p] = WK - WS * clh_ia_c1(d, a, gf, zs);
    }
  } else {
    double* o = d.nodes + d.node_gg + (size_t)zl * d.node_gg_stride;
    o[p] = a;
    o[np + p] = fK;
    o[2 * np + p] = fac;
    o[3 * np + p] = clh_nz_lens(d, z, zl) * hh;
    o[4 * np + p] = wk_pre * clh_g_lens(d, a, zl);
  }
}

// ----------------------------------------------------------------------------
// k_limber_sx: ss and gs tiles of CLH_SX_TILE rows.  One thread per GL node
// walks its rows in increasing ell, so the log10 k index of the P(k,z) lookup
// moves by at most a step between rows and the z index is found once.
// Work item (draw, kind, zl, row0).
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(CLH_BLOCK) void k_limber_sx(const DrawDesc* __restrict__ D,
                                                         const int4* __restrict__ work) {
  __shared__ double S[CLH_SX_TILE * CLH_NP_SS];
  __shared__ double W[(CLH_MAXB + 2) * CLH_NP_SS];
  __shared__ double L10[CLH_SX_TILE];
  const int4 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const int kind = w.y, zl = w.z, row0 = w.w;
  const int tid = threadIdx.x;
  const int np = CLH_NP_SS;
  const double* o = (kind == CLH_W_SS) ? d.nodes + d.node_ss
                                       : d.nodes + d.node_gs + (size_t)zl * d.node_gs_stride;
  const double* ells = (kind == CLH_W_SS) ? d.ell_ss : d.ell_gx;
  if (tid < CLH_SX_TILE) {
    const int r = row0 + tid;
    L10[tid] = log10((r < d.nrow ? ells[r] : 1.0) + 0.5);
  }
  const int nw = (kind == CLH_W_SS) ? d.ns : d.ns + 2;
  for (int i = tid; i < nw * np; i += CLH_BLOCK) W[i] = o[3 * np + i];
  __syncthreads();
  {
    const int p = tid & 63, wv = tid >> 6;
    const double a = o[p], fK = o[np + p], fac = o[2 * np + p];
    const double z = 1.0 / a - 1.0;
    const int nz = d.nzp;
    const int j = clh_bs(d.pz, nz, z);
    const double dy = (z - d.pz[j]) / (d.pz[j + 1] - d.pz[j]);
    const double l10fk = log10(fK * CLH_COVERH0);  // log10 k = log10(ell) - log10(chi)
    int i = -1;
    double inv_k = 0.0;
    constexpr int per = CLH_SX_TILE / 4;
    for (int q = 0; q < per; q++) {
      const int rr = wv * per + q;
      const int r = row0 + rr;
      double s = 0.0;
      if (r < d.nrow) {
        const double log10k = L10[rr] - l10fk;
        const int i0 = i;
        i = (i < 0) ? clh_bs(d.lk, d.nk, log10k) : clh_walk(d.lk, d.nk, log10k, i);
        if (i != i0) inv_k = 1.0 / (d.lk[i + 1] - d.lk[i]);
        const double dx = (log10k - d.lk[i]) * inv_k;
        const double* __restrict__ P = d.lnP + (size_t)i * nz + j;
        const double out = (1 - dx) * (1 - dy) * P[0] + (1 - dx) * dy * P[1] + dx * (1 - dy) * P[nz] +
                           dx * dy * P[nz + 1];
        s = exp(out) * (1.0 / (CLH_COVERH0 * CLH_COVERH0 * CLH_COVERH0)) * fac;
      }
      S[rr * np + p] = s;
    }
  }
  __syncthreads();
  if (kind == CLH_W_SS) {
    const int nout = d.n_ss * CLH_SX_TILE;
    for (int q = tid; q < nout; q += CLH_BLOCK) {
      const int pair = q / CLH_SX_TILE, rr = q % CLH_SX_TILE;
      const int r = row0 + rr;
      if (r >= d.nrow) continue;
      int i, j;
      clh_ss_pair(d.ns, pair, i, j);
      const double* Ai = W + i * np;
      const double* Aj = W + j * np;
      const double* Sr = S + rr * np;
      double acc = 0.0;
#pragma unroll 8
      for (int p = 0; p < np; p++) acc += Ai[p] * Aj[p] * Sr[p];
      const double l = ells[r], ell = l + 0.5, ell4 = ell * ell * ell * ell;
      const double pref = l * (l - 1.) * (l + 1.) * (l + 2.) / ell4;
      d.cl[(size_t)pair * d.nrow + r] = acc * pref;
    }
  } else {
    const int nout = d.ns * CLH_SX_TILE;
    const double b1 = d.par->b1[zl], bmag = d.par->bmag[zl];
    for (int q = tid; q < nout; q += CLH_BLOCK) {
      const int zs = q / CLH_SX_TILE, rr = q % CLH_SX_TILE;
      const int r = row0 + rr;
      if (r >= d.nrow) continue;
      const double l = ells[r], ell = l + 0.5;
      const double pref1 = l * (l + 1.) / (ell * ell);
      const double tmp = (l - 1.) * l * (l + 1.) * (l + 2.);
      const double pref2 = (tmp > 0) ? sqrt(tmp) / (ell * ell) : 0.0;
      const double* Bz = W + (2 + zs) * np;
      const double* Sr = S + rr * np;
      double acc = 0.0;
#pragma unroll 8
      for (int p = 0; p < np; p++) acc += Bz[p] * (W[p] * b1 + W[np + p] * pref1 * bmag) * Sr[p];
      d.cl[(size_t)(d.n_ss + zl * d.ns + zs) * d.nrow + r] = acc * pref2;
    }
  }
}

// ----------------------------------------------------------------------------
// k_limber_gg: gg tiles of CLH_GGT_TILE rows, 192 threads = two groups of 96
// (one thread per GL node, ClhGgNode), each group walking half of the rows in
// increasing ell; per-row constants in LDS.  Work item (draw, zl, row0, -).
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(192) void k_limber_gg(const DrawDesc* __restrict__ D,
                                                   const int4* __restrict__ work) {
  __shared__ double V[CLH_GGT_TILE * CLH_NP_GG];
  __shared__ ClhGgRow R[CLH_GGT_TILE];
  const int4 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const int zl = w.y, row0 = w.z;
  constexpr int np = CLH_NP_GG;
  if (threadIdx.x < CLH_GGT_TILE) {
    const int r = row0 + threadIdx.x;
    R[threadIdx.x] = clh_gg_row(r < d.nrow ? d.ell_gx[r] : 1.0);
  }
  __syncthreads();
  const int p = threadIdx.x % np, half = threadIdx.x / np;
  ClhGgNode<false> nd;
  nd.init(d, d.nodes + d.node_gg + (size_t)zl * d.node_gg_stride, p, zl);
  constexpr int per = CLH_GGT_TILE / 2;
  for (int q = 0; q < per; q++) {
    const int rr = half * per + q;
    double v = 0.0;
    if (row0 + rr < d.nrow) {
      double t, pf, dwr, dx;
      if (nd.eval(d, R[rr], t, pf, dwr, dx)) v = t * t * pf;
    }
    V[rr * np + p] = v;
  }
  __syncthreads();
  // row sums over the 96 nodes: one wave per row
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int rr = wid; rr < CLH_GGT_TILE; rr += 3) {
    double v = V[rr * np + lane] + ((lane + 64 < np) ? V[rr * np + lane + 64] : 0.0);
    v = clh_wave_sum(v);
    const int r = row0 + rr;
    if (lane == 0 && r < d.nrow) d.cl[(size_t)(d.n_ss + d.n_gs + zl) * d.nrow + r] = v;
  }
}

// ----------------------------------------------------------------------------
// k_legendre (folded kernels): work item (draw, first task); 8 tasks per block,
// 32 theta slots per task.  Task t of a draw: [0, n_ss) xi+, [n_ss, 2 n_ss)
// xi-, then gamma_t pairs, then w bins.
// ----------------------------------------------------------------------------
#define CLH_LEG_TASKS 8
__global__ __launch_bounds__(CLH_BLOCK) void k_legendre(const DrawDesc* __restrict__ D,
                                                        const int2* __restrict__ work) {
  extern __shared__ double colbuf[];  // [CLH_LEG_TASKS][nrow]
  const int2 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const int nt = d.ntheta, nrow = d.nrow;
  const int ntask = 2 * d.n_ss + d.n_gs + d.nl;
  const int tl = threadIdx.x >> 5, th = threadIdx.x & 31;
  const int task = w.y + tl;

  for (int i = threadIdx.x; i < CLH_LEG_TASKS * nrow; i += CLH_BLOCK) {
    const int tt = w.y + i / nrow;
    double v = 0.0;
    if (tt < ntask) {
      int col;
      if (tt < d.n_ss) col = tt;
      else if (tt < 2 * d.n_ss) col = tt - d.n_ss;
      else col = tt - d.n_ss;  // gs and gg columns follow the ss columns
      v = d.cl[(size_t)col * nrow + (i % nrow)];
    }
    colbuf[i] = v;
  }
  __syncthreads();
  if (task >= ntask || th >= nt) return;
  int kind, q;
  if (task < d.n_ss) { kind = 0; q = task; }
  else if (task < 2 * d.n_ss) { kind = 1; q = task - d.n_ss; }
  else if (task < 2 * d.n_ss + d.n_gs) { kind = 2; q = task - 2 * d.n_ss; }
  else { kind = 3; q = task - 2 * d.n_ss - d.n_gs; }
  const double* K = d.kfold + (size_t)kind * nrow * nt + th;
  const double* c = colbuf + tl * nrow;
  double acc = 0.0;
  for (int r = 0; r < nrow; r++) acc += K[(size_t)r * nt] * c[r];

  int idx;
  double f = 1.0;
  if (kind <= 1) {
    int i, j;
    clh_ss_pair(d.ns, q, i, j);
    if (d.apply_calib) f = (1.0 + d.par->shear_m[i]) * (1.0 + d.par->shear_m[j]);
    idx = d.dv_start[0] + nt * q + th + (kind == 1 ? nt * d.n_ss : 0);
  } else if (kind == 2) {
    if (d.apply_calib) f = 1.0 + d.par->shear_m[q % d.ns];
    idx = d.dv_start[1] + nt * q + th;
  } else {
    idx = d.dv_start[2] + nt * q + th;
  }
  d.dv[idx] = acc * f;
}

// ----------------------------------------------------------------------------
// k_legendre_full: the reference's interpolate-then-dot structure
// (cosmo2D.c:468-533): C_ell at every integer ell < LMAX is generated on the
// fly from the column's rows (ell < LMIN_tab direct, else the ln-ell fill) and
// contracted with the full-length kernel.  Same task mapping as k_legendre.
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(CLH_BLOCK) void k_legendre_full(const DrawDesc* __restrict__ D,
                                                             const int2* __restrict__ work) {
  extern __shared__ double colbuf[];
  const int2 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const int nt = d.ntheta, nrow = d.nrow;
  const int ntask = 2 * d.n_ss + d.n_gs + d.nl;
  const int tl = threadIdx.x >> 5, th = threadIdx.x & 31;
  const int task = w.y + tl;
  for (int i = threadIdx.x; i < CLH_LEG_TASKS * nrow; i += CLH_BLOCK) {
    const int tt = w.y + i / nrow;
    double v = 0.0;
    if (tt < ntask) {
      const int col = (tt < d.n_ss) ? tt : tt - d.n_ss;
      v = d.cl[(size_t)col * nrow + (i % nrow)];
    }
    colbuf[i] = v;
  }
  __syncthreads();
  if (task >= ntask || th >= nt) return;
  int kind, q;
  if (task < d.n_ss) { kind = 0; q = task; }
  else if (task < 2 * d.n_ss) { kind = 1; q = task - d.n_ss; }
  else if (task < 2 * d.n_ss + d.n_gs) { kind = 2; q = task - 2 * d.n_ss; }
  else { kind = 3; q = task - 2 * d.n_ss - d.n_gs; }
  const int L = d.lmax;
  const double* K = d.kfull + (size_t)kind * L * nt + th;
  const double* fl = d.fill + (size_t)(kind <= 1 ? 0 : 1) * L * 2;
  const double* c = colbuf + tl * nrow;
  const double* tab = c + CLH_NLOW;
  double acc = 0.0;
  for (int l = 1; l < CLH_LMIN_TAB && l < L; l++) acc += K[(size_t)l * nt] * c[l - 1];
  for (int l = CLH_LMIN_TAB; l < L; l++) {
    const int ic = (int)fl[2 * l];
    const double t = fl[2 * l + 1];
    const double cl = tab[ic] + t * (tab[ic + 1] - tab[ic]);
    acc += K[(size_t)l * nt] * cl;
  }
  int idx;
  double f = 1.0;
  if (kind <= 1) {
    int i, j;
    clh_ss_pair(d.ns, q, i, j);
    if (d.apply_calib) f = (1.0 + d.par->shear_m[i]) * (1.0 + d.par->shear_m[j]);
    idx = d.dv_start[0] + nt * q + th + (kind == 1 ? nt * d.n_ss : 0);
  } else if (kind == 2) {
    if (d.apply_calib) f = 1.0 + d.par->shear_m[q % d.ns];
    idx = d.dv_start[1] + nt * q + th;
  } else {
    idx = d.dv_start[2] + nt * q + th;
  }
  d.dv[idx] = acc * f;
}

// ----------------------------------------------------------------------------
// launch wrappers (called from the host runtime)
// ----------------------------------------------------------------------------
hipError_t clh_launch_gtab(const DrawDesc* D, const int2* work, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gtab<false>, dim3(n), dim3(CLH_BLOCK), 0, s, D, work);
  return hipGetLastError();
}
hipError_t clh_launch_nodes(const DrawDesc* D, const int2* work, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_nodes, dim3(n), dim3(128), 0, s, D, work);
  return hipGetLastError();
}
hipError_t clh_launch_legendre(const DrawDesc* D, const int2* work, int n, int max_nrow,
                               hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const size_t lds = sizeof(double) * CLH_LEG_TASKS * (size_t)max_nrow;
  hipLaunchKernelGGL(k_legendre, dim3(n), dim3(CLH_BLOCK), lds, s, D, work);
  return hipGetLastError();
}
hipError_t clh_launch_legendre_full(const DrawDesc* D, const int2* work, int n, int max_nrow,
                                    hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const size_t lds = sizeof(double) * CLH_LEG_TASKS * (size_t)max_nrow;
  hipLaunchKernelGGL(k_legendre_full, dim3(n), dim3(CLH_BLOCK), lds, s, D, work);
  return hipGetLastError();
}
hipError_t clh_launch_gtab_tangent(const DrawDesc* D, const int2* work, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gtab<true>, dim3(n), dim3(CLH_BLOCK), 0, s, D, work);
  return hipGetLastError();
}
hipError_t clh_launch_limber_sx(const DrawDesc* D, const int4* work, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_limber_sx, dim3(n), dim3(CLH_BLOCK), 0, s, D, work);
  return hipGetLastError();
}
hipError_t clh_launch_limber_gg(const DrawDesc* D, const int4* work, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_limber_gg, dim3(n), dim3(192), 0, s, D, work);
  return hipGetLastError();
}
