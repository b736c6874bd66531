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
//  k_m_grad       d/dm from cotangent-weighted uncalibrated column sums
//  k_limber_sx_adj, k_limber_gg_adj   ln P cotangent + cotangents of the node weights
//  k_nodes_adj    node weights -> photo-z shifts (n(z) and g tangents), A_IA
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <type_traits>

#include "clh_device.hpp"
#include "clh_run.hpp"
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
// k_leg_adjoint: cotangent of the C_ell rows from the data-vector cotangent,
// on MFMA (v_mfma_f64_16x16x4f64).  One block of 4 waves per work item
// (draw, kind, c0) = 16 columns of one probe (kind 0 ss, 2 gs, 3 gg):
//   cbar[c][r] = f_c sum_theta u[c][theta] K_kind[r][theta]
// (ss: the xi+ and xi- kernels with the xi+ and xi- cotangents, i.e. a
// contraction over 2 Ntheta), with u the masked data-vector cotangent and f_c
// the shear calibration.  M = columns, N = rows (16 per tile; wave w takes
// tiles w, w + 4, ...), K = theta in steps of 4, kernels read as
// [kind][theta][row] (coalesced along rows), every step of a tile loaded
// before its MFMAs.
//   T[c] = sum_r (cbar[c][r] / f_c) C[c][r]  (= dvbar . uncalibrated data
// vector of the column), summed over the waves through LDS, goes to the
// draw's `raw` buffer for the shear-calibration gradient, with the point-mass
// term of gamma_t.  (The FULL-mode forward has the same derivative with
// respect to the rows.)  TATT draws (cbar_bb set): the B-mode rows enter
// xi+ = K+ (EE + BB), xi- = K- (EE - BB), so their cotangent takes the xi-
// part with the opposite sign, and T adds their share.
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_leg_adjoint(const DrawDesc* __restrict__ D,
                                                     const int4* __restrict__ work) {
  __shared__ double vb[2][16][33];  // [xi+/xi-][column][theta], theta >= Ntheta zero
  __shared__ double cf[16];
  __shared__ double tsum[4][16];
  typedef double v4d __attribute__((ext_vector_type(4)));
  const int4 w = work[clh_xcd_item(blockIdx.x, gridDim.x)];
  const DrawDesc& d = D[w.x];
  const int kind = w.y, c0 = w.z;
  const int nt = d.ntheta, nrow = d.nrow, nss = d.n_ss, ngs = d.n_gs;
  const int ncolk = kind == 0 ? nss : (kind == 2 ? ngs : d.nl);
  const bool bbm = kind == 0 && d.cbar_bb != nullptr;  // TATT: B-mode cotangents as well
  const int colbase = kind == 0 ? 0 : (kind == 2 ? nss : nss + ngs);
  const int tid = threadIdx.x, l = tid & 63, wv = tid >> 6, lm = l & 15, lk = l >> 4;
  for (int i = tid; i < 2 * 16 * 33; i += 256) {
    const int sk = i / (16 * 33), cc = (i / 33) % 16, th = i % 33;
    const int q = c0 + cc;
    double v = 0.0;
    if (q < ncolk && th < nt && (sk == 0 || kind == 0)) {
      int canon, idx;
      if (kind == 0) {
        canon = nt * q + th + (sk ? nt * nss : 0);
        idx = d.dv_start[0] + canon;
      } else if (kind == 2) {
        const int r = d.gs_row ? d.gs_row[q] : q;  // -1: an excluded pair, no data-vector row
        canon = 2 * nt * nss + nt * r + th;
        idx = r < 0 ? -1 : d.dv_start[1] + nt * r + th;
      } else {
        canon = nt * (2 * nss + d.n_gs_dv) + nt * q + th;
        idx = d.dv_start[2] + nt * q + th;
      }
      v = (idx < 0 || clh_masked(d, canon)) ? 0.0 : d.dvbar[idx];  // masked entries are constant 0
    }
    vb[sk][cc][th] = v;
  }
  if (tid < 16) {
    const int q = c0 + tid;
    double f = 1.0;
    if (d.apply_calib && q < ncolk && kind == 0) {
      int a, b;
      clh_ss_pair(d.ns, q, a, b);
      f = (1.0 + d.par->shear_m[a]) * (1.0 + d.par->shear_m[b]);
    } else if (d.apply_calib && q < ncolk && kind == 2) {
      f = 1.0 + d.par->shear_m[q % d.ns];
    }
    cf[tid] = f;
  }
  __syncthreads();
  // the tile's K steps: up to 8 per kernel (Ntheta <= 32), two kernels for ss
  const int nst = (nt + 3) >> 2;
  const double* __restrict__ KT0 = d.kfoldT + (size_t)(kind == 0 ? 0 : kind) * nt * nrow;
  const double* __restrict__ KT1 = d.kfoldT + (size_t)1 * nt * nrow;
  double av[16];
#pragma unroll
  for (int u = 0; u < 16; u++) {
    const int sk = u >> 3, th = 4 * (u & 7) + lk;
    av[u] = (th < 33) ? vb[sk][lm][th < 33 ? th : 0] : 0.0;
  }
  double t[4] = {0.0, 0.0, 0.0, 0.0};
  for (int r0 = 16 * wv; r0 < nrow; r0 += 64) {
    const int r = r0 + lm;
    const bool rin = r < nrow;
    double bv[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int sk = u >> 3, st = u & 7, th = 4 * st + lk;
      const bool use = st < nst && (sk == 0 || kind == 0);
      bv[u] = (use && rin && th < nt) ? (sk ? KT1 : KT0)[(size_t)th * nrow + r] : 0.0;
    }
    v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int sk = u >> 3, st = u & 7;
      if (st < nst && (sk == 0 || kind == 0)) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], bv[u], acc, 0, 0, 0);
    }
    // TATT: the B-mode rows enter xi+ with + and xi- with - (cosmo2D.c:517-533)
    v4d accb = {0.0, 0.0, 0.0, 0.0};
    if (bbm) {
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int sk = u >> 3, st = u & 7;
        if (st < nst) accb = __builtin_amdgcn_mfma_f64_16x16x4f64(sk ? -av[u] : av[u], bv[u], accb, 0, 0, 0);
      }
    }
    if (!rin) continue;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int q = c0 + 4 * v + lk;
      if (q >= ncolk) continue;
      const size_t o = (size_t)(colbase + q) * nrow + r;
      d.cbar[o] = acc[v] * cf[4 * v + lk];
      if (kind != 3) t[v] += acc[v] * d.cl[o];
      if (bbm) {
        const size_t ob = (size_t)q * nrow + r;
        d.cbar_bb[ob] = accb[v] * cf[4 * v + lk];
        t[v] += accb[v] * d.cl_bb[ob];
      }
    }
  }
  // T per column: over the 16 lanes of each column group, then the 4 waves
#pragma unroll
  for (int v = 0; v < 4; v++)
    for (int off = 8; off > 0; off >>= 1) t[v] += __shfl_xor(t[v], off, 16);
  if (lm == 0)
    for (int v = 0; v < 4; v++) tsum[wv][4 * v + lk] = t[v];
  __syncthreads();
  if (tid >= 16) return;
  const int c = tid, q = c0 + c;
  if (q >= ncolk) return;
  double tt = tsum[0][c] + tsum[1][c] + tsum[2][c] + tsum[3][c];
  if (d.apply_calib && kind == 2) {
    // point mass of gamma_t: (raw + pm / theta^2)(1 + m_zs) -> m and source-shift terms
    const int zl = q / d.ns, zs = q % d.ns;
    const double pmf = clh_pm_fac(d, zl, zs);
    if (pmf != 0.0) {
      double sv = 0.0;
      for (int th = 0; th < nt; th++) sv += vb[0][c][th] / (d.theta[th] * d.theta[th]);
      tt += pmf * sv;
      const double dpm = clh_pm_fac<true>(d, zl, zs) * sv * cf[c];
      if (dpm != 0.0) atomicAdd(d.g_src + zs, dpm);
    }
  }
  d.raw[colbase + q] = tt;
}

// ----------------------------------------------------------------------------
// k_m_grad: one block per draw, one thread per source bin.  With T[col] the
// cotangent-weighted uncalibrated column sums of k_leg_adjoint (and, for
// 6x2pt, of k_cmb_place_adj's ks columns after the 3x2pt ones):
//   d/dm_i = sum_{ss pairs (a,b) with i} d[(1+m_a)(1+m_b)]/dm_i T + sum_zl T[(zl, i)] + T[ks i].
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_m_grad(const DrawDesc* __restrict__ D) {
  const DrawDesc& d = D[blockIdx.x];
  const int i = threadIdx.x;
  if (i >= d.ns) return;
  double g = 0.0;
  if (d.apply_calib) {
    const double* m = d.par->shear_m;
    for (int q = 0; q < d.n_ss; q++) {
      int a, b;
      clh_ss_pair(d.ns, q, a, b);
      if (a != i && b != i) continue;
      const double der = (a == b) ? 2.0 * (1.0 + m[i]) : 1.0 + m[a == i ? b : a];
      g += der * d.raw[q];
    }
    for (int zl = 0; zl < d.nl; zl++) g += d.raw[d.n_ss + zl * d.ns + i];
    if (d.cbar6) g += d.raw[d.n_ss + d.n_gs + d.nl + i];  // 6x2pt: the ks column of source bin i
  }
  d.g_m[i] = g;
}

// ----------------------------------------------------------------------------
// The ln P cotangent band of a gs or gg block.  The 64 nodes of a gs tile
// and the 96 of a gg tile sit in one lens bin's z range, about ten nodes per
// z column of P(k,z), so they share cells: the block folds its nodes' runs
// over (k cell, z column) in an LDS band, per column the cells its nodes
// reach over the tile's rows (from the bisection of the first and last rows'
// log10 k), packed column after column.  After the rows the band is added
// into g_lnP with atomics (the gather has stored the ss part there before
// these launches).  A column past CLH_BAND_COLS or beyond CLH_BAND_CAP
// doubles, and any cell outside its column's range, is added into g_lnP
// directly.
// ----------------------------------------------------------------------------
#define CLH_BAND_CAP 1024
#define CLH_BAND_COLS 64
struct ClhBandLds {
  double A[CLH_BAND_CAP];
  int lo[CLH_BAND_COLS], hi[CLH_BAND_COLS], start[CLH_BAND_COLS];
  int jlo, nval, used;
};

// Band set-up of a gs or gg block (every thread calls it; ends synced).
// Each reporting node (a gg node reports from the block's first group)
// gives its columns jp, jp + 1 and cells i(first row) .. i(last row) + 1;
// out: the node's ClhBand.
__device__ __forceinline__ void clh_band_setup(ClhBandLds& bd, const DrawDesc& d, const double* __restrict__ lk,
                                               double l10_first, double l10_last, int jp, double dyp, double l10fk,
                                               bool reporter, ClhBand& out) {
  const int t = threadIdx.x;
  for (int u = t; u < CLH_BAND_CAP; u += blockDim.x) bd.A[u] = 0.0;
  if (t < CLH_BAND_COLS) {
    bd.lo[t] = INT_MAX;
    bd.hi[t] = -1;
  }
  if (t == 0) bd.jlo = INT_MAX;
  __syncthreads();
  if (reporter) atomicMin(&bd.jlo, jp);
  __syncthreads();
  const int jlo = bd.jlo, c0 = jp - jlo;
  if (reporter) {
    const int i0 = clh_bs(lk, d.nk, l10_first - l10fk);
    const int i1 = clh_bs(lk, d.nk, l10_last - l10fk) + 1;
    for (int c = c0; c <= c0 + 1 && c < CLH_BAND_COLS; c++) {
      atomicMin(&bd.lo[c], i0);
      atomicMax(&bd.hi[c], i1);
    }
  }
  __syncthreads();
  if (t < 64) {  // first wave: column starts, an inclusive scan of the widths
    const int w = bd.hi[t] >= bd.lo[t] ? bd.hi[t] - bd.lo[t] + 1 : 0;
    int incl = w;
#pragma unroll
    for (int k = 1; k < 64; k <<= 1) {
      const int u = __shfl_up(incl, k, 64);
      if (t >= k) incl += u;
    }
    const bool room = incl <= CLH_BAND_CAP;
    bd.start[t] = room ? incl - w : -1;
    // the columns with room form a prefix (incl never decreases); the last
    // of them ends the band
    const int nval = (int)__popcll(__ballot(room));
    if (t == 0) bd.nval = nval;
    if (nval > 0 && t == nval - 1) bd.used = incl;
    if (nval == 0 && t == 0) bd.used = 0;
  }
  __syncthreads();
  out.G = d.g_lnP;
  out.nz = d.nzp;
  out.j = jp;
  out.dy = dyp;
  for (int c = 0; c < 2; c++) {
    const int cc = c0 + c;
    if (cc < CLH_BAND_COLS && bd.start[cc] >= 0) {
      out.lo[c] = bd.lo[cc];
      out.hi[c] = bd.hi[cc];
      out.b[c] = bd.start[cc] - bd.lo[cc];
    }
  }
}

// Adds the band into g_lnP (after a sync): cell u of the band lies in the
// last column whose start is <= u.
__device__ __forceinline__ void clh_band_flush(const ClhBandLds& bd, const DrawDesc& d) {
  const int nv = bd.nval, used = bd.used, jlo = bd.jlo, nz = d.nzp;
  for (int u = threadIdx.x; u < used; u += blockDim.x) {
    const double v = bd.A[u];
    if (v == 0.0) continue;
    int lo = 0, hi = nv - 1;  // start[lo] <= u < start[hi + 1]
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (bd.start[mid] <= u) lo = mid;
      else hi = mid - 1;
    }
    const int i = bd.lo[lo] + (u - bd.start[lo]);
    atomicAdd(d.g_lnP + (size_t)i * nz + jlo + lo, v);
  }
}

// LIM (no checkpoint): also the lens-limit part of the lens-shift gradient
// (k_nodes_tan gives the nodes' motion).  Along the node's a the integrand
// t^2 P fac changes through t (W_gal b1, W_mag pref bmag, the a_0 half c0 t0
// and the a_1 half c1 hf1 n1 of W_RSD, whose chi_1 = chi (ell+1)/ell moves
// with the node's chi) and through ln P and fac; contracted with cbar, the
// node's share is a_t (W_gal' b1 sum tb + W_mag' bmag sum tb pref + t0' sum tb
// c0 - chi' sum tb c1 d(hf1 n1)/dchi + dlog10k' sum g dlnP/dlog10k + dy' sum g
// dlnP/dy) + (ln(weight fac))_t sum g, with tb = 2 cbar t P fac and g =
// cbar t^2 P fac the row's cotangents of t and ln P.
// ----------------------------------------------------------------------------
// k_limber_sx_adj: adjoint of k_limber_sx.  One wave per work item (draw,
// kind + 4 zl, row0, nrows); thread = GL node, walking the tile's rows in
// increasing ell with the forward's P(k,z) recomputation.  Per (row, node):
//  ss: u_b = sum_j f_bj cbar_(b,j) A_j (f = 2 on the diagonal) gives both
//      abar_b += pref0 S u_b and the ln P cotangent pref0 S sum_b A_b u_b / 2
//      (the cotangents sit in LDS as the full symmetric [ns][ns][rows] array
//      with f folded in, so the b, j loops address it linearly);
//  gs: bbar_zs += pref2 cbar_zs g S, W_gal / W_mag / b1 cotangents and
//      pref2 S g sum_zs cbar_zs B_zs for ln P, with g = W_gal b1 + W_mag pref1 bmag.
// LDS: ss [ns][ns][nrows], gs [ns][nrows] (dynamic).  With the forward's
// checkpoint (d.sxk) the P(k,z) lookup is replaced by its stored S = P fac
// and only the k cell is walked again.
// LIM (gs tiles, no checkpoint): also the lens-limit part of the lens-shift
// gradient, sum over rows and bins of cbar dC/d dz through the motion of the
// nodes (k_nodes_tan): with g the row's ln P cotangent, the node's share is
// a_t (sum_b B_b' acc_b + W_gal' b1 twg + W_mag' bmag twm + dlog10k' sum g
// dlnP/dlog10k + dy' sum g dlnP/dy) + (ln(weight fac))_t sum g, where acc,
// twg, twm are the node cotangents above (the derivative of C_gs along the
// node's a, contracted with cbar, is built from them).
// ----------------------------------------------------------------------------
// GS: the launch holds gs tiles only; their ln P runs fold into the block's
// band (added into g_lnP after the gather), the ss tiles' runs go to the run
// slab.  OL (gs, no checkpoint, no lens-limit term): draws with one-loop
// galaxy bias add W_gal O per (row, node) inside the source weights' factor,
// O = g^4 fac (b2 d1d2 + bs2 d1s2 + b3 d1p3) / 2 + bK k^2 P fac
// (cosmo2D.c:2027-2057): its cotangents go to B_zs, W_gal and, through the
// bK term, ln P (bs2 held fixed: d/d b1 is taken at the uploaded bs2).
template <int NS, bool CK, bool LIM = false, bool GS = false, bool OL = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(NS < CLH_MAXB ? 3 : 1))) void k_limber_sx_adj(const DrawDesc* __restrict__ D,
                                                      const int4* __restrict__ work) {
  // NS: the source-bin count when it is at most 10 (loops fully unrolled to
  // it), else CLH_MAXB with the count checked at run time.  CK: read the
  // forward's checkpoint instead of looking P(k,z) up again (its own
  // instantiation, so the lighter path gets its own register budget)
  constexpr bool EXACT = NS < CLH_MAXB;
  extern __shared__ double Cb[];  // ss: [ns][ns][T], gs: [ns][T]
  __shared__ double L10[CLH_TILE_ROWS];
  __shared__ double PR[3][CLH_TILE_ROWS];  // per-row prefactors (ss: pref0; gs: pref1, pref2)
  __shared__ double s_lk[CLH_LDS_LK];
  __shared__ typename std::conditional<GS, ClhBandLds, int>::type bd;
  const int4 w = work[clh_xcd_item(blockIdx.x, gridDim.x)];
  const DrawDesc& d = D[w.x];
  const int kind = w.y & 3, zl = w.y >> 2, row0 = w.z, T = w.w;
  const int p = threadIdx.x;
  constexpr int np = CLH_NP_SS;
  const int nrow = d.nrow, ns = EXACT ? NS : d.ns;
  static_assert(!(LIM && CK), "the lens-limit term walks P(k,z) again");
  static_assert(GS || !LIM, "the lens-limit term is on gs tiles");
  static_assert(!OL || (GS && !CK && !LIM), "one-loop terms: gs tiles without checkpoint or lens-limit term");
  const bool ss = !GS && kind == CLH_W_SS;
  const double* o = clh_sx_nodes(d, kind, zl);
  const double* lut = o + clh_sx_lut_col(d, kind) * np;
  const double* ells = ss ? d.ell_ss : d.ell_gx;
  if (ss) {
    for (int i = p; i < ns * ns * T; i += 64) {
      const int b = i / (ns * T), jj = (i / T) % ns, r = row0 + i % T;
      Cb[i] = (b == jj ? 2.0 : 1.0) * d.cbar[(size_t)clh_ss_index(ns, b, jj) * nrow + r];
    }
  } else {
    const int col0 = d.n_ss + zl * ns;
    for (int i = p; i < ns * T; i += 64) Cb[i] = d.cbar[(size_t)(col0 + i / T) * nrow + row0 + i % T];
  }
  for (int t = p; t < T; t += 64) {
    const double l = ells[row0 + t], ell = l + 0.5;
    L10[t] = log10(ell);
    if (ss) {
      PR[0][t] = l * (l - 1.) * (l + 1.) * (l + 2.) / (ell * ell * ell * ell);
    } else {
      const double tmp = (l - 1.) * l * (l + 1.) * (l + 2.);
      PR[1][t] = l * (l + 1.) / (ell * ell);
      PR[2][t] = (tmp > 0) ? sqrt(tmp) / (ell * ell) : 0.0;
    }
  }
  const double* lk = clh_stage(d.lk, d.nk, s_lk, CLH_LDS_LK);
  __syncthreads();
  const double fac = o[2 * np + p];
  const int j = (int)lut[p];
  const double dy = lut[np + p], l10fk = lut[2 * np + p];
  const int woff = ss ? 3 : 5;
  double Wn[NS], acc[NS];
#pragma unroll
  for (int b = 0; b < NS; b++) {
    Wn[b] = (EXACT || b < ns) ? o[(woff + b) * np + p] : 0.0;
    acc[b] = 0.0;
  }
  const double WG = ss ? 0.0 : o[3 * np + p], WM = ss ? 0.0 : o[4 * np + p];
  const double b1 = ss ? 0.0 : d.par->b1[zl], bmag = ss ? 0.0 : d.par->bmag[zl];
  double twg = 0.0, twm = 0.0, tow = 0.0;  // tow: OL part of the W_gal cotangent
  const bool ol = OL && d.oneloop;
  double ol_g4f = 0.0, ol_fk = 1.0, ol_bk = 0.0;
  if (ol) {
    const double gr = clh_growfac(d, o[p]);
    ol_g4f = gr * gr * gr * gr * fac;
    ol_fk = o[np + p];
    ol_bk = d.par->bk[zl];
  }
  // the ln P cotangent is wanted when the draw has a run slab (uniform per block)
  const bool lnp = d.run != nullptr;
  ClhLnPRunT<typename std::conditional<GS, ClhBand, ClhSlab>::type> sc;
  if constexpr (GS) {
    sc.out.A = bd.A;
    if (lnp) clh_band_setup(bd, d, lk, L10[0], L10[T - 1], j, dy, l10fk, true, sc.out);
  } else {
    sc.out.R = clh_run_row(d, kind, zl, p);
  }
  // per row: S = P fac of the node (from the forward's checkpoint, or the
  // P(k,z) lookup again), its k cell (i, dx), then the cotangent terms
  auto row = [&](int rr, double S, int i, double dx) {
    double g;
    if (ss) {
      const double c = PR[0][rr] * S;
      double sb = 0.0;
      const double* cr = Cb + rr;
#pragma unroll
      for (int b = 0; b < NS; b++) {
        if (EXACT || b < ns) {
          double u = 0.0;
#pragma unroll
          for (int jj = 0; jj < NS; jj++)
            if (EXACT || jj < ns) u += cr[(b * ns + jj) * T] * Wn[jj];
          acc[b] += c * u;
          sb += Wn[b] * u;
        }
      }
      g = 0.5 * sb * c;
    } else {
      const double pref1 = PR[1][rr], pref2 = PR[2][rr];
      const double gp = WG * b1 + WM * pref1 * bmag;
      double O = 0.0, Op = 0.0;  // one-loop term and its P-proportional part
      if (ol) {
        const double k = (ells[row0 + rr] + 0.5) / ol_fk;
        O = clh_oneloop_gs(d, zl, log(k), k, ol_g4f, S);
        Op = ol_bk * k * k * S;
      }
      const double cg = pref2 * (gp * S + WG * O);
      double sv = 0.0;
#pragma unroll
      for (int b = 0; b < NS; b++) {
        if (EXACT || b < ns) {
          const double cb = Cb[b * T + rr];
          sv += cb * Wn[b];
          acc[b] += cg * cb;
        }
      }
      const double t = pref2 * S * sv;
      twg += t;
      twm += t * pref1;
      g = t * gp;
      if (ol) {
        g += pref2 * sv * WG * Op;
        tow += pref2 * sv * O;
      }
    }
    if (lnp) sc.add(i, dx, g);
    return g;
  };
  double slk = 0.0, slz = 0.0, sgl = 0.0;  // LIM: sum g dlnP/dlog10k, sum g dlnP/dy, sum g
  if constexpr (CK) {
    // the forward's checkpoint of S (d.sxk); eight rows loaded ahead of their use
    const double* sk = d.sxk + ((size_t)(ss ? 0 : 1 + zl) * nrow + row0) * np + p;
    ClhKCell kc;
    kc.nk = d.nk;
    for (int r8 = 0; r8 < T; r8 += 8) {
      double sv8[8];
#pragma unroll
      for (int u = 0; u < 8; u++) sv8[u] = sk[(size_t)(r8 + u < T ? r8 + u : T - 1) * np];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        if (r8 + u >= T) break;
        const double dx = kc.dx(lk, L10[r8 + u] - l10fk);
        row(r8 + u, sv8[u], kc.i, dx);
      }
    }
  } else {
    ClhPkNodeT<false, LIM> pkn;  // no P-row look-ahead: it would cost a wave per SIMD
    pkn.init(d, j, dy);
    for (int rr = 0; rr < T; rr++) {
      double dx;
      const double out = pkn.lnp(lk, L10[rr] - l10fk, dx);
      const double g = row(rr, exp(out) * (1.0 / (CLH_COVERH0 * CLH_COVERH0 * CLH_COVERH0)) * fac, pkn.i, dx);
      if (LIM) {
        slk += g * pkn.dlnp_dk();
        slz += g * pkn.dlnp_dy(dx);
        sgl += g;
      }
    }
  }
  if (lnp) sc.done();
  if (ss) {
#pragma unroll
    for (int b = 0; b < NS; b++)
      if ((EXACT || b < ns) && acc[b] != 0.0) atomicAdd(d.abar + (size_t)b * np + p, acc[b]);
  } else {
#pragma unroll
    for (int b = 0; b < NS; b++)
      if ((EXACT || b < ns) && acc[b] != 0.0) atomicAdd(d.bbar + ((size_t)zl * ns + b) * np + p, acc[b]);
    atomicAdd(d.wbar + ((size_t)zl * 2 + 0) * np + p, twg * b1 + tow);
    atomicAdd(d.wbar + ((size_t)zl * 2 + 1) * np + p, twm * bmag);
    const double b1p = clh_wave_sum(twg * WG);
    if (p == 0 && b1p != 0.0) atomicAdd(d.g_b1 + zl, b1p);
    if (LIM) {
      const double* tn = d.ntan + (size_t)zl * (6 + ns) * np;
      double sb = 0.0;
#pragma unroll
      for (int b = 0; b < NS; b++)
        if (EXACT || b < ns) sb += tn[(6 + b) * np + p] * acc[b];
      double lim = tn[p] * (sb + tn[4 * np + p] * b1 * twg + tn[5 * np + p] * bmag * twm + tn[2 * np + p] * slk +
                            tn[3 * np + p] * slz) +
                   tn[np + p] * sgl;
      lim = clh_wave_sum(lim);
      if (p == 0 && lim != 0.0) atomicAdd(d.g_lens + zl, lim);
    }
  }
  if constexpr (GS) {
    if (lnp) {
      __syncthreads();
      clh_band_flush(bd, d);
    }
  }
}

// ----------------------------------------------------------------------------
// k_tatt_adj: adjoint of k_limber_sx_tatt (TATT ss and gs tiles; separate
// launches, GS the gs ones).  One wave per work item (draw, kind + 4 zl,
// row0, rows), thread = GL node walking the tile's rows with the forward's
// P(k,z) lookup; the node vectors A_b = W_K - C1 W_S, B_b = W_S C1 b_TA,
// D_b = W_S C2 (k_nodes) and their cotangents stay in registers (NS source
// bins, at most 10).  Per (row, node), with S = P fac and E, M, T, X, TT, T',
// X', TT' the FAST-PT IA functions times D^4 fac at k = (l+1/2)/chi:
//  ss, F / G the EE / BB row cotangents as symmetric [ns][ns] matrices with
//  the diagonal doubled, u^V_b = sum_j F_bj V_j, v^V_b = sum_j G_bj V_j:
//    A_b: pref (u^A S - u^B E + 5 u^D M)
//    B_b: pref (-u^A E + u^B T - 5 u^D X + v^B T' - 5 v^D X')
//    D_b: pref (5 u^A M - 5 u^B X + 25 u^D TT - 5 v^B X' + 25 v^D TT')
//    ln P: pref S sum_b A_b u^A_b / 2
//  gs (lens bin l, gal = W_gal b1 + W_mag pref1 b_mag), c_s the row's gs
//  cotangents: A_s += pref2 gal S c_s, B_s -= pref2 gal E c_s, D_s += 5 pref2
//  gal M c_s; W_gal, W_mag and b1 through pref2 sum_s c_s (A_s S - B_s E +
//  5 D_s M); ln P pref2 gal S sum_s c_s A_s.
// The FAST-PT tables are inputs (no ln P dependence).  The ss runs go to the
// run slab, the gs runs to the block's band (after the gather).
// ----------------------------------------------------------------------------
template <int NS, bool GS>
__global__ __launch_bounds__(64) void k_tatt_adj(const DrawDesc* __restrict__ D, const int4* __restrict__ work) {
  extern __shared__ double Cb[];  // ss: EE [ns][ns][T] then BB [ns][ns][T]; gs: [ns][T]
  __shared__ double L10[CLH_TILE_ROWS], LN[CLH_TILE_ROWS];
  __shared__ double PR[3][CLH_TILE_ROWS];
  __shared__ double s_lk[CLH_LDS_LK];
  __shared__ typename std::conditional<GS, ClhBandLds, int>::type bd;
  const int4 w = work[clh_xcd_item(blockIdx.x, gridDim.x)];
  const DrawDesc& d = D[w.x];
  const int zl = w.y >> 2, row0 = w.z, T = w.w;
  const int kind = GS ? CLH_W_GS : CLH_W_SS;
  const int p = threadIdx.x;
  constexpr int np = CLH_NP_SS, ns = NS;
  const int nrow = d.nrow;
  const double* o = clh_sx_nodes(d, kind, zl);
  const double* lut = o + clh_sx_lut_col(d, kind) * np;
  const double* ells = GS ? d.ell_gx : d.ell_ss;
  const double* tv = d.tatt + (size_t)(GS ? 1 + zl : 0) * (2 * ns + 1) * np;
  if (!GS) {
    for (int i = p; i < ns * ns * T; i += 64) {
      const int b = i / (ns * T), jj = (i / T) % ns, r = row0 + i % T;
      const size_t q = (size_t)clh_ss_index(ns, b, jj) * nrow + r;
      const double f = b == jj ? 2.0 : 1.0;
      Cb[i] = f * d.cbar[q];
      Cb[ns * ns * T + i] = f * d.cbar_bb[q];
    }
  } else {
    const int col0 = d.n_ss + zl * ns;
    for (int i = p; i < ns * T; i += 64) Cb[i] = d.cbar[(size_t)(col0 + i / T) * nrow + row0 + i % T];
  }
  for (int t = p; t < T; t += 64) {
    const double l = ells[row0 + t], ell = l + 0.5;
    L10[t] = log10(ell);
    LN[t] = log(ell);
    PR[0][t] = l * (l - 1.) * (l + 1.) * (l + 2.) / (ell * ell * ell * ell);
    const double tmp = (l - 1.) * l * (l + 1.) * (l + 2.);
    PR[1][t] = l * (l + 1.) / (ell * ell);
    PR[2][t] = (tmp > 0) ? sqrt(tmp) / (ell * ell) : 0.0;
  }
  const double* lk = clh_stage(d.lk, d.nk, s_lk, CLH_LDS_LK);
  __syncthreads();
  const double fac = o[2 * np + p];
  const int j = (int)lut[p];
  const double dy = lut[np + p], l10fk = lut[2 * np + p], lnfk = log(o[np + p]);
  const double g4f = tv[2 * ns * np + p] * fac;
  const int woff = GS ? 5 : 3;
  double Av[NS], Bv[NS], Dv[NS], aA[NS], aB[NS], aD[NS];
#pragma unroll
  for (int b = 0; b < NS; b++) {
    Av[b] = o[(woff + b) * np + p];
    Bv[b] = tv[b * np + p];
    Dv[b] = tv[(ns + b) * np + p];
    aA[b] = aB[b] = aD[b] = 0.0;
  }
  const double WG = GS ? o[3 * np + p] : 0.0, WM = GS ? o[4 * np + p] : 0.0;
  const double b1 = GS ? d.par->b1[zl] : 0.0, bmag = GS ? d.par->bmag[zl] : 0.0;
  double twg = 0.0, twm = 0.0;
  const bool lnp = d.run != nullptr;  // uniform per block
  ClhLnPRunT<typename std::conditional<GS, ClhBand, ClhSlab>::type> sc;
  if constexpr (GS) {
    sc.out.A = bd.A;
    if (lnp) clh_band_setup(bd, d, lk, L10[0], L10[T - 1], j, dy, l10fk, true, sc.out);
  } else {
    sc.out.R = clh_run_row(d, kind, zl, p);
  }
  ClhPkNodeT<false> pkn;
  pkn.init(d, j, dy);
  constexpr double c3 = 1.0 / (CLH_COVERH0 * CLH_COVERH0 * CLH_COVERH0);
  for (int rr = 0; rr < T; rr++) {
    double dx;
    const double S = exp(pkn.lnp(lk, L10[rr] - l10fk, dx)) * c3 * fac;
    const double lnk = LN[rr] - lnfk;
    const double E = g4f * (clh_fpt(d, 2, lnk) + clh_fpt(d, 3, lnk));
    const double M = g4f * (clh_fpt(d, 6, lnk) + clh_fpt(d, 7, lnk));
    double g;
    if constexpr (!GS) {
      const double Tt = g4f * clh_fpt(d, 4, lnk), X = g4f * clh_fpt(d, 8, lnk), TT = g4f * clh_fpt(d, 0, lnk);
      const double Tp = g4f * clh_fpt(d, 5, lnk), Xp = g4f * clh_fpt(d, 9, lnk), TTp = g4f * clh_fpt(d, 1, lnk);
      const double pref = PR[0][rr];
      double sAA = 0.0;
#pragma unroll
      for (int b = 0; b < NS; b++) {
        double uA = 0.0, uB = 0.0, uD = 0.0, vB = 0.0, vD = 0.0;
        const double* F = Cb + (size_t)(b * ns) * T + rr;
        const double* G = F + (size_t)ns * ns * T;
#pragma unroll
        for (int jj = 0; jj < NS; jj++) {
          const double f = F[jj * T], gg = G[jj * T];
          uA += f * Av[jj];
          uB += f * Bv[jj];
          uD += f * Dv[jj];
          vB += gg * Bv[jj];
          vD += gg * Dv[jj];
        }
        aA[b] += pref * (uA * S - uB * E + 5.0 * uD * M);
        aB[b] += pref * (-uA * E + uB * Tt - 5.0 * uD * X + vB * Tp - 5.0 * vD * Xp);
        aD[b] += pref * (5.0 * uA * M - 5.0 * uB * X + 25.0 * uD * TT - 5.0 * vB * Xp + 25.0 * vD * TTp);
        sAA += Av[b] * uA;
      }
      g = 0.5 * pref * S * sAA;
    } else {
      const double pref1 = PR[1][rr], pref2 = PR[2][rr];
      const double gal = WG * b1 + WM * pref1 * bmag;
      const double cS = pref2 * gal * S, cE = pref2 * gal * E, cM = 5.0 * pref2 * gal * M;
      double stot = 0.0, sA = 0.0;
#pragma unroll
      for (int b = 0; b < NS; b++) {
        const double cb = Cb[b * T + rr];
        aA[b] += cS * cb;
        aB[b] -= cE * cb;
        aD[b] += cM * cb;
        stot += cb * (Av[b] * S - Bv[b] * E + 5.0 * Dv[b] * M);
        sA += cb * Av[b];
      }
      const double t = pref2 * stot;
      twg += t;
      twm += t * pref1;
      g = pref2 * gal * S * sA;
    }
    if (lnp) sc.add(pkn.i, dx, g);
  }
  if (lnp) sc.done();
  double* tb = d.tbar + (size_t)(GS ? 1 + zl : 0) * 2 * ns * np;
#pragma unroll
  for (int b = 0; b < NS; b++) {
    double* ab = GS ? d.bbar + ((size_t)zl * ns + b) * np : d.abar + (size_t)b * np;
    if (aA[b] != 0.0) atomicAdd(ab + p, aA[b]);
    if (aB[b] != 0.0) atomicAdd(tb + (size_t)b * np + p, aB[b]);
    if (aD[b] != 0.0) atomicAdd(tb + (size_t)(ns + b) * np + p, aD[b]);
  }
  if constexpr (GS) {
    atomicAdd(d.wbar + ((size_t)zl * 2 + 0) * np + p, twg * b1);
    atomicAdd(d.wbar + ((size_t)zl * 2 + 1) * np + p, twm * bmag);
    const double b1p = clh_wave_sum(twg * WG);
    if (p == 0 && b1p != 0.0) atomicAdd(d.g_b1 + zl, b1p);
    if (lnp) {
      __syncthreads();
      clh_band_flush(bd, d);
    }
  }
}

// ----------------------------------------------------------------------------
// k_limber_gg_adj: adjoint of k_limber_gg; work item (draw, zl, row0,
// nrows), two groups of 96 node threads, group h walking rows
// h ceil(nrows/2) .. of the tile (ClhGgNode from the k_nodes search results).
// Per (row, node) with integrand t^2 P fac, t = W_gal b1 + W_mag pref bmag +
// W_RSD: ln P cotangent cbar t^2 P fac; tb = 2 cbar t P fac feeds b1 (W_gal),
// the W_gal / W_mag node cotangents and the lens shift through dW_RSD/d dz.
// The ln P cotangent goes through the block's band (above).
// ----------------------------------------------------------------------------
// Three waves per SIMD are asked for: the LIM instantiation would otherwise
// take 192 VGPRs (two waves) and run 8% slower than with a few spills.
// F32: the integrand evaluated in FP32 (as k_limber_gg<F32>), the cotangent
// sums in FP64 (the mixed-precision modes' VJP).
// OL (no checkpoint, no lens-limit term): draws with one-loop galaxy bias
// add W_gal^2 (g^4 fac Q(k) + 2 b1 bK k^2 P fac) per (row, node)
// (cosmo2D.c:2560-2603): its cotangents go to W_gal, b1 (at the uploaded
// bs2) and, through the bK term, ln P.
template <bool CK, bool LIM = false, bool F32 = false, bool OL = false>
__global__ __launch_bounds__(192) __attribute__((amdgpu_waves_per_eu(3))) void k_limber_gg_adj(const DrawDesc* __restrict__ D,
                                                       const int4* __restrict__ work, int stage) {
  __shared__ ClhGgRow R[CLH_TILE_ROWS];
  __shared__ double Cr[CLH_TILE_ROWS];
  __shared__ double s_lk[CLH_LDS_LK];
  __shared__ double red[2][3];
  __shared__ ClhBandLds bd;
  // LIM: per node t0' and chi' (per unit a), dlog10k/da, d dy/da and the a_0
  // half of dW_RSD/d dz, read per row from LDS to spare registers
  __shared__ double TC[LIM ? 5 * CLH_NP_GG : 1];
  static_assert(!(LIM && CK), "the lens-limit term evaluates the integrand again");
  static_assert(!OL || (!CK && !LIM), "one-loop terms: no checkpoint, no lens-limit term");
  extern __shared__ double dyn[];  // staged chi(z), G(z), lens n(z) tables
  const int4 w = work[clh_xcd_item(blockIdx.x, gridDim.x)];
  const DrawDesc& d = D[w.x];
  const int zl = w.y, row0 = w.z, nr = w.w;
  constexpr int np = CLH_NP_GG;
  for (int t = threadIdx.x; t < nr; t += 192) {
    R[t] = clh_gg_row(d.ell_gx[row0 + t]);
    Cr[t] = d.cbar[(size_t)(d.n_ss + d.n_gs + zl) * d.nrow + row0 + t];
  }
  const ClhGgTabs T = clh_gg_tabs(d, zl, dyn, stage, s_lk);
  if (LIM) {
    const double* tn = d.ntan + (size_t)d.nl * (6 + d.ns) * CLH_NP_SS + (size_t)zl * 8 * np;
    const double* lt = clh_gg_nodes(d, zl) + 5 * np;
    for (int q = threadIdx.x; q < np; q += 192) {
      TC[q] = tn[6 * np + q];
      TC[np + q] = tn[7 * np + q];
      TC[2 * np + q] = tn[2 * np + q];
      TC[3 * np + q] = tn[3 * np + q];
      TC[4 * np + q] = lt[8 * np + q];
    }
  }
  __syncthreads();
  const int p = threadIdx.x % np, half = threadIdx.x / np;
  double tg = 0.0, tm = 0.0, b1p = 0.0, lzp = 0.0;
  const double* o = clh_gg_nodes(d, zl);
  const double* lut = o + 5 * np;
  // the ln P cotangent is wanted when the draw has a run slab (uniform per block)
  const bool lnp = d.run != nullptr;
  ClhLnPRunT<ClhBand> sc;
  sc.out.A = bd.A;  // (unconditionally: an LDS pointer throughout, ds_add_f64)
  if (lnp)
    clh_band_setup(bd, d, T.lk, R[0].l10, R[nr - 1].l10, (int)lut[p], lut[np + p], lut[2 * np + p], half == 0,
                   sc.out);
  const int hn = (nr + 1) / 2;
  const int r_lo = half * hn, r_hi = half ? nr : hn;
  const double WG = o[3 * np + p];
  double sl = 0.0, sgl = 0.0;  // LIM sums
  double tol = 0.0, b1o = 0.0;  // OL sums: W_gal and b1 cotangents of the one-loop term
  if constexpr (CK) {
    // the forward's checkpoint (d.ggk): (t, P fac, dW_RSD) per row; only the
    // k cell of the ln P cotangent is walked again
    ClhKCell kc;
    kc.nk = d.nk;
    const double l10fk = lut[2 * np + p];
    // eight rows of the checkpoint are loaded before any is used, so their
    // latencies overlap
    const double* gk = d.ggk + ((size_t)(zl * d.nrow + row0) * 3) * np + p;
    for (int r8 = r_lo; r8 < r_hi; r8 += 8) {
      double tv[8], pv[8], wv[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int rr = r8 + u < r_hi ? r8 + u : r_hi - 1;
        const double* g = gk + (size_t)rr * 3 * np;
        tv[u] = g[0];
        pv[u] = g[np];
        wv[u] = g[2 * np];
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int rr = r8 + u;
        if (rr >= r_hi) break;
        const double cb = Cr[rr];
        const double t = tv[u], pf = pv[u], dwr = wv[u];
        if (cb == 0.0 || pf == 0.0) continue;  // pf = 0: chi_1 beyond chi(a_min), zero integrand
        const double dx = kc.dx(T.lk, R[rr].l10 - l10fk);
        if (lnp) sc.add(kc.i, dx, cb * t * t * pf);
        const double tb = cb * 2.0 * t * pf;
        tg += tb;
        tm += tb * R[rr].pref;
        lzp += tb * dwr;
      }
    }
  } else {
    typedef typename std::conditional<F32, float, double>::type real;
    ClhGgNode<true, false, LIM, real> nd;  // no P-row look-ahead: it would cost a wave per SIMD
    nd.init(d, T, o, p, zl);
    const bool ol = OL && d.oneloop;
    double ol_wg2 = 0.0, ol_g4f = 0.0, ol_lnfk = 0.0, ol_fk = 1.0;
    if (ol) {
      const double gr = clh_growfac(d, o[p]);
      ol_wg2 = WG * WG;
      ol_g4f = gr * gr * gr * gr * o[2 * np + p];
      ol_fk = o[np + p];
      ol_lnfk = log(ol_fk);
    }
    for (int rr = r_lo; rr < r_hi; rr++) {
      const double cb = Cr[rr];
      if (cb == 0.0) continue;
      real t, pf, dwr, dx;
      if (!nd.eval(d, T, R[rr], t, pf, dwr, dx)) continue;
      double g = cb * t * t * pf;
      if (ol) {
        const double ell = R[rr].ell;
        double tot, pp, db1;
        clh_oneloop_gg_parts(d, zl, log(ell) - ol_lnfk, ell / ol_fk, ol_g4f, pf, tot, pp, db1);
        g += cb * ol_wg2 * pp;
        tol += cb * 2.0 * WG * tot;
        b1o += cb * ol_wg2 * db1;
      }
      if (lnp) sc.add(nd.pk.i, dx, g);
      const double tb = cb * 2.0 * t * pf;
      tg += tb;
      tm += tb * R[rr].pref;
      if (LIM) {
        const double* tc = TC + p;
        const double c0 = R[rr].c0;
        lzp += tb * (dwr + c0 * tc[4 * np]);
        sl += tb * (c0 * tc[0] - R[rr].c1 * nd.dhn * tc[np]) +
              g * (nd.pk.dlnp_dk() * tc[2 * np] + nd.pk.dlnp_dy(dx) * tc[3 * np]);
        sgl += g;
      } else {
        lzp += tb * dwr;
      }
    }
  }
  if (lnp) sc.done();
  const double b1 = d.par->b1[zl], bmag = d.par->bmag[zl];
  const size_t base = (size_t)2 * d.nl * CLH_NP_SS + (size_t)zl * 2 * np;
  if (tg != 0.0 || tol != 0.0) atomicAdd(d.wbar + base + p, tg * b1 + tol);
  if (tm != 0.0 && bmag != 0.0) atomicAdd(d.wbar + base + np + p, tm * bmag);
  if (LIM) {
    const double* tn = d.ntan + (size_t)d.nl * (6 + d.ns) * CLH_NP_SS + (size_t)zl * 8 * np;
    lzp += tn[p] * (tn[4 * np + p] * b1 * tg + tn[5 * np + p] * bmag * tm + sl) + tn[np + p] * sgl;
  }
  b1p = clh_wave_sum(WG * tg + b1o);  // d/d b1 through W_gal b1 (W_gal sum tb) and the one-loop term
  lzp = clh_wave_sum(lzp);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { red[0][wid] = b1p; red[1][wid] = lzp; }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double sb = red[0][0] + red[0][1] + red[0][2];
    const double sl = red[1][0] + red[1][1] + red[1][2];
    if (sb != 0.0) atomicAdd(d.g_b1 + zl, sb);
    if (sl != 0.0) atomicAdd(d.g_lens + zl, sl);
  }
  if (lnp) clh_band_flush(bd, d);
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
      if (d.tbar) {  // TATT: B_b = W_S C1 b_TA and D_b = W_S C2 (IA.c:238-420)
        const double* tb = d.tbar + (size_t)seg * 2 * d.ns * CLH_NP_SS;
        const double tB = tb[(size_t)b * CLH_NP_SS + p], tD = tb[(size_t)(d.ns + b) * CLH_NP_SS + p];
        if (tB != 0.0 || tD != 0.0) {
          const double C1 = clh_ia_c1(d, a, gf, b), C2 = clh_ia_c2(d, a, gf, b), bta = clh_ia_bta(d, b);
          const double WS = clh_nz_src(d, z, b) * hh;
          const double dWS = clh_dnz_src(d, z, b) * hh;
          gs += tB * dWS * C1 * bta + tD * dWS * C2;
          gA += tB * WS * c1u * bta;
        }
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
// k_nodes_tan: the motion of the gs and gg GL nodes of lens bin zl with the
// bin's photo-z shift, for the lens-limit part of the lens-shift gradient.
// amin_lens = 1/(1 + zmax + 2|dz|) and, without magnification, amax_lens =
// 1/(1 + max(zmin - 2|dz|, 0.001)) (redshift_spline.c:50-79), so node n sits
// at a = B + A x_n with weight A w_n and both move with dz.  Per node: a_t =
// da/d dz, the relative change of its weight times fac (fac = A w dchi/da /
// chi^2), and the derivatives along a (at fixed ell) of everything the gs /
// gg integrands read from the node: log10 k = log10(ell) - log10(chi c/H0),
// the P(k,z) z weight dy, W_gal, W_mag, the source weights B_zs (gs), the
// a_0 half of W_RSD and chi (gg).  Each is the derivative of the forward's
// own piecewise arithmetic on the node's table intervals (chi_all's linear
// chi and stencil dchi/dz, cosmo3D.c:40-92; the growth table; the n(z)
// spline pieces; the lens-efficiency lerp).  Segment layout as k_nodes; the
// ss segment has none.  dz = 0: zero (the limits' kink).
// ----------------------------------------------------------------------------
// slope of interpol1d (basics.c:457-480) at x: the interval's, zero where it
// extrapolates by a constant
__device__ __forceinline__ double clh_interpol1d_slope(const double* __restrict__ f, int n, double a, double dx,
                                                       double x) {
  if (x < a) return 0.0;
  const int i = (int)floor((x - a) / dx);
  if (i + 1 >= n) return 0.0;
  return (f[i + 1] - f[i]) / dx;
}

__global__ __launch_bounds__(128) void k_nodes_tan(const DrawDesc* __restrict__ D, const int2* __restrict__ work) {
  const int2 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const int seg = w.y;
  if (seg == 0) return;
  const bool gs = seg <= d.nl;
  const int zl = gs ? seg - 1 : seg - 1 - d.nl, ns = d.ns;
  const int np = gs ? CLH_NP_SS : CLH_NP_GG;
  const int p = threadIdx.x;
  if (p >= np) return;
  double* tn = gs ? d.ntan + (size_t)zl * (6 + ns) * CLH_NP_SS
                  : d.ntan + (size_t)d.nl * (6 + ns) * CLH_NP_SS + (size_t)zl * 8 * CLH_NP_GG;
  const double* o = gs ? d.nodes + d.node_gs + (size_t)zl * d.node_gs_stride
                       : d.nodes + d.node_gg + (size_t)zl * d.node_gg_stride;
  const double* lut = o + (gs ? 5 + ns : 5) * np;
  const double dz = d.par->lens_dz[zl];
  const double sg = dz > 0.0 ? 1.0 : (dz < 0.0 ? -1.0 : 0.0);
  // the limits per unit dz
  const double amin = clh_amin_lens(d, zl), amax = clh_amax_lens(d, zl);
  const double amin_t = -2.0 * amin * amin * sg;
  double amax_t = 0.0;
  if (d.par->bmag[zl] == 0) {
    const double zmin = (d.l_zmin[zl] - d.l_zmean[zl]) * d.par->lens_st[zl] + d.l_zmean[zl];
    if (zmin - 2. * clh_lim_adz(d, zl) > 0.001) amax_t = 2.0 * amax * amax * sg;
  }
  const double* glx = d.gl + (np == CLH_NP_SS ? 0 : 2 * CLH_NP_SS);
  const double A = 0.5 * (amax - amin), A_t = 0.5 * (amax_t - amin_t), B_t = 0.5 * (amax_t + amin_t);
  const double u = o[p];
  const double u_t = B_t + A_t * glx[p];
  constexpr double C = CLH_COVERH0;
  // chi_all at a = u (cosmo3D.c:40-92): chi and the stencil dchi/dz linear in z on interval j
  const double z = 1.0 / u - 1.0, z_u = -1.0 / (u * u);
  const double* __restrict__ zz = d.chi_z;
  const double* __restrict__ cc = d.chi_v;
  int j = clh_bs(zz, d.chi_n, z);
  if (j > d.chi_n - 3) j = d.chi_n - 3;
  const double hz = zz[j + 1] - zz[j];
  const double dyc = (z - zz[j]) / hz;
  const double fK = (cc[j] + dyc * (cc[j + 1] - cc[j])) / C;
  const double fK_u = (cc[j + 1] - cc[j]) / hz * z_u / C;
  const double up = (cc[j + 2] - cc[j]) / (zz[j + 2] - zz[j]);
  const double down = j > 0 ? (cc[j + 1] - cc[j - 1]) / (zz[j + 1] - zz[j - 1]) : (cc[j + 1] - cc[j]) / hz;
  const double out1 = down + dyc * (up - down);
  const double dchida = out1 / C / (u * u);
  const double dchida_u = (up - down) / hz * z_u / C / (u * u) - 2.0 * dchida / u;
  const double hh = 1.0 / ((u * u) * dchida);
  const double hh_u = -hh * (2.0 / u + dchida_u / dchida);
  const int jp = (int)lut[0 * np + p];
  tn[p] = u_t;
  tn[np + p] = A_t / A + (dchida_u / dchida - 2.0 * fK_u / fK) * u_t;
  tn[2 * np + p] = -fK_u / (fK * M_LN10);
  tn[3 * np + p] = z_u / (d.pz[jp + 1] - d.pz[jp]);
  // W_gal = n_lens(z) hh, W_mag = wk_pre g_lens (n_lens' = -d n_lens / d dz)
  const double nlv = clh_nz_lens(d, z, zl), nl_z = -clh_dnz_lens(d, z, zl);
  tn[4 * np + p] = nl_z * z_u * hh + nlv * hh_u;
  const double wk_pre = 1.5 * d.omega_m * fK / u;
  const double wk_u = 1.5 * d.omega_m * (fK_u / u - fK / (u * u));
  const double gla = 1.0 / (d.l_zmax_all + 1.0), gda = (0.999999 - gla) / ((double)d.n_a - 1.0);
  const bool gl_in = !(u < gla || u > 1.0 - gda);
  const double* gtl = d.gtab + (size_t)(ns + zl) * d.n_a;
  const double gL = clh_g_lens(d, u, zl), gL_u = gl_in ? clh_interpol1d_slope(gtl, d.n_a, gla, gda, u) : 0.0;
  tn[5 * np + p] = wk_u * gL + wk_pre * gL_u;
  if (gs) {
    // B_zs = wk_pre g_tomo - n_src hh C1 (NLA, IA.c:238-296)
    const int jg = clh_bs(d.G_z, d.G_n, z);
    const double slG = (d.G_v[jg + 1] - d.G_v[jg]) / (d.G_z[jg + 1] - d.G_z[jg]);
    const double Gz = d.G_v[jg] + (z - d.G_z[jg]) * slG;
    const double gf = Gz * u / d.G0, gf_u = (slG * z_u * u + Gz) / d.G0;
    const double gsa = 1.0 / (d.s_zmax_all + 1.0), gsd = (0.999999 - gsa) / ((double)d.n_a - 1.0);
    const bool gs_in = !(u <= gsa || u > 1.0 - gsd);
    for (int zs = 0; zs < ns; zs++) {
      const double gT = clh_g_tomo(d, u, zs);
      const double gT_u = gs_in ? clh_interpol1d_slope(d.gtab + (size_t)zs * d.n_a, d.n_a, gsa, gsd, u) : 0.0;
      const double nsv = clh_nz_src(d, z, zs), ns_z = -clh_dnz_src(d, z, zs);
      const double WS = nsv * hh, WS_u = ns_z * z_u * hh + nsv * hh_u;
      const double C1 = clh_ia_c1(d, u, gf, zs);
      const double C1_u = -C1 * gf_u / gf - (d.ia_mode == 3 ? C1 * d.par->ia_eta / u : 0.0);
      tn[(6 + zs) * np + p] = wk_u * gT + wk_pre * gT_u - WS_u * C1 - WS * C1_u;
    }
  } else {
    // t0 = n_lens(z0) hf0 at a_0 = a_chi(chi) (radial_weights.c:98-99):
    // z0 linear in chi on the chi(z) interval jc, hf0 = (c/H0) f(z0) / out1(z0)
    // with the stencil out1 = dchi/dz and f = 1 - slG (1 + z) / G (cosmo3D.c:200-228)
    const double chi = fK * C;
    const int jc = (int)lut[3 * np + p];
    const double sz = (zz[jc + 1] - zz[jc]) / (cc[jc + 1] - cc[jc]);
    const double z0 = zz[jc] + (chi - cc[jc]) * sz;
    const double z0_u = C * fK_u * sz;
    int j0 = clh_bs(zz, d.chi_n, z0);
    if (j0 > d.chi_n - 3) j0 = d.chi_n - 3;
    const double hz0 = zz[j0 + 1] - zz[j0];
    const double up0 = (cc[j0 + 2] - cc[j0]) / (zz[j0 + 2] - zz[j0]);
    const double dn0 = j0 > 0 ? (cc[j0 + 1] - cc[j0 - 1]) / (zz[j0 + 1] - zz[j0 - 1]) : (cc[j0 + 1] - cc[j0]) / hz0;
    const double o1 = dn0 + (z0 - zz[j0]) / hz0 * (up0 - dn0), o1_z = (up0 - dn0) / hz0;
    const int jg = clh_bs(d.G_z, d.G_n, z0);
    const double slG = (d.G_v[jg + 1] - d.G_v[jg]) / (d.G_z[jg + 1] - d.G_z[jg]);
    const double Gz = d.G_v[jg] + (z0 - d.G_z[jg]) * slG;
    const double f = 1.0 - slG * (1.0 + z0) / Gz;
    const double f_z = -slG / Gz * (1.0 - (1.0 + z0) * slG / Gz);
    const double hf0 = C * f / o1, hf0_z = C * (f_z / o1 - f * o1_z / (o1 * o1));
    const double n0 = clh_nz_lens(d, z0, zl), n0_z = -clh_dnz_lens(d, z0, zl);
    tn[6 * np + p] = (n0_z * hf0 + n0 * hf0_z) * z0_u;
    tn[7 * np + p] = fK_u;
  }
}

// ----------------------------------------------------------------------------
// k_lim_grad: the part of the lens-shift gradient carried by the integration
// limits.  amin_lens / amax_lens move with |dz| (redshift_spline.c:50-79), so
// the GL nodes of the bin's gs and gg integrals move with it; the adjoint
// kernels differentiate at fixed nodes.  P / M: views of the draws whose
// limits take |dz| + eps / |dz| - eps (DrawDesc::lim_eps) with their gs and gg
// C_ell rows.  The integrands jump where a node crosses a knot of a table
// whose slope enters them (the growth rate from G(z), the dchi/dz stencil):
// a row's forward and backward differences then disagree, and the one of
// smaller magnitude is the row's slope (a jump adds |jump| / eps to the
// other); else their mean.  One block per (draw, lens bin):
// sign(dz) sum_(c, r) cbar[c][r] slope[c][r] over the bin's gs columns and
// its gg column (and its C_gk column under 6x2pt), into g_lens; nothing for |dz| <= 2 eps (the forward's kink
// at dz = 0, where the dependence on |dz| is symmetric).
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_lim_grad(const DrawDesc* __restrict__ D, const DrawDesc* __restrict__ P,
                                                  const DrawDesc* __restrict__ M, const int2* __restrict__ work) {
  __shared__ double red[4];
  const int2 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const double* __restrict__ cp = P[w.x].cl;
  const double* __restrict__ cm = M[w.x].cl;
  const double eps = P[w.x].lim_eps;
  const int zl = w.y, ns = d.ns, nrow = d.nrow;
  const double dz = d.par->lens_dz[zl];
  if (!(fabs(dz) > 2.0 * eps)) return;
  double acc = 0.0;
  const int n = (ns + 1) * nrow;
  for (int t = threadIdx.x; t < n; t += 256) {
    const int k = t / nrow, r = t % nrow;
    const int col = k < ns ? d.n_ss + zl * ns + k : d.n_ss + d.n_gs + zl;
    const size_t o = (size_t)col * nrow + r;
    const double c0 = d.cl[o], sp = cp[o] - c0, sm = c0 - cm[o];
    const double s = fabs(sp - sm) <= 0.05 * (fabs(sp) + fabs(sm)) ? 0.5 * (sp + sm)
                   : (fabs(sp) < fabs(sm) ? sp : sm);
    acc += d.cbar[o] * s;
  }
  if (d.cbar6) {  // 6x2pt: the bin's C_gk column (its limits are the gs / gg ones)
    const double *c6 = d.cl6 + (size_t)zl * d.nrow6, *cb6 = d.cbar6 + (size_t)zl * d.nrow6;
    const double *p6 = P[w.x].cl6 + (size_t)zl * d.nrow6, *m6 = M[w.x].cl6 + (size_t)zl * d.nrow6;
    for (int r = threadIdx.x; r < d.nrow6; r += 256) {
      const double c0 = c6[r], sp = p6[r] - c0, sm = c0 - m6[r];
      const double s = fabs(sp - sm) <= 0.05 * (fabs(sp) + fabs(sm)) ? 0.5 * (sp + sm)
                     : (fabs(sp) < fabs(sm) ? sp : sm);
      acc += cb6[r] * s;
    }
  }
  const double sum = clh_block_sum(acc, red);
  if (threadIdx.x == 0 && sum != 0.0) atomicAdd(d.g_lens + zl, (dz > 0 ? sum : -sum) / eps);
}

// ----------------------------------------------------------------------------
// k_lnp_gather: the ss part of a draw's ln P cotangent from the per-node
// runs.  Work item (draw, i0): lnP rows i0 .. i0 + ic - 1, every z column.
// ss node q with z column j and weight dy adds R_q[i] (1 - dy) to cell
// (i, j) and R_q[i] dy to (i, j + 1) in an LDS tile (LDS atomics: nodes of
// two groups may share a column); the tile then goes out with coalesced
// stores and covers every cell, so g_lnP needs no zeroing (the gs and gg
// adjoints add their bands on top afterwards).  The nodes' (j, dy) are
// staged in LDS first; group g of ic lanes takes nodes g, g + 256/ic, ...,
// lane l reading R_q[i0 + l], four nodes per step.
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_lnp_gather(const DrawDesc* __restrict__ D,
                                                    const int2* __restrict__ work, int ic) {
  extern __shared__ double A[];  // [ic][nz], then dy[total], then j[total]
  const int2 w = work[clh_xcd_item(blockIdx.x, gridDim.x)];
  const DrawDesc& d = D[w.x];
  const int nz = d.nzp, nk = d.nk, i0 = w.y;
  const int total = CLH_NP_SS;  // the ss nodes (gs and gg fold theirs in LDS bands)
  double* s_dy = A + ic * nz;
  int* s_j = (int*)(s_dy + total);
  for (int t = threadIdx.x; t < ic * nz; t += 256) A[t] = 0.0;
  for (int q = threadIdx.x; q < total; q += 256) {
    const double* lut = clh_sx_nodes(d, CLH_W_SS, 0) + clh_sx_lut_col(d, CLH_W_SS) * CLH_NP_SS;
    const int p = q;
    s_j[q] = (int)lut[p];
    s_dy[q] = lut[CLH_NP_SS + p];
  }
  __syncthreads();
  const int li = threadIdx.x % ic, grp = threadIdx.x / ic, ng = 256 / ic;
  const int i = i0 + li;
  const bool in = i < nk;
  const double* __restrict__ R = d.run + i;
  double* Al = A + li * nz;
  int q = grp;
  for (; q + 3 * ng < total; q += 4 * ng) {
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) v[u] = in ? R[(size_t)(q + u * ng) * nk] : 0.0;
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (v[u] != 0.0) {
        const int j = s_j[q + u * ng];
        const double dy = s_dy[q + u * ng];
        atomicAdd(Al + j, v[u] * (1 - dy));
        atomicAdd(Al + j + 1, v[u] * dy);
      }
    }
  }
  for (; q < total; q += ng) {
    const double v = in ? R[(size_t)q * nk] : 0.0;
    if (v != 0.0) {
      const int j = s_j[q];
      const double dy = s_dy[q];
      atomicAdd(Al + j, v * (1 - dy));
      atomicAdd(Al + j + 1, v * dy);
    }
  }
  __syncthreads();
  const int rows = nk - i0 < ic ? nk - i0 : ic;
  double* out = d.g_lnP + (size_t)i0 * nz;
  for (int t = threadIdx.x; t < rows * nz; t += 256) out[t] = A[t];
}

// ----------------------------------------------------------------------------
hipError_t clh_launch_lim_grad(const DrawDesc* D, const DrawDesc* P, const DrawDesc* M, const int2* work, int n,
                               hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_lim_grad, dim3(n), dim3(256), 0, s, D, P, M, work);
  return hipGetLastError();
}
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
hipError_t clh_launch_leg_adjoint(const DrawDesc* D, const int4* work, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_leg_adjoint, dim3(n), dim3(256), 0, s, D, work);
  return hipGetLastError();
}
hipError_t clh_launch_m_grad(const DrawDesc* D, int B, hipStream_t s) {
  hipLaunchKernelGGL(k_m_grad, dim3(B), dim3(64), 0, s, D);
  return hipGetLastError();
}
hipError_t clh_launch_limber_sx_adj(const DrawDesc* D, const int4* work, int n, int max_npair, int ns,
                                   int ck, int lim, int gs, int ol, hipStream_t s) {
  // ol: gs tiles of a batch with one-loop bias (no checkpoint, no lens-limit term)
  // max_npair: largest npair x rows of any work item (doubles of LDS); gs:
  // the items are gs tiles (LDS band); lim: with the lens-limit term (no
  // checkpoint)
  if (n <= 0) return hipSuccess;
  const size_t lds = sizeof(double) * (size_t)max_npair;  // max over items of npair x rows
#define CLH_SXADJ(NSV)                                                                                        \
  do {                                                                                                        \
    if (gs && ol) hipLaunchKernelGGL((k_limber_sx_adj<NSV, false, false, true, true>), dim3(n), dim3(64), lds, s, D, work); \
    else if (gs && lim) hipLaunchKernelGGL((k_limber_sx_adj<NSV, false, true, true>), dim3(n), dim3(64), lds, s, D, work); \
    else if (gs && ck) hipLaunchKernelGGL((k_limber_sx_adj<NSV, true, false, true>), dim3(n), dim3(64), lds, s, D, work); \
    else if (gs) hipLaunchKernelGGL((k_limber_sx_adj<NSV, false, false, true>), dim3(n), dim3(64), lds, s, D, work); \
    else if (ck) hipLaunchKernelGGL((k_limber_sx_adj<NSV, true>), dim3(n), dim3(64), lds, s, D, work);          \
    else hipLaunchKernelGGL((k_limber_sx_adj<NSV, false>), dim3(n), dim3(64), lds, s, D, work);                \
  } while (0)
  switch (ns) {
    case 1: CLH_SXADJ(1); break;
    case 2: CLH_SXADJ(2); break;
    case 3: CLH_SXADJ(3); break;
    case 4: CLH_SXADJ(4); break;
    case 5: CLH_SXADJ(5); break;
    case 6: CLH_SXADJ(6); break;
    case 7: CLH_SXADJ(7); break;
    case 8: CLH_SXADJ(8); break;
    case 9: CLH_SXADJ(9); break;
    case 10: CLH_SXADJ(10); break;
    default: CLH_SXADJ(CLH_MAXB); break;
  }
#undef CLH_SXADJ
  return hipGetLastError();
}
hipError_t clh_launch_tatt_adj(const DrawDesc* D, const int4* work, int n, int lds_doubles, int ns, int gs,
                               hipStream_t s) {
  // lds_doubles: largest cotangent tile of the items (ss: 2 ns^2 rows, gs: ns rows)
  if (n <= 0) return hipSuccess;
  const size_t lds = sizeof(double) * (size_t)lds_doubles;
#define CLH_TADJ(NSV)                                                                           \
  do {                                                                                          \
    if (gs) hipLaunchKernelGGL((k_tatt_adj<NSV, true>), dim3(n), dim3(64), lds, s, D, work);      \
    else hipLaunchKernelGGL((k_tatt_adj<NSV, false>), dim3(n), dim3(64), lds, s, D, work);        \
  } while (0)
  switch (ns) {
    case 1: CLH_TADJ(1); break;
    case 2: CLH_TADJ(2); break;
    case 3: CLH_TADJ(3); break;
    case 4: CLH_TADJ(4); break;
    case 5: CLH_TADJ(5); break;
    case 6: CLH_TADJ(6); break;
    case 7: CLH_TADJ(7); break;
    case 8: CLH_TADJ(8); break;
    case 9: CLH_TADJ(9); break;
    case 10: CLH_TADJ(10); break;
    default: return hipErrorInvalidValue;  // the host refuses more than 10 source bins
  }
#undef CLH_TADJ
  return hipGetLastError();
}
hipError_t clh_launch_limber_gg_adj(const DrawDesc* D, const int4* work, int n, int stage_doubles, int ck,
                                   int lim, int f32, int ol, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const size_t lds = sizeof(double) * (size_t)stage_doubles;
  const int st = stage_doubles > 0 ? 1 : 0;
  if (ol) {  // a batch with one-loop bias: no checkpoint, no lens-limit term
    if (f32) hipLaunchKernelGGL((k_limber_gg_adj<false, false, true, true>), dim3(n), dim3(192), lds, s, D, work, st);
    else hipLaunchKernelGGL((k_limber_gg_adj<false, false, false, true>), dim3(n), dim3(192), lds, s, D, work, st);
    return hipGetLastError();
  }
  if (lim && f32) hipLaunchKernelGGL((k_limber_gg_adj<false, true, true>), dim3(n), dim3(192), lds, s, D, work, st);
  else if (lim) hipLaunchKernelGGL((k_limber_gg_adj<false, true>), dim3(n), dim3(192), lds, s, D, work, st);
  else if (f32 && !ck)
    hipLaunchKernelGGL((k_limber_gg_adj<false, false, true>), dim3(n), dim3(192), lds, s, D, work, st);
  else if (ck) hipLaunchKernelGGL((k_limber_gg_adj<true>), dim3(n), dim3(192), lds, s, D, work, stage_doubles > 0 ? 1 : 0);
  else hipLaunchKernelGGL((k_limber_gg_adj<false>), dim3(n), dim3(192), lds, s, D, work, stage_doubles > 0 ? 1 : 0);
  return hipGetLastError();
}
hipError_t clh_launch_lnp_gather(const DrawDesc* D, const int2* work, int n, int ic, int nz, int nl,
                                 hipStream_t s) {
  if (n <= 0) return hipSuccess;
  (void)nl;
  const size_t total = CLH_NP_SS;  // the ss nodes
  const size_t lds = sizeof(double) * ((size_t)ic * nz + total) + sizeof(int) * total;
  hipLaunchKernelGGL(k_lnp_gather, dim3(n), dim3(256), lds, s, D, work, ic);
  return hipGetLastError();
}
hipError_t clh_launch_nodes_tan(const DrawDesc* D, const int2* work, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_nodes_tan, dim3(n), dim3(128), 0, s, D, work);
  return hipGetLastError();
}
hipError_t clh_launch_nodes_adj(const DrawDesc* D, const int2* work, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_nodes_adj, dim3(n), dim3(128), 0, s, D, work);
  return hipGetLastError();
}
