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

#include <type_traits>

#include "clh_device.hpp"
#include "clh_internal.hpp"

#define CLH_BLOCK 256

typedef double clh_v4d __attribute__((ext_vector_type(4)));  // an FP64 MFMA 16x16 accumulator

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


// ----------------------------------------------------------------------------
// k_gtab: one block per (draw, bin).  bin < ns -> g_tomo of source bin,
// otherwise g_lens of lens bin (bin - ns).  Thread t integrates interval t of
// the fine grid (x = 25 points, trapezoids of n/a^2 and n/(chi a^2)), walking
// the n(z) spline and chi(z) indices along its points; a block scan turns
// the interval sums into the table.
// TANGENT: also the same prefix sums over d(n)/d(photo-z shift of the bin),
// the derivative of the table with respect to that shift (into dgtab).
// ----------------------------------------------------------------------------
// The interval records of a draw's chi(z) and G(z) tables (DrawDesc::crec,
// grec) that the gg integrand's a_1 walks read instead of recomputing the
// interval slopes and the dchi/dz stencil (cosmo3D.c:40-133, 200-228) each
// time a walk changes interval.  One block.
__device__ void clh_cell_records(const DrawDesc& d) {
  const int n = d.chi_n;
  const double* zz = d.chi_z;
  const double* cc = d.chi_v;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    double* r = d.crec + 8 * (size_t)j;
    r[0] = cc[j];
    r[1] = zz[j];
    if (j > n - 2) continue;
    const double zS = (zz[j + 1] - zz[j]) / (cc[j + 1] - cc[j]);
    r[2] = zz[j] - cc[j] * zS;
    r[3] = zS;
    const int jj = j > n - 3 ? n - 3 : j;  // chi_all's stencil clamp (cosmo3D.c:58-67)
    const double inv_z = 1.0 / (zz[jj + 1] - zz[jj]);
    const double up = (cc[jj + 2] - cc[jj]) / (zz[jj + 2] - zz[jj]);
    const double down = (jj > 0) ? (cc[jj + 1] - cc[jj - 1]) / (zz[jj + 1] - zz[jj - 1]) : (cc[jj + 1] - cc[jj]) * inv_z;
    const double D1 = inv_z * (up - down);
    r[4] = down - zz[jj] * D1;
    r[5] = D1;
    if (d.zg_same) {
      const double slG = (d.G_v[j + 1] - d.G_v[j]) / (d.G_z[j + 1] - d.G_z[j]);
      r[6] = d.G_v[j] - d.G_z[j] * slG;
      r[7] = slG;
    }
  }
  if (!d.zg_same)
    for (int j = threadIdx.x; j < d.G_n; j += blockDim.x) {
      double* r = d.grec + 3 * (size_t)j;
      r[0] = d.G_z[j];
      if (j > d.G_n - 2) continue;
      const double slG = (d.G_v[j + 1] - d.G_v[j]) / (d.G_z[j + 1] - d.G_z[j]);
      r[1] = d.G_v[j] - d.G_z[j] * slG;
      r[2] = slG;
    }
}

// MODE 0: the tables; 1: the tables and their shift tangents; 2: the
// tangents only (the tables and interval records of these parameters are
// already in place).
template <int MODE>
__global__ __launch_bounds__(CLH_BLOCK) void k_gtab(const DrawDesc* __restrict__ D,
                                                    const int2* __restrict__ work) {
  constexpr bool TANGENT = MODE >= 1, TABLES = MODE <= 1;
  __shared__ double sP[TABLES ? 1024 : 1], sQ[TABLES ? 1024 : 1];
  __shared__ double sPd[TANGENT ? 1024 : 1], sQd[TANGENT ? 1024 : 1];
  __shared__ double red[8];
  const int2 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  if (TABLES && w.y == 0) clh_cell_records(d);  // once per draw: the block of its first bin
  if (TABLES && w.y >= d.ns) {  // the lens bin's n(z) interval records (ClhGgNode)
    const int zl = w.y - d.ns, nr = d.l_rows;
    const double inv_st = 1.0 / d.par->lens_st[zl];
    const double* cf = d.l_coef + (size_t)zl * 4 * (nr - 1);
    double* r = d.nrec + (size_t)zl * 5 * nr;
    for (int j = threadIdx.x; j < nr; j += blockDim.x) {
      r[5 * j] = d.l_zc[j];
      if (j < nr - 1)
        for (int k = 0; k < 4; k++) r[5 * j + 1 + k] = cf[4 * j + k] * inv_st;
    }
  }
  const bool src = w.y < d.ns;
  const int b = src ? w.y : w.y - d.ns;
  const int x = d.gx;  // 25 (1 + |high_def_integration|)
  const int nint = d.n_a - 1;
  const int Na = x * nint + 1;
  const double amax = 0.999999;
  const double amin = 1.0 / ((src ? d.s_zmax_all : d.l_zmax_all) + 1.0);
  const double da = (amax - amin) / ((double)Na - 1.0);

  // g_lens pre-integral from a_min of the shear sample (redshift_spline.c:1052-1071)
  double P0 = 0.0, Q0 = 0.0, P0d = 0.0, Q0d = 0.0;
  if (!src) {
    const double amin_shear = 1.0 / (d.s_zmax_all + 1.0);
    const int Na0 = (amin_shear < amin) ? (int)ceil((amin - amin_shear) / da) + 1 : 1;
    const double da0 = (Na0 > 1) ? (amin - amin_shear) / ((double)Na0 - 1.) : 0.0;
    double p = 0.0, q = 0.0, pd = 0.0, qd = 0.0;
    for (int i = 1 + threadIdx.x; i < Na0; i += CLH_BLOCK) {
      const double apl = amin_shear + (i - 1) * da0;
      const double aph = amin_shear + i * da0;
      if (TABLES) {
        const double fl = clh_nz_lens(d, 1.0 / apl - 1.0, b) / (apl * apl);
        const double fh = clh_nz_lens(d, 1.0 / aph - 1.0, b) / (aph * aph);
        p += 0.5 * da0 * (fl + fh);
        // beyond the bin's n(z) support both ends are zero and so is their
        // n / (chi a^2) term (+0 exactly): chi(a), a bisection, only where it counts
        if (fl != 0.0 || fh != 0.0) q += 0.5 * da0 * (fl / clh_chi(d, apl) + fh / clh_chi(d, aph));
      }
      if (TANGENT) {
        const double gl = clh_dnz_lens(d, 1.0 / apl - 1.0, b) / (apl * apl);
        const double gh = clh_dnz_lens(d, 1.0 / aph - 1.0, b) / (aph * aph);
        pd += 0.5 * da0 * (gl + gh);
        if (gl != 0.0 || gh != 0.0) qd += 0.5 * da0 * (gl / clh_chi(d, apl) + gh / clh_chi(d, aph));
      }
    }
    if (TABLES) {
      P0 = clh_block_sum(p, red);
      Q0 = clh_block_sum(q, red);
    }
    if (TANGENT) {
      P0d = clh_block_sum(pd, red);
      Q0d = clh_block_sum(qd, red);
    }
  }

  // per-interval trapezoid sums of n/a^2 and n/(chi a^2)
  int jlast = -1;  // the thread's last chi(z) interval: where the table loop's walks start
  for (int t = threadIdx.x; t < nint; t += CLH_BLOCK) {
    double p = 0.0, q = 0.0, pd = 0.0, qd = 0.0;
    double fprev = 0.0, gprev = 0.0, fdprev = 0.0, gdprev = 0.0;
    ClhNzRun<TANGENT> nw;
    nw.init(d, src, b);
    ClhChiWalk cw;
    for (int s2 = 0; s2 <= x; s2++) {
      const int i = t * x + s2;
      const double a = amin + i * da;
      // per-point divisions as the reciprocal with two Newton steps (clh_rcp):
      // a and chi are normal and positive on the grid
      const double ia = clh_rcp(a), z = ia - 1.0;
      double n, dn;
      nw.eval(z, n, dn);
      const double chi = cw.chi(d, z);
      const double ichi = clh_rcp(chi);
      if (TABLES) {
        const double f = n * ia * ia;
        const double g = f * ichi;
        if (s2 > 0) {
          p += 0.5 * da * (fprev + f);
          q += 0.5 * da * (gprev + g);
        }
        fprev = f;
        gprev = g;
      }
      if (TANGENT) {
        const double fd = dn * ia * ia;
        const double gd = fd * ichi;
        if (s2 > 0) {
          pd += 0.5 * da * (fdprev + fd);
          qd += 0.5 * da * (gdprev + gd);
        }
        fdprev = fd;
        gdprev = gd;
      }
    }
    jlast = cw.j;
    if (TABLES) {
      sP[t] = p;
      sQ[t] = q;
    }
    if (TANGENT) {
      sPd[t] = pd;
      sQd[t] = qd;
    }
  }
  __syncthreads();

  // chunked prefix scan over intervals
  const int cpt = (nint + CLH_BLOCK - 1) / CLH_BLOCK;
  const int t0 = threadIdx.x * cpt;
  const int t1 = min(t0 + cpt, nint);
  double cp = 0.0, cq = 0.0, cpd = 0.0, cqd = 0.0;
  for (int t = t0; t < t1; t++) {
    if (TABLES) {
      cp += sP[t];
      cq += sQ[t];
    }
    if (TANGENT) {
      cpd += sPd[t];
      cqd += sQd[t];
    }
  }
  double P = 0.0, Q = 0.0;
  if (TABLES) {
    const double ip = clh_block_scan(cp, red);
    const double iq = clh_block_scan(cq, red);
    P = P0 + ip - cp;
    Q = Q0 + iq - cq;
  }
  double Pd = 0.0, Qd = 0.0;
  if (TANGENT) {
    const double ipd = clh_block_scan(cpd, red);
    const double iqd = clh_block_scan(cqd, red);
    Pd = P0d + ipd - cpd;
    Qd = Q0d + iqd - cqd;
  }
  double* tab = d.gtab + (size_t)w.y * d.n_a;
  double* dtab = TANGENT ? d.dgtab + (size_t)w.y * d.n_a : nullptr;
  for (int t = t0; t < t1; t++) {
    const int i = (t + 1) * x;
    const double chi = clh_chi_near(d, amin + i * da, jlast);
    if (TABLES) {
      P += sP[t];
      Q += sQ[t];
      tab[t + 1] = P - chi * Q;
    }
    if (TANGENT) {
      Pd += sPd[t];
      Qd += sQd[t];
      dtab[t + 1] = Pd - chi * Qd;
    }
  }
  if (threadIdx.x == 0) {
    const double chi0 = clh_chi(d, amin);
    if (TABLES) tab[0] = src ? 0.0 : P0 - chi0 * Q0;
    if (TANGENT) dtab[0] = src ? 0.0 : P0d - chi0 * Q0d;
  }
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
  // search results every row tile of this segment starts from
  double* lut;
  {
    const int jp = clh_bs(d.pz, d.nzp, z);
    lut = (seg == 0) ? d.nodes + d.node_ss + (3 + d.ns) * np
        : (seg <= d.nl) ? d.nodes + d.node_gs + (size_t)zl * d.node_gs_stride + (5 + d.ns) * np
                        : d.nodes + d.node_gg + (size_t)zl * d.node_gg_stride + 5 * np;
    lut[p] = (double)jp;
    lut[np + p] = (z - d.pz[jp]) / (d.pz[jp + 1] - d.pz[jp]);
    lut[2 * np + p] = log10(fK * CLH_COVERH0);
  }
  if (seg > d.nl) {
    // a_0 side of W_RSD at chi_0 = chi (radial_weights.c:98-99; a_chi cosmo3D.c:109-133)
    const double chi = fK * CLH_COVERH0;
    const int jc = clh_bs(d.chi_v, d.chi_n, chi);
    const double z0 = d.chi_z[jc] + (chi - d.chi_v[jc]) * (1.0 / (d.chi_v[jc + 1] - d.chi_v[jc])) *
                                        (d.chi_z[jc + 1] - d.chi_z[jc]);
    const double a0 = 1.0 / (1.0 + z0);
    double c0, dc0;
    clh_chi_all(d, a0, c0, dc0);
    const double zz0 = 1.0 / a0 - 1.0;
    const double hf0 = 1.0 / ((a0 * a0) * dc0) * clh_f_growth(d, zz0);
    lut[3 * np + p] = (double)jc;
    lut[4 * np + p] = (double)clh_bs(d.chi_z, d.chi_n, zz0);
    lut[5 * np + p] = (double)clh_bs(d.G_z, d.G_n, zz0);
    lut[6 * np + p] = (double)clh_bs(d.l_zc, d.l_rows, zz0);
    lut[7 * np + p] = clh_nz_lens(d, zz0, zl) * hf0;
    lut[8 * np + p] = clh_dnz_lens(d, zz0, zl) * hf0;
  }
  if (d.tatt && seg <= d.nl) {
    // TATT node vectors (B_b = W_S C1 b_TA, D_b = W_S C2) and D(a)^4
    double* t = d.tatt + (size_t)seg * (2 * d.ns + 1) * np;
    for (int bb = 0; bb < d.ns; bb++) {
      const double WS = clh_nz_src(d, z, bb) * hh;
      t[bb * np + p] = WS * clh_ia_c1(d, a, gf, bb) * clh_ia_bta(d, bb);
      t[(d.ns + bb) * np + p] = WS * clh_ia_c2(d, a, gf, bb);
    }
    t[2 * d.ns * np + p] = gf * gf * gf * gf;
  }
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
// k_limber_sx: ss and gs tiles.  Work item (draw, kind + 4 zl, row0, nrows):
// up to CLH_TILE_ROWS consecutive rows, processed in sub-tiles of
// CLH_SUB_ROWS.  One thread per GL node (4 waves x 64 nodes) walks 4 rows of
// every sub-tile in increasing ell, so the log10 k index of the P(k,z) lookup
// moves by a step or two between rows; the z index, its weight and
// log10(chi c/H0) come from k_nodes.  S = P fac of the sub-tile goes to LDS.
// The pair sums are then products on FP64 MFMA (v_mfma_f64_16x16x4f64) over
// the 64 nodes: ss C[pair][row] = sum_p (A_i A_j)[p] S[row][p] in tiles of 16
// pairs x 16 rows (one tile per wave in turn); gs C[zs][row] = pref2 (b1 G1 +
// pref1 bmag G2 (+ G3)), G1 = (B_zs W_gal) S^T, G2 = (B_zs W_mag) S^T and,
// with the one-loop bias, G3 = (B_zs W_gal) O^T.
// F32: the ss rows' P(k,z) as expf of the FP64 bilinear value (products and
// sums stay FP64).
// KEEP: S = P fac of every (row, node) also goes to the draw's checkpoint
// (d.sxk) for the VJP.
// ----------------------------------------------------------------------------
template <bool F32, bool KEEP, bool OL = false>
__global__ __launch_bounds__(CLH_BLOCK) void k_limber_sx(const DrawDesc* __restrict__ D,
                                                         const int4* __restrict__ work) {
  // rows of S / O padded by one element: the MFMA B operand reads 16 rows at
  // 4 node indices per wave, which a 64-element stride maps onto few banks
  constexpr int sp = CLH_NP_SS + 1;
  __shared__ double S[CLH_SUB_ROWS * sp];
  __shared__ double W[(CLH_MAXB + 2) * CLH_NP_SS];
  __shared__ double L10[CLH_TILE_ROWS];
  __shared__ double s_lk[CLH_LDS_LK];
  __shared__ double O[OL ? CLH_SUB_ROWS * sp : 1];  // gs one-loop bias term per (row, node)
  __shared__ unsigned char PB[2][CLH_MAXB * (CLH_MAXB + 1) / 2];  // ss pair -> (bin i, bin j)
  const int4 w = work[clh_xcd_item(blockIdx.x, gridDim.x)];
  const DrawDesc& d = D[w.x];
  const int kind = w.y & 3, zl = w.y >> 2, row0 = w.z, nr = w.w;
  const bool ss = kind == CLH_W_SS;
  const bool f32 = F32 && ss;
  const bool ol = OL && !ss && d.oneloop;
  const int tid = threadIdx.x;
  constexpr int np = CLH_NP_SS;
  const double* o = clh_sx_nodes(d, kind, zl);
  const double* lut = o + clh_sx_lut_col(d, kind) * np;
  const double* ells = ss ? d.ell_ss : d.ell_gx;
  for (int t = tid; t < nr; t += CLH_BLOCK) L10[t] = log10(ells[row0 + t] + 0.5);
  const int nw = ss ? d.ns : d.ns + 2;
  for (int i = tid; i < nw * np; i += CLH_BLOCK) W[i] = o[3 * np + i];
  if (ss)
    for (int q = tid; q < d.n_ss; q += CLH_BLOCK) {
      int bi, bj;
      clh_ss_pair(d.ns, q, bi, bj);
      PB[0][q] = (unsigned char)bi;
      PB[1][q] = (unsigned char)bj;
    }
  const double* lk = clh_stage(d.lk, d.nk, s_lk, CLH_LDS_LK);
  __syncthreads();
  const int p = tid & 63, wv = tid >> 6;
  const int lm = p & 15, lq = p >> 4;  // MFMA lane coordinates
  const double l10fk = lut[2 * np + p], fac = o[2 * np + p];
  ClhPkNode pkn;
  pkn.init(d, (int)lut[p], lut[np + p]);
  constexpr double c3 = 1.0 / (CLH_COVERH0 * CLH_COVERH0 * CLH_COVERH0);
  double ol_g4f = 0.0, ol_fk = 1.0;
  if (ol) {
    const double g = clh_growfac(d, o[p]);
    ol_g4f = g * g * g * g * fac;
    ol_fk = o[np + p];
  }
  constexpr int per = CLH_SUB_ROWS / 4;
  for (int sub = 0; sub * CLH_SUB_ROWS < nr; sub++) {
    const int base = sub * CLH_SUB_ROWS;
    for (int q = 0; q < per; q++) {
      const int rl = wv * per + q;  // row within the sub-tile
      const int rr = base + rl;     // row within the tile
      double s = 0.0;
      if (rr < nr) {
        double dx;
        const double out = pkn.lnp(lk, L10[rr] - l10fk, dx);
        s = f32 ? (double)(expf((float)out) * (float)(c3 * fac)) : exp(out) * c3 * fac;
      }
      S[rl * sp + p] = s;
      if (OL) {
        double ov = 0.0;
        if (ol && rr < nr) {
          const double k = (ells[row0 + rr] + 0.5) / ol_fk;
          ov = clh_oneloop_gs(d, zl, log(k), k, ol_g4f, s);
        }
        O[rl * sp + p] = ov;
      }
      if (KEEP && rr < nr) d.sxk[((size_t)(ss ? 0 : 1 + zl) * d.nrow + row0 + rr) * np + p] = s;
    }
    __syncthreads();
    // MFMA operands: A[m][k] on lane m + 16 k, B[k][n] on lane n + 16 k; the
    // result D[4 v + lq][lm] is value v of lane p.  n = lm is the row.
    const double* Sb = S + lm * sp + lq;
    const int rr = base + lm;
    if (ss) {
      const int ntile = (d.n_ss + 15) >> 4;
      for (int t = wv; t < ntile; t += 4) {
        const int pr = 16 * t + lm;
        const bool pin = pr < d.n_ss;
        const double* Ai = W + (pin ? PB[0][pr] : 0) * np + lq;
        const double* Aj = W + (pin ? PB[1][pr] : 0) * np + lq;
        clh_v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
        for (int st = 0; st < np / 4; st++) {
          const double a = pin ? Ai[4 * st] * Aj[4 * st] : 0.0;
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Sb[4 * st], acc, 0, 0, 0);
        }
        if (rr < nr) {
          const double l = ells[row0 + rr], ell = l + 0.5, ell4 = ell * ell * ell * ell;
          const double pref = l * (l - 1.) * (l + 1.) * (l + 2.) / ell4;
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int pair = 16 * t + 4 * v + lq;
            if (pair < d.n_ss) d.cl[(size_t)pair * d.nrow + row0 + rr] = acc[v] * pref;
          }
        }
      }
    } else {
      const int ntile = (d.ns + 15) >> 4;
      const double b1 = d.par->b1[zl], bmag = d.par->bmag[zl];
      const double* Ob = O + lm * sp + lq;
      for (int t = wv; t < ntile; t += 4) {
        const int zs = 16 * t + lm;
        const bool zin = zs < d.ns;
        const double* Bz = W + (2 + (zin ? zs : 0)) * np + lq;
        clh_v4d g1 = {0.0, 0.0, 0.0, 0.0}, g2 = g1, g3 = g1;
#pragma unroll 4
        for (int st = 0; st < np / 4; st++) {
          const double wb = zin ? Bz[4 * st] : 0.0;
          const double a1 = wb * W[4 * st + lq], a2 = wb * W[np + 4 * st + lq];
          const double bs = Sb[4 * st];
          g1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bs, g1, 0, 0, 0);
          g2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, bs, g2, 0, 0, 0);
          if (ol) g3 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, Ob[4 * st], g3, 0, 0, 0);
        }
        if (rr < nr) {
          const double l = ells[row0 + rr], ell = l + 0.5;
          const double pref1 = l * (l + 1.) / (ell * ell);
          const double tmp = (l - 1.) * l * (l + 1.) * (l + 2.);
          const double pref2 = (tmp > 0) ? sqrt(tmp) / (ell * ell) : 0.0;
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int z = 16 * t + 4 * v + lq;
            if (z < d.ns)
              d.cl[(size_t)(d.n_ss + zl * d.ns + z) * d.nrow + row0 + rr] =
                  (g1[v] * b1 + g2[v] * pref1 * bmag + g3[v]) * pref2;
          }
        }
      }
    }
    __syncthreads();
  }
}

// ----------------------------------------------------------------------------
// k_limber_sx_tatt: ss and gs tiles under TATT intrinsic alignments
// (int_for_C_ss_tomo_limber_core, cosmo2D.c:1395-1487; gs, :1919-2006),
// FP64.  Work item as k_limber_sx.  Per node (k_nodes): A_b = W_K - C1 W_S
// (the NLA weight), B_b = W_S C1 b_TA, D_b = W_S C2, g4 = D(a)^4.  Per
// (row, node) the P(k,z) product S = P fac and the FAST-PT IA functions at
// k = (l + 1/2)/chi, each times g4 fac: E = ta_dE1 + ta_dE2, M = mixA + mixB,
// T = ta_EE, X = mix_EE, TT = tt_EE and, for the B mode, T' = ta_BB,
// X' = mix_BB, TT' = tt_BB.  The reference integrands, expanded per node:
//   EE_ij = A_iA_j S - (B_iA_j + A_iB_j) E + 5 (D_iA_j + A_iD_j) M + B_iB_j T
//           - 5 (B_iD_j + D_iB_j) X + 25 D_iD_j TT
//   BB_ij = B_iB_j T' - 5 (B_iD_j + D_iB_j) X' + 25 D_iD_j TT'
//   gs_(l,s) = gal_l (A_s S - B_s E + 5 D_s M),  gal = W_gal b1 + W_mag pref bmag
// Block of 256 threads: 4 rows x 64 nodes per step (each thread walks rows
// r, r + 4, ... in increasing ell), then the pair sums from LDS.  A draw of a
// ragged batch without TATT (d.tatt null) gets B = D = 0, which leaves the
// NLA sums.
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(CLH_BLOCK) void k_limber_sx_tatt(const DrawDesc* __restrict__ D,
                                                              const int4* __restrict__ work) {
  constexpr int np = CLH_NP_SS, sp = CLH_NP_SS + 1, NK = 9;
  __shared__ double K[4 * NK * sp];        // [row][function][node]
  __shared__ double V[3 * CLH_MAXB * np];  // A, B, D per source bin and node
  __shared__ double G[2 * np];             // gs: W_gal, W_mag
  __shared__ double L10[CLH_TILE_ROWS], LN[CLH_TILE_ROWS];
  __shared__ double s_lk[CLH_LDS_LK];
  const int4 w = work[clh_xcd_item(blockIdx.x, gridDim.x)];
  const DrawDesc& d = D[w.x];
  const int kind = w.y & 3, zl = w.y >> 2, row0 = w.z, nr = w.w;
  const bool ss = kind == CLH_W_SS;
  const int tid = threadIdx.x, ns = d.ns;
  const double* o = clh_sx_nodes(d, kind, zl);
  const double* lut = o + clh_sx_lut_col(d, kind) * np;
  const bool tatt = d.tatt != nullptr;
  const double* tv = tatt ? d.tatt + (size_t)(ss ? 0 : 1 + zl) * (2 * ns + 1) * np : nullptr;
  const double* ells = ss ? d.ell_ss : d.ell_gx;
  const int woff = ss ? 3 : 5;
  for (int i = tid; i < ns * np; i += CLH_BLOCK) {
    V[i] = o[woff * np + i];
    V[CLH_MAXB * np + i] = tatt ? tv[i] : 0.0;
    V[2 * CLH_MAXB * np + i] = tatt ? tv[ns * np + i] : 0.0;
  }
  if (!ss)
    for (int i = tid; i < 2 * np; i += CLH_BLOCK) G[i] = o[3 * np + i];
  for (int t = tid; t < nr; t += CLH_BLOCK) {
    const double ell = ells[row0 + t] + 0.5;
    L10[t] = log10(ell);
    LN[t] = log(ell);
  }
  const double* lk = clh_stage(d.lk, d.nk, s_lk, CLH_LDS_LK);
  __syncthreads();
  const int p = tid & 63, wv = tid >> 6;
  const double l10fk = lut[2 * np + p], fac = o[2 * np + p], lnfk = log(o[np + p]);
  const double g4f = tatt ? tv[2 * ns * np + p] * fac : 0.0;
  ClhPkNode pkn;
  pkn.init(d, (int)lut[p], lut[np + p]);
  constexpr double c3 = 1.0 / (CLH_COVERH0 * CLH_COVERH0 * CLH_COVERH0);
  // FAST-PT rows of E, M, T, X, TT, T', X', TT' (E and M are sums of two rows)
  const double b1 = ss ? 0.0 : d.par->b1[zl], bmag = ss ? 0.0 : d.par->bmag[zl];
  const double* Av = V;
  const double* Bv = V + CLH_MAXB * np;
  const double* Dv = V + 2 * CLH_MAXB * np;
  for (int base = 0; base < nr; base += 4) {
    const int rr = base + wv;
    double* k = K + wv * NK * sp;
    if (rr < nr) {
      double dx;
      k[p] = exp(pkn.lnp(lk, L10[rr] - l10fk, dx)) * c3 * fac;
      const double lnk = LN[rr] - lnfk;
      if (!tatt) {
        for (int f = 1; f < NK; f++) k[f * sp + p] = 0.0;
      } else {
        k[sp + p] = g4f * (clh_fpt(d, 2, lnk) + clh_fpt(d, 3, lnk));
        k[2 * sp + p] = g4f * (clh_fpt(d, 6, lnk) + clh_fpt(d, 7, lnk));
      }
      if (ss && tatt) {
        k[3 * sp + p] = g4f * clh_fpt(d, 4, lnk);
        k[4 * sp + p] = g4f * clh_fpt(d, 8, lnk);
        k[5 * sp + p] = g4f * clh_fpt(d, 0, lnk);
        k[6 * sp + p] = g4f * clh_fpt(d, 5, lnk);
        k[7 * sp + p] = g4f * clh_fpt(d, 9, lnk);
        k[8 * sp + p] = g4f * clh_fpt(d, 1, lnk);
      }
    }
    __syncthreads();
    if (ss) {
      const int nout = (d.cl_bb ? 2 : 1) * d.n_ss * 4;
      for (int q = tid; q < nout; q += CLH_BLOCK) {
        const int bb_mode = q >= d.n_ss * 4;
        const int qq = bb_mode ? q - d.n_ss * 4 : q;
        const int pair = qq >> 2, r = qq & 3, row = base + r;
        if (row >= nr) continue;
        int bi, bj;
        clh_ss_pair(ns, pair, bi, bj);
        const double* kr = K + r * NK * sp;
        const double *Ai = Av + bi * np, *Aj = Av + bj * np, *Bi = Bv + bi * np, *Bj = Bv + bj * np;
        const double *Di = Dv + bi * np, *Dj = Dv + bj * np;
        double acc = 0.0;
        if (!bb_mode) {
          for (int n = 0; n < np; n++) {
            const double ai = Ai[n], aj = Aj[n], bi_ = Bi[n], bj_ = Bj[n], di = Di[n], dj = Dj[n];
            acc += ai * aj * kr[n] - (bi_ * aj + ai * bj_) * kr[sp + n] + 5.0 * (di * aj + ai * dj) * kr[2 * sp + n] +
                   bi_ * bj_ * kr[3 * sp + n] - 5.0 * (bi_ * dj + di * bj_) * kr[4 * sp + n] +
                   25.0 * di * dj * kr[5 * sp + n];
          }
        } else {
          for (int n = 0; n < np; n++) {
            const double bi_ = Bi[n], bj_ = Bj[n], di = Di[n], dj = Dj[n];
            acc += bi_ * bj_ * kr[6 * sp + n] - 5.0 * (bi_ * dj + di * bj_) * kr[7 * sp + n] +
                   25.0 * di * dj * kr[8 * sp + n];
          }
        }
        const double l = ells[row0 + row], ell = l + 0.5;
        const double pref = l * (l - 1.) * (l + 1.) * (l + 2.) / (ell * ell * ell * ell);
        double* out = bb_mode ? d.cl_bb : d.cl;
        out[(size_t)pair * d.nrow + row0 + row] = acc * pref;
      }
    } else {
      const int nout = ns * 4;
      for (int q = tid; q < nout; q += CLH_BLOCK) {
        const int zs = q >> 2, r = q & 3, row = base + r;
        if (row >= nr) continue;
        const double l = ells[row0 + row], ell = l + 0.5;
        const double pref1 = l * (l + 1.) / (ell * ell);
        const double tmp = (l - 1.) * l * (l + 1.) * (l + 2.);
        const double pref2 = (tmp > 0) ? sqrt(tmp) / (ell * ell) : 0.0;
        const double* kr = K + r * NK * sp;
        const double *As = Av + zs * np, *Bs = Bv + zs * np, *Ds = Dv + zs * np;
        double acc = 0.0;
        for (int n = 0; n < np; n++) {
          const double gal = G[n] * b1 + G[np + n] * pref1 * bmag;
          acc += gal * (As[n] * kr[n] - Bs[n] * kr[sp + n] + 5.0 * Ds[n] * kr[2 * sp + n]);
        }
        d.cl[(size_t)(d.n_ss + zl * ns + zs) * d.nrow + row0 + row] = acc * pref2;
      }
    }
    __syncthreads();
  }
}

// ----------------------------------------------------------------------------
// k_limber_gg: gg tiles.  Work item (draw, zl, row0, nrows); 192 threads = two
// groups of 96 (one thread per GL node, ClhGgNode started from the k_nodes
// search results; chi(z), G(z) and the lens n(z) spline staged in LDS when
// stage != 0).  Sub-tile s covers rows 16 s .. 16 s + 15 of the tile:
// group h takes rows 16 s + 8 h .. + 7, so each thread still walks its rows in
// increasing ell.  Row sums over the 96 nodes: one wave per row.
// F32: the integrand in FP32 arithmetic (walks, interpolations, W_RSD, the
// P(k,z) bilinear value and its exp; tables stay FP64 in memory); the t^2 P
// fac of each node goes into the FP64 row sums.
// KEEP: also store the bracket t, P fac and dW_RSD / d dz of every (row, node)
// in the draw's checkpoint (d.ggk), which the VJP then reads instead of
// evaluating the integrand again.
// ----------------------------------------------------------------------------
template <bool F32, bool KEEP, bool OL = false>
__global__ __launch_bounds__(192) void k_limber_gg(const DrawDesc* __restrict__ D,
                                                   const int4* __restrict__ work, int stage) {
  __shared__ double V[CLH_SUB_ROWS * CLH_NP_GG];
  __shared__ ClhGgRow R[CLH_TILE_ROWS];
  __shared__ double s_lk[CLH_LDS_LK];
  extern __shared__ double dyn[];  // staged chi(z), G(z), lens n(z) tables
  const int4 w = work[clh_xcd_item(blockIdx.x, gridDim.x)];
  const DrawDesc& d = D[w.x];
  const int zl = w.y, row0 = w.z, nr = w.w;
  constexpr int np = CLH_NP_GG;
  for (int t = threadIdx.x; t < nr; t += 192) R[t] = clh_gg_row(d.ell_gx[row0 + t]);
  const ClhGgTabs T = clh_gg_tabs(d, zl, dyn, stage, s_lk);
  __syncthreads();
  const int p = threadIdx.x % np, half = threadIdx.x / np;
  typedef typename std::conditional<F32, float, double>::type real;
  ClhGgNode<KEEP, !KEEP && !OL, false, real> nd;  // the look-ahead only where it keeps 3 waves per SIMD
  nd.init(d, T, clh_gg_nodes(d, zl), p, zl);
  // OL: the one-loop bias term W_gal^2 (...) of draws with some b2 != 0
  double ol_wg2 = 0.0, ol_g4f = 0.0, ol_lnfk = 0.0, ol_fk = 1.0;
  if (OL && d.oneloop) {
    const double* o = clh_gg_nodes(d, zl);
    const double g = clh_growfac(d, o[p]);
    ol_wg2 = o[3 * np + p] * o[3 * np + p];
    ol_g4f = g * g * g * g * o[2 * np + p];
    ol_fk = o[np + p];
    ol_lnfk = log(ol_fk);
  }
  constexpr int per = CLH_SUB_ROWS / 2;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int sub = 0; sub * CLH_SUB_ROWS < nr; sub++) {
    const int base = sub * CLH_SUB_ROWS;
    for (int q = 0; q < per; q++) {
      const int rl = half * per + q;
      double v = 0.0;
      if (base + rl < nr) {
        real t = 0, pf = 0, dwr = 0, dx;
        if (nd.template eval<false>(d, T, R[base + rl], t, pf, dwr, dx)) {
          v = (double)t * t * pf;
          if (OL && d.oneloop) {
            const double ell = R[base + rl].ell;
            v += ol_wg2 * clh_oneloop_gg(d, zl, log(ell) - ol_lnfk, ell / ol_fk, ol_g4f, pf);
          }
        } else {
          t = pf = dwr = 0.0;
        }
        if (KEEP) {
          double* g = d.ggk + ((size_t)(zl * d.nrow + row0 + base + rl) * 3) * np + p;
          g[0] = t;
          g[np] = pf;
          g[2 * np] = dwr;
        }
      }
      V[rl * np + p] = v;
    }
    __syncthreads();
    for (int rl = wid; rl < CLH_SUB_ROWS; rl += 3) {
      double v = V[rl * np + lane] + ((lane + 64 < np) ? V[rl * np + lane + 64] : 0.0);
      v = clh_wave_sum(v);
      const int rr = base + rl;
      if (lane == 0 && rr < nr) d.cl[(size_t)(d.n_ss + d.n_gs + zl) * d.nrow + row0 + rr] = v;
    }
    __syncthreads();
  }
}

// ----------------------------------------------------------------------------
// Data-vector placement of one (kind, pair, theta) value, as
// compute_Mx2pt_N_masked / add_calib_and_set_mask_X_N do it
// (generic_interface.hpp:800-862, 1137-1159): gamma_t gets the point-mass
// term, then shear calibration (1+m_i)(1+m_j) or (1+m_zs); masked entries are
// set to 0.  Without apply_calib the raw value is stored (the reference's
// xi_pm_tomo / w_gammat_tomo / w_gg_tomo).
// ----------------------------------------------------------------------------
__device__ __forceinline__ void clh_place(const DrawDesc& d, int kind, int q, int th, double acc) {
  const int nt = d.ntheta;
  int idx, canon;
  double f = 1.0;
  if (kind <= 1) {
    int i, j;
    clh_ss_pair(d.ns, q, i, j);
    if (d.apply_calib) f = (1.0 + d.par->shear_m[i]) * (1.0 + d.par->shear_m[j]);
    const int o = nt * q + th + (kind == 1 ? nt * d.n_ss : 0);
    idx = d.dv_start[0] + o;
    canon = o;
  } else if (kind == 2) {
    const int zs = q % d.ns, r = d.gs_row ? d.gs_row[q] : q;
    if (r < 0) return;  // an excluded pair has no data-vector row
    if (d.apply_calib) {
      f = 1.0 + d.par->shear_m[zs];
      const double pmf = clh_pm_fac(d, q / d.ns, zs);
      if (pmf != 0.0) acc += pmf / (d.theta[th] * d.theta[th]);
    }
    idx = d.dv_start[1] + nt * r + th;
    canon = 2 * nt * d.n_ss + nt * r + th;
  } else {
    idx = d.dv_start[2] + nt * q + th;
    canon = nt * (2 * d.n_ss + d.n_gs_dv) + nt * q + th;
  }
  d.dv[idx] = clh_masked(d, canon) ? 0.0 : acc * f;
}

// ----------------------------------------------------------------------------
// k_legendre (FP64, folded kernels) on MFMA (v_mfma_f64_16x16x4f64): one block
// of 4 waves per work item (draw, kind, c0) = 16 C_ell columns of one probe
// kind (kind 0 xi+, 1 xi-, 2 gamma_t, 3 w) against every theta bin:
//   out[c][theta] = sum_r C[c][r] K_kind[r][theta]
// M = columns, N = theta (two 16-wide tiles when Ntheta > 16), K = rows in
// steps of 4; wave w takes the steps w, w + 4, ... and the four partial tiles
// are summed through LDS.  Operands come straight from the C_ell rows and the
// layout's kernel rows, eight steps loaded ahead of their MFMAs.  Lane layout:
// A[m][k] on lane m + 16 k, B[k][n] on lane n + 16 k, D[4 v + l / 16][l % 16]
// in value v of lane l.
// ----------------------------------------------------------------------------
#define CLH_LEG_UNROLL 8
__global__ __launch_bounds__(256) void k_legendre(const DrawDesc* __restrict__ D,
                                                  const int4* __restrict__ work) {
  __shared__ clh_v4d part[3][2][64];
  const int4 w = work[clh_xcd_item(blockIdx.x, gridDim.x)];
  const DrawDesc& d = D[w.x];
  const int kind = w.y, c0 = w.z;
  const int nt = d.ntheta, nrow = d.nrow;
  const int ncolk = kind <= 1 ? d.n_ss : (kind == 2 ? d.n_gs : d.nl);
  const int colbase = kind <= 1 ? 0 : (kind == 2 ? d.n_ss : d.n_ss + d.n_gs);
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, lm = l & 15, lk = l >> 4;
  const double* __restrict__ K = d.kfold + (size_t)kind * nrow * nt;
  const int q = c0 + lm;
  const double* __restrict__ C = d.cl + (size_t)(colbase + (q < ncolk ? q : 0)) * nrow;
  // xi+- read EE +- BB when the B mode is present (TATT), cosmo2D.c:517-533
  const double* __restrict__ CB = (kind <= 1 && d.cl_bb) ? d.cl_bb + (size_t)(q < ncolk ? q : 0) * nrow : nullptr;
  const double sb = kind == 0 ? 1.0 : -1.0;
  const bool qin = q < ncolk, t0in = lm < nt, t1in = 16 + lm < nt, two = nt > 16;
  clh_v4d acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
  const int nstep = (nrow + 3) >> 2;
  for (int s0 = wv; s0 < nstep; s0 += 4 * CLH_LEG_UNROLL) {
    double a[CLH_LEG_UNROLL], b0[CLH_LEG_UNROLL], b1[CLH_LEG_UNROLL];
#pragma unroll
    for (int u = 0; u < CLH_LEG_UNROLL; u++) {
      const int r = 4 * (s0 + 4 * u) + lk;
      const bool rin = r < nrow;
      a[u] = (rin && qin) ? (CB ? C[r] + sb * CB[r] : C[r]) : 0.0;
      b0[u] = (rin && t0in) ? K[(size_t)r * nt + lm] : 0.0;
      b1[u] = (rin && t1in) ? K[(size_t)r * nt + 16 + lm] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < CLH_LEG_UNROLL; u++) {
      acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b0[u], acc0, 0, 0, 0);
      if (two) acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b1[u], acc1, 0, 0, 0);
    }
  }
  if (wv > 0) {
    part[wv - 1][0][l] = acc0;
    part[wv - 1][1][l] = acc1;
  }
  __syncthreads();
  if (wv > 0) return;
  for (int k = 0; k < 3; k++) {
    acc0 += part[k][0][l];
    acc1 += part[k][1][l];
  }
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int qq = c0 + 4 * v + lk;
    if (qq >= ncolk) continue;
    if (lm < nt) clh_place(d, kind, qq, lm, acc0[v]);
    if (two && 16 + lm < nt) clh_place(d, kind, qq, 16 + lm, acc1[v]);
  }
}

#define CLH_LEG_TASKS 8  // k_legendre_full: tasks per block, 32 theta slots each

// ----------------------------------------------------------------------------
// k_legendre_full: the reference's interpolate-then-dot structure
// (cosmo2D.c:468-533): C_ell at every integer ell < LMAX is generated on the
// fly from the column's rows (ell < LMIN_tab direct, else the ln-ell fill) and
// contracted with the full-length kernel.  Work item (draw, first task);
// CLH_LEG_TASKS tasks per block, 32 theta slots per task.  Task t of a draw:
// [0, n_ss) xi+, [n_ss, 2 n_ss) xi-, then gamma_t pairs, then w bins.
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
      if (d.cl_bb && tt < 2 * d.n_ss)  // xi+- from EE +- BB (TATT)
        v += (tt < d.n_ss ? 1.0 : -1.0) * d.cl_bb[(size_t)col * nrow + (i % nrow)];
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
  clh_place(d, kind, q, th, acc);
}

// ----------------------------------------------------------------------------
// launch wrappers (called from the host runtime)
// ----------------------------------------------------------------------------
hipError_t clh_launch_gtab(const DrawDesc* D, const int2* work, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gtab<0>, dim3(n), dim3(CLH_BLOCK), 0, s, D, work);
  return hipGetLastError();
}
hipError_t clh_launch_nodes(const DrawDesc* D, const int2* work, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_nodes, dim3(n), dim3(128), 0, s, D, work);
  return hipGetLastError();
}
hipError_t clh_launch_legendre(const DrawDesc* D, const int4* work, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_legendre, dim3(n), dim3(256), 0, s, D, work);
  return hipGetLastError();
}
hipError_t clh_launch_legendre_full(const DrawDesc* D, const int2* work, int n, int max_nrow,
                                    hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const size_t lds = sizeof(double) * CLH_LEG_TASKS * (size_t)max_nrow;
  hipLaunchKernelGGL(k_legendre_full, dim3(n), dim3(CLH_BLOCK), lds, s, D, work);
  return hipGetLastError();
}
hipError_t clh_launch_gtab_tangent(const DrawDesc* D, const int2* work, int n, int tables, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (tables) hipLaunchKernelGGL(k_gtab<1>, dim3(n), dim3(CLH_BLOCK), 0, s, D, work);
  else hipLaunchKernelGGL(k_gtab<2>, dim3(n), dim3(CLH_BLOCK), 0, s, D, work);
  return hipGetLastError();
}
hipError_t clh_launch_limber_sx_tatt(const DrawDesc* D, const int4* work, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_limber_sx_tatt, dim3(n), dim3(CLH_BLOCK), 0, s, D, work);
  return hipGetLastError();
}
hipError_t clh_launch_limber_sx(const DrawDesc* D, const int4* work, int n, int f32, int keep, hipStream_t s,
                                int ol) {
  if (n <= 0) return hipSuccess;
  if (ol && f32) hipLaunchKernelGGL((k_limber_sx<true, false, true>), dim3(n), dim3(CLH_BLOCK), 0, s, D, work);
  else if (ol) hipLaunchKernelGGL((k_limber_sx<false, false, true>), dim3(n), dim3(CLH_BLOCK), 0, s, D, work);
  else if (f32 && keep) hipLaunchKernelGGL((k_limber_sx<true, true>), dim3(n), dim3(CLH_BLOCK), 0, s, D, work);
  else if (f32) hipLaunchKernelGGL((k_limber_sx<true, false>), dim3(n), dim3(CLH_BLOCK), 0, s, D, work);
  else if (keep) hipLaunchKernelGGL((k_limber_sx<false, true>), dim3(n), dim3(CLH_BLOCK), 0, s, D, work);
  else hipLaunchKernelGGL((k_limber_sx<false, false>), dim3(n), dim3(CLH_BLOCK), 0, s, D, work);
  return hipGetLastError();
}
hipError_t clh_launch_limber_gg(const DrawDesc* D, const int4* work, int n, int f32, int stage_doubles,
                                int keep, hipStream_t s, int ol) {
  if (n <= 0) return hipSuccess;
  const size_t lds = sizeof(double) * (size_t)stage_doubles;
  const int stage = stage_doubles > 0;
  if (ol && f32) hipLaunchKernelGGL((k_limber_gg<true, false, true>), dim3(n), dim3(192), lds, s, D, work, stage);
  else if (ol) hipLaunchKernelGGL((k_limber_gg<false, false, true>), dim3(n), dim3(192), lds, s, D, work, stage);
  else if (f32 && keep) hipLaunchKernelGGL((k_limber_gg<true, true>), dim3(n), dim3(192), lds, s, D, work, stage);
  else if (f32) hipLaunchKernelGGL((k_limber_gg<true, false>), dim3(n), dim3(192), lds, s, D, work, stage);
  else if (keep) hipLaunchKernelGGL((k_limber_gg<false, true>), dim3(n), dim3(192), lds, s, D, work, stage);
  else hipLaunchKernelGGL((k_limber_gg<false, false>), dim3(n), dim3(192), lds, s, D, work, stage);
  return hipGetLastError();
}
