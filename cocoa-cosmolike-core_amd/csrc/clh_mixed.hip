// Mixed-precision kernels of the projection path (gfx950).
//
// Same work decomposition as the FP64 kernels (k_limber_sx, k_limber_gg,
// k_legendre) with the per-(ell, node) arithmetic in FP32: the P(k,z)
// bilinear lookup and exponential, the RSD weight, the node sums (Kahan
// compensated) and the Legendre contraction (FP32 kernels, FP32
// accumulation).  Table indices come from the same FP64 searches, and the
// lens-efficiency tables, chi(z), growth and node weights stay FP64 (computed
// by k_gtab / k_nodes).
#include <hip/hip_runtime.h>

#include "clh_device.hpp"
#include "clh_internal.hpp"

#define CLH_BLOCK 256

namespace {

struct Kahan {
  float s = 0.0f, c = 0.0f;
  __device__ __forceinline__ void add(float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

constexpr float kC3inv = (float)(1.0 / (CLH_COVERH0 * CLH_COVERH0 * CLH_COVERH0));

// bilinear ln P at (log10 k, node z index/weight) in FP32, exponentiated
__device__ __forceinline__ float pk32(const DrawDesc& d, float log10k, int& i, int j, float dy) {
  const double lq = (double)log10k;
  i = (i < 0) ? clh_bs(d.lk, d.nk, lq) : clh_walk(d.lk, d.nk, lq, i);
  const float l0 = (float)d.lk[i], l1 = (float)d.lk[i + 1];
  const float dx = (log10k - l0) / (l1 - l0);
  const double* __restrict__ P = d.lnP + (size_t)i * d.nzp + j;
  const float p00 = (float)P[0], p01 = (float)P[1], p10 = (float)P[d.nzp], p11 = (float)P[d.nzp + 1];
  const float out = (1.f - dx) * (1.f - dy) * p00 + (1.f - dx) * dy * p01 + dx * (1.f - dy) * p10 + dx * dy * p11;
  return expf(out) * kC3inv;
}

}  // namespace

// ----------------------------------------------------------------------------
__global__ __launch_bounds__(CLH_BLOCK) void k_limber_sx32(const DrawDesc* __restrict__ D,
                                                           const int4* __restrict__ work) {
  __shared__ float S[CLH_SX_TILE * CLH_NP_SS];
  __shared__ float W[(CLH_MAXB + 2) * CLH_NP_SS];
  __shared__ float L10[CLH_SX_TILE];
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
    L10[tid] = (r < d.nrow) ? log10f((float)(ells[r] + 0.5)) : 0.0f;
  }
  const int nw = (kind == CLH_W_SS) ? d.ns : d.ns + 2;
  for (int i = tid; i < nw * np; i += CLH_BLOCK) W[i] = (float)o[3 * np + i];
  __syncthreads();
  {
    const int p = tid & 63, wv = tid >> 6;
    const double a = o[p];
    const float l10fk = log10f((float)(o[np + p] * CLH_COVERH0));
    const float fac = (float)o[2 * np + p];
    const double z = 1.0 / a - 1.0;
    const int j = clh_bs(d.pz, d.nzp, z);
    const float dy = (float)((z - d.pz[j]) / (d.pz[j + 1] - d.pz[j]));
    int i = -1;
    constexpr int per = CLH_SX_TILE / 4;
    for (int q = 0; q < per; q++) {
      const int rr = wv * per + q;
      const int r = row0 + rr;
      float s = 0.0f;
      if (r < d.nrow) s = pk32(d, L10[rr] - l10fk, i, j, dy) * fac;
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
      const float* Ai = W + i * np;
      const float* Aj = W + j * np;
      const float* Sr = S + rr * np;
      Kahan acc;
#pragma unroll 8
      for (int p = 0; p < np; p++) acc.add(Ai[p] * Aj[p] * Sr[p]);
      const double l = ells[r], ell = l + 0.5, ell4 = ell * ell * ell * ell;
      const double pref = l * (l - 1.) * (l + 1.) * (l + 2.) / ell4;
      d.cl[(size_t)pair * d.nrow + r] = (double)acc.s * pref;
    }
  } else {
    const int nout = d.ns * CLH_SX_TILE;
    const float b1 = (float)d.par->b1[zl], bmag = (float)d.par->bmag[zl];
    for (int q = tid; q < nout; q += CLH_BLOCK) {
      const int zs = q / CLH_SX_TILE, rr = q % CLH_SX_TILE;
      const int r = row0 + rr;
      if (r >= d.nrow) continue;
      const double l = ells[r], ell = l + 0.5;
      const float pref1 = (float)(l * (l + 1.) / (ell * ell));
      const double tmp = (l - 1.) * l * (l + 1.) * (l + 2.);
      const double pref2 = (tmp > 0) ? sqrt(tmp) / (ell * ell) : 0.0;
      const float* Bz = W + (2 + zs) * np;
      const float* Sr = S + rr * np;
      Kahan acc;
#pragma unroll 8
      for (int p = 0; p < np; p++) acc.add(Bz[p] * (W[p] * b1 + W[np + p] * pref1 * bmag) * Sr[p]);
      d.cl[(size_t)(d.n_ss + zl * d.ns + zs) * d.nrow + r] = (double)acc.s * pref2;
    }
  }
}

// ----------------------------------------------------------------------------
__global__ __launch_bounds__(192) void k_limber_gg32(const DrawDesc* __restrict__ D,
                                                     const int4* __restrict__ work) {
  __shared__ float V[CLH_GGT_TILE * CLH_NP_GG];
  __shared__ float R[CLH_GGT_TILE][6];  // per row: log10(ell), (ell+1)/ell, c0, c1, pref
  const int4 w = work[blockIdx.x];
  const DrawDesc& d = D[w.x];
  const int zl = w.y, row0 = w.z;
  const int np = CLH_NP_GG;
  if (threadIdx.x < CLH_GGT_TILE) {
    const int r = row0 + threadIdx.x;
    if (r < d.nrow) {
      const double l = d.ell_gx[r], ell = l + 0.5;
      R[threadIdx.x][0] = log10f((float)ell);
      R[threadIdx.x][1] = (float)((ell + 1.0) / ell);
      R[threadIdx.x][2] = (float)((1 + 8. * ell) / ((2. * ell + 1.) * (2. * ell + 1.)));
      R[threadIdx.x][3] = (float)(4. / (2 * ell + 3.) * sqrt((2 * ell + 1.) / (2 * ell + 3.)));
      R[threadIdx.x][4] = (float)(l * (l + 1.) / (ell * ell));
    }
  }
  __syncthreads();
  const int p = threadIdx.x % np, half = threadIdx.x / np;
  const double* o = d.nodes + d.node_gg + (size_t)zl * d.node_gg_stride;
  const double a = o[p], fK = o[np + p];
  const float fac = (float)o[2 * np + p], WG = (float)o[3 * np + p], WM = (float)o[4 * np + p];
  const float b1 = (float)d.par->b1[zl], bmag = (float)d.par->bmag[zl];
  const float dzl = (float)d.par->lens_dz[zl], zm = (float)d.l_zmean[zl];
  const float inv_st = (float)(1.0 / d.par->lens_st[zl]);
  const double z = 1.0 / a - 1.0;
  const int j = clh_bs(d.pz, d.nzp, z);
  const float dy = (float)((z - d.pz[j]) / (d.pz[j + 1] - d.pz[j]));
  const float l10fk = log10f((float)(fK * CLH_COVERH0));
  // a_0 side of W_RSD once per node (FP64, as the tables)
  float t0;
  int jc, jz, jg, jn;
  {
    const double chi = fK * CLH_COVERH0;
    jc = clh_bs(d.chi_v, d.chi_n, chi);
    const double dyc = (chi - d.chi_v[jc]) / (d.chi_v[jc + 1] - d.chi_v[jc]);
    const double z0 = d.chi_z[jc] + dyc * (d.chi_z[jc + 1] - d.chi_z[jc]);
    const double a0 = 1.0 / (1.0 + z0);
    double c0, dc0;
    clh_chi_all(d, a0, c0, dc0);
    const double zz0 = 1.0 / a0 - 1.0;
    t0 = (float)(clh_nz_lens(d, zz0, zl) * (1.0 / ((a0 * a0) * dc0)) * clh_f_growth(d, zz0));
    jz = clh_bs(d.chi_z, d.chi_n, zz0);
    jg = clh_bs(d.G_z, d.G_n, zz0);
    jn = clh_bs(d.l_zc, d.l_rows, zz0);
  }
  const float chi_amin = (float)d.chi_amin;
  const float fKf = (float)fK;
  int i = -1;
  constexpr int per = CLH_GGT_TILE / 2;
  for (int q = 0; q < per; q++) {
    const int rr = half * per + q;
    const int r = row0 + rr;
    float v = 0.0f;
    if (r < d.nrow) {
      const float chi_1 = fKf * R[rr][1];
      if (!(chi_1 > chi_amin)) {
        const float c1v = chi_1 * (float)CLH_COVERH0;
        jc = clh_walk(d.chi_v, d.chi_n, (double)c1v, jc);
        const float cv0 = (float)d.chi_v[jc], cv1 = (float)d.chi_v[jc + 1];
        const float zv0 = (float)d.chi_z[jc], zv1 = (float)d.chi_z[jc + 1];
        const float z1 = zv0 + (c1v - cv0) / (cv1 - cv0) * (zv1 - zv0);
        // dchi/dz at z1 (3-point stencils of cosmo3D.c:64-81), hoverh0 = c/H0 / (dchi/dz)
        jz = clh_walk(d.chi_z, d.chi_n, (double)z1, jz);
        const int jj = jz > d.chi_n - 3 ? d.chi_n - 3 : jz;
        const float zz0 = (float)d.chi_z[jj], zz1 = (float)d.chi_z[jj + 1], zz2 = (float)d.chi_z[jj + 2];
        const float cc0 = (float)d.chi_v[jj], cc1 = (float)d.chi_v[jj + 1], cc2 = (float)d.chi_v[jj + 2];
        const float dyz = (z1 - zz0) / (zz1 - zz0);
        const float up = (cc2 - cc0) / (zz2 - zz0);
        const float down = (jj > 0) ? (cc1 - (float)d.chi_v[jj - 1]) / (zz1 - (float)d.chi_z[jj - 1])
                                    : (cc1 - cc0) / (zz1 - zz0);
        const float h1 = (float)CLH_COVERH0 / (down + dyz * (up - down));
        // f_growth(z1) = 1 - (dG/dz)(1+z)/G
        jg = clh_walk(d.G_z, d.G_n, (double)z1, jg);
        const float gz0 = (float)d.G_z[jg], gz1 = (float)d.G_z[jg + 1];
        const float g0 = (float)d.G_v[jg], g1 = (float)d.G_v[jg + 1];
        const float slope = (g1 - g0) / (gz1 - gz0);
        const float G = g0 + (z1 - gz0) * slope;
        const float f1 = 1.0f - slope * (1.0f + z1) / G;
        // nz_lens(z1)
        float n1 = 0.0f;
        const float zs = (z1 - dzl - zm) * inv_st + zm;
        const int nr = d.l_rows;
        if (!(zs <= (float)d.l_zc[0] || zs >= (float)d.l_zc[nr - 1])) {
          jn = clh_walk(d.l_zc, nr, (double)zs, jn);
          const int jn2 = jn > nr - 2 ? nr - 2 : jn;
          const double* cf = d.l_coef + (size_t)zl * 4 * (nr - 1) + 4 * jn2;
          const float dxs = zs - (float)d.l_zc[jn2];
          n1 = ((float)cf[0] + dxs * ((float)cf[1] + dxs * ((float)cf[2] + dxs * (float)cf[3]))) * inv_st;
        }
        const float wr = R[rr][2] * t0 - R[rr][3] * n1 * h1 * f1;
        const float t = WG * b1 + WM * R[rr][4] * bmag + wr;
        v = t * t * pk32(d, R[rr][0] - l10fk, i, j, dy) * fac;
      }
    }
    V[rr * np + p] = v;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int rr = wid; rr < CLH_GGT_TILE; rr += 3) {
    // compensated: each lane folds two nodes, then a tree sum across the wave
    float v = V[rr * np + lane];
    float c = 0.0f;
    if (lane + 64 < np) {
      const float x = V[rr * np + lane + 64];
      const float t = v + x;
      c = (fabsf(v) >= fabsf(x)) ? (v - t) + x : (x - t) + v;
      v = t;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const float x = __shfl_xor(v, off, 64);
      const float xc = __shfl_xor(c, off, 64);
      const float t = v + x;
      c += xc + ((fabsf(v) >= fabsf(x)) ? (v - t) + x : (x - t) + v);
      v = t;
    }
    const int r = row0 + rr;
    if (lane == 0 && r < d.nrow) d.cl[(size_t)(d.n_ss + d.n_gs + zl) * d.nrow + r] = (double)v + (double)c;
  }
}

// ----------------------------------------------------------------------------
// k_legendre_mx: folded Legendre contraction on MFMA.  One wave per work item
// (draw, kind, first column): a 32 (theta) x 32 (C_ell column) output tile
// X[theta][q] = sum_r K_kind[r][theta] C_q[r], reduced over the nrow rows in
// groups of 8.  Lane l = j + 32 h reads rows 8g + 4h .. 8g + 4h + 3 of column
// j, which is both the B-operand slice of v_mfma_f32_32x32x8_bf16 and, one
// row per instruction, of four v_mfma_f32_32x32x2_f32.
//   MODE 0: FP32 operands.
//   MODE 1: BF16 operands, 3-term split a*b ~ ah*bh + ah*bl + al*bh with
//           a = ah + al, b = bh + bl (each part rounded to nearest even).
// The accumulator tile of lane l, register v holds theta = 8(v/4) + 4h + v%4.
// ----------------------------------------------------------------------------
typedef float clh_v16f __attribute__((ext_vector_type(16)));
typedef short clh_v4s __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned clh_bf16_rne(float f) {
  const unsigned u = __float_as_uint(f);
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

template <int MODE>
__global__ __launch_bounds__(CLH_BLOCK) void k_legendre_mx(const DrawDesc* __restrict__ D,
                                                           const int4* __restrict__ work, int nwork) {
  const int item = blockIdx.x * (CLH_BLOCK / 64) + (threadIdx.x >> 6);
  if (item >= nwork) return;  // whole waves; the kernel has no block barrier
  const int lane = threadIdx.x & 63;
  const int4 w = work[item];
  const DrawDesc& d = D[w.x];
  const int kind = w.y, col0 = w.z;
  const int nrow = d.nrow, nt = d.ntheta, ngrp = d.kmx_ngrp;
  const int ncol = kind <= 1 ? d.n_ss : (kind == 2 ? d.n_gs : d.nl);
  const int cbase = kind <= 1 ? 0 : (kind == 2 ? d.n_ss : d.n_ss + d.n_gs);
  const int j = lane & 31, h = lane >> 5;
  const int q = col0 + j;
  const bool colok = q < ncol;
  const double* __restrict__ C = d.cl + (size_t)(cbase + (colok ? q : 0)) * nrow + 4 * h;
  clh_v16f acc;
#pragma unroll
  for (int v = 0; v < 16; v++) acc[v] = 0.0f;
  const int gfull = colok ? (nrow - 4 * h - 4) / 8 + 1 : 0;  // groups with all 4 rows in range
  for (int g = 0; g < ngrp; g++) {
    double c[4];
    if (g < gfull) {
#pragma unroll
      for (int t = 0; t < 4; t++) c[t] = C[8 * g + t];
    } else {
#pragma unroll
      for (int t = 0; t < 4; t++) c[t] = (colok && 8 * g + 4 * h + t < nrow) ? C[8 * g + t] : 0.0;
    }
    if (MODE == 0) {
      const float4 a = *reinterpret_cast<const float4*>(d.kmx32 + (((size_t)kind * ngrp + g) * 64 + lane) * 4);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, (float)c[0], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, (float)c[1], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.z, (float)c[2], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.w, (float)c[3], acc, 0, 0, 0);
    } else {
      const uint4 a = *reinterpret_cast<const uint4*>(d.kmxbf + (((size_t)kind * ngrp + g) * 64 + lane) * 8);
      const float sc = d.kmxsc[kind * ngrp + g];
      clh_v4s bh, bl;
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const float f = (float)c[t] * sc;
        const unsigned hi = clh_bf16_rne(f);
        const float rem = f - __uint_as_float(hi << 16);
        bh[t] = (short)hi;
        bl[t] = (short)clh_bf16_rne(rem);
      }
      clh_v4s ah, al;
      ah[0] = (short)(a.x & 0xffffu); ah[1] = (short)(a.x >> 16);
      ah[2] = (short)(a.y & 0xffffu); ah[3] = (short)(a.y >> 16);
      al[0] = (short)(a.z & 0xffffu); al[1] = (short)(a.z >> 16);
      al[2] = (short)(a.w & 0xffffu); al[3] = (short)(a.w >> 16);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(al, bh, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ah, bl, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ah, bh, acc, 0, 0, 0);
    }
  }
  if (!colok) return;
  int base;
  double f = 1.0;
  if (kind <= 1) {
    int bi, bj;
    clh_ss_pair(d.ns, q, bi, bj);
    if (d.apply_calib) f = (1.0 + d.par->shear_m[bi]) * (1.0 + d.par->shear_m[bj]);
    base = d.dv_start[0] + nt * q + (kind == 1 ? nt * d.n_ss : 0);
  } else if (kind == 2) {
    if (d.apply_calib) f = 1.0 + d.par->shear_m[q % d.ns];
    base = d.dv_start[1] + nt * q;
  } else {
    base = d.dv_start[2] + nt * q;
  }
#pragma unroll
  for (int v = 0; v < 16; v++) {
    const int th = 8 * (v >> 2) + 4 * h + (v & 3);
    if (th < nt) d.dv[base + th] = (double)acc[v] * f;
  }
}

// ----------------------------------------------------------------------------
hipError_t clh_launch_limber_sx32(const DrawDesc* D, const int4* work, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_limber_sx32, dim3(n), dim3(CLH_BLOCK), 0, s, D, work);
  return hipGetLastError();
}
hipError_t clh_launch_limber_gg32(const DrawDesc* D, const int4* work, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_limber_gg32, dim3(n), dim3(192), 0, s, D, work);
  return hipGetLastError();
}
hipError_t clh_launch_legendre_mx(const DrawDesc* D, const int4* work, int n, int mode, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int blocks = (n + CLH_BLOCK / 64 - 1) / (CLH_BLOCK / 64);
  if (mode == CLH_PREC_BF16X3)
    hipLaunchKernelGGL(k_legendre_mx<1>, dim3(blocks), dim3(CLH_BLOCK), 0, s, D, work, n);
  else
    hipLaunchKernelGGL(k_legendre_mx<0>, dim3(blocks), dim3(CLH_BLOCK), 0, s, D, work, n);
  return hipGetLastError();
}
