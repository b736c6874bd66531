// Mixed-precision Legendre contraction of the projection path (gfx950):
// the folded kernels against the FP64 C_ell rows on MFMA, in FP32 or as a
// 3-term split-BF16 product with FP32 accumulation.  (The FP32 Limber
// variants are the F32 instantiations of k_limber_sx / k_limber_gg.)
#include <hip/hip_runtime.h>

#include "clh_device.hpp"
#include "clh_internal.hpp"

#define CLH_BLOCK 256

// ----------------------------------------------------------------------------
// k_legendre_mx: folded Legendre contraction on MFMA.  One wave per work item
// (draw, kind, first column): a 32 (theta) x 32 (C_ell column) output tile
// X[theta][q] = sum_r K_kind[r][theta] C_q[r], reduced over the nrow rows in
// groups of 8.  Lane l = j + 32 h reads rows 8g + 4h .. 8g + 4h + 3 of column
// j, which is both the B-operand slice of v_mfma_f32_32x32x8_bf16 and, one
// row per instruction, of four v_mfma_f32_32x32x2_f32.
//   MODE 0: FP32 operands.
//   MODE 1: BF16 operands, each split into three parts a = a0 + a1 + a2,
//           b = b0 + b1 + b2 (each part the nearest-even BF16 of what the
//           previous parts left), and a*b ~ a2 b0 + a1 b1 + a0 b2 + a1 b0 +
//           a0 b1 + a0 b0 (six MFMAs, smallest first; the dropped products
//           are below FP32 rounding).
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
  // xi+- read EE +- BB when the B mode is present (TATT)
  const double* __restrict__ CB =
      (kind <= 1 && d.cl_bb) ? d.cl_bb + (size_t)(colok ? q : 0) * nrow + 4 * h : nullptr;
  const double sb = kind == 0 ? 1.0 : -1.0;
  clh_v16f acc;
#pragma unroll
  for (int v = 0; v < 16; v++) acc[v] = 0.0f;
  const int gfull = colok ? (nrow - 4 * h - 4) / 8 + 1 : 0;  // groups with all 4 rows in range
  for (int g = 0; g < ngrp; g++) {
    double c[4];
    if (g < gfull) {
#pragma unroll
      for (int t = 0; t < 4; t++) c[t] = C[8 * g + t];
      if (CB)
        for (int t = 0; t < 4; t++) c[t] += sb * CB[8 * g + t];
    } else {
#pragma unroll
      for (int t = 0; t < 4; t++) c[t] = (colok && 8 * g + 4 * h + t < nrow) ? C[8 * g + t] : 0.0;
      if (CB)
        for (int t = 0; t < 4; t++) c[t] += (colok && 8 * g + 4 * h + t < nrow) ? sb * CB[8 * g + t] : 0.0;
    }
    if (MODE == 0) {
      const float4 a = *reinterpret_cast<const float4*>(d.kmx32 + (((size_t)kind * ngrp + g) * 64 + lane) * 4);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, (float)c[0], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, (float)c[1], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.z, (float)c[2], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.w, (float)c[3], acc, 0, 0, 0);
    } else {
      const uint2* ap = reinterpret_cast<const uint2*>(d.kmxbf + (((size_t)kind * ngrp + g) * 64 + lane) * 12);
      const uint2 u0 = ap[0], u1 = ap[1], u2 = ap[2];
      const float sc = d.kmxsc[kind * ngrp + g];
      clh_v4s b0, b1, b2;
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const float f = (float)c[t] * sc;
        const unsigned h0 = clh_bf16_rne(f);
        const float r1 = f - __uint_as_float(h0 << 16);
        const unsigned h1 = clh_bf16_rne(r1);
        const float r2 = r1 - __uint_as_float(h1 << 16);
        b0[t] = (short)h0;
        b1[t] = (short)h1;
        b2[t] = (short)clh_bf16_rne(r2);
      }
      clh_v4s a0, a1, a2;
      a0[0] = (short)(u0.x & 0xffffu); a0[1] = (short)(u0.x >> 16);
      a0[2] = (short)(u0.y & 0xffffu); a0[3] = (short)(u0.y >> 16);
      a1[0] = (short)(u1.x & 0xffffu); a1[1] = (short)(u1.x >> 16);
      a1[2] = (short)(u1.y & 0xffffu); a1[3] = (short)(u1.y >> 16);
      a2[0] = (short)(u2.x & 0xffffu); a2[1] = (short)(u2.x >> 16);
      a2[2] = (short)(u2.y & 0xffffu); a2[3] = (short)(u2.y >> 16);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a2, b0, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a1, b1, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a0, b2, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a1, b0, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a0, b1, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a0, b0, acc, 0, 0, 0);
    }
  }
  if (!colok) return;
  int base, canon0;
  double f = 1.0, pmf = 0.0;
  if (kind <= 1) {
    int bi, bj;
    clh_ss_pair(d.ns, q, bi, bj);
    if (d.apply_calib) f = (1.0 + d.par->shear_m[bi]) * (1.0 + d.par->shear_m[bj]);
    canon0 = nt * q + (kind == 1 ? nt * d.n_ss : 0);
    base = d.dv_start[0] + canon0;
  } else if (kind == 2) {
    const int zs = q % d.ns, r = d.gs_row ? d.gs_row[q] : q;
    if (r < 0) return;  // an excluded pair has no data-vector row
    if (d.apply_calib) {
      f = 1.0 + d.par->shear_m[zs];
      pmf = clh_pm_fac(d, q / d.ns, zs);  // point mass (generic_interface.cpp:2626-2641)
    }
    base = d.dv_start[1] + nt * r;
    canon0 = 2 * nt * d.n_ss + nt * r;
  } else {
    base = d.dv_start[2] + nt * q;
    canon0 = nt * (2 * d.n_ss + d.n_gs_dv) + nt * q;
  }
#pragma unroll
  for (int v = 0; v < 16; v++) {
    const int th = 8 * (v >> 2) + 4 * h + (v & 3);
    if (th < nt) {
      double x = (double)acc[v];
      if (pmf != 0.0) x += pmf / (d.theta[th] * d.theta[th]);
      d.dv[base + th] = clh_masked(d, canon0 + th) ? 0.0 : x * f;
    }
  }
}

// ----------------------------------------------------------------------------
hipError_t clh_launch_legendre_mx(const DrawDesc* D, const int4* work, int n, int mode, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int blocks = (n + CLH_BLOCK / 64 - 1) / (CLH_BLOCK / 64);
  if (mode == CLH_PREC_BF16X3)
    hipLaunchKernelGGL(k_legendre_mx<1>, dim3(blocks), dim3(CLH_BLOCK), 0, s, D, work, n);
  else
    hipLaunchKernelGGL(k_legendre_mx<0>, dim3(blocks), dim3(CLH_BLOCK), 0, s, D, work, n);
  return hipGetLastError();
}
