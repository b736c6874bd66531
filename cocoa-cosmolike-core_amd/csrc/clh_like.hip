// Likelihood end of the projection path (gfx950): masked residuals of a
// batch of data vectors against the data, and chi^2 = r^T C^-1 r with its
// gradient 2 C^-1 r (IP::get_chi2, generic_interface.cpp:2288-2325).  The
// dense product with the inverse covariance is a plain rocBLAS dgemm over the
// whole batch; these kernels form the residuals and reduce the products.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "clh_internal.hpp"

// R[b][i] = mask[i] (dv[b][i] - data[i])
__global__ __launch_bounds__(256) void k_chi2_resid(int B, int n, const double* __restrict__ dv,
                                                    const double* __restrict__ data,
                                                    const double* __restrict__ mask, double* __restrict__ R) {
  const size_t total = (size_t)B * n;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < total; k += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(k % n);
    R[k] = mask[i] * (dv[k] - data[i]);
  }
}

// one block per draw: chi2[b] = sum_i R[b][i] Y[b][i]; grad[b][i] = 2 mask[i] Y[b][i]
__global__ __launch_bounds__(256) void k_chi2_reduce(int n, const double* __restrict__ R,
                                                     const double* __restrict__ Y,
                                                     const double* __restrict__ mask,
                                                     double* __restrict__ chi2, double* __restrict__ grad) {
  __shared__ double red[4];
  const size_t o = (size_t)blockIdx.x * n;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double y = Y[o + i];
    s += R[o + i] * y;
    if (grad) grad[o + i] = 2.0 * mask[i] * y;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) chi2[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

hipError_t clh_launch_chi2_resid(int B, int n, const double* dv, const double* data, const double* mask,
                                 double* R, hipStream_t s) {
  const size_t total = (size_t)B * n;
  const int blocks = (int)std::min<size_t>((total + 255) / 256, 8192);
  hipLaunchKernelGGL(k_chi2_resid, dim3(blocks), dim3(256), 0, s, B, n, dv, data, mask, R);
  return hipGetLastError();
}

hipError_t clh_launch_chi2_reduce(int B, int n, const double* R, const double* Y, const double* mask,
                                  double* chi2, double* grad, hipStream_t s) {
  hipLaunchKernelGGL(k_chi2_reduce, dim3(B), dim3(256), 0, s, n, R, Y, mask, chi2, grad);
  return hipGetLastError();
}
