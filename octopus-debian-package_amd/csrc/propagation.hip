// Streaming kernels of the propagators: complex and real axpy over packed rows
// (batch_axpy, src/grid/batch_ops_inc.F90:23-106), the density of a packed batch
// (density_calc_accumulate, src/states/density.F90:375-417) and the fixed-order sum of
// the density partials the extended Taylor epilogue writes.
#include <algorithm>

#include "kernels.h"

namespace octhip {

namespace {

// y[i] += a * x[i] (the real-constant axpy Octopus uses for real(zfact))
__global__ __launch_bounds__(BLOCK) void k_daxpy(double* __restrict__ y, const double* __restrict__ x,
                                                 size_t count, double a) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    y[i] = y[i] + a * x[i];
}

// One group of tpp consecutive threads per point: each thread sums w |psi|^2 over the
// states it owns (lane pairs l, l + 2 tpp, ... for complex rows), the group adds its
// partial sums with shuffles, the first thread adds the total to density[p].
__global__ __launch_bounds__(BLOCK) void k_density(const double* __restrict__ psi, int L, int tshift, int np,
                                                   int is_complex, const double* __restrict__ w,
                                                   double* __restrict__ density) {
  const int tpp = 1 << tshift;
  const int slot = threadIdx.x >> tshift, lane = threadIdx.x & (tpp - 1);
  const int slots = BLOCK >> tshift;
  const int step = is_complex ? 2 : 1;
  for (int p = blockIdx.x * slots + slot; p < np; p += gridDim.x * slots) {
    const double* row = psi + ((size_t)p << __builtin_ctz((unsigned)L));
    double t = 0.0;
    for (int l = lane * step; l < L; l += tpp * step) {
      const double re = row[l];
      const double a2 = is_complex ? re * re + row[l + 1] * row[l + 1] : re * re;
      t = t + w[l / step] * a2;
    }
    for (int off = tpp >> 1; off > 0; off >>= 1) t += __shfl_xor(t, off);
    if (lane == 0) density[p] = density[p] + t;
  }
}

__global__ __launch_bounds__(BLOCK) void k_density_reduce(double* __restrict__ density, const double* __restrict__ dpart,
                                                          int nchunk, int np) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
    double t = 0.0;
    for (int c = 0; c < nchunk; ++c) t = t + dpart[(size_t)c * np + p];
    density[p] = density[p] + t;
  }
}

unsigned grid_for(size_t count) { return (unsigned)std::min<size_t>((count + BLOCK - 1) / BLOCK, 8192); }

}  // namespace

int launch_zaxpy(double* y, const double* x, size_t count, double zr, double zi, hipStream_t s) {
  if (count == 0) return OCTHIP_SUCCESS;
  ProfScope prof("k_zaxpy", s);
  hipLaunchKernelGGL(k_zaxpy, dim3(grid_for(count)), dim3(BLOCK), 0, s, y, x, count, zr, zi);
  OCTHIP_HIP(hipGetLastError());
  return OCTHIP_SUCCESS;
}

int launch_daxpy(double* y, const double* x, size_t count, double a, hipStream_t s) {
  if (count == 0) return OCTHIP_SUCCESS;
  ProfScope prof("k_daxpy", s);
  hipLaunchKernelGGL(k_daxpy, dim3(grid_for(count)), dim3(BLOCK), 0, s, y, x, count, a);
  OCTHIP_HIP(hipGetLastError());
  return OCTHIP_SUCCESS;
}

int launch_density(const double* psi, int L, int np, int is_complex, const double* w, double* density,
                   hipStream_t s) {
  if (np <= 0) return OCTHIP_SUCCESS;
  const int per_thread = is_complex ? 2 : 1;
  int tpp = std::min(std::max(L / per_thread, 1), 64);
  int tshift = 0;
  while ((1 << tshift) < tpp) ++tshift;
  const int slots = BLOCK >> tshift;
  const unsigned grid = (unsigned)std::min((np + slots - 1) / slots, 16384);
  ProfScope prof("k_density", s);
  hipLaunchKernelGGL(k_density, dim3(grid), dim3(BLOCK), 0, s, psi, L, tshift, np, is_complex, w, density);
  OCTHIP_HIP(hipGetLastError());
  return OCTHIP_SUCCESS;
}

int launch_density_reduce(double* density, const double* dpart, int nchunk, int np, hipStream_t s) {
  if (np <= 0) return OCTHIP_SUCCESS;
  ProfScope prof("k_density", s);
  hipLaunchKernelGGL(k_density_reduce, dim3(grid_for((size_t)np)), dim3(BLOCK), 0, s, density, dpart, nchunk, np);
  OCTHIP_HIP(hipGetLastError());
  return OCTHIP_SUCCESS;
}

}  // namespace octhip
