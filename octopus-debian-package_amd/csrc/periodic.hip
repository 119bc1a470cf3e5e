// Periodic boundaries and Bloch phases of a k-point (complex batches):
//  * k_periodic_bc: boundary row np+b takes the value of its periodic image bnd_src[b],
//    times phase_corr[b] for a batch that carries the Bloch phase (the periodic branch of
//    boundaries_set with phase_correction, src/grid/boundaries_inc.F90:366-743);
//  * k_phase: dst(p) = phase(p) src(p), or conj(phase(p)) src(p), over rows p < n
//    (hamiltonian_elec_base_phase, src/hamiltonian/hamiltonian_elec_base.F90:979-1139).
// One thread per (re, im) pair: in place is safe.
#include <algorithm>

#include "kernels.h"

namespace octhip {
namespace {

// corr == nullptr: plain copy of every double of the row (real or complex rows);
// otherwise complex rows, one thread per (re, im) pair
__global__ __launch_bounds__(BLOCK) void k_periodic_bc(double* __restrict__ f, const int* __restrict__ src, int np,
                                                       int nb, int L, const double* __restrict__ corr) {
  const int w = corr ? L >> 1 : L;  // work units per row
  const size_t total = (size_t)nb * w;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const int b = (int)(t / w), c = (int)(t - (size_t)b * w);
    if (!corr) {
      f[(size_t)(np + b) * L + c] = f[(size_t)src[b] * L + c];
      continue;
    }
    const double* in = f + (size_t)src[b] * L + 2 * c;
    double* out = f + (size_t)(np + b) * L + 2 * c;
    const double re = in[0], im = in[1];
    const double cr = corr[2 * (size_t)b], ci = corr[2 * (size_t)b + 1];
    out[0] = cr * re - ci * im;
    out[1] = cr * im + ci * re;
  }
}

__global__ __launch_bounds__(BLOCK) void k_phase(double* __restrict__ dst, const double* __restrict__ src,
                                                 const double* __restrict__ phase, int n, int L, double sgn) {
  const int half = L >> 1;
  const size_t total = (size_t)n * half;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const int p = (int)(t / half), c = (int)(t - (size_t)p * half);
    const double pr = phase[2 * (size_t)p], pi = sgn * phase[2 * (size_t)p + 1];
    const double re = src[(size_t)p * L + 2 * c], im = src[(size_t)p * L + 2 * c + 1];
    dst[(size_t)p * L + 2 * c] = pr * re - pi * im;
    dst[(size_t)p * L + 2 * c + 1] = pr * im + pi * re;
  }
}

unsigned grid_for(size_t count) { return (unsigned)std::min<size_t>((count + BLOCK - 1) / BLOCK, 8192); }

}  // namespace

int launch_periodic_bc(double* f, const int* src, int np, int nb, int L, const double* corr, hipStream_t s) {
  if (nb <= 0) return OCTHIP_SUCCESS;
  ProfScope prof("k_periodic_bc", s);
  hipLaunchKernelGGL(k_periodic_bc, dim3(grid_for((size_t)nb * L)), dim3(BLOCK), 0, s, f, src, np, nb, L, corr);
  OCTHIP_HIP(hipGetLastError());
  return OCTHIP_SUCCESS;
}

int launch_phase(double* dst, const double* src, const double* phase, int n, int L, bool conjugate, hipStream_t s) {
  if (n <= 0) return OCTHIP_SUCCESS;
  ProfScope prof("k_phase", s);
  hipLaunchKernelGGL(k_phase, dim3(grid_for((size_t)n * (L / 2))), dim3(BLOCK), 0, s, dst, src, phase, n, L,
                     conjugate ? -1.0 : 1.0);
  OCTHIP_HIP(hipGetLastError());
  return OCTHIP_SUCCESS;
}

}  // namespace octhip
