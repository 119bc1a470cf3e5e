// Batched mesh reductions and per-state BLAS-1 on packed batches (mesh_batch_dotp_vector,
// mesh_batch_nrm2, mesh_batch_orthogonalization: src/grid/mesh_batch_inc.F90:379-549,
// 970-1274; batch_axpy / batch_scal with one factor per state: src/grid/batch_ops_inc.F90),
// and the two Krylov-free alternatives to Taylor built on them: the Lanczos and Chebyshev
// exponentials (src/electrons/exponential.F90:766-971).
//
// Dot products are deterministic two-stage sums: every workgroup sums a fixed set of rows
// per state column (rows p = b*G + g, b*G + g + nblk*G, ... for its G row groups, then the
// G groups in order through LDS) into one partial per (workgroup, column); a second kernel
// adds the partials in workgroup order and scales by the volume element.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <mutex>
#include <vector>

#include "kernels.h"

namespace octhip {

namespace {

constexpr int DOT_BLOCKS = 512;   // workgroups of the first stage (partials per column)

// columns of a row: complex states (lane pairs) or real lanes
struct Cols {
  int n;   // columns per row
  int cw;  // columns handled side by side by the threads of a workgroup (<= BLOCK)
  int g;   // row groups per workgroup = BLOCK / cw
  int per; // columns per thread = n / cw
};
Cols cols_of(int L, int is_complex) {
  Cols c;
  c.n = is_complex ? L / 2 : L;
  c.cw = std::min(c.n, BLOCK);
  c.g = BLOCK / c.cw;
  c.per = c.n / c.cw;
  return c;
}

// partial[b][c] = sum over the rows of workgroup b of conj(a) * b (complex: two doubles)
__global__ __launch_bounds__(BLOCK) void k_dotp_partial(const double* __restrict__ a, const double* __restrict__ bb,
                                                        int L, int np, int is_complex, int cw, int per,
                                                        double* __restrict__ partial) {
  __shared__ double sre[BLOCK * 4], sim[BLOCK * 4];
  const int G = BLOCK / cw;
  const int c0 = threadIdx.x % cw, g = threadIdx.x / cw;
  const int lsh = __builtin_ctz((unsigned)L);
  double re[4] = {0, 0, 0, 0}, im[4] = {0, 0, 0, 0};
  for (int p = blockIdx.x * G + g; p < np; p += gridDim.x * G) {
    const double* ra = a + ((size_t)p << lsh);
    const double* rb = bb + ((size_t)p << lsh);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (k >= per) break;
      const int c = c0 + k * cw;
      if (is_complex) {
        const double ar = ra[2 * c], ai = ra[2 * c + 1], br = rb[2 * c], bi = rb[2 * c + 1];
        re[k] = re[k] + (ar * br + ai * bi);
        im[k] = im[k] + (ar * bi - ai * br);
      } else {
        re[k] = re[k] + ra[c] * rb[c];
      }
    }
  }
  for (int k = 0; k < per && k < 4; ++k) {
    sre[k * BLOCK + threadIdx.x] = re[k];
    sim[k * BLOCK + threadIdx.x] = im[k];
  }
  __syncthreads();
  // column c = c0 + k*cw: sum its G row groups in order
  for (int t = threadIdx.x; t < cw * per; t += BLOCK) {
    const int k = t / cw, cc = t % cw;
    double r = 0.0, i = 0.0;
    for (int gg = 0; gg < G; ++gg) {
      r = r + sre[k * BLOCK + gg * cw + cc];
      i = i + sim[k * BLOCK + gg * cw + cc];
    }
    const int c = cc + k * cw;
    partial[((size_t)blockIdx.x * (cw * per) + c) * 2] = r;
    partial[((size_t)blockIdx.x * (cw * per) + c) * 2 + 1] = i;
  }
}

// out[c] = vol * sum_b partial[b][c], b in order (complex pairs)
__global__ __launch_bounds__(BLOCK) void k_dotp_final(const double* __restrict__ partial, int nblk, int ncol,
                                                      double vol, double* __restrict__ out) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncol; c += gridDim.x * blockDim.x) {
    double r = 0.0, i = 0.0;
    for (int b = 0; b < nblk; ++b) {
      r = r + partial[((size_t)b * ncol + c) * 2];
      i = i + partial[((size_t)b * ncol + c) * 2 + 1];
    }
    out[2 * c] = vol * r;
    out[2 * c + 1] = vol * i;
  }
}

// y(p, c) = y(p, c) + coef[c] * x(p, c) over the np rows (complex factor per state, or a
// real factor per lane: coef holds (re, im) pairs either way, im unused for real rows)
__global__ __launch_bounds__(BLOCK) void k_axpy_cols(double* __restrict__ y, const double* __restrict__ x,
                                                     size_t count, int L, int is_complex,
                                                     const double* __restrict__ coef) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
    const int l = (int)(i & (size_t)(L - 1));
    if (is_complex) {
      const int c = l >> 1;
      y[i] = y[i] + zmul_lane(coef[2 * c], coef[2 * c + 1], x[i], x[i ^ 1], (l & 1) != 0);
    } else {
      y[i] = y[i] + coef[2 * l] * x[i];
    }
  }
}

// x(p, c) = coef[c] * x(p, c)
__global__ __launch_bounds__(BLOCK) void k_scal_cols(double* __restrict__ x, size_t count, int L, int is_complex,
                                                     const double* __restrict__ coef) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
    const int l = (int)(i & (size_t)(L - 1));
    if (is_complex) {
      // both lanes of a pair read both values before either is written: pairs are handled
      // by one thread each (even lanes), the odd lane is skipped
      if (l & 1) continue;
      const int c = l >> 1;
      const double xr = x[i], xi = x[i + 1], ar = coef[2 * c], ai = coef[2 * c + 1];
      x[i] = ar * xr - ai * xi;
      x[i + 1] = ar * xi + ai * xr;
    } else {
      x[i] = coef[2 * l] * x[i];
    }
  }
}

// Chebyshev update in the reference's operation order (exponential.F90:940-947):
// t = Hpsi1 + (-mid) psi1; t = t * (2/span); t = t + zfact psib; t = t + (-1) psi2
__global__ __launch_bounds__(BLOCK) void k_cheby_update(double* __restrict__ out, const double* __restrict__ psi1,
                                                        const double* __restrict__ b, const double* __restrict__ psi2,
                                                        size_t count, double mmid, double two_span, double zr,
                                                        double zi) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
    double t = out[i] + mmid * psi1[i];
    t = t * two_span;
    t = t + zmul_lane(zr, zi, b[i], b[i ^ 1], (i & 1) != 0);
    out[i] = t + (-1.0) * psi2[i];
  }
}

// psib = (psi0 + (-1) psi2) * z  (exponential.F90:950-952)
__global__ __launch_bounds__(BLOCK) void k_cheby_final(double* __restrict__ psib, const double* __restrict__ psi0,
                                                       const double* __restrict__ psi2, size_t count, double zr,
                                                       double zi) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
    if (i & 1) continue;
    const double tr = psi0[i] + (-1.0) * psi2[i];
    const double ti = psi0[i + 1] + (-1.0) * psi2[i + 1];
    psib[i] = zr * tr - zi * ti;
    psib[i + 1] = zr * ti + zi * tr;
  }
}

unsigned grid_for(size_t count) { return (unsigned)std::min<size_t>((count + BLOCK - 1) / BLOCK, 8192); }

}  // namespace

// Scratch bytes of a dot product of rows of L doubles: DOT_BLOCKS partials + the result.
size_t dotp_scratch_bytes(int L, int is_complex) {
  const Cols c = cols_of(L, is_complex);
  return sizeof(double) * 2 * ((size_t)DOT_BLOCKS * c.n + c.n);
}

// dot[c] = vol * sum_p conj(a(p, c)) b(p, c): partials in scratch, the ncol complex results
// in dest (device) or, when dest is null, at the end of scratch; *result = their address.
int launch_dotp(const double* a, const double* b, int L, int np, int is_complex, double vol, double* scratch,
                hipStream_t s, double** result, double* dest) {
  const Cols c = cols_of(L, is_complex);
  if (c.per > 4) return fail(OCTHIP_ERR_UNSUPPORTED, "dot product: more than %d columns", 4 * BLOCK);
  double* partial = scratch;
  double* out = dest ? dest : scratch + (size_t)2 * DOT_BLOCKS * c.n;
  const int nblk = std::max(1, std::min(DOT_BLOCKS, (np + c.g - 1) / c.g));
  ProfScope prof("k_dotp", s);
  hipLaunchKernelGGL(k_dotp_partial, dim3(nblk), dim3(BLOCK), 0, s, a, b, L, np, is_complex, c.cw, c.per, partial);
  OCTHIP_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_dotp_final, dim3((c.n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, partial, nblk, c.n, vol, out);
  OCTHIP_HIP(hipGetLastError());
  *result = out;
  return OCTHIP_SUCCESS;
}

int launch_axpy_cols(double* y, const double* x, int L, int np, int is_complex, const double* coef, hipStream_t s) {
  const size_t count = (size_t)np * L;
  if (count == 0) return OCTHIP_SUCCESS;
  hipLaunchKernelGGL(k_axpy_cols, dim3(grid_for(count)), dim3(BLOCK), 0, s, y, x, count, L, is_complex, coef);
  OCTHIP_HIP(hipGetLastError());
  return OCTHIP_SUCCESS;
}

int launch_scal_cols(double* x, int L, int np, int is_complex, const double* coef, hipStream_t s) {
  const size_t count = (size_t)np * L;
  if (count == 0) return OCTHIP_SUCCESS;
  hipLaunchKernelGGL(k_scal_cols, dim3(grid_for(count)), dim3(BLOCK), 0, s, x, count, L, is_complex, coef);
  OCTHIP_HIP(hipGetLastError());
  return OCTHIP_SUCCESS;
}

int launch_cheby_update(double* out, const double* psi1, const double* b, const double* psi2, size_t count,
                        double mid, double span, std::complex<double> z, hipStream_t s) {
  hipLaunchKernelGGL(k_cheby_update, dim3(grid_for(count)), dim3(BLOCK), 0, s, out, psi1, b, psi2, count, -mid,
                     2.0 / span, z.real(), z.imag());
  OCTHIP_HIP(hipGetLastError());
  return OCTHIP_SUCCESS;
}

int launch_cheby_final(double* psib, const double* psi0, const double* psi2, size_t count, std::complex<double> z,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_cheby_final, dim3(grid_for(count)), dim3(BLOCK), 0, s, psib, psi0, psi2, count, z.real(),
                     z.imag());
  OCTHIP_HIP(hipGetLastError());
  return OCTHIP_SUCCESS;
}

}  // namespace octhip

using namespace octhip;

extern "C" int octhip_mesh_batch_dotp_vector(const double* a, const double* b, int ldf, int np, int is_complex,
                                             double vol, double* dot, void* stream) {
  OCTHIP_RANGE("octhip_mesh_batch_dotp_vector");
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(a && b && dot, "null argument");
  OCTHIP_REQUIRE(ldf >= (is_complex ? 1 : 0) && ldf <= MAX_LDF, "ldf=%d outside the packed range", ldf);
  OCTHIP_REQUIRE(np >= 0, "np < 0");
  const int L = 1 << ldf;
  const size_t bytes = dotp_scratch_bytes(L, is_complex);
  double* scratch = nullptr;
  int rc = dalloc((void**)&scratch, bytes);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  double* res = nullptr;
  rc = launch_dotp(a, b, L, np, is_complex, vol, scratch, s, &res, nullptr);
  const int n = is_complex ? L / 2 : L;
  std::vector<double> host(2 * (size_t)n);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(host.data(), res, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = fail(OCTHIP_ERR_HIP, "mesh_batch_dotp_vector: %s", hipGetErrorString(e));
  }
  dfree(scratch);
  if (rc) return rc;
  // complex results as (re, im) pairs, real ones as one double per lane
  for (int c = 0; c < n; ++c) {
    if (is_complex) {
      dot[2 * c] = host[2 * c];
      dot[2 * c + 1] = host[2 * c + 1];
    } else {
      dot[c] = host[2 * c];
    }
  }
  return OCTHIP_SUCCESS;
}


// ------------------------------------------------------------ no allocation --
// octhip_mesh_batch_dotp_vector without any allocation: the partials and the results go
// to a fixed device array of this library instead of a per-call buffer.  Its capacity is
// the largest row the library handles (2^MAX_LDF real lanes), 8 MiB + 16 KiB.
namespace {
constexpr size_t NOALLOC_DOUBLES = 2 * ((size_t)DOT_BLOCKS * (1 << MAX_LDF) + (1 << MAX_LDF));
__device__ __align__(256) double scratch[NOALLOC_DOUBLES];

// device address of `scratch` on each device, looked up once with that device current
// (in host code the array's name is not a device address)
constexpr int MAX_DEVICES = 64;
double* scratch_ptr[MAX_DEVICES] = {};
std::mutex scratch_mu;

int scratch_of_current_device(double** out) {
  int dev = 0;
  OCTHIP_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return fail(OCTHIP_ERR_UNSUPPORTED, "device %d beyond %d", dev, MAX_DEVICES);
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return fail(OCTHIP_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = scratch_ptr[dev];
  return OCTHIP_SUCCESS;
}
}  // namespace

extern "C" int octhip_mesh_batch_dotp_vector_noalloc(const double* a, const double* b, int ldf, int np, int is_complex,
                                                     double vol, double* dot, void* stream) {
  OCTHIP_RANGE("octhip_mesh_batch_dotp_vector_noalloc");
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(a && b && dot, "null argument");
  OCTHIP_REQUIRE(ldf >= (is_complex ? 1 : 0) && ldf <= MAX_LDF, "ldf=%d outside the packed range", ldf);
  OCTHIP_REQUIRE(np >= 0, "np < 0");
  const int L = 1 << ldf;
  OCTHIP_REQUIRE(dotp_scratch_bytes(L, is_complex) <= sizeof(double) * NOALLOC_DOUBLES,
                 "rows of %d doubles need %zu bytes of scratch, more than the %zu available", L,
                 dotp_scratch_bytes(L, is_complex), sizeof(double) * NOALLOC_DOUBLES);
  double* sp = nullptr;
  int rc = scratch_of_current_device(&sp);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  // results as (re, im) pairs straight into the caller's device array (im = 0 for real rows)
  double* res = nullptr;
  return launch_dotp(a, b, L, np, is_complex, vol, sp, s, &res, dot);
}
