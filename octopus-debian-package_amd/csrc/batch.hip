// batch_t pack/unpack on the device: the transposes between Octopus's unpacked storage
// (one column of np elements per state, batch_t%dff_linear / zff_linear) and the packed
// layout (one row of 2^ldf doubles per point, state index fastest), which
// batch_write_unpacked_to_device / batch_read_device_to_unpacked do with the dpack/zpack
// and dunpack/zunpack kernels (src/grid/batch.F90:625-760, share/opencl/pack.cl:23-67).
//
// One workgroup moves a tile of 64 points x 16 states through LDS: reads are coalesced
// along points on the unpacked side (64 consecutive elements per state), writes along
// states on the packed side (16 consecutive elements of one row), and the other way round
// for unpack.  Elements are doubles (real batches) or re/im pairs moved as one 16-byte
// value (complex batches).
#include "octhip_internal.h"

namespace octhip {
namespace {

constexpr int TP = 64;  // points per tile
constexpr int TS = 16;  // states per tile

template <typename E>
__global__ __launch_bounds__(BLOCK) void k_pack(const E* __restrict__ src, size_t ld, int nst, int np,
                                                E* __restrict__ dst, int row, int first) {
  __shared__ E tile[TS][TP + 1];
  const int p0 = blockIdx.x * TP, s0 = blockIdx.y * TS;
  const int t = threadIdx.x;
#pragma unroll
  for (int r = 0; r < TS * TP / BLOCK; ++r) {
    const int i = t % TP, k = t / TP + r * (BLOCK / TP);
    const int p = p0 + i, st = s0 + k;
    if (p < np && st < nst) tile[k][i] = src[(size_t)st * ld + p];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < TS * TP / BLOCK; ++r) {
    const int k = t % TS, i = t / TS + r * (BLOCK / TS);
    const int p = p0 + i, st = s0 + k;
    if (p < np && st < nst) dst[(size_t)p * row + first + st] = tile[k][i];
  }
}

template <typename E>
__global__ __launch_bounds__(BLOCK) void k_unpack(const E* __restrict__ src, int row, int first, int nst, int np,
                                                  E* __restrict__ dst, size_t ld) {
  __shared__ E tile[TS][TP + 1];
  const int p0 = blockIdx.x * TP, s0 = blockIdx.y * TS;
  const int t = threadIdx.x;
#pragma unroll
  for (int r = 0; r < TS * TP / BLOCK; ++r) {
    const int k = t % TS, i = t / TS + r * (BLOCK / TS);
    const int p = p0 + i, st = s0 + k;
    if (p < np && st < nst) tile[k][i] = src[(size_t)p * row + first + st];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < TS * TP / BLOCK; ++r) {
    const int i = t % TP, k = t / TP + r * (BLOCK / TP);
    const int p = p0 + i, st = s0 + k;
    if (p < np && st < nst) dst[(size_t)st * ld + p] = tile[k][i];
  }
}

int check_pack_args(const void* a, const void* b, size_t ld, int nst, int np, int is_complex, int ldf, int first) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE((nst == 0 || np == 0) || (a && b && a != b), "bad buffer pointers");
  OCTHIP_REQUIRE(ldf >= (is_complex ? 1 : 0) && ldf <= MAX_LDF, "ldf=%d outside the packed range", ldf);
  OCTHIP_REQUIRE(nst >= 0 && np >= 0 && ld >= (size_t)np, "bad sizes nst=%d np=%d ld=%zu", nst, np, ld);
  const int row = is_complex ? (1 << (ldf - 1)) : (1 << ldf);
  OCTHIP_REQUIRE(first >= 0 && first + nst <= row, "states %d..%d do not fit a row of %d", first, first + nst - 1,
                 row);
  return OCTHIP_SUCCESS;
}

}  // namespace
}  // namespace octhip

using namespace octhip;

extern "C" int octhip_batch_pack(const double* unpacked, size_t ld, int nst, int np, int is_complex, double* packed,
                                 int ldf, int first, void* stream) {
  OCTHIP_RANGE("octhip_batch_pack");
  int rc = check_pack_args(unpacked, packed, ld, nst, np, is_complex, ldf, first);
  if (rc || nst == 0 || np == 0) return rc;
  const dim3 grid((unsigned)((np + TP - 1) / TP), (unsigned)((nst + TS - 1) / TS));
  hipStream_t s = (hipStream_t)stream;
  ProfScope prof("k_pack", s);
  if (is_complex)
    hipLaunchKernelGGL(k_pack<double2>, grid, dim3(BLOCK), 0, s, reinterpret_cast<const double2*>(unpacked), ld,
                       nst, np, reinterpret_cast<double2*>(packed), 1 << (ldf - 1), first);
  else
    hipLaunchKernelGGL(k_pack<double>, grid, dim3(BLOCK), 0, s, unpacked, ld, nst, np, packed, 1 << ldf, first);
  OCTHIP_HIP(hipGetLastError());
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_batch_unpack(const double* packed, int ldf, int first, int nst, int np, int is_complex,
                                   double* unpacked, size_t ld, void* stream) {
  OCTHIP_RANGE("octhip_batch_unpack");
  int rc = check_pack_args(packed, unpacked, ld, nst, np, is_complex, ldf, first);
  if (rc || nst == 0 || np == 0) return rc;
  const dim3 grid((unsigned)((np + TP - 1) / TP), (unsigned)((nst + TS - 1) / TS));
  hipStream_t s = (hipStream_t)stream;
  ProfScope prof("k_unpack", s);
  if (is_complex)
    hipLaunchKernelGGL(k_unpack<double2>, grid, dim3(BLOCK), 0, s, reinterpret_cast<const double2*>(packed),
                       1 << (ldf - 1), first, nst, np, reinterpret_cast<double2*>(unpacked), ld);
  else
    hipLaunchKernelGGL(k_unpack<double>, grid, dim3(BLOCK), 0, s, packed, 1 << ldf, first, nst, np, unpacked, ld);
  OCTHIP_HIP(hipGetLastError());
  return OCTHIP_SUCCESS;
}
