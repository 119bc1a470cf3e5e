// Device kernels of the H-apply path (gfx950, wave64).
//
// Data layout (Octopus batch_t packed, src/grid/batch.F90:1017-1045): row p of a batch
// holds L = 2^ldf doubles (states fastest, complex as re/im pairs).  A "point" is one
// row.  Every kernel streams whole rows: a group of tpp = min(L, 64) consecutive
// threads owns one point and walks its lanes, so each load instruction of a wave
// covers 64 consecutive doubles of one or more rows (512 B when L >= 64).
#pragma once

#include "octhip_internal.h"

namespace octhip {

enum { IDX_LATTICE = 0, IDX_RI = 1 };
enum { EPI_STORE = 0, EPI_TAYLOR = 1 };

struct ApplyArgs {
  const double* __restrict__ fi;   // input batch (>= lim rows)
  double* __restrict__ fo;         // output batch (np rows), EPI_STORE
  const double* __restrict__ zrow; // zeros standing for boundary rows >= lim
  int np;                          // rows computed
  int lim;                         // neighbour rows >= lim read as zero
  int L;                           // doubles per row
  int tpp, tshift, ppb;            // threads per point, log2(tpp), points per block
  int chunk;                       // blockIdx -> tile remap granularity (XCD locality)
  int ntiles;
  int n;                           // stencil size (runtime-size path)
  const int* __restrict__ cidx;    // lattice: site of point p
  const int* __restrict__ cube;    // lattice: site -> point
  const int* __restrict__ ri;      // run-compressed offsets
  const int* __restrict__ rimap;   // run of point p
  int dc[MAX_STENCIL];             // lattice: site delta of stencil point j
  double w[MAX_STENCIL];           // weights already multiplied by the factor
  // local potential: v_eff = v + f . x
  const double* __restrict__ v;
  const double* __restrict__ xyz;
  double f0, f1, f2;
  int has_field;
  // nonlocal ket: per point, entries (global sphere point, matrix) in matrix order
  const int* __restrict__ pos;
  const int2* __restrict__ ent;
  const double* __restrict__ pt;   // transposed projector rows, stride pjs
  const int* __restrict__ mat_pbase;
  const int* __restrict__ mat_npj;
  int pjs;
  const double* __restrict__ proj; // projections (full_projection_size x L)
  // Taylor epilogue: yout = Hy ; acc += c1 * y + c2 * Hy (complex lanes)
  double* yout;
  double* acc;
  double c1r, c1i, c2r, c2i;
  int has_c1;
};

__device__ __forceinline__ int tile_of(int b, int chunk) {
  if (chunk <= 1) return b;
  const int g = b / (8 * chunk);
  const int r = b - g * 8 * chunk;
  return g * 8 * chunk + (r & 7) * chunk + (r >> 3);
}

// Lane-pair complex axpy term: for lane l of a (re, im) pair, the l-component of z*x,
// x = own lane value, xq = partner lane value.
__device__ __forceinline__ double zmul_lane(double zr, double zi, double x, double xq, bool odd) {
  return odd ? (zr * x + zi * xq) : (zr * x - zi * xq);
}

template <int N, int IDX, bool UNI, int EPI>
__global__ __launch_bounds__(BLOCK) void k_apply(const ApplyArgs a) {
  const int tile = tile_of(blockIdx.x, a.chunk);
  if (tile >= a.ntiles) return;
  int lp, lane0, tpp;
  if constexpr (UNI) {
    lp = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    lane0 = threadIdx.x & 63;
    tpp = 64;
  } else {
    lp = threadIdx.x >> a.tshift;
    lane0 = threadIdx.x & (a.tpp - 1);
    tpp = a.tpp;
  }
  const int p = tile * a.ppb + lp;
  if (p >= a.np) return;
  const size_t L = (size_t)a.L;
  constexpr int NQ = N > 0 ? N : 1;
  int q[NQ];
  if constexpr (N > 0) {
    if constexpr (IDX == IDX_LATTICE) {
      const int c = a.cidx[p];
#pragma unroll
      for (int j = 0; j < N; ++j) q[j] = a.cube[c + a.dc[j]];
    } else {
      const int* row = a.ri + (size_t)a.rimap[p] * N;
#pragma unroll
      for (int j = 0; j < N; ++j) q[j] = p + row[j];
    }
  }
  double vv = 0.0;
  if (a.v) {
    vv = a.v[p];
    if (a.has_field)
      vv = vv + (a.f0 * a.xyz[3 * (size_t)p] + a.f1 * a.xyz[3 * (size_t)p + 1] +
                 a.f2 * a.xyz[3 * (size_t)p + 2]);
  }
  int e0 = 0, e1 = 0;
  if (a.pos) {
    e0 = a.pos[p];
    e1 = a.pos[p + 1];
  }
  const double* own = a.fi + (size_t)p * L;
  for (int l = lane0; l < a.L; l += tpp) {
    double s = 0.0;
    if constexpr (N > 0) {
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const int qq = q[j];
        const double* src = ((unsigned)qq < (unsigned)a.lim) ? a.fi + (size_t)qq * L : a.zrow;
        s = fma(a.w[j], src[l], s);
      }
    } else {
      const int c = (IDX == IDX_LATTICE) ? a.cidx[p] : 0;
      const int* row = (IDX == IDX_RI) ? a.ri + (size_t)a.rimap[p] * a.n : nullptr;
      for (int j = 0; j < a.n; ++j) {
        const int qq = (IDX == IDX_LATTICE) ? a.cube[c + a.dc[j]] : p + row[j];
        const double* src = ((unsigned)qq < (unsigned)a.lim) ? a.fi + (size_t)qq * L : a.zrow;
        s = fma(a.w[j], src[l], s);
      }
    }
    const double x = own[l];
    if (a.v) s = fma(vv, x, s);
    for (int e = e0; e < e1; ++e) {
      const int2 en = a.ent[e];
      const int pb = a.mat_pbase[en.y];
      const int nj = a.mat_npj[en.y];
      const double* prow = a.pt + (size_t)en.x * a.pjs;
      double t = 0.0;
      for (int jj = 0; jj < nj; ++jj) t = fma(prow[jj], a.proj[(size_t)(pb + jj) * L + l], t);
      s = s + t;
    }
    if constexpr (EPI == EPI_STORE) {
      a.fo[(size_t)p * L + l] = s;
    } else {
      if (a.yout) a.yout[(size_t)p * L + l] = s;
      if (a.acc) {
        const bool odd = (l & 1) != 0;
        double av = a.acc[(size_t)p * L + l];
        if (a.has_c1) {
          const double xq = __shfl_xor(x, 1);
          av = av + zmul_lane(a.c1r, a.c1i, x, xq, odd);
        }
        const double sq = __shfl_xor(s, 1);
        av = av + zmul_lane(a.c2r, a.c2i, s, sq, odd);
        a.acc[(size_t)p * L + l] = av;
      }
    }
  }
}

// Zero rows [row0, row0+nrows) of a batch (boundaries_set for zero Dirichlet).
static __global__ void k_zero_rows(double* __restrict__ f, size_t first, size_t count) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
    f[first + i] = 0.0;
}

// y[i] += z * x[i] over complex lane pairs, i < count (count even).
static __global__ void k_zaxpy(double* __restrict__ y, const double* __restrict__ x, size_t count,
                        double zr, double zi) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x) {
    const double xv = x[i];
    const double xq = x[i ^ 1];
    y[i] = y[i] + zmul_lane(zr, zi, xv, xq, (i & 1) != 0);
  }
}

// launch helper: picks the template instance for (n, index kind, lanes, epilogue)
int launch_apply(ApplyArgs& a, bool lattice, int epi, hipStream_t s);

// ---------------------------------------------------------------- projectors --
struct BraArgs {
  const double* __restrict__ psi;
  int L;
  const int4* __restrict__ chunks;    // {matrix, q0, q1, chunk-in-matrix}
  const int4* __restrict__ mat_info;  // {npoints, nprojs, matrix offset, map offset}
  const double* __restrict__ mats;    // column-major blocks
  const int* __restrict__ maps;       // 1-based
  double* __restrict__ partial;       // (chunk, pjmax, L)
  int pjmax;
};

struct BraReduceArgs {
  const double* __restrict__ partial;
  const int* __restrict__ proj_mat;   // projection index -> matrix
  const int* __restrict__ mat_pbase;
  const int* __restrict__ mat_chunk0; // nmat+1
  const double* __restrict__ scals;
  double* __restrict__ proj;          // (full_projection_size, L)
  int L, pjmax, nproj;
};

int launch_bra(const BraArgs& b, int nchunk, bool mfma, hipStream_t s);
int launch_bra_reduce(const BraReduceArgs& r, hipStream_t s);

// sphere points per bra chunk
constexpr int BRA_CHUNK = 256;

}  // namespace octhip
