// Kernel instances and launchers: fused apply, projector bra (FP64 MFMA and plain).
#include <algorithm>

#include "kernels.h"

namespace octhip {

namespace {

template <int N, int IDX, int EPI, int LPT>
int launch_apply_n(ApplyArgs& a, hipStream_t s) {
  const int per = 8 * (a.chunk > 1 ? a.chunk : 1);
  const int grid = ((a.ntiles + per - 1) / per) * per;
  const size_t shmem = (size_t)a.tpts * a.sstride * sizeof(int);
  ProfScope prof("k_apply", s);
  hipLaunchKernelGGL((k_apply<N, IDX, EPI, LPT>), dim3(grid), dim3(BLOCK), shmem, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(OCTHIP_ERR_HIP, "k_apply launch: %s", hipGetErrorString(e));
  return OCTHIP_SUCCESS;
}

template <int IDX, int EPI, int LPT>
int launch_apply_lpt(ApplyArgs& a, hipStream_t s) {
  switch (a.n) {
    case 25: return launch_apply_n<25, IDX, EPI, LPT>(a, s);
    case 19: return launch_apply_n<19, IDX, EPI, LPT>(a, s);
    case 13: return launch_apply_n<13, IDX, EPI, LPT>(a, s);
    case 7: return launch_apply_n<7, IDX, EPI, LPT>(a, s);
    default: return launch_apply_n<0, IDX, EPI, LPT>(a, s);
  }
}

template <int IDX>
int launch_apply_idx(ApplyArgs& a, int epi, int lpt, hipStream_t s) {
  if (epi == EPI_TAYLOR) return launch_apply_lpt<IDX, EPI_TAYLOR, 2>(a, s);
  if (epi == EPI_TAYLOR_X) return launch_apply_lpt<IDX, EPI_TAYLOR_X, 2>(a, s);
  return lpt == 2 ? launch_apply_lpt<IDX, EPI_STORE, 2>(a, s) : launch_apply_lpt<IDX, EPI_STORE, 1>(a, s);
}

}  // namespace

int launch_apply(ApplyArgs& a, bool lattice, int epi, hipStream_t s) {
  if (a.np <= 0) return OCTHIP_SUCCESS;
  const int lpt = a.L >= 2 ? 2 : 1;
  if (epi != EPI_STORE && lpt != 2) return fail(OCTHIP_ERR_INVALID, "Taylor epilogue needs complex rows");
  int tpp = a.L / lpt;
  if (tpp > 64) tpp = 64;
  int tshift = 0;
  while ((1 << tshift) < tpp) ++tshift;
  a.tpp = tpp;
  a.tshift = tshift;
  const int slots = BLOCK / tpp;
  int iters = a.tile_iters > 0 ? a.tile_iters : 4;
  while (iters > 1 && slots * iters > 256) --iters;
  a.tpts = slots * iters;
  a.sstride = a.n > 0 ? ((a.n + 3) & ~3) : 1;
  a.ntiles = (a.np + a.tpts - 1) / a.tpts;
  if (a.chunk < 1) a.chunk = 1;
  return lattice ? launch_apply_idx<IDX_LATTICE>(a, epi, lpt, s) : launch_apply_idx<IDX_RI>(a, epi, lpt, s);
}

// -------------------------------------------------------------- bra (plain) --
// One workgroup per chunk of <= BRA_CHUNK sphere points of one matrix.  Threads are
// (slice, lane): each slice sums every nsl-th sphere point of the chunk for one lane,
// slices are then added in fixed order through LDS, so results are deterministic.
constexpr int BRA_PJB = 8;  // projectors per register block

__device__ __forceinline__ void bra_plain_body(const BraArgs& b, int c) {
  __shared__ double red[BLOCK * BRA_PJB];
  const int4 ch = b.chunks[c];
  const int4 mi = b.mat_info[ch.x];
  const int npt = mi.x, npj = mi.y, moff = mi.z, mapoff = mi.w;
  const int Lc = b.L < BLOCK ? b.L : BLOCK;
  const int nsl = BLOCK / Lc;
  const int sl = threadIdx.x / Lc;
  const int lc = threadIdx.x - sl * Lc;
  for (int lg = 0; lg < b.L; lg += Lc) {
    const int l = lg + lc;
    for (int jb = 0; jb < npj; jb += BRA_PJB) {
      double acc[BRA_PJB];
#pragma unroll
      for (int jj = 0; jj < BRA_PJB; ++jj) acc[jj] = 0.0;
      for (int q = ch.y + sl; q < ch.z; q += nsl) {
        const int row = b.maps[mapoff + q] - 1;
        double x = b.psi[(size_t)row * b.L + l];
        if (b.pphase) {
          const double xq = b.psi[(size_t)row * b.L + (l ^ 1)];
          x = zmul_lane(b.pphase[2 * (size_t)(mapoff + q)], b.pphase[2 * (size_t)(mapoff + q) + 1], x, xq,
                        (l & 1) != 0);
        }
#pragma unroll
        for (int jj = 0; jj < BRA_PJB; ++jj)
          if (jb + jj < npj) acc[jj] = fma(b.mats[moff + (size_t)(jb + jj) * npt + q], x, acc[jj]);
      }
#pragma unroll
      for (int jj = 0; jj < BRA_PJB; ++jj) red[(sl * BRA_PJB + jj) * Lc + lc] = acc[jj];
      __syncthreads();
      if (sl == 0) {
        for (int jj = 0; jj < BRA_PJB && jb + jj < npj; ++jj) {
          double t = 0.0;
          for (int s2 = 0; s2 < nsl; ++s2) t += red[(s2 * BRA_PJB + jj) * Lc + lc];
          b.partial[((size_t)c * b.pjmax + jb + jj) * b.L + l] = t;
        }
      }
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(BLOCK) void k_bra_plain(const BraArgs b) { bra_plain_body(b, blockIdx.x); }

// --------------------------------------------------------------- bra (MFMA) --
// D(16 projectors x 16 lanes) += A(16 x 4 sphere points) * B(4 x 16 lanes) with
// v_mfma_f64_16x16x4_f64.  Operand layout (wave64): lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; accumulator element r of lane l is D[4*r + l/16][l%16] (checked on the
// device with an in-bounds probe).
// A comes from the transposed projector table (16 lanes read consecutive doubles of one
// sphere point's row), B from the state rows (16 lanes read 128 contiguous bytes, 256
// with two columns per lane).
typedef double d4 __attribute__((ext_vector_type(4)));


// One workgroup per (chunk, lane tile of 16 TW lanes): grids of nchunk * L/(16 TW) blocks
// keep every CU busy; the 4 waves take interleaved groups of 4 sphere points and are
// summed through LDS in a fixed order.  Every load is issued unconditionally (clamped
// indices, zero operands past the chunk end) so the compiler counts them exactly.
// TW = 2: lane ii of a wave loads the 16 bytes of columns 2ii and 2ii+1 of its row (one
// complex value) and feeds two MFMAs, one per column, that share the projector operand:
// each column's sum is the TW = 1 sum, in the same order.
constexpr int BRA_RED = 4 * 16 * 16;  // LDS doubles per accumulator of the wave reduction
template <int TW>
__device__ __forceinline__ void bra_mfma_body(const BraMfmaArgs& m, int c, int lt, double (*red)[BRA_RED]) {
  constexpr int KB = 16;                    // MFMA steps (4 sphere points each) per load block
  const BraArgs& b = m.b;
  const int4 ch = b.chunks[c];
  const int4 mi = b.mat_info[ch.x];
  const int npj = mi.y, mapoff = mi.w;
  const int wave = threadIdx.x >> 6;
  const int ln = threadIdx.x & 63;
  const int kk = ln >> 4, ii = ln & 15;
  const int col = lt * 16 * TW + ii * TW;   // first column of this lane
  for (int jb = 0; jb < npj; jb += 16) {
    d4 acc[TW];
#pragma unroll
    for (int w = 0; w < TW; ++w) acc[w] = d4{0.0, 0.0, 0.0, 0.0};
    const int jc = min(jb + ii, npj - 1);
    const bool jok = jb + ii < npj;
    for (int q0 = ch.y + 4 * wave; q0 < ch.z; q0 += KB * 16) {
      int row[KB];
      double av[KB];
      Vec<TW> bv[KB];
#pragma unroll
      for (int k = 0; k < KB; ++k) {
        const int qk = q0 + 16 * k + kk;
        const int qs = qk < ch.z ? qk : ch.y;
        row[k] = b.maps[mapoff + qs] - 1;
        av[k] = m.pt[(size_t)(mapoff + qs) * m.pjs + jc];
      }
#pragma unroll
      for (int k = 0; k < KB; ++k) bv[k] = ldv<TW>(b.psi + (size_t)row[k] * b.L + col);
      if (b.pphase) {
        // Bloch phase of each sphere point on the (re, im) pair: with TW = 2 the lane
        // holds both, otherwise the other one is the neighbouring thread's (lanes ii and
        // ii ^ 1 hold the same point)
#pragma unroll
        for (int k = 0; k < KB; ++k) {
          const int qk = q0 + 16 * k + kk;
          const size_t pi2 = 2 * (size_t)(mapoff + (qk < ch.z ? qk : ch.y));
          if constexpr (TW == 2) {
            const double re = zmul_lane(b.pphase[pi2], b.pphase[pi2 + 1], bv[k].v[0], bv[k].v[1], false);
            const double im = zmul_lane(b.pphase[pi2], b.pphase[pi2 + 1], bv[k].v[1], bv[k].v[0], true);
            bv[k].v[0] = re;
            bv[k].v[1] = im;
          } else {
            const double xq = __shfl_xor(bv[k].v[0], 1);
            bv[k].v[0] = zmul_lane(b.pphase[pi2], b.pphase[pi2 + 1], bv[k].v[0], xq, (col & 1) != 0);
          }
        }
      }
#pragma unroll
      for (int k = 0; k < KB; ++k) {
        const bool ok = q0 + 16 * k + kk < ch.z;
#pragma unroll
        for (int w = 0; w < TW; ++w)
          acc[w] = __builtin_amdgcn_mfma_f64_16x16x4f64((ok && jok) ? av[k] : 0.0, ok ? bv[k].v[w] : 0.0, acc[w],
                                                        0, 0, 0);
      }
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < TW; ++w)
#pragma unroll
      for (int r = 0; r < 4; ++r) red[w][(wave * 16 + 4 * r + kk) * 16 + ii] = acc[w][r];
    __syncthreads();
    if (wave == 0) {
#pragma unroll
      for (int w = 0; w < TW; ++w)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int jl = 4 * r + kk;
          const double t = (red[w][(0 * 16 + jl) * 16 + ii] + red[w][(1 * 16 + jl) * 16 + ii]) +
                           (red[w][(2 * 16 + jl) * 16 + ii] + red[w][(3 * 16 + jl) * 16 + ii]);
          const int j = jb + jl;
          if (j < npj) b.partial[((size_t)c * b.pjmax + j) * b.L + col + w] = t;
        }
    }
  }
}

__global__ __launch_bounds__(BLOCK) void k_bra_mfma(const BraMfmaArgs m) {
  const int nz = (int)gridDim.x - m.nbra;
  if ((int)blockIdx.x < nz) {
    // the first blocks clear the boundary rows of the batch (boundaries_set), fused into
    // this launch (the bra reads inner rows only); dispatched first, the streaming stores
    // overlap the latency-bound bra blocks instead of trailing them (-4% for the launch)
    for (size_t t = blockIdx.x * (size_t)BLOCK + threadIdx.x; t < m.zcount; t += (size_t)nz * BLOCK)
      m.zf[m.zfirst + t] = 0.0;
    return;
  }
  __shared__ double red[2][BRA_RED];
  const int b = (int)blockIdx.x - nz;
  if ((m.b.L & 31) == 0) {
    const int nlt = m.b.L >> 5;
    bra_mfma_body<2>(m, b / nlt, b % nlt, red);
  } else {
    const int nlt = m.b.L >> 4;
    bra_mfma_body<1>(m, b / nlt, b % nlt, red);
  }
}

// proj[i, l] = scal[i] * sum_{chunks of matrix(i)} partial[chunk, i - pbase, l]
__device__ __forceinline__ void bra_reduce_body(const BraReduceArgs& r, int bid, int nblk) {
  const size_t total = (size_t)r.nproj * r.L;
  for (size_t t = bid * (size_t)BLOCK + threadIdx.x; t < total; t += (size_t)nblk * BLOCK) {
    const int i = (int)(t / r.L);
    const int l = (int)(t - (size_t)i * r.L);
    const int a = r.proj_mat[i];
    const int j = i - r.mat_pbase[a];
    double s = 0.0;
    const int c0 = r.mat_chunk0[a], c1 = r.mat_chunk0[a + 1];
    // a matrix's partials (one per chunk of <= 256 sphere points: 29 for a radius-3
    // sphere at h = 0.25) are loaded together, clamped indices, then added in chunk order
    for (int c = c0; c < c1; c += 32) {
      double v[32];
#pragma unroll
      for (int u = 0; u < 32; ++u) v[u] = r.partial[((size_t)min(c + u, c1 - 1) * r.pjmax + j) * r.L + l];
#pragma unroll
      for (int u = 0; u < 32; ++u)
        if (c + u < c1) s += v[u];
    }
    r.proj[t] = s * r.scals[i];
  }
}

__global__ __launch_bounds__(BLOCK) void k_bra_reduce(const BraReduceArgs r) {
  bra_reduce_body(r, blockIdx.x, gridDim.x);
}

}  // namespace octhip

namespace octhip {

int launch_bra(const BraArgs& b, int nchunk, bool mfma, hipStream_t s) {
  if (nchunk <= 0) return OCTHIP_SUCCESS;
  ProfScope prof("k_bra", s);
  hipLaunchKernelGGL(k_bra_plain, dim3(nchunk), dim3(BLOCK), 0, s, b);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(OCTHIP_ERR_HIP, "k_bra launch: %s", hipGetErrorString(e));
  (void)mfma;
  return OCTHIP_SUCCESS;
}

int launch_bra_mfma(const BraArgs& b, const double* pt, int pjs, int nchunk, hipStream_t s, double* zf,
                    size_t zfirst, size_t zcount) {
  if (nchunk <= 0) return OCTHIP_SUCCESS;
  // 32-lane tiles (two columns per lane) when the row length allows, else 16-lane tiles
  BraMfmaArgs m{b, pt, pjs, zf, zfirst, zf ? zcount : 0, nchunk * mfma_lane_tiles(b.L)};
  const int zblocks = zf ? (int)std::min<size_t>((zcount + BLOCK - 1) / BLOCK, 2048) : 0;
  ProfScope prof("k_bra", s);
  hipLaunchKernelGGL(k_bra_mfma, dim3(m.nbra + zblocks), dim3(BLOCK), 0, s, m);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(OCTHIP_ERR_HIP, "k_bra_mfma launch: %s", hipGetErrorString(e));
  return OCTHIP_SUCCESS;
}

// One workgroup per bra chunk (<= BRA_CHUNK sphere points of one matrix).  MFMA form:
// D(16 points x 16 lanes) = A(16 points x 4 projectors) B(4 projectors x 16 lanes); lane
// (kk, ii) supplies A[ii][kk] = P[q0+ii, jb+kk] and B[kk][ii] = proj[pbase+jb+kk, col];
// accumulator element r of lane (kk, ii) is D[4r+kk][ii].
// One workgroup per (chunk, lane tile of 16 TW lanes); the 4 waves take the chunk's groups
// of 16 sphere points in turn.  Loads are unconditional (clamped indices, zero operands).
// TW = 2: 32-lane tiles, lane ii computes columns 2ii and 2ii+1 with two MFMAs that share
// the projector operand and stores them as one 16-byte piece (one complex value); each
// column's sum is the TW = 1 sum, in the same order.
template <int TW>
__device__ __forceinline__ void ket_mfma_body(const KetArgs& k, int c, int lt, double* sproj) {
  const int4 ch = k.chunks[c];
  const int4 mi = k.mat_info[ch.x];
  const int npj = mi.y, mapoff = mi.w;
  const int pb = k.mat_pbase[ch.x];
  const int wave = threadIdx.x >> 6, ln = threadIdx.x & 63;
  const int kk = ln >> 4, ii = ln & 15;
  constexpr int TL = 16 * TW;               // lanes of the tile
  const int col = lt * TL + ii * TW;        // first column of this lane
  const int ngrp = (ch.z - ch.y + 15) >> 4;  // <= 16: each wave takes at most 4 groups
  static_assert(BRA_CHUNK <= 256, "4 groups of 16 sphere points per wave");
  // projector values of the first 4 projectors of this wave's groups, loaded before the
  // projections are formed so that both sets of loads are in flight together
  double apre[4];
#pragma unroll
  for (int gi = 0; gi < 4; ++gi) {
    const int q = ch.y + (wave + 4 * gi) * 16 + ii;
    apre[gi] = k.pt[(size_t)(mapoff + (q < ch.z ? q : ch.y)) * k.pjs + (kk < npj ? kk : npj - 1)];
  }
  if (k.partial) {
    // proj[pb + j, lt * TL + i] for j < npj, i < TL: the bra_reduce_body sum, same order
    const int c0 = k.mat_chunk0[ch.x], c1 = k.mat_chunk0[ch.x + 1];
    for (int t = threadIdx.x; t < npj * TL; t += BLOCK) {
      const int j = t / TL, l = lt * TL + (t % TL);
      double s = 0.0;
      for (int c = c0; c < c1; c += 32) {
        double v[32];
#pragma unroll
        for (int u = 0; u < 32; ++u) v[u] = k.partial[((size_t)min(c + u, c1 - 1) * k.pjmax + j) * k.L + l];
#pragma unroll
        for (int u = 0; u < 32; ++u)
          if (c + u < c1) s += v[u];
      }
      sproj[t] = s * k.scals[pb + j];
    }
    __syncthreads();
  }
#pragma unroll
  for (int gi = 0; gi < 4; ++gi) {
    const int g = wave + 4 * gi;
    if (g >= ngrp) break;
    const int pt0 = ch.y + g * 16;
    const int q = pt0 + ii;
    const int qs = q < ch.z ? q : ch.y;
    d4 acc[TW];
#pragma unroll
    for (int w = 0; w < TW; ++w) acc[w] = d4{0.0, 0.0, 0.0, 0.0};
    for (int jb = 0; jb < npj; jb += 4) {
      const bool jok = jb + kk < npj;
      const int jj = jok ? jb + kk : npj - 1;
      const double a = jb == 0 ? apre[gi] : k.pt[(size_t)(mapoff + qs) * k.pjs + jj];
      const Vec<TW> bb = k.partial ? ldv<TW>(sproj + jj * TL + ii * TW) : ldv<TW>(k.proj + (size_t)(pb + jj) * k.L + col);
#pragma unroll
      for (int w = 0; w < TW; ++w)
        acc[w] = __builtin_amdgcn_mfma_f64_16x16x4f64((q < ch.z && jok) ? a : 0.0, jok ? bb.v[w] : 0.0, acc[w], 0, 0, 0);
    }
    if (k.pphase) {
      // ket row of point qo times conj(its Bloch phase): with TW = 2 the lane holds the
      // (re, im) pair, otherwise lanes ii ^ 1 hold the pair
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int qo = pt0 + 4 * r + kk;
        const size_t pi2 = 2 * (size_t)(mapoff + (qo < ch.z ? qo : ch.y));
        if constexpr (TW == 2) {
          const double re = zmul_lane(k.pphase[pi2], -k.pphase[pi2 + 1], acc[0][r], acc[1][r], false);
          const double im = zmul_lane(k.pphase[pi2], -k.pphase[pi2 + 1], acc[1][r], acc[0][r], true);
          acc[0][r] = re;
          acc[1][r] = im;
        } else {
          const double xq = __shfl_xor(acc[0][r], 1);
          acc[0][r] = zmul_lane(k.pphase[pi2], -k.pphase[pi2 + 1], acc[0][r], xq, (col & 1) != 0);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int qo = pt0 + 4 * r + kk;
      if (qo < ch.z) {
        Vec<TW> o;
#pragma unroll
        for (int w = 0; w < TW; ++w) o.v[w] = acc[w][r];
        stv<TW>(k.kbuf + (size_t)(mapoff + qo) * k.L + col, o);
      }
    }
  }
}

__global__ __launch_bounds__(BLOCK) void k_ket_mfma(const KetArgs k) {
  __shared__ double2 sproj2[KET_FUSE_PJMAX * 16];  // 16-byte aligned: read as lane pairs
  double* const sproj = reinterpret_cast<double*>(sproj2);
  if ((k.L & 31) == 0) {
    const int nlt = k.L >> 5;
    ket_mfma_body<2>(k, blockIdx.x / nlt, blockIdx.x % nlt, sproj);
  } else {
    const int nlt = k.L >> 4;
    ket_mfma_body<1>(k, blockIdx.x / nlt, blockIdx.x % nlt, sproj);
  }
}

// plain form, any row length: one thread per (sphere point, lane)
__device__ __forceinline__ void ket_plain_body(const KetArgs& k, int c) {
  const int4 ch = k.chunks[c];
  const int4 mi = k.mat_info[ch.x];
  const int npj = mi.y, mapoff = mi.w;
  const int pb = k.mat_pbase[ch.x];
  const int total = (ch.z - ch.y) * k.L;
  for (int t = threadIdx.x; t < total; t += BLOCK) {
    const int q = ch.y + t / k.L, l = t % k.L;
    const double* prow = k.pt + (size_t)(mapoff + q) * k.pjs;
    double acc = 0.0;
    for (int jj = 0; jj < npj; ++jj) acc = fma(prow[jj], k.proj[(size_t)(pb + jj) * k.L + l], acc);
    if (k.pphase) {
      double accq = 0.0;  // the other lane of the (re, im) pair
      for (int jj = 0; jj < npj; ++jj) accq = fma(prow[jj], k.proj[(size_t)(pb + jj) * k.L + (l ^ 1)], accq);
      const size_t pi2 = 2 * (size_t)(mapoff + q);
      acc = zmul_lane(k.pphase[pi2], -k.pphase[pi2 + 1], acc, accq, (l & 1) != 0);
    }
    k.kbuf[(size_t)(mapoff + q) * k.L + l] = acc;
  }
}

// one thread per (multi-sphere mesh point, LPT consecutive lanes of its row)
template <int LPT>
__device__ __forceinline__ void ket_sum_lanes(const KetArgs& k, int bid, int nblk) {
  const int lp = k.L / LPT;
  const size_t total = (size_t)k.nmulti * lp;
  for (size_t t = bid * (size_t)BLOCK + threadIdx.x; t < total; t += (size_t)nblk * BLOCK) {
    const int m = (int)(t / lp);
    const int l = (int)(t - (size_t)m * lp) * LPT;
    Vec<LPT> s;
#pragma unroll
    for (int q = 0; q < LPT; ++q) s.v[q] = 0.0;
    const int e0 = k.multi_pos[m], e1 = k.multi_pos[m + 1];
    // rows in groups of 4 loaded together (clamped indices), added in entry order
    for (int e = e0; e < e1; e += 4) {
      Vec<LPT> v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = ldv<LPT>(k.kbuf + (size_t)k.multi_g[min(e + u, e1 - 1)] * k.L + l);
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (e + u < e1) {
#pragma unroll
          for (int q = 0; q < LPT; ++q) s.v[q] = s.v[q] + v[u].v[q];
        }
    }
    stv<LPT>(k.kbuf + ((size_t)k.total_points + m) * k.L + l, s);
  }
}

__device__ __forceinline__ void ket_sum_body(const KetArgs& k, int bid, int nblk) {
  if (k.L & 1) ket_sum_lanes<1>(k, bid, nblk);
  else ket_sum_lanes<2>(k, bid, nblk);
}

__global__ __launch_bounds__(BLOCK) void k_ket_plain(const KetArgs k) { ket_plain_body(k, blockIdx.x); }
__global__ __launch_bounds__(BLOCK) void k_ket_sum(const KetArgs k) { ket_sum_body(k, blockIdx.x, gridDim.x); }

int launch_ket(const KetArgs& k, int nchunk, bool mfma, hipStream_t s) {
  if (nchunk <= 0) return OCTHIP_SUCCESS;
  if (k.partial && !(mfma && k.L >= 16 && k.pjmax <= KET_FUSE_PJMAX))
    return fail(OCTHIP_ERR_INVALID, "k_ket: projections from partials need the MFMA form and <= %d projectors",
                KET_FUSE_PJMAX);
  {
    ProfScope prof("k_ket", s);
    if (mfma && k.L >= 16)  // 32-lane tiles when the row length allows, else 16-lane tiles
      hipLaunchKernelGGL(k_ket_mfma, dim3(nchunk * mfma_lane_tiles(k.L)), dim3(BLOCK), 0, s, k);
    else
      hipLaunchKernelGGL(k_ket_plain, dim3(nchunk), dim3(BLOCK), 0, s, k);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(OCTHIP_ERR_HIP, "k_ket launch: %s", hipGetErrorString(e));
  }
  // one thread per lane pair of each multi-sphere point (per lane for odd row lengths)
  const size_t total = (size_t)k.nmulti * ((k.L & 1) ? k.L : k.L / 2);
  if (total == 0) return OCTHIP_SUCCESS;
  const int grid = (int)std::min<size_t>((total + BLOCK - 1) / BLOCK, 8192);
  ProfScope prof("k_ket_sum", s);
  hipLaunchKernelGGL(k_ket_sum, dim3(grid), dim3(BLOCK), 0, s, k);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(OCTHIP_ERR_HIP, "k_ket_sum launch: %s", hipGetErrorString(e));
  return OCTHIP_SUCCESS;
}

int launch_bra_reduce(const BraReduceArgs& r, hipStream_t s) {
  const size_t total = (size_t)r.nproj * r.L;
  if (total == 0) return OCTHIP_SUCCESS;
  int grid = (int)((total + BLOCK - 1) / BLOCK);
  if (grid > 4096) grid = 4096;
  ProfScope prof("k_bra_reduce", s);
  hipLaunchKernelGGL(k_bra_reduce, dim3(grid), dim3(BLOCK), 0, s, r);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(OCTHIP_ERR_HIP, "k_bra_reduce launch: %s", hipGetErrorString(e));
  return OCTHIP_SUCCESS;
}

}  // namespace octhip

// ------------------------------------------------------- multi-segment launches --
// One launch over several (system, batch) pairs: segment i owns blocks off[i] ..
// off[i+1]-1 and its argument block args[i]; a block finds its segment by bisection over
// the (uniform) offset table.  Empty segments are skipped by taking the last segment
// whose first block is <= blockIdx.
namespace octhip {
namespace {
__device__ __forceinline__ int segment_of(const int* __restrict__ off, int nseg, int b) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(BLOCK) void k_bra_mfma_multi(const BraMfmaArgs* __restrict__ a,
                                                          const int* __restrict__ off, int nseg) {
  const int i = segment_of(off, nseg, blockIdx.x);
  __shared__ double red[2][BRA_RED];
  const int b = blockIdx.x - off[i], nlt = mfma_lane_tiles(a[i].b.L);
  if ((a[i].b.L & 31) == 0) bra_mfma_body<2>(a[i], b / nlt, b % nlt, red);
  else bra_mfma_body<1>(a[i], b / nlt, b % nlt, red);
}
__global__ __launch_bounds__(BLOCK) void k_bra_plain_multi(const BraMfmaArgs* __restrict__ a,
                                                           const int* __restrict__ off, int nseg) {
  const int i = segment_of(off, nseg, blockIdx.x);
  bra_plain_body(a[i].b, blockIdx.x - off[i]);
}
__global__ __launch_bounds__(BLOCK) void k_bra_reduce_multi(const BraReduceArgs* __restrict__ a,
                                                            const int* __restrict__ off, int nseg) {
  const int i = segment_of(off, nseg, blockIdx.x);
  bra_reduce_body(a[i], blockIdx.x - off[i], off[i + 1] - off[i]);
}
__global__ __launch_bounds__(BLOCK) void k_ket_mfma_multi(const KetArgs* __restrict__ a,
                                                          const int* __restrict__ off, int nseg) {
  const int i = segment_of(off, nseg, blockIdx.x);
  __shared__ double2 sproj2[KET_FUSE_PJMAX * 16];  // 16-byte aligned: read as lane pairs
  double* const sproj = reinterpret_cast<double*>(sproj2);
  const int b = blockIdx.x - off[i], nlt = mfma_lane_tiles(a[i].L);
  if ((a[i].L & 31) == 0) ket_mfma_body<2>(a[i], b / nlt, b % nlt, sproj);
  else ket_mfma_body<1>(a[i], b / nlt, b % nlt, sproj);
}
__global__ __launch_bounds__(BLOCK) void k_ket_plain_multi(const KetArgs* __restrict__ a,
                                                           const int* __restrict__ off, int nseg) {
  const int i = segment_of(off, nseg, blockIdx.x);
  ket_plain_body(a[i], blockIdx.x - off[i]);
}
__global__ __launch_bounds__(BLOCK) void k_ket_sum_multi(const KetArgs* __restrict__ a,
                                                         const int* __restrict__ off, int nseg) {
  const int i = segment_of(off, nseg, blockIdx.x);
  ket_sum_body(a[i], blockIdx.x - off[i], off[i + 1] - off[i]);
}
__global__ __launch_bounds__(BLOCK) void k_zero_multi(const ZeroArgs* __restrict__ a,
                                                      const int* __restrict__ off, int nseg) {
  const int i = segment_of(off, nseg, blockIdx.x);
  const ZeroArgs z = a[i];
  const size_t nb = (size_t)(off[i + 1] - off[i]);
  for (size_t t = (blockIdx.x - off[i]) * (size_t)BLOCK + threadIdx.x; t < z.count; t += nb * BLOCK)
    z.f[z.first + t] = 0.0;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(OCTHIP_ERR_HIP, "%s launch: %s", what, hipGetErrorString(e));
  return OCTHIP_SUCCESS;
}
}  // namespace

int launch_multi(int kind, const void* args, const int* off, int nseg, int nblocks, hipStream_t s) {
  if (nblocks <= 0) return OCTHIP_SUCCESS;
  const dim3 g(nblocks), b(BLOCK);
  switch (kind) {
    case MULTI_BRA_MFMA: {
      ProfScope prof("k_bra", s);
      hipLaunchKernelGGL(k_bra_mfma_multi, g, b, 0, s, (const BraMfmaArgs*)args, off, nseg);
      return check_launch("k_bra_mfma_multi");
    }
    case MULTI_BRA_PLAIN: {
      ProfScope prof("k_bra", s);
      hipLaunchKernelGGL(k_bra_plain_multi, g, b, 0, s, (const BraMfmaArgs*)args, off, nseg);
      return check_launch("k_bra_plain_multi");
    }
    case MULTI_BRA_REDUCE: {
      ProfScope prof("k_bra_reduce", s);
      hipLaunchKernelGGL(k_bra_reduce_multi, g, b, 0, s, (const BraReduceArgs*)args, off, nseg);
      return check_launch("k_bra_reduce_multi");
    }
    case MULTI_KET_MFMA: {
      ProfScope prof("k_ket", s);
      hipLaunchKernelGGL(k_ket_mfma_multi, g, b, 0, s, (const KetArgs*)args, off, nseg);
      return check_launch("k_ket_mfma_multi");
    }
    case MULTI_KET_PLAIN: {
      ProfScope prof("k_ket", s);
      hipLaunchKernelGGL(k_ket_plain_multi, g, b, 0, s, (const KetArgs*)args, off, nseg);
      return check_launch("k_ket_plain_multi");
    }
    case MULTI_KET_SUM: {
      ProfScope prof("k_ket_sum", s);
      hipLaunchKernelGGL(k_ket_sum_multi, g, b, 0, s, (const KetArgs*)args, off, nseg);
      return check_launch("k_ket_sum_multi");
    }
    case MULTI_ZERO: {
      ProfScope prof("k_zero_rows", s);
      hipLaunchKernelGGL(k_zero_multi, g, b, 0, s, (const ZeroArgs*)args, off, nseg);
      return check_launch("k_zero_multi");
    }
    default:
      return fail(OCTHIP_ERR_INVALID, "unknown multi-segment kernel %d", kind);
  }
}
}  // namespace octhip
