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
// EPI_TAYLOR_X: the Taylor epilogue plus an optional second accumulator (psib2 at deltat2)
// and optional density partials of the final states (separate instances, so the plain
// Taylor epilogue keeps its register budget)
enum { EPI_STORE = 0, EPI_TAYLOR = 1, EPI_TAYLOR_X = 2 };

struct ApplyArgs {
  const double* __restrict__ fi;   // input batch (>= lim rows)
  double* __restrict__ fo;         // output batch (np rows), EPI_STORE
  const double* __restrict__ zrow; // zeros standing for boundary rows >= lim
  int np;                          // rows computed
  int lim;                         // neighbour rows >= lim read as zero
  int L;                           // doubles per row
  int tpp, tshift;                 // threads per point (power of 2), log2(tpp)
  int tpts;                        // points per tile (one workgroup)
  int sstride;                     // LDS row stride of the tile's neighbour table
  int chunk;                       // blockIdx -> tile remap granularity (XCD locality)
  int tile_iters;                  // requested passes per tile (0 = default)
  int ntiles;
  int n;                           // stencil size (runtime-size path)
  const int* __restrict__ cidx;    // lattice: site of point p
  const int* __restrict__ cube;    // lattice: site -> point
  const int* __restrict__ dcp;     // lattice: site delta of stencil point j (device copy)
  int dc[MAX_STENCIL];             // lattice: the same deltas as kernel arguments
  const int* __restrict__ ri;      // run-compressed offsets
  const int* __restrict__ rimap;   // run of point p
  double w[MAX_STENCIL];           // weights already multiplied by the factor
  // local potential (the dipole term, if any, is already folded in)
  const double* __restrict__ v;
  // nonlocal ket: kslot[p] = row of ksum holding sum_a P_a proj_a at point p, or -1
  const int* __restrict__ kslot;
  const double* __restrict__ ksum;
  int nk;                          // rows of ksum
  // Taylor epilogue: yout = Hy ; acc = acc_in + ca * ya + cb * yb + c1 * y + c2 * Hy
  // (complex lanes, added in this order; acc_in = acc when null, i.e. in place; ya, yb
  // optional: earlier powers added at the last order)
  double* yout;
  double* acc;
  const double* acc_in;
  double c1r, c1i, c2r, c2i;
  int has_c1;
  const double* ya;
  const double* yb;
  double car, cai, cbr, cbi;
  // EPI_TAYLOR_X: acc2 = acc_in + da ya + db yb + d1 y + d2 Hy (same inputs, own factors;
  // d1 used when has_c1), and density partials dpart[chunk * np + p] = sum over the
  // chunk's lanes of dw[state] |final value|^2, the final value being acc2 when given,
  // else acc (dw: one weight per complex state of the row)
  double* acc2;
  double d1r, d1i, d2r, d2i, dar, dai, dbr, dbi;
  double* dpart;
  const double* dw;
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

template <int LPT>
struct Vec;
template <>
struct Vec<1> {
  double v[1];
};
template <>
struct Vec<2> {
  double v[2];
};

template <int LPT>
__device__ __forceinline__ Vec<LPT> ldv(const double* p) {
  Vec<LPT> r;
  if constexpr (LPT == 2) {
    const double2 t = *reinterpret_cast<const double2*>(p);
    r.v[0] = t.x;
    r.v[1] = t.y;
  } else {
    r.v[0] = *p;
  }
  return r;
}

template <int LPT>
__device__ __forceinline__ void stv(double* p, const Vec<LPT>& r) {
  if constexpr (LPT == 2) {
    *reinterpret_cast<double2*>(p) = make_double2(r.v[0], r.v[1]);
  } else {
    *p = r.v[0];
  }
}

// base row piece of the Taylor accumulators: acc_in, or acc itself when in place
template <int LPT>
__device__ __forceinline__ Vec<LPT> stv_base(const ApplyArgs& a, int p, int l) {
  return ldv<LPT>((a.acc_in ? a.acc_in : a.acc) + (size_t)p * a.L + l);
}

// Fused apply over a tile of `tpts` consecutive points per workgroup.
//  phase 1: the tile's neighbour indices (site map or run-compressed ri) are gathered
//           into LDS by all threads with vector loads -- one memory round trip per tile;
//  phase 2: threads stream rows; a thread owns LPT consecutive doubles of one point (one
//           complex value for LPT == 2), tpp threads cover a point's row, BLOCK/tpp points
//           are processed per pass.  Per lane: stencil sum in stencil order with fma,
//           then fma(v, psi), then the KB ket terms matrix by matrix, then the epilogue.
template <int N, int IDX, int EPI, int LPT>
__global__ __launch_bounds__(BLOCK) void k_apply(const ApplyArgs a) {
  extern __shared__ int sidx[];
  const int tile = tile_of(blockIdx.x, a.chunk);
  if (tile >= a.ntiles) return;
  const int n = N > 0 ? N : a.n;
  const int ss = a.sstride;
  const int p0 = tile * a.tpts;
  const int cnt = min(a.tpts, a.np - p0);
  if constexpr (N > 0) {
    // one thread per tile point: its N index loads are independent and issued together
    for (int t = threadIdx.x; t < cnt; t += BLOCK) {
      const int p = p0 + t;
      int q[N];
      if constexpr (IDX == IDX_LATTICE) {
        const int c = a.cidx[p];
#pragma unroll
        for (int j = 0; j < N; ++j) q[j] = a.cube[c + a.dc[j]];
      } else {
        const int* row = a.ri + (size_t)a.rimap[p] * N;
#pragma unroll
        for (int j = 0; j < N; ++j) q[j] = p + row[j];
      }
#pragma unroll
      for (int j = 0; j < N; ++j) sidx[t * ss + j] = q[j];
    }
    __syncthreads();
  } else if (n > 0) {
    for (int e = threadIdx.x; e < cnt * n; e += BLOCK) {
      const int t = e / n;
      const int j = e - t * n;
      const int p = p0 + t;
      int q;
      if constexpr (IDX == IDX_LATTICE) {
        q = a.cube[a.cidx[p] + a.dcp[j]];
      } else {
        q = p + a.ri[(size_t)a.rimap[p] * n + j];
      }
      sidx[t * ss + j] = q;
    }
    __syncthreads();
  }
  const int tpp = a.tpp;
  const int slots = BLOCK >> a.tshift;
  const int slot = threadIdx.x >> a.tshift;
  const int lane0 = (threadIdx.x & (tpp - 1)) * LPT;
  const size_t L = (size_t)a.L;
  for (int t = slot; t < cnt; t += slots) {
    const int p = p0 + t;
    const int* qi = sidx + t * ss;
    const double vv = a.v ? a.v[p] : 0.0;
    const int ks = a.kslot ? a.kslot[p] : -1;
    double dsum = 0.0;  // EPI_TAYLOR_X: this thread's share of the point's density
    for (int l = lane0; l < a.L; l += tpp * LPT) {
      Vec<LPT> s;
#pragma unroll
      for (int k = 0; k < LPT; ++k) s.v[k] = 0.0;
      if constexpr (N > 0) {
        int q[N];
#pragma unroll
        for (int j = 0; j < N; ++j) q[j] = qi[j];
        Vec<LPT> x[N];
#pragma unroll
        for (int j = 0; j < N; ++j) {
          const double* src = ((unsigned)q[j] < (unsigned)a.lim) ? a.fi + (size_t)q[j] * L : a.zrow;
          x[j] = ldv<LPT>(src + l);
        }
#pragma unroll
        for (int j = 0; j < N; ++j)
#pragma unroll
          for (int k = 0; k < LPT; ++k) s.v[k] = fma(a.w[j], x[j].v[k], s.v[k]);
      } else {
        for (int j = 0; j < n; ++j) {
          const int qq = qi[j];
          const double* src = ((unsigned)qq < (unsigned)a.lim) ? a.fi + (size_t)qq * L : a.zrow;
          const Vec<LPT> x = ldv<LPT>(src + l);
#pragma unroll
          for (int k = 0; k < LPT; ++k) s.v[k] = fma(a.w[j], x.v[k], s.v[k]);
        }
      }
      const Vec<LPT> x = ldv<LPT>(a.fi + (size_t)p * L + l);
      if (a.v) {
#pragma unroll
        for (int k = 0; k < LPT; ++k) s.v[k] = fma(vv, x.v[k], s.v[k]);
      }
      if (ks >= 0) {
        const Vec<LPT> kv = ldv<LPT>(a.ksum + (size_t)ks * L + l);
#pragma unroll
        for (int k = 0; k < LPT; ++k) s.v[k] = s.v[k] + kv.v[k];
      }
      if constexpr (EPI == EPI_STORE) {
        stv<LPT>(a.fo + (size_t)p * L + l, s);
      } else {
        static_assert(EPI != EPI_TAYLOR || LPT == 2, "Taylor epilogue needs complex lane pairs");
        if (a.yout) stv<LPT>(a.yout + (size_t)p * L + l, s);
        if (a.acc) {
          Vec<LPT> av = ldv<LPT>((a.acc_in ? a.acc_in : a.acc) + (size_t)p * L + l);
          if (a.ya) {
            const Vec<LPT> u = ldv<LPT>(a.ya + (size_t)p * L + l);
            av.v[0] = av.v[0] + zmul_lane(a.car, a.cai, u.v[0], u.v[1], false);
            av.v[1] = av.v[1] + zmul_lane(a.car, a.cai, u.v[1], u.v[0], true);
          }
          if (a.yb) {
            const Vec<LPT> u = ldv<LPT>(a.yb + (size_t)p * L + l);
            av.v[0] = av.v[0] + zmul_lane(a.cbr, a.cbi, u.v[0], u.v[1], false);
            av.v[1] = av.v[1] + zmul_lane(a.cbr, a.cbi, u.v[1], u.v[0], true);
          }
          if (a.has_c1) {
            av.v[0] = av.v[0] + zmul_lane(a.c1r, a.c1i, x.v[0], x.v[1], false);
            av.v[1] = av.v[1] + zmul_lane(a.c1r, a.c1i, x.v[1], x.v[0], true);
          }
          av.v[0] = av.v[0] + zmul_lane(a.c2r, a.c2i, s.v[0], s.v[1], false);
          av.v[1] = av.v[1] + zmul_lane(a.c2r, a.c2i, s.v[1], s.v[0], true);
          if constexpr (EPI == EPI_TAYLOR_X) {
            if (a.acc2) {
              // psib2: the same terms with the deltat2 factors, from the same base
              const Vec<LPT> b0 = stv_base<LPT>(a, p, l);
              Vec<LPT> a2 = b0;
              if (a.ya) {
                const Vec<LPT> u = ldv<LPT>(a.ya + (size_t)p * L + l);
                a2.v[0] = a2.v[0] + zmul_lane(a.dar, a.dai, u.v[0], u.v[1], false);
                a2.v[1] = a2.v[1] + zmul_lane(a.dar, a.dai, u.v[1], u.v[0], true);
              }
              if (a.yb) {
                const Vec<LPT> u = ldv<LPT>(a.yb + (size_t)p * L + l);
                a2.v[0] = a2.v[0] + zmul_lane(a.dbr, a.dbi, u.v[0], u.v[1], false);
                a2.v[1] = a2.v[1] + zmul_lane(a.dbr, a.dbi, u.v[1], u.v[0], true);
              }
              if (a.has_c1) {
                a2.v[0] = a2.v[0] + zmul_lane(a.d1r, a.d1i, x.v[0], x.v[1], false);
                a2.v[1] = a2.v[1] + zmul_lane(a.d1r, a.d1i, x.v[1], x.v[0], true);
              }
              a2.v[0] = a2.v[0] + zmul_lane(a.d2r, a.d2i, s.v[0], s.v[1], false);
              a2.v[1] = a2.v[1] + zmul_lane(a.d2r, a.d2i, s.v[1], s.v[0], true);
              stv<LPT>(a.acc2 + (size_t)p * L + l, a2);
              if (a.dpart) dsum = fma(a.dw[l >> 1], a2.v[0] * a2.v[0] + a2.v[1] * a2.v[1], dsum);
            } else if (a.dpart) {
              dsum = fma(a.dw[l >> 1], av.v[0] * av.v[0] + av.v[1] * av.v[1], dsum);
            }
          }
          stv<LPT>(a.acc + (size_t)p * L + l, av);
        }
      }
    }
    if constexpr (EPI == EPI_TAYLOR_X) {
      // the tpp threads of this point are consecutive lanes of one wave
      if (a.dpart) {
        for (int off = tpp >> 1; off > 0; off >>= 1) dsum += __shfl_xor(dsum, off);
        if ((threadIdx.x & (tpp - 1)) == 0) a.dpart[p] = dsum;
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

// streaming helpers of the propagators (csrc/propagation.hip)
int launch_zaxpy(double* y, const double* x, size_t count, double zr, double zi, hipStream_t s);
int launch_daxpy(double* y, const double* x, size_t count, double a, hipStream_t s);
// density[p] += sum_states w[state] |psi(state, p)|^2 over the np inner rows
int launch_density(const double* psi, int L, int np, int is_complex, const double* w, double* density,
                   hipStream_t s);
// density[p] += sum_{c < nchunk} dpart[c * np + p], in chunk order
int launch_density_reduce(double* density, const double* dpart, int nchunk, int np, hipStream_t s);
// lane chunks of the density partials an EPI_TAYLOR_X pass of this operator writes
int apply_lane_chunks(const octhip_nl_operator_s* op, const ApplyArgs& a);

// launch helper: picks the template instance for (n, index kind, lanes, epilogue)
int launch_apply(ApplyArgs& a, bool lattice, int epi, hipStream_t s);
// zero rows np..np_part-1 of a batch of L lanes (boundaries_set, zero Dirichlet)
int zero_boundary(double* f, int np, int np_part, int L, hipStream_t s);

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
  // Bloch phases of the sphere points (complex pairs, indexed like the maps; complex rows
  // only): the gathered values are multiplied by them (hamiltonian_elec_base_inc.F90:
  // 683-692); null: none
  const double* __restrict__ pphase = nullptr;
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

// Ket of X(hamiltonian_elec_base_nlocal_finish) (src/hamiltonian/hamiltonian_elec_base_inc.F90:
// 914-943) in two passes:
//  k_ket:     kbuf[g, l] = sum_j P[g, j] proj[pbase(g) + j, l] for every global sphere point g
//             (matrix by matrix: a (npoints x nprojs) x (nprojs x L) product, FP64 MFMA);
//  k_ket_sum: kbuf[total_points + m, l] = sum of the rows of the m-th mesh point that lies in
//             several spheres, in matrix order.
// A mesh point p then adds row kref[p] of kbuf (its single sphere row or its sum), or nothing.
struct KetArgs {
  const int4* __restrict__ chunks;    // bra chunks {matrix, q0, q1, -}
  const int4* __restrict__ mat_info;  // {npoints, nprojs, matrix offset, map offset}
  const int* __restrict__ mat_pbase;
  const double* __restrict__ pt;      // transposed projector rows, stride pjs
  int pjs;
  const double* __restrict__ proj;
  double* __restrict__ kbuf;
  int L;
  // multi-sphere sums
  const int* __restrict__ multi_pos;  // nmulti+1 CSR into multi_g
  const int* __restrict__ multi_g;    // global sphere points, matrix order
  int nmulti, total_points;
  // optional (MFMA form only): the bra's chunk partials instead of proj; each workgroup
  // then forms the projections of its matrix and lane tile itself, summed and scaled as
  // k_bra_reduce does (so bit-identical), and k_bra_reduce is not launched
  const double* __restrict__ partial = nullptr;
  const int* __restrict__ mat_chunk0 = nullptr;  // nmat+1
  const double* __restrict__ scals = nullptr;
  int pjmax = 0;
  // Bloch phases of the sphere points: the ket row of point q is multiplied by
  // conj(pphase[q]) (hamiltonian_elec_base_inc.F90:955-966); null: none
  const double* __restrict__ pphase = nullptr;
};

// largest projector count per matrix for which the ket forms its projections itself
constexpr int KET_FUSE_PJMAX = 32;

struct BraMfmaArgs {
  BraArgs b;
  const double* __restrict__ pt;  // total_points x pjs
  int pjs;
  // optional: rows [zfirst, zfirst + zcount) of zf cleared by blocks nbra.. of the launch
  double* zf;
  size_t zfirst, zcount;
  int nbra;                       // blocks doing the bra: chunks x 16-lane tiles
};

struct ZeroArgs {
  double* f;
  size_t first, count;            // zero f[first .. first+count)
};

// kernels of a multi-segment launch (several systems/batches in one grid)
enum { MULTI_BRA_MFMA = 1, MULTI_BRA_PLAIN, MULTI_BRA_REDUCE, MULTI_KET_MFMA, MULTI_KET_PLAIN,
       MULTI_KET_SUM, MULTI_ZERO };
// args: device array of nseg argument blocks; off: device array of nseg+1 block offsets
int launch_multi(int kind, const void* args, const int* off, int nseg, int nblocks, hipStream_t s);

int launch_bra(const BraArgs& b, int nchunk, bool mfma, hipStream_t s);
int launch_ket(const KetArgs& k, int nchunk, bool mfma, hipStream_t s);
// lane tiles of the MFMA bra and ket over rows of L lanes (L >= 16): 32-lane tiles (two
// columns per lane) when 32 divides L, else 16-lane tiles
__host__ __device__ inline int mfma_lane_tiles(int L) { return (L & 31) == 0 ? L >> 5 : L >> 4; }
int launch_bra_reduce(const BraReduceArgs& r, hipStream_t s);

// sphere points per bra chunk
constexpr int BRA_CHUNK = 256;

}  // namespace octhip
