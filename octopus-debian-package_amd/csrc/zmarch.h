// 2.5-D "z-march" apply for the order-4 star stencil on a lattice (gfx950).
//
// The gather kernel (k_apply) reads 25 full rows per point, so L2 has to deliver 25x the
// HBM traffic.  Here a workgroup owns a ZM_TX x ZM_TY tile of lattice columns (8x8 by
// default) and one chunk of LC lanes, and walks a segment of the tile's z range plane by
// plane:
//   * the current plane (tile + 4-site halo strips on the four sides) is staged in LDS
//     (double-buffered): x and y neighbours are LDS reads;
//   * each thread keeps its column's values at z-4..z+4 in registers: z neighbours cost
//     nothing;
//   * the site -> point indices of the item (precomputed per item on the host) and the
//     ket slots of its own sites are staged in LDS once per item;
//   * loads run a few planes ahead (own column values, halo values, and the potential,
//     ket row and Taylor accumulator of the output point; depths in ZRings), into ring buffers
//     with compile-time slots (the z loop is unrolled by the ring period); every load is
//     unconditional, so the compiler's wait counts stay exact and a step waits only on
//     loads issued planes earlier.
// Per output point an 8x8 tile reads 1 own row chunk + 2 halo chunks (128 halo sites for
// 64 columns) instead of 25 rows; rows go through raw buffer resources (below).  The lattice is a padded "march cube" of point indices;
// sites that are not points hold -1.  Values follow the gather kernel exactly: point
// q < lim -> its row, anything else -> 0, and the 25 terms are summed in stencil order
// (x -4..4, y -4..-1,1..4, z -4..-1,1..4) with fma, so results are bit-identical to it.
#pragma once

#include "kernels.h"

namespace octhip {

#ifndef ZM_TX
#define ZM_TX 8
#endif
#ifndef ZM_TY
#define ZM_TY 8
#endif
constexpr int ZTX = ZM_TX;            // tile columns along x
constexpr int ZTY = ZM_TY;            // tile columns along y
constexpr int ZR = 4;                 // stencil reach
constexpr int ZNC = ZTX * ZTY;        // own columns of a tile
constexpr int ZNH = 8 * (ZTX + ZTY);  // halo sites of the cross-shaped plane window
constexpr int ZS = ZNC + ZNH;         // sites of the plane window
constexpr int ZHU = (ZNH + ZNC - 1) / ZNC;  // halo sites loaded per column thread group
static_assert((ZTX & (ZTX - 1)) == 0 && (ZTY & (ZTY - 1)) == 0 && ZS % 4 == 0, "tile shape");
constexpr int ZPADZ = 8;
// ints of the index table of an item of zs planes: own columns of zs + 8 planes, halo sites
// of zs planes (a multiple of 4: the table is copied as int4)
__host__ __device__ constexpr int zmarch_tab_ints(int zs) { return (zs + 8) * ZNC + zs * ZNH; }
static_assert(ZNC % 4 == 0 && ZNH % 4 == 0, "index table copied as int4");             // march-cube padding below/above the inner points in z
// Prefetch depths (planes loaded ahead) of the own column values, the halo values and the
// output-point data, and the own-value ring length (= z loop unroll, >= own depth + 5).
// A tile's own values of plane P are the halo of its neighbours, which load them HD
// planes ahead of P, against own-depth planes ahead here: the smaller that gap, the more
// of the halo reads hit in L2 (own 5 rather than 6 for the plain store: 2% less time on
// the R=12 sphere).  The Taylor epilogue carries an accumulator row per plane more, so it
// takes shallow rings (127 VGPRs: two 512-thread workgroups per CU, whose loads cover each
// other's latency; 2.3% faster Taylor steps than with the deep rings at 143 VGPRs); the
// plain store keeps two planes of halo and point data (126 VGPRs with ket rows, still two
// per CU; the shallow ones cost the 8-lane Laplacian 12%).
template <int EPI>
struct ZRings {
  static constexpr int own = 5, halo = 2, aux = 2, ring = 12;
};
template <>
struct ZRings<EPI_TAYLOR> {
  static constexpr int own = 5, halo = 1, aux = 1, ring = 10;
};
template <>
struct ZRings<EPI_TAYLOR_X> : ZRings<EPI_TAYLOR> {};

// Raw buffer resources over the row arrays (gfx9 range checking, stride 0): a load whose
// byte offset lies at or beyond num_records returns zeros and such a store is dropped.
// Offsets are 32-bit, row q at q << log2(row bytes), so the point index -1 (and any
// index >= the array's row count) falls out of range: this replaces the zero-row pointer
// selects and keeps every load and store unconditional.  The host sends a batch here only
// when (rows + 1) * row bytes <= 2^32 for every array.
constexpr int ZM_RSRC_DW3 = 0x00020000;  // dword 3 of a raw buffer resource: 32-bit data format
__device__ __forceinline__ __amdgpu_buffer_rsrc_t zrsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, ZM_RSRC_DW3);
}
template <int LPT>
__device__ __forceinline__ Vec<LPT> bld(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  Vec<LPT> v;
  if constexpr (LPT == 2) {
    const double2 t = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
    v.v[0] = t.x;
    v.v[1] = t.y;
  } else {
    v.v[0] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  }
  return v;
}
// Stores are marked non-temporal (cache policy NT): output rows are not read again in
// the pass, and keeping them out of the way leaves the XCD's L2 to the input rows that
// the neighbouring tiles read a few planes later as their halo (7% less read traffic and
// 4% less time for the H-apply pass on the R=12 sphere).
constexpr int ZM_STORE_NT = 2;  // cache-policy operand: NT
template <int LPT>
__device__ __forceinline__ void bst(__amdgpu_buffer_rsrc_t r, uint32_t off, const Vec<LPT>& v) {
  if constexpr (LPT == 2) {
    using U4 = decltype(__builtin_amdgcn_raw_buffer_load_b128(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(U4, make_double2(v.v[0], v.v[1])), r, off, 0,
                                           ZM_STORE_NT);
  } else {
    using U2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v.v[0]), r, off, 0, ZM_STORE_NT);
  }
}

// smallest divisor of rq above d: the length of a prefetch ring d planes deep whose slots
// stay compile-time constants in a z loop unrolled rq times
constexpr int ring_above(int rq, int d) {
  for (int r = d + 1; r < rq; ++r)
    if (rq % r == 0) return r;
  return rq;
}

struct ZArgs {
  ApplyArgs a;
  const int4* __restrict__ items;    // {tile x, tile y, first plane, end plane} (cube z)
  const int* __restrict__ tab;       // per item: index table (zmarch_tab_ints ints, layout below)
  const int* __restrict__ tab_off;   // per item: first int of its table (a multiple of 4)
  const int2* __restrict__ sched;    // blockIdx -> {item, lane chunk}
};

// One work item: tile it = {tile x, tile y, first plane, end plane}, lane chunk `chunk`.
// za may live in kernel arguments (single operator) or in global memory (one descriptor
// per system of a batched launch): every field is copied to a local once.
// SHALLOW: the shallow prefetch rings whatever the epilogue (the batched instances, whose
// descriptor handling costs a few registers more).
template <int LPT, int LC, int EPI, bool NL, bool SHALLOW = false>
__device__ __forceinline__ void zmarch_body(const ZArgs& za, const int4 it, const int item, const int chunk) {
  constexpr int PAIRS = LC / LPT;    // threads per column
  constexpr int NT = ZNC * PAIRS;
  extern __shared__ double2 zsh2[];  // 16-byte aligned: row chunks move as ds_*_b128
  double* const zsh = reinterpret_cast<double*>(zsh2);
  double* const lds0 = zsh;                              // plane buffer 0
  double* const lds1 = zsh + ZS * LC;                    // plane buffer 1
  int* const tidx = reinterpret_cast<int*>(zsh + 2 * ZS * LC);  // item index table
  const ApplyArgs& A = za.a;
  const int a_np = A.np;
  const double* const a_v = A.v;
  const int* const a_kslot = A.kslot;
  double* const a_acc = A.acc;
  const double a_c1r = A.c1r, a_c1i = A.c1i, a_c2r = A.c2r, a_c2i = A.c2i;
  const int a_has_c1 = A.has_c1;
  const double a_car = A.car, a_cai = A.cai, a_cbr = A.cbr, a_cbi = A.cbi;
  double a_w[25];
#pragma unroll
  for (int j = 0; j < 25; ++j) a_w[j] = A.w[j];
  const int kp = threadIdx.x % PAIRS;
  const int col = threadIdx.x / PAIRS;
  const int cx = col % ZTX, cy = col / ZTX;
  const int l = chunk * LC + kp * LPT;         // first lane of this thread
  constexpr bool TAY = EPI != EPI_STORE;
  const bool has_acc = TAY && a_acc != nullptr;
  // row q of an array starts at byte q << sh; arrays without a role here get no records
  const int sh = __builtin_ctz((unsigned)A.L) + 3;
  const uint32_t lbyte = (uint32_t)l * 8u;
  const uint32_t np_bytes = (uint32_t)a_np << sh;
  const auto r_fi = zrsrc(A.fi, (uint32_t)A.lim << sh);
  const auto r_v = zrsrc(a_v, a_v ? (uint32_t)a_np * 8u : 0u);
  const auto r_k = zrsrc(A.ksum, NL ? (uint32_t)A.nk << sh : 0u);
  const auto r_ai = zrsrc(A.acc_in ? A.acc_in : a_acc, has_acc ? np_bytes : 0u);
  const auto r_fo = zrsrc(A.fo, EPI == EPI_STORE ? np_bytes : 0u);
  const auto r_y = zrsrc(A.yout, A.yout ? np_bytes : 0u);
  const auto r_acc = zrsrc(a_acc, has_acc ? np_bytes : 0u);
  const auto r_ya = zrsrc(A.ya, (has_acc && A.ya) ? np_bytes : 0u);  // earlier powers, last order
  const auto r_yb = zrsrc(A.yb, (has_acc && A.yb) ? np_bytes : 0u);
  // EPI_TAYLOR_X: second accumulator and density partials of this thread's state
  const auto r_a2 = zrsrc(A.acc2, (EPI == EPI_TAYLOR_X && has_acc && A.acc2) ? np_bytes : 0u);
  double* const a_dpart = EPI == EPI_TAYLOR_X ? A.dpart : nullptr;
  const double wl = (EPI == EPI_TAYLOR_X && A.dw) ? A.dw[l >> 1] : 0.0;
  auto roff = [&](int q) { return ((uint32_t)q << sh) + lbyte; };
  // halo units of this thread: halo site h = col + ZNC*i (if h < ZNH), same lane pair
  int hlds[ZHU];
#pragma unroll
  for (int i = 0; i < ZHU; ++i) hlds[i] = (ZNC + col + ZNC * i) * LC + kp * LPT;
  auto halo_ok = [&](int i) { return ZNC * (i + 1) <= ZNH || col + ZNC * i < ZNH; };
  // Cross-shaped window in LDS, one LC-lane chunk per site slot ((x0, y0) = first tile
  // column, W = ZTX, H = ZTY, N = W H):
  //   slot c (0..N-1)               tile column c = cx + W cy
  //   slot N + 4 r + k              left strip,   x = x0-4+k, y = y0+r   (k < 4, r < H)
  //   slot N + 4 H + 4 r + k        right strip,  x = x0+W+k, y = y0+r
  //   slot N + 8 H + W k + c        bottom strip, x = x0+c,   y = y0-4+k (c < W)
  //   slot N + 8 H + 4 W + W k + c  top strip,    x = x0+c,   y = y0+H+k
  const int own_lds = col * LC + kp * LPT;
  auto xnb = [&](int dx) {  // LDS offset of the x-neighbour dx of this thread's column
    const int xx = cx + dx;
    const int slot = xx < 0 ? ZNC + 4 * cy + (xx + 4)
                            : (xx >= ZTX ? ZNC + 4 * ZTY + 4 * cy + (xx - ZTX) : xx + ZTX * cy);
    return slot * LC + kp * LPT;
  };
  auto ynb = [&](int dy) {  // LDS offset of the y-neighbour dy
    const int yy = cy + dy;
    const int slot = yy < 0 ? ZNC + 8 * ZTY + ZTX * (yy + 4) + cx
                            : (yy >= ZTY ? ZNC + 8 * ZTY + 4 * ZTX + ZTX * (yy - ZTY) + cx : cx + ZTX * yy);
    return slot * LC + kp * LPT;
  };

  // value of point q (zero for boundary/non-points when q >= lim or q = -1)
  auto val = [&](int q) { return bld<LPT>(r_fi, roff(q)); };
  auto lds_st = [&](int buf, int off, const Vec<LPT>& v) {
    double* b = buf ? lds1 : lds0;
    stv<LPT>(b + off, v);
  };
  auto lds_ld = [&](int buf, int off) {
    const double* b = buf ? lds1 : lds0;
    return ldv<LPT>(b + off);
  };

  const int zb = it.z, ze = it.w;
  // ---- index table of the item in LDS: the ZNC own columns of planes zb-4 .. ze+3, then
  //      the ZNH halo sites of planes zb .. ze-1 (the only planes whose halo is read).
  //      Precomputed per item, so the prologue is one coalesced block copy; index reads
  //      then cost LDS, not the in-order global-load counter.
  const int nown = (ze - zb + 8) * ZNC;
  int* const tks = tidx + nown + (ze - zb) * ZNH;  // ket slots of the own sites
  {
    // (through a buffer resource: the descriptors of a batched launch live in global
    // memory, and a generic pointer loaded from there would be read with flat loads that
    // the compiler must order against the LDS stores)
    const int n4 = zmarch_tab_ints(ze - zb) / 4;
    const auto r_tab = zrsrc(za.tab + za.tab_off[item], (uint32_t)n4 * 16u);
    int4* const t4 = reinterpret_cast<int4*>(tidx);
    for (int base = 0; base < n4; base += NT * 8) {
      int4 tq[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = base + threadIdx.x + u * NT;
        tq[u] = __builtin_bit_cast(int4, __builtin_amdgcn_raw_buffer_load_b128(r_tab, (uint32_t)e * 16u, 0, 0));
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = base + threadIdx.x + u * NT;
        if (e < n4) t4[e] = tq[u];
      }
    }
  }
  __syncthreads();
  auto own_idx = [&](int P) { return tidx[(P - zb + 4) * ZNC + col]; };
  // (zb <= P < ze)
  auto halo_idx = [&](int P, int i) { return halo_ok(i) ? tidx[nown + (P - zb) * ZNH + col + ZNC * i] : -1; };

  // Rings with static slots (the z loop is unrolled by RQ, so every slot index is a
  // compile-time constant and values stay in the registers their loads wrote):
  //   ov[(P - zb + 4) % RQ]  own value of plane P,  P in z-4 .. z+OA (loaded OA planes ahead)
  //   hv[(P - zb) % HR]      halo values of plane P, P in z .. z+HD (loaded HD planes ahead)
  //   vr/kr/ar[(P - zb) % AR] output point data,     P in z .. z+AD (loaded AD planes ahead)
  // The load counter retires in order, so the shortest distance sets how many planes of
  // loads stay in flight: halo and point data go as deep as the own values allow.
  using Rings = ZRings<SHALLOW ? EPI_TAYLOR : EPI>;
  constexpr int RQ = Rings::ring, OA = Rings::own;
  constexpr int HD = Rings::halo, AD = Rings::aux;
  constexpr int HR = ring_above(RQ, HD);  // ring sizes dividing RQ
  constexpr int AR = ring_above(RQ, AD);
  static_assert(OA >= 5 && RQ >= OA + 5 && HR > HD && AR > AD, "ring sizes");
  Vec<LPT> ov[RQ];
  Vec<LPT> hv[HR][ZHU];
  double vr[AR] = {};
  Vec<LPT> kr[AR];
  Vec<LPT> ar[AR];
  Vec<LPT> ra[TAY ? AR : 1], rb[TAY ? AR : 1];
  // Every global load below is issued unconditionally (out-of-range requests read zeros):
  // with loads behind branches the compiler cannot count the loads still in flight and
  // drains the whole prefetch queue at every plane.
  // per-point data of plane P (< ze + 4) for its output point: potential, ket sum row,
  // accumulator row
  auto load_aux = [&](int slot, int P) {
    const int q = own_idx(P);
    const int qo = (P < ze && (unsigned)q < (unsigned)a_np) ? q : -1;
    vr[slot] = bld<1>(r_v, (uint32_t)qo << 3).v[0];
    if constexpr (NL) {
      const int ks = tks[(P - zb + 4) * ZNC + col];
      kr[slot] = bld<LPT>(r_k, roff(qo >= 0 ? ks : -1));
    }
    if constexpr (TAY) {
      ar[slot] = bld<LPT>(r_ai, roff(qo));
      ra[slot] = bld<LPT>(r_ya, roff(qo));  // zeros (no memory access) without ya / yb
      rb[slot] = bld<LPT>(r_yb, roff(qo));
    }
  };
#pragma unroll
  for (int k = 0; k < OA + 4; ++k) ov[k] = val(zb - 4 + k < ze + 4 ? own_idx(zb - 4 + k) : -1);
#pragma unroll
  for (int k = 0; k < HD; ++k)
#pragma unroll
    for (int i = 0; i < ZHU; ++i) hv[k][i] = val(zb + k < ze ? halo_idx(zb + k, i) : -1);
  // ket slots of the own sites of the output planes zb .. ze-1 (entries (P - zb + 4) ZNC +
  // col, as the own indices), looked up while the first planes' values are in flight (one
  // unconditional range-checked load per entry; -1 where the site has no point)
  if constexpr (NL) {
    const auto r_ks = zrsrc(a_kslot, (uint32_t)a_np * 4u);
    const int totk = (ze - zb + 4) * ZNC;
    for (int base = 4 * ZNC; base < totk; base += NT * 8) {
      int kq[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = min(base + (int)threadIdx.x + u * NT, totk - 1);
        const int q = tidx[e];
        const int k = (int)__builtin_amdgcn_raw_buffer_load_b32(r_ks, (uint32_t)q << 2, 0, 0);
        kq[u] = (unsigned)q < (unsigned)a_np ? k : -1;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = base + threadIdx.x + u * NT;
        if (e < totk) tks[e] = kq[u];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < AD; ++k) load_aux(k, zb + k);

  for (int z0 = zb; z0 < ze; z0 += RQ) {
#pragma unroll
    for (int kk = 0; kk < RQ; ++kk) {
      const int z = z0 + kk;
      if (z >= ze) break;
      const int buf = kk & 1;
      // prefetch: output-point data of plane z+AD, own column z+OA, halo of plane z+HD
      load_aux((kk + AD) % AR, z + AD);
      ov[(kk + 4 + OA) % RQ] = val(z + OA < ze + 4 ? own_idx(z + OA) : -1);
#pragma unroll
      for (int i = 0; i < ZHU; ++i) hv[(kk + HD) % HR][i] = val(z + HD < ze ? halo_idx(z + HD, i) : -1);
      // stage plane z: tile centre and halo strips
      lds_st(buf, own_lds, ov[(kk + 4) % RQ]);
#pragma unroll
      for (int i = 0; i < ZHU; ++i)
        if (halo_ok(i)) lds_st(buf, hlds[i], hv[kk % HR][i]);
      __syncthreads();
      // output point (x0+cx, y0+cy, z)
      const int q0 = own_idx(z);
      if ((unsigned)q0 < (unsigned)a_np) {
        Vec<LPT> s;
#pragma unroll
        for (int k = 0; k < LPT; ++k) s.v[k] = 0.0;
#pragma unroll
        for (int j = 0; j < 25; ++j) {
          Vec<LPT> x;
          if (j < 9) {
            const int dx = j - 4;
            x = (dx == 0) ? ov[(kk + 4) % RQ] : lds_ld(buf, xnb(dx));
          } else if (j < 17) {
            const int dy = (j < 13) ? j - 13 : j - 12;
            x = lds_ld(buf, ynb(dy));
          } else {
            const int dz = (j < 21) ? j - 21 : j - 20;
            x = ov[(kk + 4 + dz + RQ) % RQ];
          }
#pragma unroll
          for (int k = 0; k < LPT; ++k) s.v[k] = fma(a_w[j], x.v[k], s.v[k]);
        }
        const Vec<LPT> x = ov[(kk + 4) % RQ];
        if (a_v) {
          const double vv = vr[kk % AR];
#pragma unroll
          for (int k = 0; k < LPT; ++k) s.v[k] = fma(vv, x.v[k], s.v[k]);
        }
        if constexpr (NL) {
#pragma unroll
          for (int k = 0; k < LPT; ++k) s.v[k] = s.v[k] + kr[kk % AR].v[k];
        }
        if constexpr (EPI == EPI_STORE) {
          bst<LPT>(r_fo, roff(q0), s);
        } else {
          bst<LPT>(r_y, roff(q0), s);  // dropped when there is no yout
          if (has_acc) {
            Vec<LPT> av = ar[kk % AR];
            const Vec<LPT> ua = ra[kk % AR], ub = rb[kk % AR];
            if constexpr (LPT == 2) {
              av.v[0] = av.v[0] + zmul_lane(a_car, a_cai, ua.v[0], ua.v[1], false);
              av.v[1] = av.v[1] + zmul_lane(a_car, a_cai, ua.v[1], ua.v[0], true);
              av.v[0] = av.v[0] + zmul_lane(a_cbr, a_cbi, ub.v[0], ub.v[1], false);
              av.v[1] = av.v[1] + zmul_lane(a_cbr, a_cbi, ub.v[1], ub.v[0], true);
              if (a_has_c1) {
                av.v[0] = av.v[0] + zmul_lane(a_c1r, a_c1i, x.v[0], x.v[1], false);
                av.v[1] = av.v[1] + zmul_lane(a_c1r, a_c1i, x.v[1], x.v[0], true);
              }
              av.v[0] = av.v[0] + zmul_lane(a_c2r, a_c2i, s.v[0], s.v[1], false);
              av.v[1] = av.v[1] + zmul_lane(a_c2r, a_c2i, s.v[1], s.v[0], true);
            } else {
              // the partner lane of a (re, im) pair is held by the neighbouring thread
              const bool odd = (l & 1) != 0;
              av.v[0] = av.v[0] + zmul_lane(a_car, a_cai, ua.v[0], __shfl_xor(ua.v[0], 1), odd);
              av.v[0] = av.v[0] + zmul_lane(a_cbr, a_cbi, ub.v[0], __shfl_xor(ub.v[0], 1), odd);
              if (a_has_c1) av.v[0] = av.v[0] + zmul_lane(a_c1r, a_c1i, x.v[0], __shfl_xor(x.v[0], 1), odd);
              av.v[0] = av.v[0] + zmul_lane(a_c2r, a_c2i, s.v[0], __shfl_xor(s.v[0], 1), odd);
            }
            bst<LPT>(r_acc, roff(q0), av);
            if constexpr (EPI == EPI_TAYLOR_X) {
              Vec<LPT> dv = av;
              if (A.acc2) {
                // psib2: the same terms with the deltat2 factors, from the same base
                Vec<LPT> a2 = ar[kk % AR];
                if constexpr (LPT == 2) {
                  a2.v[0] = a2.v[0] + zmul_lane(A.dar, A.dai, ua.v[0], ua.v[1], false);
                  a2.v[1] = a2.v[1] + zmul_lane(A.dar, A.dai, ua.v[1], ua.v[0], true);
                  a2.v[0] = a2.v[0] + zmul_lane(A.dbr, A.dbi, ub.v[0], ub.v[1], false);
                  a2.v[1] = a2.v[1] + zmul_lane(A.dbr, A.dbi, ub.v[1], ub.v[0], true);
                  if (a_has_c1) {
                    a2.v[0] = a2.v[0] + zmul_lane(A.d1r, A.d1i, x.v[0], x.v[1], false);
                    a2.v[1] = a2.v[1] + zmul_lane(A.d1r, A.d1i, x.v[1], x.v[0], true);
                  }
                  a2.v[0] = a2.v[0] + zmul_lane(A.d2r, A.d2i, s.v[0], s.v[1], false);
                  a2.v[1] = a2.v[1] + zmul_lane(A.d2r, A.d2i, s.v[1], s.v[0], true);
                } else {
                  const bool odd = (l & 1) != 0;
                  a2.v[0] = a2.v[0] + zmul_lane(A.dar, A.dai, ua.v[0], __shfl_xor(ua.v[0], 1), odd);
                  a2.v[0] = a2.v[0] + zmul_lane(A.dbr, A.dbi, ub.v[0], __shfl_xor(ub.v[0], 1), odd);
                  if (a_has_c1) a2.v[0] = a2.v[0] + zmul_lane(A.d1r, A.d1i, x.v[0], __shfl_xor(x.v[0], 1), odd);
                  a2.v[0] = a2.v[0] + zmul_lane(A.d2r, A.d2i, s.v[0], __shfl_xor(s.v[0], 1), odd);
                }
                bst<LPT>(r_a2, roff(q0), a2);
                dv = a2;
              }
              if (a_dpart) {
                // |value|^2 of this thread's lanes times its state's weight, summed over
                // the PAIRS consecutive threads of the column (one lane chunk)
                double t = dv.v[0] * dv.v[0];
                if constexpr (LPT == 2) t = t + dv.v[1] * dv.v[1];
                t = wl * t;
#pragma unroll
                for (int off = PAIRS >> 1; off > 0; off >>= 1) t += __shfl_xor(t, off);
                if (kp == 0) a_dpart[(size_t)chunk * a_np + q0] = t;
              }
            }
          }
        }
      }
    }
  }
}

template <int LPT, int LC, int EPI, bool NL>
__global__ __launch_bounds__(ZNC * LC / LPT) void k_zmarch(const ZArgs za) {
  const int2 sc = za.sched[blockIdx.x];
  if (sc.x < 0) return;              // padding entry of the schedule (whole workgroup)
  zmarch_body<LPT, LC, EPI, NL>(za, za.items[sc.x], sc.x, sc.y);
}

// Several operators (systems) and batches in one launch: sched[b] = {descriptor, item,
// lane chunk, -}; descriptors zs[] are ZArgs in device memory (their sched is unused).
template <int LPT, int LC, int EPI, bool NL>
__global__ __launch_bounds__(ZNC * LC / LPT) void k_zmarch_multi(const ZArgs* __restrict__ zs,
                                                                const int4* __restrict__ sched) {
  const int4 sc = sched[blockIdx.x];
  if (sc.x < 0) return;
  const ZArgs& za = zs[sc.x];
  zmarch_body<LPT, LC, EPI, NL, true>(za, za.items[sc.y], sc.y, sc.z);
}

// dynamic LDS bytes of k_zmarch for a lane chunk and segment length: two plane buffers,
// the item's index table and the ket slots of its own sites
inline size_t zmarch_shmem(int lc, int zseg) {
  return sizeof(double) * 2 * ZS * lc + sizeof(int) * ((size_t)zmarch_tab_ints(zseg) + (size_t)(zseg + 8) * ZNC);
}

}  // namespace octhip
