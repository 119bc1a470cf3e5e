// Batched launch over several independent (system, batch) pairs of uneven size.
//
// The reference runs the Taylor propagator batch by batch and system by system
// (src/td/propagation_ops_elec.F90:232-249 over batches, one propagation per system),
// each H-apply being ~5 kernel launches.  Here every kernel kind of one Taylor order runs
// once for all pairs: the boundary clear, the projector bra, its reduction, the ket and
// the multi-sphere sums go through multi-segment launches (a block finds its pair from a
// block-offset table), and the fused stencil + local + ket + Taylor-epilogue pass is one
// z-march launch over a (pair, tile segment, lane chunk) work list, cut into 8
// contiguous parts for the 8 XCDs.  A Taylor step is then 1 + 5 * order launches
// whatever the number of systems and batches.  Every pair has its own work buffers;
// results are bit-identical to per-pair octhip_exponential_taylor_batch /
// octhip_hamiltonian_apply_batch calls (same kernel bodies on the same data).
#include <algorithm>
#include <complex>
#include <cstring>
#include <vector>

#include "kernels.h"
#include "octhip_internal.h"
#include "zmarch.h"

namespace octhip {
void h_apply_args(const octhip_hamiltonian_s* h, const double* in, int L, double* ksum, ApplyArgs& a);
int zmarch_items(octhip_nl_operator_s* op, int zs);
int zmarch_zseg(const octhip_nl_operator_s* op, int nchunk, int lc, int resident, bool alone);
int zmarch_reg_blocks(int lc, int lpt, int epi, bool nl, bool multi);
bool zmarch_fits(const octhip_nl_operator_s* op, const ApplyArgs& a);
int launch_zmarch_multi(const ZArgs* zs, const int4* sched, int len, int lc, int lpt, int epi, bool nl, int zseg,
                        hipStream_t s);
}  // namespace octhip

using namespace octhip;

namespace {

struct Pair {
  octhip_hamiltonian_s* h = nullptr;
  double* psi = nullptr;
  double* hpsi = nullptr;
  int L = 0;
  double* partial = nullptr;
  double* proj = nullptr;
  double* ksum = nullptr;
  double* y[3] = {nullptr, nullptr, nullptr};  // Taylor powers (three for order 4)
  // the operator's z-march items and their index tables, one plan per epilogue ([0] plain
  // store, [1] Taylor): the two instances keep different numbers of workgroups resident
  const int4* items[2] = {nullptr, nullptr};
  const int* tab[2] = {nullptr, nullptr};
  const int* tab_off[2] = {nullptr, nullptr};
  int nitems[2] = {0, 0};
};

// a device table of argument blocks plus block offsets for one multi-segment kernel
struct Seg {
  void* args = nullptr;
  int* off = nullptr;
  int nseg = 0, nblocks = 0;
};

template <typename T>
int make_seg(const std::vector<T>& args, const std::vector<int>& blocks, Seg& out) {
  std::vector<int> off(blocks.size() + 1, 0);
  for (size_t i = 0; i < blocks.size(); ++i) off[i + 1] = off[i] + blocks[i];
  out.nseg = (int)blocks.size();
  out.nblocks = off.back();
  T* d = nullptr;
  int rc = upload(&d, args.data(), args.size(), nullptr);
  if (rc) return rc;
  out.args = d;
  return upload(&out.off, off.data(), off.size(), nullptr);
}

void free_seg(Seg& s) {
  dfree(s.args);
  dfree(s.off);
  s = Seg();
}

}  // namespace

struct octhip_batched_s {
  std::vector<Pair> pairs;
  int order = 4;
  double dt = 0.0;
  bool mfma = true, nl = false;
  int lc = 16, lpt = 2;
  int zseg[2] = {8, 8};                 // per plan: longest segment (LDS size)
  std::complex<double> zf[33];
  Seg zero, reduce, ket, ksum;
  std::vector<Seg> bra;                 // [0] H-apply input psi, [k] Taylor order k input
  std::vector<ZArgs*> zdesc;            // [0] H-apply (store to hpsi), [k] Taylor order k
  int4* sched[2] = {nullptr, nullptr};  // per plan: blockIdx -> (pair, item, lane chunk)
  int sched_len[2] = {0, 0};
};

namespace {

void destroy(octhip_batched_s* b) {
  for (auto& p : b->pairs) {
    dfree(p.partial);
    dfree(p.proj);
    dfree(p.ksum);
    dfree(p.y[0]);
    dfree(p.y[1]);
    dfree(p.y[2]);
  }
  free_seg(b->zero);
  free_seg(b->reduce);
  free_seg(b->ket);
  free_seg(b->ksum);
  for (auto& s : b->bra) free_seg(s);
  for (auto* z : b->zdesc) dfree(z);
  dfree(b->sched[0]);
  dfree(b->sched[1]);
  delete b;
}

int blocks_for(size_t total, int cap) {
  return (int)std::min<size_t>((total + BLOCK - 1) / BLOCK, (size_t)cap);
}

// bra argument blocks of every pair for input buffers in[p]
int build_bra(octhip_batched_s* b, const std::vector<const double*>& in, Seg& out) {
  std::vector<BraMfmaArgs> args;
  std::vector<int> blocks;
  for (size_t i = 0; i < b->pairs.size(); ++i) {
    const Pair& p = b->pairs[i];
    const octhip_hamiltonian_s* h = p.h;
    BraMfmaArgs m;
    std::memset(&m, 0, sizeof(m));
    m.b.psi = in[i];
    m.b.L = p.L;
    m.b.chunks = h->d_chunks;
    m.b.mat_info = h->d_mat_info;
    m.b.mats = h->d_matrices;
    m.b.maps = h->d_maps;
    m.b.partial = p.partial;
    m.b.pjmax = h->max_nprojs;
    m.pt = h->d_pt;
    m.pjs = h->pjs;
    if (b->nl && !(m.b.psi && m.b.partial && m.b.chunks && m.b.maps && m.pt))
      return fail(OCTHIP_ERR_STATE, "batched: incomplete bra arguments for pair %zu", i);
    args.push_back(m);
    // MFMA bra: one block per (chunk, lane tile of 32, or of 16 when 32 does not divide
    // the row); plain bra: one per chunk
    blocks.push_back(h->nmat > 0 ? h->nchunk * (b->mfma ? mfma_lane_tiles(p.L) : 1) : 0);
  }
  return make_seg(args, blocks, out);
}

// input of pass k (0 = the H-apply): y_{k-1} (y_0 = psi); orders 2..4 keep y_j in y[j-1],
// higher orders ping-pong two buffers
const double* pass_input(const Pair& p, int k, int order) {
  if (k <= 1) return p.psi;
  return taylor_deferred(order) ? p.y[k - 2] : p.y[(k - 2) & 1];
}

// epilogue of pass k (0 = the H-apply), which is also its plan index
int pass_epi(const octhip_batched_s* b, int k) {
  if (k == 0) return EPI_STORE;
  return (taylor_deferred(b->order) && k < b->order) ? EPI_STORE : EPI_TAYLOR;
}

// z-march descriptors of every pair for one pass
int build_zdesc(octhip_batched_s* b, int k, ZArgs** out) {
  std::vector<ZArgs> zs(b->pairs.size());
  for (size_t i = 0; i < b->pairs.size(); ++i) {
    const Pair& p = b->pairs[i];
    const octhip_nl_operator_s* op = p.h->kin;
    ZArgs& z = zs[i];
    std::memset(&z, 0, sizeof(z));
    const double* in = pass_input(p, k, b->order);
    h_apply_args(p.h, in, p.L, p.ksum, z.a);
    const bool deferred = taylor_deferred(b->order);
    if (k == 0) {
      z.a.fo = p.hpsi;
    } else if (deferred && k < b->order) {
      z.a.fo = p.y[k - 1];  // y_k, stored (no accumulator in between)
    } else if (deferred) {
      // last order: psi + z_1 y_1 + ... + z_order H y_{order-1}, terms added in order k
      z.a.acc = p.psi;
      if (k >= 3) {
        z.a.ya = p.y[0];
        z.a.car = b->zf[1].real();
        z.a.cai = b->zf[1].imag();
      }
      if (k >= 4) {
        z.a.yb = p.y[1];
        z.a.cbr = b->zf[2].real();
        z.a.cbi = b->zf[2].imag();
      }
      z.a.has_c1 = 1;
      z.a.c1r = b->zf[k - 1].real();
      z.a.c1i = b->zf[k - 1].imag();
      z.a.c2r = b->zf[k].real();
      z.a.c2i = b->zf[k].imag();
    } else {
      z.a.yout = (k < b->order) ? p.y[(k - 1) & 1] : nullptr;
      if (k == 1) {
        z.a.acc = nullptr;
      } else {
        z.a.acc = p.psi;
        z.a.has_c1 = (k == 2);
        z.a.c1r = b->zf[1].real();
        z.a.c1i = b->zf[1].imag();
        z.a.c2r = b->zf[k].real();
        z.a.c2i = b->zf[k].imag();
      }
    }
    const int plan = pass_epi(b, k);
    z.items = p.items[plan];
    z.tab = p.tab[plan];
    z.tab_off = p.tab_off[plan];
    z.sched = nullptr;
    // every pointer the kernel dereferences must exist
    // (the H-apply descriptor may lack fo: octhip_batched_apply refuses to run without hpsi)
    const bool ok = z.a.fi && z.items && z.tab && z.tab_off && (!z.a.kslot || z.a.ksum) &&
                    (deferred ? (k == 0 || (k < b->order ? z.a.fo != nullptr : z.a.acc != nullptr))
                              : ((k < 2 || z.a.acc) && (k == 0 || k == b->order || z.a.yout)));
    if (!ok) return fail(OCTHIP_ERR_STATE, "batched: incomplete descriptor for pair %zu, pass %d", i, k);
    if (!zmarch_fits(op, z.a))
      return fail(OCTHIP_ERR_INVALID, "batched: pair %zu is too large for 32-bit row offsets (%d rows of %d doubles)",
                  i, op->np_part, p.L);
  }
  return upload(out, zs.data(), zs.size(), nullptr);
}

int run_bra(octhip_batched_s* b, const Seg& bra, hipStream_t s) {
  if (!b->nl) return OCTHIP_SUCCESS;
  int rc = launch_multi(b->mfma ? MULTI_BRA_MFMA : MULTI_BRA_PLAIN, bra.args, bra.off, bra.nseg, bra.nblocks, s);
  if (!rc) rc = launch_multi(MULTI_BRA_REDUCE, b->reduce.args, b->reduce.off, b->reduce.nseg, b->reduce.nblocks, s);
  if (!rc) rc = launch_multi(b->mfma ? MULTI_KET_MFMA : MULTI_KET_PLAIN, b->ket.args, b->ket.off, b->ket.nseg,
                             b->ket.nblocks, s);
  if (!rc) rc = launch_multi(MULTI_KET_SUM, b->ksum.args, b->ksum.off, b->ksum.nseg, b->ksum.nblocks, s);
  return rc;
}

int run_pass(octhip_batched_s* b, int k, hipStream_t s) {
  int rc = run_bra(b, b->bra[k], s);
  if (rc) return rc;
  const int epi = pass_epi(b, k);
  return launch_zmarch_multi(b->zdesc[k], b->sched[epi], b->sched_len[epi], b->lc, b->lpt, epi, b->nl,
                             b->zseg[epi], s);
}

}  // namespace

extern "C" int octhip_batched_create(int npairs, const octhip_hamiltonian_t* hms, double* const* psi,
                                     double* const* hpsi, const int* ldf, double deltat, int order,
                                     octhip_batched_t* out) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(out && hms && psi && ldf && npairs > 0, "bad argument");
  *out = nullptr;
  OCTHIP_REQUIRE(order >= 2 && order <= 32, "batched Taylor needs order in 2..32 (got %d)", order);
  int minL = 1 << 30;
  bool any_nl = false, all_nl = true;
  for (int i = 0; i < npairs; ++i) {
    OCTHIP_REQUIRE(hms[i] && psi[i], "pair %d: null handle or batch", i);
    if (is_periodic(hms[i]))
      return fail(OCTHIP_ERR_UNSUPPORTED, "pair %d: the batched launch covers finite systems (zero boundary)", i);
    OCTHIP_REQUIRE(ldf[i] >= 1 && ldf[i] <= MAX_LDF, "pair %d: ldf %d outside 1..%d (complex batches)", i, ldf[i],
                   MAX_LDF);
    const octhip_nl_operator_s* op = hms[i]->kin;
    OCTHIP_REQUIRE(op->zmarch_ok && op->n == 25,
                   "pair %d: the batched launch needs the order-4 star on a lattice (z-march)", i);
    minL = std::min(minL, 1 << ldf[i]);
    const bool nl = hms[i]->nmat > 0 && hms[i]->nslot > 0;
    any_nl = any_nl || nl;
    all_nl = all_nl && nl;
  }
  OCTHIP_REQUIRE(any_nl == all_nl, "the batched launch needs every system with projectors or none");
  auto* b = new octhip_batched_s();
  b->order = order;
  b->dt = deltat;
  b->nl = all_nl;
  b->lc = std::min(std::min(16, 2048 / ZNC), minL);  // lanes per workgroup: <= 1024 threads
  b->lpt = 2;
  b->mfma = minL >= 16;
  for (int i = 0; i < npairs; ++i) b->mfma = b->mfma && hms[i]->use_mfma;
  taylor_coefs(deltat, 0, order, b->zf);
  const int resident[2] = {zmarch_reg_blocks(b->lc, b->lpt, EPI_STORE, b->nl, true),
                           zmarch_reg_blocks(b->lc, b->lpt, EPI_TAYLOR, b->nl, true)};
  int rc = OCTHIP_SUCCESS;
  // pairs and their work buffers
  for (int i = 0; i < npairs && !rc; ++i) {
    Pair p;
    p.h = hms[i];
    p.psi = psi[i];
    p.hpsi = hpsi ? hpsi[i] : nullptr;
    p.L = 1 << ldf[i];
    const octhip_hamiltonian_s* h = p.h;
    const size_t L = (size_t)p.L;
    if (b->nl) {
      rc = dalloc((void**)&p.partial, sizeof(double) * h->nchunk * h->max_nprojs * L);
      if (!rc) rc = dalloc((void**)&p.proj, sizeof(double) * h->full_projection_size * L);
      if (!rc) rc = dalloc((void**)&p.ksum, sizeof(double) * ((size_t)h->total_points + h->nmulti) * L);
    }
    if (!rc) rc = dalloc((void**)&p.y[0], sizeof(double) * h->np * L);
    if (!rc) rc = dalloc((void**)&p.y[1], sizeof(double) * h->np * L);
    if (!rc && taylor_deferred(order) && order == 4) rc = dalloc((void**)&p.y[2], sizeof(double) * h->np * L);
    // one plan per epilogue, segments for the residency of that instance
    for (int plan = 0; plan < 2 && !rc; ++plan) {
      rc = zmarch_items(h->kin, zmarch_zseg(h->kin, p.L / b->lc, b->lc, resident[plan], false));
      if (!rc) {
        p.items[plan] = h->kin->d_items;
        p.tab[plan] = h->kin->d_tab;
        p.tab_off[plan] = h->kin->d_tab_off;
        p.nitems[plan] = h->kin->nitems;
        b->zseg[plan] = std::max(b->zseg[plan], h->kin->zseg_built);
      }
    }
    b->pairs.push_back(p);
  }
  // boundary clear, reduce, ket, ket sums
  if (!rc) {
    std::vector<ZeroArgs> za;
    std::vector<BraReduceArgs> ra;
    std::vector<KetArgs> ka;
    std::vector<int> zb, rb, kb, sb;
    for (const Pair& p : b->pairs) {
      const octhip_hamiltonian_s* h = p.h;
      const int np = h->kin->np, np_part = h->kin->np_part;
      za.push_back(ZeroArgs{p.psi, (size_t)np * p.L, (size_t)(np_part - np) * p.L});
      zb.push_back(blocks_for((size_t)(np_part - np) * p.L, 8192));
      BraReduceArgs r;
      std::memset(&r, 0, sizeof(r));
      r.partial = p.partial;
      r.proj_mat = h->d_proj_mat;
      r.mat_pbase = h->d_mat_pbase;
      r.mat_chunk0 = h->d_mat_chunk0;
      r.scals = h->d_scals;
      r.proj = p.proj;
      r.L = p.L;
      r.pjmax = h->max_nprojs;
      r.nproj = b->nl ? h->full_projection_size : 0;
      ra.push_back(r);
      rb.push_back(blocks_for((size_t)r.nproj * p.L, 4096));
      KetArgs k;
      std::memset(&k, 0, sizeof(k));
      k.chunks = h->d_chunks;
      k.mat_info = h->d_mat_info;
      k.mat_pbase = h->d_mat_pbase;
      k.pt = h->d_pt;
      k.pjs = h->pjs;
      k.proj = p.proj;
      k.kbuf = p.ksum;
      k.L = p.L;
      k.multi_pos = h->d_multi_pos;
      k.multi_g = h->d_multi_g;
      k.nmulti = b->nl ? h->nmulti : 0;
      k.total_points = h->total_points;
      ka.push_back(k);
      kb.push_back(b->nl ? h->nchunk * (b->mfma ? mfma_lane_tiles(p.L) : 1) : 0);
      sb.push_back(blocks_for((size_t)k.nmulti * ((p.L & 1) ? p.L : p.L / 2), 8192));
    }
    for (size_t i = 0; i < b->pairs.size() && b->nl && !rc; ++i)
      if (!(ra[i].partial && ra[i].proj && ka[i].proj && ka[i].kbuf))
        rc = fail(OCTHIP_ERR_STATE, "batched: missing work buffer for pair %zu", i);
    if (!rc) rc = make_seg(za, zb, b->zero);
    if (!rc) rc = make_seg(ra, rb, b->reduce);
    if (!rc) rc = make_seg(ka, kb, b->ket);
    if (!rc) rc = make_seg(ka, sb, b->ksum);
  }
  // bra tables and z-march descriptors: [0] H-apply, [k] Taylor order k
  for (int k = 0; k <= order && !rc; ++k) {
    std::vector<const double*> in;
    for (const Pair& p : b->pairs) in.push_back(pass_input(p, k, order));
    b->bra.emplace_back();
    rc = build_bra(b, in, b->bra.back());
    ZArgs* zd = nullptr;
    if (!rc) rc = build_zdesc(b, k, &zd);
    b->zdesc.push_back(zd);
  }
  // (pair, item, lane chunk) work list of each plan, cut into 8 contiguous parts dealt to
  // the 8 XCDs
  for (int plan = 0; plan < 2 && !rc; ++plan) {
    std::vector<int4> all;
    for (size_t i = 0; i < b->pairs.size(); ++i) {
      const int nchunk = b->pairs[i].L / b->lc;
      for (int c = 0; c < nchunk; ++c)
        for (int it = 0; it < b->pairs[i].nitems[plan]; ++it) all.push_back(make_int4((int)i, it, c, 0));
    }
    const int n = (int)all.size();
    const int per = (n + 7) / 8;
    std::vector<int4> sc((size_t)per * 8, make_int4(-1, 0, 0, 0));
    for (int k = 0; k < 8; ++k)
      for (int i = 0; i < per; ++i) {
        const int g = k * per + i;
        if (g >= n) break;
        sc[(size_t)i * 8 + k] = all[g];
      }
    rc = upload(&b->sched[plan], sc.data(), sc.size(), nullptr);
    b->sched_len[plan] = (int)sc.size();
  }
  if (!rc) {
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) rc = fail(OCTHIP_ERR_HIP, "batched_create: %s", hipGetErrorString(e));
  }
  if (rc) {
    destroy(b);
    return rc;
  }
  *out = b;
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_batched_apply(octhip_batched_t b, void* stream) {
  OCTHIP_RANGE("octhip_batched_apply");
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(b != nullptr, "null handle");
  for (const Pair& p : b->pairs) OCTHIP_REQUIRE(p.hpsi != nullptr, "batched_apply: created without hpsi");
  hipStream_t s = (hipStream_t)stream;
  int rc = launch_multi(MULTI_ZERO, b->zero.args, b->zero.off, b->zero.nseg, b->zero.nblocks, s);
  if (rc) return rc;
  return run_pass(b, 0, s);
}

extern "C" int octhip_batched_taylor_step(octhip_batched_t b, void* stream) {
  OCTHIP_RANGE("octhip_batched_taylor_step");
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(b != nullptr, "null handle");
  hipStream_t s = (hipStream_t)stream;
  int rc = launch_multi(MULTI_ZERO, b->zero.args, b->zero.off, b->zero.nseg, b->zero.nblocks, s);
  for (int k = 1; k <= b->order && !rc; ++k) rc = run_pass(b, k, s);
  return rc;
}

extern "C" int octhip_batched_graph_create(octhip_batched_t b, void* stream, octhip_taylor_graph_t* out) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(b && out, "null argument");
  *out = nullptr;
  return capture_graph((hipStream_t)stream, nullptr, [&](hipStream_t cs) { return octhip_batched_taylor_step(b, cs); }, out);
}

extern "C" int octhip_batched_destroy(octhip_batched_t b) {
  if (!b) return OCTHIP_SUCCESS;
  (void)hipDeviceSynchronize();
  destroy(b);
  return OCTHIP_SUCCESS;
}
