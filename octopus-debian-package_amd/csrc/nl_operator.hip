// nl_operator handle: upload of the run-compressed stencil, lattice rebuild, operate.
#include <algorithm>
#include <climits>
#include <cstring>
#include <deque>

#include <map>

#include "zmarch.h"

using namespace octhip;

namespace {

// Rebuild lattice coordinates from the run-compressed neighbour table: start at point 0,
// walk every stencil edge p -> p + ri[run(p)][j] assigning coords(p) + points[j], and
// reject the table if any point would receive two different coordinates.
bool rebuild_lattice(int n, const int* pts, const int* ri, const std::vector<int>& rimap, int np,
                     int np_part, std::vector<int>& coords) {
  coords.assign(3 * (size_t)np_part, INT_MIN);
  std::vector<char> seen(np_part, 0);
  std::deque<int> work;
  int visited = 0;
  for (int start = 0; start < np; ++start) {
    if (seen[start]) continue;
    if (start != 0) return false;  // disconnected inner set: no single lattice
    seen[start] = 1;
    coords[0] = coords[1] = coords[2] = 0;
    work.push_back(start);
    while (!work.empty()) {
      const int p = work.front();
      work.pop_front();
      ++visited;
      const int* row = ri + (size_t)rimap[p] * n;
      for (int j = 0; j < n; ++j) {
        const long qq = (long)p + row[j];
        if (qq < 0 || qq >= np_part) return false;
        const int q = (int)qq;
        int c[3];
        for (int d = 0; d < 3; ++d) c[d] = coords[3 * (size_t)p + d] + pts[3 * j + d];
        if (coords[3 * (size_t)q] == INT_MIN) {
          for (int d = 0; d < 3; ++d) coords[3 * (size_t)q + d] = c[d];
        } else {
          for (int d = 0; d < 3; ++d)
            if (coords[3 * (size_t)q + d] != c[d]) return false;
        }
        if (q < np && !seen[q]) {
          seen[q] = 1;
          work.push_back(q);
        }
      }
    }
  }
  if (visited != np) return false;
  for (int q = np; q < np_part; ++q)  // boundary points never reached by any stencil
    if (coords[3 * (size_t)q] == INT_MIN) coords[3 * (size_t)q] = coords[3 * (size_t)q + 1] = coords[3 * (size_t)q + 2] = INT_MIN;
  return true;
}

int build_site_map(octhip_nl_operator_s* op, const int* pts, hipStream_t s) {
  const int np = op->np, np_part = op->np_part, n = op->n;
  const std::vector<int>& c = op->coords;
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  for (int q = 0; q < np_part; ++q) {
    if (c[3 * (size_t)q] == INT_MIN) continue;
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], c[3 * (size_t)q + d]);
      hi[d] = std::max(hi[d], c[3 * (size_t)q + d]);
    }
  }
  // pad by the stencil reach so every site + delta of an inner point is inside
  int reach = 0;
  for (int j = 0; j < n; ++j)
    for (int d = 0; d < 3; ++d) reach = std::max(reach, std::abs(pts[3 * j + d]));
  long sz[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] -= reach;
    hi[d] += reach;
    sz[d] = (long)hi[d] - lo[d] + 1;
  }
  const long nsite = sz[0] * sz[1] * sz[2];
  if (nsite > (long)INT_MAX / 2) return fail(OCTHIP_ERR_UNSUPPORTED, "lattice too large (%ld sites)", nsite);
  std::vector<int> cube((size_t)nsite, -1);
  std::vector<int> cidx(np);
  for (int q = 0; q < np_part; ++q) {
    if (c[3 * (size_t)q] == INT_MIN) continue;
    const long site = (c[3 * (size_t)q] - lo[0]) + sz[0] * ((c[3 * (size_t)q + 1] - lo[1]) + sz[1] * (long)(c[3 * (size_t)q + 2] - lo[2]));
    if (cube[site] != -1) return fail(OCTHIP_ERR_INVALID, "two points share lattice site");
    cube[site] = q;
    if (q < np) cidx[q] = (int)site;
  }
  for (int j = 0; j < n; ++j)
    op->dcube[j] = (int)(pts[3 * j] + sz[0] * (pts[3 * j + 1] + sz[1] * (long)pts[3 * j + 2]));
  op->ncube = (size_t)nsite;
  int rc = upload(&op->d_cube, cube.data(), cube.size(), s);
  if (!rc) rc = upload(&op->d_dcube, op->dcube, (size_t)n, s);
  if (rc) return rc;
  return upload(&op->d_cidx, cidx.data(), cidx.size(), s);
}

// The z-march kernel needs the order-4 star in Octopus's point order: x -4..4, then
// y -4..-1,1..4, then z -4..-1,1..4 (src/grid/stencil_star.F90:59-93).
bool is_star4(int n, const int* pts) {
  if (n != 25) return false;
  int j = 0;
  for (int d = 0; d < 3; ++d)
    for (int k = -4; k <= 4; ++k) {
      if (k == 0 && d > 0) continue;
      for (int e = 0; e < 3; ++e)
        if (pts[3 * j + e] != (e == d ? k : 0)) return false;
      ++j;
    }
  return true;
}

// March cube: the inner points' bounding box rounded up to whole ZTX x ZTY column tiles in x
// and y, padded by the stencil reach in x/y and by ZPADZ planes in z; site -> point index.
int build_march(octhip_nl_operator_s* op, hipStream_t s) {
  const int np = op->np, np_part = op->np_part;
  const std::vector<int>& c = op->coords;
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  for (int q = 0; q < np; ++q)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], c[3 * (size_t)q + d]);
      hi[d] = std::max(hi[d], c[3 * (size_t)q + d]);
    }
  op->ntx = (hi[0] - lo[0] + ZTX) / ZTX;
  op->nty = (hi[1] - lo[1] + ZTY) / ZTY;
  op->mx = op->ntx * ZTX + 2 * ZR;
  op->my = op->nty * ZTY + 2 * ZR;
  op->mz = (hi[2] - lo[2] + 1) + 2 * ZPADZ;
  const int ox = lo[0] - ZR, oy = lo[1] - ZR, oz = lo[2] - ZPADZ;
  const size_t nsite = (size_t)op->mx * op->my * op->mz;
  if (nsite > (size_t)INT_MAX / 2) return fail(OCTHIP_ERR_UNSUPPORTED, "march cube too large");
  std::vector<int> mc(nsite, -1);
  op->tile_zlo.assign((size_t)op->ntx * op->nty, INT_MAX);
  op->tile_zhi.assign((size_t)op->ntx * op->nty, -1);
  for (int q = 0; q < np_part; ++q) {
    const int* p = &c[3 * (size_t)q];
    if (p[0] == INT_MIN) continue;
    const long x = p[0] - ox, y = p[1] - oy, z = p[2] - oz;
    if (x < 0 || y < 0 || z < 0 || x >= op->mx || y >= op->my || z >= op->mz) continue;
    mc[(size_t)x + (size_t)op->mx * ((size_t)y + (size_t)op->my * z)] = q;
    if (q < np) {
      const int t = (int)((x - ZR) / ZTX + op->ntx * ((y - ZR) / ZTY));
      op->tile_zlo[t] = std::min(op->tile_zlo[t], (int)z);
      op->tile_zhi[t] = std::max(op->tile_zhi[t], (int)z);
    }
  }
  (void)s;
  op->mcube = std::move(mc);
  return OCTHIP_SUCCESS;
}

}  // namespace

namespace octhip {
int zmarch_reg_blocks(int lc, int lpt, int epi, bool nl, bool multi);

// Segment length of a launch with nchunk lane chunks of lc lanes: the operator's setting,
// or the candidate (48, 32, 28, 24, 22, 21, 20, 16, 12, 8 planes) with the smallest
// estimated time, in plane steps of one workgroup:
//   work  = nchunk * sum over items (planes + 8 overlap planes + ZM_PROLOGUE)
//   time  = max(work / (CUs * g), longest item * r / g) * (1 + ZM_TAIL * idle)
// where r = workgroups resident per CU (2 when the instance's registers allow it and the
// segment's LDS fits twice, else 1) and g = the throughput of r resident workgroups
// relative to one (ZM_TWO_GAIN for two).  The first term is the whole launch spread over
// the chip, the second its critical path: long segments cost little overlap but leave
// CUs idle on small meshes.  idle = the unused fraction of the last wave of workgroups on
// an XCD (items per XCD over its resident slots, rounded up, minus the unrounded value;
// counted when there is at least one full wave): near-equal items leave that many slots
// empty at the end.  Measured on MI355X: R=12 sphere, 32 complex states: 24 planes with
// two resident workgroups beat 48 with one by 7%, and 21-22 planes (3.7-3.8 waves per
// XCD) beat 24 (3.4) and 28 (3.1) by 1%; the 64^3 Laplacian with 8 lanes: 8 planes
// (24 us against 38 us at 32).
constexpr double ZM_PROLOGUE = 4.0;  // item prologue, in plane steps
constexpr double ZM_TWO_GAIN = 1.3;  // throughput of two resident workgroups against one
constexpr double ZM_TAIL = 0.25;     // cost of the idle slots of a partial last wave
// alone: the launch holds this operator's items only (a batched launch shares its grid
// with other systems, so the last-wave term of one system does not apply)
int zmarch_zseg(const octhip_nl_operator_s* op, int nchunk, int lc, int resident, bool alone) {
  if (op->zseg > 0) return op->zseg;
  const int rkey = resident | (alone ? 0 : 0x100);
  for (const auto& c : op->zseg_cache)
    if (c.x == nchunk && c.y == lc && c.z == rkey) return c.w;
  static int ncu = 0;
  if (ncu <= 0 && (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, current_device()) != hipSuccess ||
                   ncu <= 0))
    ncu = 256;
  const bool two_ok = resident >= 2;
  double best_t = 1e300;
  int best = 8;
  for (int zs : {48, 32, 28, 24, 22, 21, 20, 16, 12, 8}) {
    const size_t shm = zmarch_shmem(lc, zs);
    if (shm > 160 * 1024) continue;
    const int r = (two_ok && shm <= 80 * 1024) ? 2 : 1;
    const double g = r == 2 ? ZM_TWO_GAIN : 1.0;
    double work = 0.0, cmax = 0.0;
    long items = 0;
    for (size_t t = 0; t < op->tile_zlo.size(); ++t) {
      if (op->tile_zhi[t] < 0) continue;
      const int len = op->tile_zhi[t] - op->tile_zlo[t] + 1;
      const int n = (len + zs - 1) / zs;
      work += len + n * (8.0 + ZM_PROLOGUE);
      cmax = std::max(cmax, (len + n - 1) / n + 8.0 + ZM_PROLOGUE);
      items += n;
    }
    double tm = std::max(nchunk * work / (ncu * g), cmax * r / g);
    const double waves = std::ceil((double)items * nchunk / 8.0) / (std::max(ncu / 8, 1) * (double)r);
    if (alone && waves >= 1.0) tm *= 1.0 + ZM_TAIL * (std::ceil(waves) - waves);
    if (tm < best_t) {
      best_t = tm;
      best = zs;
    }
  }
  op->zseg_cache.push_back(make_int4(nchunk, lc, rkey, best));
  return best;
}

// lane chunk and lanes per thread of a z-march launch over rows of L doubles
void zmarch_shape(const octhip_nl_operator_s* op, int L, int* lc, int* lpt) {
  int c = op->lane_chunk > 0 ? op->lane_chunk : 16;
  int t = op->lanes_per_thread > 0 ? op->lanes_per_thread : 2;
  while (c > 1 && ZNC * c / 2 > 1024) c /= 2;  // 1024 threads per workgroup at most
  if (c > L) c = L;
  if (t > c) t = c;
  if (ZNC * c / t > 1024) t = 2;
  *lc = c;
  *lpt = t;
}

// segment length of a launch of this epilogue (with or without ket rows)
int zmarch_zseg_for(const octhip_nl_operator_s* op, int L, int epi, bool nl) {
  int lc, lpt;
  zmarch_shape(op, L, &lc, &lpt);
  return zmarch_zseg(op, L / lc, lc, zmarch_reg_blocks(lc, lpt, epi, nl, false), true);
}

// (tile, z-segment) work items for segment length zs; every list built stays allocated
// until the operator is destroyed (graphs and batched launches may point at it)
int zmarch_items(octhip_nl_operator_s* op, int zs) {
  if (op->zseg_built == zs && op->d_items) return OCTHIP_SUCCESS;
  // save the current schedules with their plan, then switch
  for (auto& pl : op->plans)
    if (pl.zseg == op->zseg_built) {
      pl.sched = op->sched;
      pl.sched_len = op->sched_len;
    }
  for (auto& pl : op->plans)
    if (pl.zseg == zs) {
      op->d_items = pl.d_items;
      op->d_tab = pl.d_tab;
      op->d_tab_off = pl.d_tab_off;
      op->items = pl.items;
      op->nitems = (int)pl.items.size();
      op->sched = pl.sched;
      op->sched_len = pl.sched_len;
      op->zseg_built = zs;
      return OCTHIP_SUCCESS;
    }
  std::vector<int4> items;
  // A tile's z range of len planes is cut into ceil(len / zs) segments of equal length
  // (+-1), so no item is a short remainder that pays a whole prologue and 8 overlap
  // planes for a few outputs.  Segment-major, then tile rows: concurrently running items
  // share halo rows in L2.
  int maxseg = 0;
  for (size_t t = 0; t < op->tile_zlo.size(); ++t)
    if (op->tile_zhi[t] >= 0) maxseg = std::max(maxseg, (op->tile_zhi[t] - op->tile_zlo[t] + zs) / zs);
  for (int sg = 0; sg < maxseg; ++sg)
    for (int ty = 0; ty < op->nty; ++ty)
      for (int tx = 0; tx < op->ntx; ++tx) {
        const int t = tx + op->ntx * ty;
        if (op->tile_zhi[t] < 0) continue;
        const int len = op->tile_zhi[t] - op->tile_zlo[t] + 1;
        const int n = (len + zs - 1) / zs;
        if (sg >= n) continue;
        const int base = len / n, rem = len % n;
        const int zb = op->tile_zlo[t] + sg * base + std::min(sg, rem);
        const int ze = zb + base + (sg < rem ? 1 : 0);
        items.push_back(make_int4(tx, ty, zb, ze));
      }
  // Index table of every item, read by the kernel's prologue as one coalesced block: the
  // point index (or -1) at the ZNC tile columns of planes zb-4 .. ze+3, then at the ZNH
  // halo sites of the cross-shaped window (the kernel's slot order) of planes zb .. ze-1.
  std::vector<int> tab, toff(items.size());
  auto site = [&](long x, long y, long z) {
    return op->mcube[(size_t)x + (size_t)op->mx * ((size_t)y + (size_t)op->my * (size_t)z)];
  };
  for (size_t i = 0; i < items.size(); ++i) {
    const int4 it = items[i];
    toff[i] = (int)tab.size();
    const long x0 = (long)it.x * ZTX + ZR, y0 = (long)it.y * ZTY + ZR;
    for (int P = it.z - 4; P < it.w + 4; ++P)
      for (int t = 0; t < ZNC; ++t) tab.push_back(site(x0 + t % ZTX, y0 + t / ZTX, P));
    for (int P = it.z; P < it.w; ++P)
      for (int h = 0; h < ZNH; ++h) {
        long hx, hy;
        if (h < 4 * ZTY) { hx = x0 - 4 + h % 4; hy = y0 + h / 4; }
        else if (h < 8 * ZTY) { hx = x0 + ZTX + (h - 4 * ZTY) % 4; hy = y0 + (h - 4 * ZTY) / 4; }
        else if (h < 8 * ZTY + 4 * ZTX) { hx = x0 + (h - 8 * ZTY) % ZTX; hy = y0 - 4 + (h - 8 * ZTY) / ZTX; }
        else { hx = x0 + (h - 8 * ZTY - 4 * ZTX) % ZTX; hy = y0 + ZTY + (h - 8 * ZTY - 4 * ZTX) / ZTX; }
        tab.push_back(site(hx, hy, P));
      }
  }
  if (tab.size() > (size_t)INT_MAX) return fail(OCTHIP_ERR_UNSUPPORTED, "z-march index tables too large");
  int4* d = nullptr;
  int* dt = nullptr;
  int* dto = nullptr;
  int rc = upload(&d, items.data(), items.size(), nullptr);
  if (!rc) rc = upload(&dt, tab.data(), tab.size(), nullptr);
  if (!rc) rc = upload(&dto, toff.data(), toff.size(), nullptr);
  if (rc) {
    dfree(d);
    dfree(dt);
    dfree(dto);
    return rc;
  }
  op->plans.push_back({zs, d, dt, dto, items, {}, {}});
  op->d_items = d;
  op->d_tab = dt;
  op->d_tab_off = dto;
  op->items = items;
  op->nitems = (int)items.size();
  op->sched.clear();
  op->sched_len.clear();
  op->zseg_built = zs;
  return OCTHIP_SUCCESS;
}

// blockIdx -> (item, lane chunk): the (chunk, item) list is cut into 8 contiguous parts,
// part k runs on XCD k (workgroups are dealt to XCDs round-robin by blockIdx).
int zmarch_sched(octhip_nl_operator_s* op, int nchunk, const int2** out, int* len) {
  for (size_t i = 0; i < op->sched.size(); ++i)
    if (op->sched[i].first == nchunk) {
      *out = op->sched[i].second;
      *len = op->sched_len[i];
      return OCTHIP_SUCCESS;
    }
  const int n = op->nitems * nchunk;
  const int per = (n + 7) / 8;
  std::vector<int2> sc((size_t)per * 8, make_int2(-1, 0));
  for (int k = 0; k < 8; ++k)
    for (int i = 0; i < per; ++i) {
      const int g = k * per + i;
      if (g >= n) break;
      sc[(size_t)i * 8 + k] = make_int2(g % op->nitems, g / op->nitems);
    }
  int2* d = nullptr;
  int rc = upload(&d, sc.data(), sc.size(), nullptr);
  if (rc) return rc;
  op->sched.push_back({nchunk, d});
  op->sched_len.push_back((int)sc.size());
  for (auto& pl : op->plans)
    if (pl.zseg == op->zseg_built) {
      pl.sched = op->sched;
      pl.sched_len = op->sched_len;
    }
  *out = d;
  *len = (int)sc.size();
  return OCTHIP_SUCCESS;
}

int launch_zmarch(const ZArgs& z, int lc, int lpt, int epi, int grid, int zseg, hipStream_t s);

// The z-march addresses rows with 32-bit byte offsets: every row array it touches (psi
// over np_part rows, the outputs over np, the ket rows) must satisfy
// (rows + 1) * row bytes <= 2^32; larger batches take the gather kernel.
bool zmarch_fits(const octhip_nl_operator_s* op, const ApplyArgs& a) {
  const uint64_t rowb = (uint64_t)a.L * sizeof(double);
  const uint64_t lim = 1ull << 32;
  return (uint64_t)(op->np_part + 1) * rowb <= lim && (!a.kslot || (uint64_t)(a.nk + 1) * rowb <= lim);
}

// Build every lazily created device structure an apply with rows of L doubles needs (both
// epilogues; nl = with ket rows), so that later applies allocate nothing (required before
// a graph capture).
int prepare_operator(octhip_nl_operator_s* op, int L, bool nl) {
  const bool zm = op->zmarch_ok && op->kernel != 1 && op->n == 25;
  if (!zm) return OCTHIP_SUCCESS;
  int lc, lpt;
  zmarch_shape(op, L, &lc, &lpt);
  for (int epi : {EPI_STORE, EPI_TAYLOR, EPI_TAYLOR_X}) {
    if (epi != EPI_STORE && lc < 2) continue;
    int rc = zmarch_items(op, zmarch_zseg_for(op, L, epi, nl));
    if (rc) return rc;
    const int2* sched = nullptr;
    int len = 0;
    rc = zmarch_sched(op, L / lc, &sched, &len);
    if (rc) return rc;
  }
  return OCTHIP_SUCCESS;
}

// The z-march writes one density partial per lane chunk of a row, the gather kernel one
// per row (keep in step with launch_operator's kernel choice).
int apply_lane_chunks(const octhip_nl_operator_s* op, const ApplyArgs& a) {
  const bool zm = op->zmarch_ok && op->kernel != 1 && a.n == 25 && zmarch_fits(op, a);
  if (!zm) return 1;
  int lc, lpt;
  zmarch_shape(op, a.L, &lc, &lpt);
  return a.L / lc;
}

// Apply through the best kernel for this operator: z-march for the order-4 star on a
// lattice, the tiled gather otherwise.
int launch_operator(octhip_nl_operator_s* op, ApplyArgs& a, int epi, hipStream_t s) {
  if (!a.fi || !a.zrow || (a.kslot && !a.ksum) || (epi == EPI_STORE && !a.fo))
    return fail(OCTHIP_ERR_STATE, "apply: incomplete kernel arguments");
  const bool zm = op->zmarch_ok && op->kernel != 1 && a.n == 25 && zmarch_fits(op, a);
  if (!zm) return launch_apply(a, op->lattice && a.n > 0, epi, s);
  int lc, lpt;
  zmarch_shape(op, a.L, &lc, &lpt);
  int rc = zmarch_items(op, zmarch_zseg_for(op, a.L, epi, a.kslot != nullptr));
  if (rc) return rc;
  if (op->nitems == 0 || a.np <= 0) return OCTHIP_SUCCESS;
  if (zmarch_shmem(lc, op->zseg_built) > 160 * 1024)
    return fail(OCTHIP_ERR_INVALID, "z-march: lane chunk %d with z segment %d needs more than 160 KB of LDS", lc,
                op->zseg_built);
  const int2* sched = nullptr;
  int len = 0;
  rc = zmarch_sched(op, a.L / lc, &sched, &len);
  if (rc) return rc;
  ZArgs z;
  z.a = a;
  z.tab = op->d_tab;
  z.tab_off = op->d_tab_off;
  z.items = op->d_items;
  z.sched = sched;
  return launch_zmarch(z, lc, lpt, epi, len, op->zseg_built, s);
}
}  // namespace octhip

extern "C" int octhip_nl_operator_create(int n, const double* w, const int* stencil_points, int nri,
                                         const int* ri, const int* rimap_inv, int np, int np_part,
                                         const int* coords, octhip_nl_operator_t* out) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(out != nullptr, "nl_operator_create: null output handle");
  *out = nullptr;
  OCTHIP_REQUIRE(n > 0 && n <= MAX_STENCIL, "stencil size %d outside 1..%d", n, MAX_STENCIL);
  OCTHIP_REQUIRE(w && ri && rimap_inv, "nl_operator_create: null descriptor array");
  OCTHIP_REQUIRE(np > 0 && np_part >= np && nri > 0, "bad sizes np=%d np_part=%d nri=%d", np, np_part, nri);
  // rimap from rimap_inv, with the checks nl_operator_build guarantees
  OCTHIP_REQUIRE(rimap_inv[0] == 0 && rimap_inv[nri] == np, "rimap_inv must start at 0 and end at np");
  std::vector<int> rimap(np);
  for (int l = 0; l < nri; ++l) {
    OCTHIP_REQUIRE(rimap_inv[l + 1] > rimap_inv[l], "run %d is empty", l);
    for (int p = rimap_inv[l]; p < rimap_inv[l + 1]; ++p) rimap[p] = l;
  }
  for (int l = 0; l < nri; ++l)
    for (int j = 0; j < n; ++j) {
      const long lo = (long)rimap_inv[l] + ri[(size_t)l * n + j];
      const long hi = (long)rimap_inv[l + 1] - 1 + ri[(size_t)l * n + j];
      OCTHIP_REQUIRE(lo >= 0 && hi < np_part, "ri[%d][%d] reaches outside 0..np_part-1", l, j);
    }
  auto* op = new octhip_nl_operator_s();
  op->n = n;
  op->np = np;
  op->np_part = np_part;
  op->nri = nri;
  op->w.assign(w, w + n);
  if (stencil_points) {
    for (int j = 0; j < n; ++j)
      if (!stencil_points[3 * j] && !stencil_points[3 * j + 1] && !stencil_points[3 * j + 2]) op->center = j;
  }
  hipStream_t s = nullptr;
  int rc = upload(&op->d_ri, ri, (size_t)nri * n, s);
  if (!rc) rc = upload(&op->d_rimap, rimap.data(), rimap.size(), s);
  // inner / outer runs: a run is inner when its last point plus the largest offset is
  // still an inner point (maxp <= np in the reference's 1-based terms)
  {
    std::vector<int> pri[2], pmin[2], pmax[2];
    for (int l = 0; l < nri; ++l) {
      int mx = ri[(size_t)l * n];
      for (int j = 1; j < n; ++j) mx = std::max(mx, ri[(size_t)l * n + j]);
      const int part = (long)rimap_inv[l + 1] + mx <= np ? 0 : 1;
      pri[part].insert(pri[part].end(), ri + (size_t)l * n, ri + (size_t)(l + 1) * n);
      pmin[part].push_back(rimap_inv[l]);
      pmax[part].push_back(rimap_inv[l + 1]);
    }
    for (int part = 0; part < 2 && !rc; ++part) {
      op->part_nri[part] = (int)pmin[part].size();
      if (pmin[part].empty()) continue;
      rc = upload(&op->d_part_ri[part], pri[part].data(), pri[part].size(), s);
      if (!rc) rc = upload(&op->d_part_imin[part], pmin[part].data(), pmin[part].size(), s);
      if (!rc) rc = upload(&op->d_part_imax[part], pmax[part].data(), pmax[part].size(), s);
    }
  }
  if (!rc && stencil_points) {
    bool ok = false;
    if (coords) {
      op->coords.assign(coords, coords + 3 * (size_t)np_part);
      ok = true;
      // verify the caller's coordinates against ri before trusting them
      for (int p = 0; p < np && ok; ++p) {
        const int* row = ri + (size_t)rimap[p] * n;
        for (int j = 0; j < n && ok; ++j) {
          const int q = p + row[j];
          for (int d = 0; d < 3; ++d)
            if (coords[3 * (size_t)q + d] != coords[3 * (size_t)p + d] + stencil_points[3 * j + d]) ok = false;
        }
      }
    } else {
      ok = rebuild_lattice(n, stencil_points, ri, rimap, np, np_part, op->coords);
    }
    if (ok) {
      rc = build_site_map(op, stencil_points, s);
      if (!rc) op->lattice = true;
      if (!rc && is_star4(n, stencil_points)) {
        rc = build_march(op, s);
        if (!rc) op->zmarch_ok = true;
      }
    }
  }
  if (rc) {
    octhip_nl_operator_destroy(op);
    return rc;
  }
  *out = op;
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_nl_operator_destroy(octhip_nl_operator_t op) {
  if (!op) return OCTHIP_SUCCESS;
  dfree(op->d_cidx);
  dfree(op->d_cube);
  dfree(op->d_dcube);
  for (auto& pl : op->plans)
    if (pl.zseg == op->zseg_built) pl.sched = op->sched;
  for (auto& pl : op->plans) {
    dfree(pl.d_items);
    dfree(pl.d_tab);
    dfree(pl.d_tab_off);
    for (auto& e : pl.sched) dfree(e.second);
  }
  for (void* p : op->retired) dfree(p);
  dfree(op->d_ri);
  dfree(op->d_rimap);
  for (int part = 0; part < 2; ++part) {
    dfree(op->d_part_ri[part]);
    dfree(op->d_part_imin[part]);
    dfree(op->d_part_imax[part]);
  }
  delete op;
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_nl_operator_set_tuning(octhip_nl_operator_t op, int key, int value) {
  OCTHIP_REQUIRE(op != nullptr, "null operator handle");
  switch (key) {
    case OCTHIP_TUNE_KERNEL:
      OCTHIP_REQUIRE(value >= 0 && value <= 2, "kernel %d outside 0..2", value);
      OCTHIP_REQUIRE(value != 2 || op->zmarch_ok, "z-march needs the order-4 star on a lattice");
      op->kernel = value;
      return OCTHIP_SUCCESS;
    case OCTHIP_TUNE_XCD_CHUNK:
      OCTHIP_REQUIRE(value >= 1 && value <= 4096, "xcd_chunk %d outside 1..4096", value);
      op->xcd_chunk = value;
      return OCTHIP_SUCCESS;
    case OCTHIP_TUNE_GATHER_TILE:
      OCTHIP_REQUIRE(value >= 1 && value <= 8, "gather tile %d outside 1..8", value);
      op->tile_iters = value;
      return OCTHIP_SUCCESS;
    case OCTHIP_TUNE_ZSEG:
      OCTHIP_REQUIRE(value >= 0 && value <= 64, "z segment %d outside 0..64 (0 = automatic)", value);
      op->zseg = value;
      return OCTHIP_SUCCESS;
    case OCTHIP_TUNE_LANE_CHUNK:
      OCTHIP_REQUIRE(value == 0 || value == 1 || value == 2 || value == 4 || value == 8 || value == 16 ||
                         value == 32,
                     "lane chunk %d not in {0,1,2,4,8,16,32}", value);
      op->lane_chunk = value;
      return OCTHIP_SUCCESS;
    case OCTHIP_TUNE_LANES_PER_THREAD:
      OCTHIP_REQUIRE(value >= 0 && value <= 2, "lanes per thread %d outside 0..2", value);
      op->lanes_per_thread = value;
      return OCTHIP_SUCCESS;
    default:
      return fail(OCTHIP_ERR_INVALID, "unknown operator tuning key %d", key);
  }
}

extern "C" int octhip_nl_operator_get_tuning(octhip_nl_operator_t op, int key, int* value) {
  OCTHIP_REQUIRE(op != nullptr && value != nullptr, "null argument");
  switch (key) {
    case OCTHIP_TUNE_KERNEL: *value = op->kernel; break;
    case OCTHIP_TUNE_XCD_CHUNK: *value = op->xcd_chunk; break;
    case OCTHIP_TUNE_GATHER_TILE: *value = op->tile_iters; break;
    case OCTHIP_TUNE_ZSEG: *value = op->zseg; break;
    case OCTHIP_TUNE_LANE_CHUNK: *value = op->lane_chunk; break;
    case OCTHIP_TUNE_LANES_PER_THREAD: *value = op->lanes_per_thread; break;
    default: return fail(OCTHIP_ERR_INVALID, "unknown operator tuning key %d", key);
  }
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_nl_operator_uses_lattice(octhip_nl_operator_t op, int* flag) {
  OCTHIP_REQUIRE(op && flag, "null argument");
  *flag = op->lattice ? 1 : 0;
  return OCTHIP_SUCCESS;
}

namespace octhip {
// fill the stencil part of ApplyArgs from an operator handle
void stencil_args(const octhip_nl_operator_s* op, double factor, ApplyArgs& a) {
  a.n = op->n;
  a.np = op->np;
  a.cidx = op->d_cidx;
  a.cube = op->d_cube;
  a.dcp = op->d_dcube;
  a.ri = op->d_ri;
  a.rimap = op->d_rimap;
  for (int j = 0; j < op->n; ++j) {
    a.w[j] = op->w[j] * factor;  // wre = w * factor, src/grid/nl_operator_inc.F90:72-80
    a.dc[j] = op->dcube[j];
  }
  a.zrow = zero_row();
  a.chunk = op->xcd_chunk;
  a.tile_iters = op->tile_iters;
}
}  // namespace octhip

namespace octhip {
// the run-compressed leaf over nri runs given by their ri rows and point ranges
int launch_ri_runs(int n, const double* w, int nri, const int* ri, const int* imin, const int* imax,
                   const double* fi, int ldf, double* fo, hipStream_t s);
}  // namespace octhip

extern "C" int octhip_nl_operator_operate_batch(octhip_nl_operator_t op, const double* fi, double* fo,
                                                int ldf, double factor, int points, void* stream) {
  OCTHIP_RANGE("octhip_nl_operator_operate_batch");
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(op != nullptr, "null operator handle");
  OCTHIP_REQUIRE(fi && fo, "null batch pointer");
  OCTHIP_REQUIRE(ldf >= 0 && ldf <= MAX_LDF, "ldf=%d outside 0..%d", ldf, MAX_LDF);
  OCTHIP_REQUIRE(fi != fo, "operate_batch cannot work in place");
  OCTHIP_REQUIRE(points == OCTHIP_OP_ALL || points == OCTHIP_OP_INNER || points == OCTHIP_OP_OUTER,
                 "points=%d is none of OP_INNER, OP_OUTER, OP_ALL", points);
  if (points != OCTHIP_OP_ALL) {
    // the inner or the outer runs only (nl_operator_inc.F90:257-469 with op%inner / op%outer),
    // through the run-compressed leaf: same weights, same summation order per lane
    const int part = points == OCTHIP_OP_INNER ? 0 : 1;
    std::vector<double> wf(op->n);
    for (int j = 0; j < op->n; ++j) wf[j] = op->w[j] * factor;
    return launch_ri_runs(op->n, wf.data(), op->part_nri[part], op->d_part_ri[part], op->d_part_imin[part],
                          op->d_part_imax[part], fi, ldf, fo, (hipStream_t)stream);
  }
  ApplyArgs a;
  std::memset(&a, 0, sizeof(a));
  stencil_args(op, factor, a);
  a.fi = fi;
  a.fo = fo;
  a.L = 1 << ldf;
  a.lim = op->np_part;  // boundary rows are read as the caller set them
  return launch_operator(op, a, EPI_STORE, (hipStream_t)stream);
}

// ------------------------------------------------------- stateless leaf --
namespace {
struct RiWeights {
  double w[MAX_STENCIL];
};

// tpr = min(L, 64) consecutive threads own one run and walk its points, each thread a
// lane (stride tpr) of every point; a 256-thread workgroup covers 256 / tpr runs, so
// short runs (two points on average in Octopus's blocked order) still fill the waves.
// Per lane the sum runs over j in stencil order with fused multiply-adds, as the C leaf.
__global__ __launch_bounds__(BLOCK) void k_operate_ri(int n, const RiWeights w, int nri, const int* __restrict__ ri,
                                                      const int* __restrict__ imin, const int* __restrict__ imax,
                                                      const double* __restrict__ fi, int L, int tshift,
                                                      double* __restrict__ fo) {
  const int tpr = 1 << tshift;
  const int l = (int)(blockIdx.x * (BLOCK >> tshift)) + (int)(threadIdx.x >> tshift);
  if (l >= nri) return;
  const int* row = ri + (size_t)l * n;
  const int i0 = imin[l], i1 = imax[l];
  for (int i = i0; i < i1; ++i)
    for (int k = threadIdx.x & (tpr - 1); k < L; k += tpr) {
      double a = 0.0;
      for (int j = 0; j < n; ++j) a = fma(w.w[j], fi[(size_t)(row[j] + i) * L + k], a);
      fo[(size_t)i * L + k] = a;
    }
}
}  // namespace

namespace octhip {
int launch_ri_runs(int n, const double* w, int nri, const int* ri, const int* imin, const int* imax,
                   const double* fi, int ldf, double* fo, hipStream_t s) {
  if (nri == 0) return OCTHIP_SUCCESS;
  // the weights travel as a kernel argument: nothing is allocated, nothing synchronised
  RiWeights rw;
  for (int j = 0; j < MAX_STENCIL; ++j) rw.w[j] = j < n ? w[j] : 0.0;
  const int tshift = ldf < 6 ? ldf : 6;
  const int runs_per_block = BLOCK >> tshift;
  hipLaunchKernelGGL(k_operate_ri, dim3((nri + runs_per_block - 1) / runs_per_block), dim3(BLOCK), 0, s, n, rw,
                     nri, ri, imin, imax, fi, 1 << ldf, tshift, fo);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(OCTHIP_ERR_HIP, "operate_ri_vec: %s", hipGetErrorString(e));
  return OCTHIP_SUCCESS;
}
}  // namespace octhip

extern "C" int octhip_operate_ri_vec(int n, const double* w, int nri, const int* ri, const int* rimap_inv,
                                     const int* rimap_inv_max, const double* fi, int ldf, double* fo,
                                     void* stream) {
  OCTHIP_RANGE("octhip_operate_ri_vec");
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(n > 0 && n <= MAX_STENCIL && nri >= 0, "bad n=%d or nri=%d", n, nri);
  OCTHIP_REQUIRE(ldf >= 0 && ldf <= MAX_LDF, "ldf=%d outside 0..%d", ldf, MAX_LDF);
  if (nri == 0) return OCTHIP_SUCCESS;
  OCTHIP_REQUIRE(w && ri && rimap_inv && rimap_inv_max && fi && fo, "null argument");
  return launch_ri_runs(n, w, nri, ri, rimap_inv, rimap_inv_max, fi, ldf, fo, (hipStream_t)stream);
}
