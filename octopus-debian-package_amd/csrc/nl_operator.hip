// nl_operator handle: upload of the run-compressed stencil, lattice rebuild, operate.
#include <algorithm>
#include <climits>
#include <cstring>
#include <deque>

#include "kernels.h"

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
  if (rc) return rc;
  return upload(&op->d_cidx, cidx.data(), cidx.size(), s);
}

}  // namespace

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
  dfree(op->d_ri);
  dfree(op->d_rimap);
  delete op;
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_nl_operator_set_schedule(octhip_nl_operator_t op, int xcd_chunk) {
  OCTHIP_REQUIRE(op != nullptr, "null operator handle");
  OCTHIP_REQUIRE(xcd_chunk >= 1 && xcd_chunk <= 4096, "xcd_chunk %d outside 1..4096", xcd_chunk);
  op->xcd_chunk = xcd_chunk;
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
  a.ri = op->d_ri;
  a.rimap = op->d_rimap;
  for (int j = 0; j < op->n; ++j) {
    a.w[j] = op->w[j] * factor;  // wre = w * factor, src/grid/nl_operator_inc.F90:72-80
    a.dc[j] = op->dcube[j];
  }
  a.zrow = zero_row();
  a.chunk = op->xcd_chunk;
}
}  // namespace octhip

extern "C" int octhip_nl_operator_operate_batch(octhip_nl_operator_t op, const double* fi, double* fo,
                                                int ldf, double factor, int points, void* stream) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(op != nullptr, "null operator handle");
  OCTHIP_REQUIRE(fi && fo, "null batch pointer");
  OCTHIP_REQUIRE(ldf >= 0 && ldf <= MAX_LDF, "ldf=%d outside 0..%d", ldf, MAX_LDF);
  OCTHIP_REQUIRE(fi != fo, "operate_batch cannot work in place");
  if (points != OCTHIP_OP_ALL)
    return fail(OCTHIP_ERR_UNSUPPORTED, "only OP_ALL is supported (no domain parallelisation)");
  ApplyArgs a;
  std::memset(&a, 0, sizeof(a));
  stencil_args(op, factor, a);
  a.fi = fi;
  a.fo = fo;
  a.L = 1 << ldf;
  a.lim = op->np_part;  // boundary rows are read as the caller set them
  return launch_apply(a, op->lattice, EPI_STORE, (hipStream_t)stream);
}

// ------------------------------------------------------- stateless leaf --
namespace {
// one wave per run; the wave walks the run's points, lanes walk the row
__global__ __launch_bounds__(BLOCK) void k_operate_ri(int n, const double* __restrict__ w, int nri,
                                                      const int* __restrict__ ri,
                                                      const int* __restrict__ imin,
                                                      const int* __restrict__ imax,
                                                      const double* __restrict__ fi, int L,
                                                      double* __restrict__ fo) {
  const int l = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6)));
  if (l >= nri) return;
  const int* row = ri + (size_t)l * n;
  for (int i = imin[l]; i < imax[l]; ++i)
    for (int k = threadIdx.x & 63; k < L; k += 64) {
      double a = 0.0;
      for (int j = 0; j < n; ++j) a = fma(w[j], fi[(size_t)(row[j] + i) * L + k], a);
      fo[(size_t)i * L + k] = a;
    }
}
}  // namespace

extern "C" int octhip_operate_ri_vec(int n, const double* w, int nri, const int* ri, const int* rimap_inv,
                                     const int* rimap_inv_max, const double* fi, int ldf, double* fo,
                                     void* stream) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(n > 0 && n <= MAX_STENCIL && nri >= 0, "bad n=%d or nri=%d", n, nri);
  OCTHIP_REQUIRE(ldf >= 0 && ldf <= MAX_LDF, "ldf=%d outside 0..%d", ldf, MAX_LDF);
  if (nri == 0) return OCTHIP_SUCCESS;
  OCTHIP_REQUIRE(w && ri && rimap_inv && rimap_inv_max && fi && fo, "null argument");
  hipStream_t s = (hipStream_t)stream;
  double* dw = nullptr;
  int rc = dalloc((void**)&dw, sizeof(double) * n);
  if (rc) return rc;
  OCTHIP_HIP(hipMemcpyAsync(dw, w, sizeof(double) * n, hipMemcpyHostToDevice, s));
  const int waves = BLOCK / 64;
  hipLaunchKernelGGL(k_operate_ri, dim3((nri + waves - 1) / waves), dim3(BLOCK), 0, s, n, dw, nri, ri,
                     rimap_inv, rimap_inv_max, fi, 1 << ldf, fo);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dfree(dw);
  if (e != hipSuccess) return fail(OCTHIP_ERR_HIP, "operate_ri_vec: %s", hipGetErrorString(e));
  return OCTHIP_SUCCESS;
}
