// Hamiltonian handle: local potential, KB projectors, fused apply, Taylor propagator.
#include <algorithm>
#include <complex>
#include <functional>
#include <cstring>

#include "kernels.h"

using namespace octhip;

namespace octhip {
void stencil_args(const octhip_nl_operator_s* op, double factor, ApplyArgs& a);
int launch_bra_mfma(const BraArgs& b, const double* pt, int pjs, int nchunk, hipStream_t s, double* zf = nullptr,
                    size_t zfirst = 0, size_t zcount = 0);
int launch_operator(octhip_nl_operator_s* op, ApplyArgs& a, int epi, hipStream_t s);
int prepare_operator(octhip_nl_operator_s* op, int L, bool nl);
}  // namespace octhip

namespace {

enum { O_POINTS = 0, O_PROJS, O_MATRIX, O_MAP, O_SCAL, O_MIX };

// a graph whose Hamiltonian is destroyed no longer owns work sets (they are freed with it)
void detach_graph(octhip_taylor_graph_t g);

// the work-buffer set of stream s, created empty on first use, and mirrored into h
octhip_hamiltonian_s::WorkSet& workset(octhip_hamiltonian_s* h, hipStream_t s) {
  for (auto& w : h->sets)
    if (w.s == s) return w;
  h->sets.push_back({s, -1, -1, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0, nullptr, -1});
  return h->sets.back();
}

void bind(octhip_hamiltonian_s* h, const octhip_hamiltonian_s::WorkSet& w) {
  h->d_partial = w.partial;
  h->d_proj = w.proj;
  h->d_ksum = w.ksum;
  h->d_y[0] = w.y0;
  h->d_y[1] = w.y1;
  h->d_y[2] = w.y2;
  h->work_ldf = w.work_ldf;
  h->y_ldf = w.y_ldf;
}

// projector work buffers of stream s for rows of 2^ldf lanes (outgrown ones are retired)
int ensure_work(octhip_hamiltonian_s* h, int ldf, hipStream_t s) {
  auto& w = workset(h, s);
  if (w.work_ldf < ldf) {
    for (void* p : {(void*)w.partial, (void*)w.proj, (void*)w.ksum})
      if (p) h->retired.push_back(p);
    w.partial = w.proj = w.ksum = nullptr;
    const size_t L = (size_t)1 << ldf;
    int rc = OCTHIP_SUCCESS;
    if (h->nmat > 0) {
      rc = dalloc((void**)&w.partial, sizeof(double) * h->nchunk * h->max_nprojs * L);
      if (!rc) rc = dalloc((void**)&w.proj, sizeof(double) * h->full_projection_size * L);
      if (!rc) rc = dalloc((void**)&w.ksum, sizeof(double) * ((size_t)h->total_points + h->nmulti) * L);
    }
    if (rc) return rc;
    w.work_ldf = ldf;
  }
  bind(h, w);
  return OCTHIP_SUCCESS;
}

// Taylor intermediates of stream s (np rows each)
int ensure_taylor(octhip_hamiltonian_s* h, int ldf, hipStream_t s) {
  auto& w = workset(h, s);
  if (w.y_ldf < ldf) {
    for (void* p : {(void*)w.y0, (void*)w.y1, (void*)w.y2})
      if (p) h->retired.push_back(p);
    w.y0 = w.y1 = w.y2 = nullptr;
    const size_t bytes = sizeof(double) * ((size_t)h->np << ldf);
    int rc = dalloc((void**)&w.y0, bytes);
    if (!rc) rc = dalloc((void**)&w.y1, bytes);
    if (!rc) rc = dalloc((void**)&w.y2, bytes);
    if (rc) return rc;
    w.y_ldf = ldf;
  }
  bind(h, w);
  return OCTHIP_SUCCESS;
}

// projections of batch f (L lanes) into h->d_proj; zero_bc (optional) = a batch whose
// boundary rows are cleared on the way (fused into the bra launch when it can be)
int run_bra(octhip_hamiltonian_s* h, const double* f, int L, hipStream_t s, double* zero_bc = nullptr) {
  const int np = h->kin->np, np_part = h->kin->np_part;
  if (h->nmat == 0) return zero_bc ? zero_boundary(zero_bc, np, np_part, L, s) : OCTHIP_SUCCESS;
  const bool mfma = h->use_mfma && L >= 16;
  if (zero_bc && !mfma) {
    int rc = zero_boundary(zero_bc, np, np_part, L, s);
    if (rc) return rc;
    zero_bc = nullptr;
  }
  BraArgs b;
  b.psi = f;
  b.L = L;
  b.chunks = h->d_chunks;
  b.mat_info = h->d_mat_info;
  b.mats = h->d_matrices;
  b.maps = h->d_maps;
  b.partial = h->d_partial;
  b.pjmax = h->max_nprojs;
  b.pphase = h->has_kpoint ? h->d_pphase : nullptr;
  int rc = mfma ? launch_bra_mfma(b, h->d_pt, h->pjs, h->nchunk, s, zero_bc, (size_t)np * L,
                                  (size_t)(np_part - np) * L)
                : launch_bra(b, h->nchunk, false, s);
  if (rc) return rc;
  // with the MFMA ket, each ket workgroup sums the partials of its matrix itself
  const bool fuse = mfma && h->max_nprojs <= KET_FUSE_PJMAX;
  BraReduceArgs r;
  r.partial = h->d_partial;
  r.proj_mat = h->d_proj_mat;
  r.mat_pbase = h->d_mat_pbase;
  r.mat_chunk0 = h->d_mat_chunk0;
  r.scals = h->d_scals;
  r.proj = h->d_proj;
  r.L = L;
  r.pjmax = h->max_nprojs;
  r.nproj = h->full_projection_size;
  if (!fuse) rc = launch_bra_reduce(r, s);
  if (rc) return rc;
  KetArgs k;
  if (fuse) {
    k.partial = h->d_partial;
    k.mat_chunk0 = h->d_mat_chunk0;
    k.scals = h->d_scals;
    k.pjmax = h->max_nprojs;
  }
  k.chunks = h->d_chunks;
  k.mat_info = h->d_mat_info;
  k.mat_pbase = h->d_mat_pbase;
  k.pt = h->d_pt;
  k.pjs = h->pjs;
  k.proj = h->d_proj;
  k.kbuf = h->d_ksum;
  k.L = L;
  k.multi_pos = h->d_multi_pos;
  k.multi_g = h->d_multi_g;
  k.nmulti = h->nmulti;
  k.total_points = h->total_points;
  k.pphase = h->has_kpoint ? h->d_pphase : nullptr;
  return launch_ket(k, h->nchunk, h->use_mfma, s);
}

// v_eff(p) = v(p) + (f0 x + f1 y + f2 z): the dipole term of laser_potential
// (src/hamiltonian/lasers.F90:823-853) folded into the local potential once per field.
__global__ void k_veff(double* __restrict__ veff, const double* __restrict__ v,
                       const double* __restrict__ xyz, double f0, double f1, double f2, int np) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x)
    veff[p] = v[p] + (f0 * xyz[3 * (size_t)p] + f1 * xyz[3 * (size_t)p + 1] + f2 * xyz[3 * (size_t)p + 2]);
}

}  // namespace

namespace octhip {
int zero_boundary(double* f, int np, int np_part, int L, hipStream_t s) {
  const size_t count = (size_t)(np_part - np) * L;
  if (count == 0) return OCTHIP_SUCCESS;
  size_t grid = (count + BLOCK - 1) / BLOCK;
  if (grid > 8192) grid = 8192;
  ProfScope prof("k_zero_rows", s);
  hipLaunchKernelGGL(k_zero_rows, dim3((unsigned)grid), dim3(BLOCK), 0, s, f, (size_t)np * L, count);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(OCTHIP_ERR_HIP, "k_zero_rows: %s", hipGetErrorString(e));
  return OCTHIP_SUCCESS;
}

// out = v + field . x over the np points (a caller's buffer, e.g. one per stream)
int launch_veff(const octhip_hamiltonian_s* h, double* out, const double field[3], hipStream_t s) {
  if (!h->d_v || !h->d_xyz) return fail(OCTHIP_ERR_STATE, "launch_veff: no local potential or positions");
  const int grid = std::min((h->np + BLOCK - 1) / BLOCK, 4096);
  hipLaunchKernelGGL(k_veff, dim3(grid), dim3(BLOCK), 0, s, out, h->d_v, h->d_xyz, field[0], field[1], field[2],
                     h->np);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(OCTHIP_ERR_HIP, "k_veff: %s", hipGetErrorString(e));
  return OCTHIP_SUCCESS;
}
}  // namespace octhip

namespace {

int update_veff(octhip_hamiltonian_s* h, hipStream_t s) {
  if (!h->has_field || !h->d_v || !h->d_xyz) return OCTHIP_SUCCESS;
  if (!h->d_veff) {
    int rc = dalloc((void**)&h->d_veff, sizeof(double) * h->np);
    if (rc) return rc;
  }
  const int grid = std::min((h->np + BLOCK - 1) / BLOCK, 4096);
  hipLaunchKernelGGL(k_veff, dim3(grid), dim3(BLOCK), 0, s, h->d_veff, h->d_v, h->d_xyz, h->field[0],
                     h->field[1], h->field[2], h->np);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(OCTHIP_ERR_HIP, "k_veff: %s", hipGetErrorString(e));
  return OCTHIP_SUCCESS;
}

void fill_terms(const octhip_hamiltonian_s* h, int terms, ApplyArgs& a) {
  if (terms & OCTHIP_TERM_KINETIC) {
    stencil_args(h->kin, -0.5 / h->mass, a);  // factor -1/(2m), hamiltonian_elec_inc.F90:155
  } else {
    stencil_args(h->kin, 0.0, a);
    a.n = 0;  // no stencil term: runtime-size path with zero points
  }
  if ((terms & OCTHIP_TERM_LOCAL_POTENTIAL) && h->d_v) {
    a.v = (h->has_field && h->d_veff) ? h->d_veff : h->d_v;
  }
  if ((terms & OCTHIP_TERM_NON_LOCAL_POTENTIAL) && h->nmat > 0 && h->nslot > 0) {
    a.kslot = h->d_kslot;
    a.ksum = h->d_ksum;
    a.nk = h->total_points + h->nmulti;
  }
}

}  // namespace

namespace {
// the phased copy of a k-point apply's input (np_part rows) of stream s
int ensure_ephase(octhip_hamiltonian_s* h, int ldf, hipStream_t s, double** out) {
  auto& w = workset(h, s);
  if (w.ephase_ldf < ldf) {
    if (w.ephase) h->retired.push_back(w.ephase);
    w.ephase = nullptr;
    w.ephase_ldf = -1;
    int rc = dalloc((void**)&w.ephase, sizeof(double) * ((size_t)h->kin->np_part << ldf));
    if (rc) return rc;
    w.ephase_ldf = ldf;
  }
  *out = w.ephase;
  return OCTHIP_SUCCESS;
}

// X(hamiltonian_elec_apply_batch) for a periodic handle or one with Bloch phases
// (src/hamiltonian/hamiltonian_elec_inc.F90:116-154, 243-247): boundaries_set fills the
// boundary rows from their periodic images (times phase_corr when psi carries the phase);
// a psi without the phase is copied with it (epsib, np_part rows) and H psi is multiplied
// by the conjugate phase at the end; the projectors use their Bloch phases; the stencil
// reads the boundary rows as they are.
int periodic_apply(octhip_hamiltonian_s* h, double* psi, double* hpsi, int ldf, int is_complex, int terms,
                   int set_bc, bool has_phase, hipStream_t s) {
  const int L = 1 << ldf;
  const int np = h->np, np_part = h->kin->np_part;
  if (h->has_kpoint && !is_complex) return fail(OCTHIP_ERR_INVALID, "Bloch phases need a complex batch");
  int rc = ensure_work(h, ldf, s);
  if (rc) return rc;
  if (set_bc) {
    rc = h->periodic ? launch_periodic_bc(psi, h->d_bnd_src, np, h->nbnd, L,
                                          (has_phase && h->has_kpoint) ? h->d_phase_corr : nullptr, s)
                     : zero_boundary(psi, np, np_part, L, s);
    if (rc) return rc;
  }
  const double* src = psi;
  const bool set_phase = h->has_kpoint && !has_phase;
  if (set_phase) {
    double* e = nullptr;
    rc = ensure_ephase(h, ldf, s, &e);
    if (!rc) rc = launch_phase(e, psi, h->d_phase, np_part, L, false, s);
    if (rc) return rc;
    src = e;
  }
  if ((terms & OCTHIP_TERM_NON_LOCAL_POTENTIAL) && h->nmat > 0) {
    rc = run_bra(h, src, L, s, nullptr);
    if (rc) return rc;
  }
  ApplyArgs a;
  std::memset(&a, 0, sizeof(a));
  fill_terms(h, terms, a);
  a.fi = src;
  a.fo = hpsi;
  a.L = L;
  a.lim = np_part;  // boundary rows as set above (or by the caller)
  rc = launch_operator(h->kin, a, EPI_STORE, s);
  if (!rc && set_phase) rc = launch_phase(hpsi, hpsi, h->d_phase, np, L, true, s);
  return rc;
}
}  // namespace

namespace octhip {
int propagator_apply(octhip_hamiltonian_s* h, double* psi, double* hpsi, int ldf, hipStream_t s) {
  if (is_periodic(h)) return periodic_apply(h, psi, hpsi, ldf, 1, OCTHIP_TERM_ALL, 1, true, s);
  return octhip_hamiltonian_apply_batch(h, psi, hpsi, ldf, 1, OCTHIP_TERM_ALL, 1, s);
}
}  // namespace octhip

extern "C" int octhip_hamiltonian_create(octhip_nl_operator_t kinetic, double mass, const double* vlocal,
                                         int nmat, const int* offsets, const double* matrices,
                                         const int* maps, const double* scals, octhip_hamiltonian_t* out) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(out != nullptr, "hamiltonian_create: null output handle");
  *out = nullptr;
  OCTHIP_REQUIRE(kinetic != nullptr, "hamiltonian_create: null kinetic operator");
  OCTHIP_REQUIRE(mass > 0.0, "mass must be positive");
  OCTHIP_REQUIRE(nmat >= 0, "nmat < 0");
  OCTHIP_REQUIRE(nmat == 0 || (offsets && matrices && maps && scals), "null projector arrays");
  const int np = kinetic->np;
  // sizes and validation of build_accel's offsets
  size_t matrix_size = 0, total_points = 0, nproj = 0;
  int max_nprojs = 0;
  for (int a = 0; a < nmat; ++a) {
    const int* o = offsets + 6 * (size_t)a;
    OCTHIP_REQUIRE(o[O_POINTS] >= 0 && o[O_PROJS] >= 1, "matrix %d: bad POINTS/PROJS", a);
    OCTHIP_REQUIRE(o[O_MIX] == -1, "matrix %d: mixing matrices (HGH/SOC) are not supported", a);
    OCTHIP_REQUIRE(o[O_MATRIX] >= 0 && o[O_MAP] >= 0 && o[O_SCAL] >= 0, "matrix %d: negative offset", a);
    matrix_size = std::max(matrix_size, (size_t)o[O_MATRIX] + (size_t)o[O_POINTS] * o[O_PROJS]);
    total_points = std::max(total_points, (size_t)o[O_MAP] + o[O_POINTS]);
    nproj = std::max(nproj, (size_t)o[O_SCAL] + o[O_PROJS]);
    max_nprojs = std::max(max_nprojs, o[O_PROJS]);
  }
  for (size_t i = 0; i < total_points; ++i)
    OCTHIP_REQUIRE(maps[i] >= 1 && maps[i] <= np, "map entry %zu = %d outside 1..np", i, maps[i]);
  auto* h = new octhip_hamiltonian_s();
  h->kin = kinetic;
  h->mass = mass;
  h->np = np;
  h->nmat = nmat;
  h->max_nprojs = max_nprojs;
  h->full_projection_size = (int)nproj;
  h->total_points = (int)total_points;
  h->pjs = max_nprojs;
  h->offsets.assign(offsets ? offsets : nullptr, offsets ? offsets + 6 * (size_t)nmat : nullptr);
  hipStream_t s = nullptr;
  int rc = OCTHIP_SUCCESS;
  if (vlocal) rc = upload(&h->d_v, vlocal, (size_t)np, s);
  if (!rc && nmat > 0) {
    std::vector<double> pt(total_points * (size_t)h->pjs, 0.0);
    std::vector<int> cnt(np + 1, 0), pbase(nmat), npj(nmat), proj_mat(nproj, 0), chunk0(nmat + 1);
    std::vector<int4> info(nmat), chunks;
    for (int a = 0; a < nmat; ++a) {
      const int* o = offsets + 6 * (size_t)a;
      const int npt = o[O_POINTS], nj = o[O_PROJS];
      pbase[a] = o[O_SCAL];
      npj[a] = nj;
      info[a] = make_int4(npt, nj, o[O_MATRIX], o[O_MAP]);
      for (int j = 0; j < nj; ++j) proj_mat[o[O_SCAL] + j] = a;
      for (int q = 0; q < npt; ++q) {
        for (int j = 0; j < nj; ++j)
          pt[(size_t)(o[O_MAP] + q) * h->pjs + j] = matrices[o[O_MATRIX] + (size_t)j * npt + q];
        cnt[maps[o[O_MAP] + q]]++;  // 1-based -> counts at index p+1
      }
      chunk0[a] = (int)chunks.size();
      for (int q0 = 0; q0 < npt; q0 += BRA_CHUNK)
        chunks.push_back(make_int4(a, q0, std::min(q0 + BRA_CHUNK, npt), (int)chunks.size() - chunk0[a]));
    }
    chunk0[nmat] = (int)chunks.size();
    // CSR of (sphere point, matrix) per mesh point, matrix order preserved
    std::vector<int> pos(np + 1, 0);
    for (int p = 0; p < np; ++p) pos[p + 1] = pos[p] + cnt[p + 1];
    std::vector<int2> ent(pos[np]);
    std::vector<int> fillp(pos.begin(), pos.end() - 1);
    for (int a = 0; a < nmat; ++a) {
      const int* o = offsets + 6 * (size_t)a;
      for (int q = 0; q < o[O_POINTS]; ++q) {
        const int p = maps[o[O_MAP] + q] - 1;
        ent[fillp[p]++] = make_int2(o[O_MAP] + q, a);
      }
    }
    h->nchunk = (int)chunks.size();
    // ket row per mesh point: its sphere point if it lies in one sphere, else a summed row
    std::vector<int> kslot(np, -1), multi_pos(1, 0), multi_g;
    int nslot = 0;
    for (int p = 0; p < np; ++p) {
      const int ne = pos[p + 1] - pos[p];
      if (ne == 0) continue;
      ++nslot;
      if (ne == 1) {
        kslot[p] = ent[pos[p]].x;
      } else {
        kslot[p] = (int)total_points + (int)multi_pos.size() - 1;
        for (int e = pos[p]; e < pos[p + 1]; ++e) multi_g.push_back(ent[e].x);
        multi_pos.push_back((int)multi_g.size());
      }
    }
    h->nslot = nslot;
    h->nmulti = (int)multi_pos.size() - 1;
    rc = upload(&h->d_kslot, kslot.data(), kslot.size(), s);
    if (!rc) rc = upload(&h->d_multi_pos, multi_pos.data(), multi_pos.size(), s);
    if (!rc && !multi_g.empty()) rc = upload(&h->d_multi_g, multi_g.data(), multi_g.size(), s);
    if (!rc) rc = upload(&h->d_matrices, matrices, matrix_size, s);
    if (!rc) rc = upload(&h->d_maps, maps, total_points, s);
    if (!rc) rc = upload(&h->d_scals, scals, nproj, s);
    if (!rc) rc = upload(&h->d_pt, pt.data(), pt.size(), s);
    if (!rc) rc = upload(&h->d_mat_pbase, pbase.data(), pbase.size(), s);
    if (!rc) rc = upload(&h->d_mat_npj, npj.data(), npj.size(), s);
    if (!rc) rc = upload(&h->d_chunks, chunks.data(), chunks.size(), s);
    if (!rc) rc = upload(&h->d_mat_chunk0, chunk0.data(), chunk0.size(), s);
    if (!rc) rc = upload(&h->d_mat_info, info.data(), info.size(), s);
    if (!rc) rc = upload(&h->d_proj_mat, proj_mat.data(), proj_mat.size(), s);
  }
  if (rc) {
    octhip_hamiltonian_destroy(h);
    return rc;
  }
  *out = h;
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_hamiltonian_destroy(octhip_hamiltonian_t h) {
  if (!h) return OCTHIP_SUCCESS;
  (void)hipDeviceSynchronize();
  dfree(h->d_mat_info);
  dfree(h->d_proj_mat);
  dfree(h->d_kslot);
  dfree(h->d_multi_pos);
  dfree(h->d_multi_g);
  dfree(h->d_v);
  dfree(h->d_xyz);
  dfree(h->d_veff);
  dfree(h->d_matrices);
  dfree(h->d_maps);
  dfree(h->d_scals);
  dfree(h->d_pt);
  dfree(h->d_mat_pbase);
  dfree(h->d_mat_npj);
  dfree(h->d_chunks);
  dfree(h->d_mat_chunk0);
  for (auto& w : h->sets)
    for (void* p : {(void*)w.partial, (void*)w.proj, (void*)w.ksum, (void*)w.y0, (void*)w.y1, (void*)w.y2,
                    (void*)w.dpart, (void*)w.dw, (void*)w.ephase})
      dfree(p);
  for (void* p : h->retired) dfree(p);
  dfree(h->d_bnd_src);
  dfree(h->d_phase);
  dfree(h->d_phase_corr);
  dfree(h->d_pphase);
  drop_krylov(h, nullptr, true);
  for (octhip_taylor_graph_t g : h->graphs) detach_graph(g);
  delete h;
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_hamiltonian_set_tuning(octhip_hamiltonian_t h, int key, int value) {
  OCTHIP_REQUIRE(h != nullptr, "null hamiltonian handle");
  if (key == OCTHIP_TUNE_BRA_MODE) {
    OCTHIP_REQUIRE(value == 0 || value == 1, "bra mode must be 0 or 1");
    h->use_mfma = value;
    return OCTHIP_SUCCESS;
  }
  return octhip_nl_operator_set_tuning(h->kin, key, value);  // stencil knobs go to the kinetic operator
}

extern "C" int octhip_hamiltonian_set_potential(octhip_hamiltonian_t h, const double* vlocal, void* stream) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(h && vlocal, "null argument");
  hipStream_t s = (hipStream_t)stream;
  if (!h->d_v) {
    int rc = dalloc((void**)&h->d_v, sizeof(double) * h->np);
    if (rc) return rc;
  }
  OCTHIP_HIP(hipMemcpyAsync(h->d_v, vlocal, sizeof(double) * h->np, hipMemcpyHostToDevice, s));
  OCTHIP_HIP(hipStreamSynchronize(s));
  return update_veff(h, s);
}

extern "C" int octhip_hamiltonian_set_dipole_field(octhip_hamiltonian_t h, const double* positions,
                                                   const double* field, void* stream) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(h && field, "null argument");
  hipStream_t s = (hipStream_t)stream;
  if (positions) {
    // d_veff keeps its address (captured graphs read it); only the positions are replaced
    dfree(h->d_xyz);
    h->d_xyz = nullptr;
    int rc = upload(&h->d_xyz, positions, 3 * (size_t)h->np, s);
    if (rc) return rc;
  }
  OCTHIP_REQUIRE(h->d_xyz != nullptr, "set_dipole_field: positions never uploaded");
  for (int d = 0; d < 3; ++d) h->field[d] = field[d];
  h->has_field = true;
  return update_veff(h, s);
}

extern "C" int octhip_hamiltonian_release_stream(octhip_hamiltonian_t h, void* stream) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(h != nullptr, "null hamiltonian handle");
  const hipStream_t s = (hipStream_t)stream;
  OCTHIP_HIP(hipStreamSynchronize(s));
  drop_workset(h, s);
  drop_krylov(h, s, false);
  return OCTHIP_SUCCESS;
}

namespace {
size_t dev_bytes(const void* p) {
  if (!p) return 0;
  size_t n = 0;
  if (hipMemPtrGetInfo(const_cast<void*>(p), &n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}
}  // namespace

extern "C" int octhip_hamiltonian_memory(octhip_hamiltonian_t h, size_t* bytes) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(h != nullptr && bytes != nullptr, "null argument");
  size_t data = 0, work = 0, graph = 0, old = 0;
  for (const void* p : {(const void*)h->d_v, (const void*)h->d_xyz, (const void*)h->d_veff,
                        (const void*)h->d_matrices, (const void*)h->d_maps, (const void*)h->d_scals,
                        (const void*)h->d_pt,
                        (const void*)h->d_mat_pbase, (const void*)h->d_mat_npj, (const void*)h->d_chunks,
                        (const void*)h->d_mat_chunk0, (const void*)h->d_mat_info, (const void*)h->d_proj_mat,
                        (const void*)h->d_kslot, (const void*)h->d_multi_pos, (const void*)h->d_multi_g,
                        (const void*)h->d_bnd_src, (const void*)h->d_phase, (const void*)h->d_phase_corr,
                        (const void*)h->d_pphase})
    data += dev_bytes(p);
  for (const auto& w : h->sets) {
    size_t b = 0;
    for (const void* p : {(const void*)w.partial, (const void*)w.proj, (const void*)w.ksum, (const void*)w.y0,
                          (const void*)w.y1, (const void*)w.y2, (const void*)w.dpart, (const void*)w.dw,
                          (const void*)w.ephase})
      b += dev_bytes(p);
    // sets handed to Taylor graphs are kept under negative keys
    if (reinterpret_cast<intptr_t>(w.s) < 0) graph += b;
    else work += b;
  }
  for (const auto& k : h->krylov) work += dev_bytes(k.vec) + dev_bytes(k.small);
  for (const void* p : h->retired) old += dev_bytes(p);
  bytes[0] = data;
  bytes[1] = work;
  bytes[2] = graph;
  bytes[3] = old;
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_hamiltonian_trim(octhip_hamiltonian_t h) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(h != nullptr, "null hamiltonian handle");
  // retired buffers belong to no set and no graph; work queued before they were retired
  // may still read them, so wait for the device first
  OCTHIP_HIP(hipDeviceSynchronize());
  for (void* p : h->retired) dfree(p);
  h->retired.clear();
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_hamiltonian_apply_batch(octhip_hamiltonian_t h, double* psi, double* hpsi, int ldf,
                                              int is_complex, int terms, int set_bc, void* stream) {
  OCTHIP_RANGE("octhip_hamiltonian_apply_batch");
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(h != nullptr, "null hamiltonian handle");
  OCTHIP_REQUIRE(psi && hpsi && psi != hpsi, "bad batch pointers");
  OCTHIP_REQUIRE(ldf >= 0 && ldf <= MAX_LDF, "ldf=%d outside 0..%d", ldf, MAX_LDF);
  OCTHIP_REQUIRE(!is_complex || ldf >= 1, "complex batch needs ldf >= 1");
  hipStream_t s = (hipStream_t)stream;
  if (is_periodic(h)) return periodic_apply(h, psi, hpsi, ldf, is_complex, terms, set_bc, false, s);
  const int L = 1 << ldf;
  const octhip_nl_operator_s* op = h->kin;
  int rc = ensure_work(h, ldf, s);
  if (rc) return rc;
  // boundaries_set: the boundary rows are cleared inside the bra launch when there is one
  if ((terms & OCTHIP_TERM_NON_LOCAL_POTENTIAL) && h->nmat > 0) {
    rc = run_bra(h, psi, L, s, set_bc ? psi : nullptr);
    if (rc) return rc;
  } else if (set_bc) {
    rc = zero_boundary(psi, op->np, op->np_part, L, s);
    if (rc) return rc;
  }
  ApplyArgs a;
  std::memset(&a, 0, sizeof(a));
  fill_terms(h, terms, a);
  a.fi = psi;
  a.fo = hpsi;
  a.L = L;
  a.lim = set_bc ? op->np : op->np_part;
  return launch_operator(h->kin, a, EPI_STORE, s);
}

namespace octhip {
// zfact_k, k = 1..order: zfact *= (-i * deltat) / k (exponential.F90:720-724); deltat ->
// i*deltat in imaginary time (609-619).  z[0] = 1.
void taylor_coefs(double deltat, int imag_time, int order, std::complex<double>* z) {
  const std::complex<double> dt_c = imag_time ? std::complex<double>(0.0, deltat) : std::complex<double>(deltat, 0.0);
  const std::complex<double> mi(-0.0, -1.0);
  const std::complex<double> step(mi.real() * dt_c.real() - mi.imag() * dt_c.imag(),
                                  mi.real() * dt_c.imag() + mi.imag() * dt_c.real());
  z[0] = std::complex<double>(1.0, 0.0);
  for (int k = 1; k <= order; ++k) {
    const std::complex<double> p = z[k - 1];
    const std::complex<double> zz(p.real() * step.real() - p.imag() * step.imag(),
                                  p.real() * step.imag() + p.imag() * step.real());
    z[k] = std::complex<double>(zz.real() / k, zz.imag() / k);
  }
}

// Arguments of the fused pass H `in` for every term, boundary rows of `in` read as zero,
// ket rows taken from `ksum` (a work buffer of the caller's choice).
void h_apply_args(const octhip_hamiltonian_s* h, const double* in, int L, double* ksum, ApplyArgs& a) {
  std::memset(&a, 0, sizeof(a));
  fill_terms(h, OCTHIP_TERM_ALL, a);
  a.fi = in;
  a.L = L;
  a.lim = h->kin->np;
  // fill_terms points ksum at the handle's own buffer, which may not exist yet: the
  // caller's buffer replaces it whenever the nonlocal rows are used
  if (a.kslot) a.ksum = ksum;
}

int prepare_taylor(octhip_hamiltonian_s* h, int ldf, hipStream_t s) {
  int rc = ensure_work(h, ldf, s);
  if (!rc) rc = ensure_taylor(h, ldf, s);
  if (!rc) rc = prepare_operator(h->kin, 1 << ldf, h->nmat > 0 && h->nslot > 0);
  return rc;
}

int set_field(octhip_hamiltonian_s* h, const double field[3], hipStream_t s) {
  for (int d = 0; d < 3; ++d) h->field[d] = field[d];
  h->has_field = true;
  return update_veff(h, s);
}
}  // namespace octhip

extern "C" int octhip_exponential_taylor_batch(octhip_hamiltonian_t h, double* psi, int ldf, double deltat,
                                               int order, int imag_time, void* stream) {
  OCTHIP_RANGE("octhip_exponential_taylor_batch");
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(h && psi, "null argument");
  OCTHIP_REQUIRE(ldf >= 1 && ldf <= MAX_LDF, "Taylor needs a complex batch (ldf in 1..%d)", MAX_LDF);
  OCTHIP_REQUIRE(order >= 1 && order <= 32, "order %d outside 1..32", order);
  if (is_periodic(h)) return taylor_apply(h, psi, ldf, deltat, order, imag_time, TaylorOptions(), (hipStream_t)stream);
  int rc = prepare_taylor(h, ldf, (hipStream_t)stream);
  if (rc) return rc;
  return taylor_step(h, psi, ldf, deltat, order, imag_time, nullptr, (hipStream_t)stream, nullptr);
}

// ---------------------------------------------------------------- graphs --
struct octhip_taylor_graph_s {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // the Hamiltonian whose work sets this graph was captured with, and their keys there
  octhip_hamiltonian_s* h = nullptr;
  std::vector<hipStream_t> set_keys;
};

namespace {
void detach_graph(octhip_taylor_graph_t g) {
  g->h = nullptr;
  g->set_keys.clear();
}
}  // namespace

extern "C" int octhip_taylor_graph_create(octhip_hamiltonian_t h, double* const* psi, int nbatch, int ldf,
                                          double deltat, int order, int imag_time, void* stream,
                                          octhip_taylor_graph_t* out) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(h && psi && out && nbatch > 0, "bad argument");
  OCTHIP_REQUIRE(ldf >= 1 && ldf <= MAX_LDF, "Taylor needs a complex batch (ldf in 1..%d)", MAX_LDF);
  OCTHIP_REQUIRE(order >= 1 && order <= 32, "order %d outside 1..32", order);
  *out = nullptr;
  if (is_periodic(h))
    return fail(OCTHIP_ERR_UNSUPPORTED, "taylor_graph: periodic handles and Bloch phases run eagerly");
  // Batch b runs on branch b mod nbr of the graph (the capturing stream and up to three
  // more, forked and joined with events), as the reference gives every batch its own
  // stream (src/td/propagation_ops_elec.F90:236): the kernels of different batches
  // overlap, each branch with its own work buffers (measured, 4 batches of 32 states on
  // the R=12 sphere: 4.34 ms per step on one branch, 3.84 on two, 3.75 on four).
  constexpr int MAX_BRANCHES = 4;
  const int nbr = std::min(nbatch, MAX_BRANCHES);
  hipStream_t br[MAX_BRANCHES] = {};      // br[0] is the capturing stream, set below
  hipEvent_t fork = nullptr, join[MAX_BRANCHES] = {};
  auto cleanup = [&]() {
    if (fork) (void)hipEventDestroy(fork);
    for (int k = 1; k < nbr; ++k) {
      if (join[k]) (void)hipEventDestroy(join[k]);
      if (br[k]) (void)hipStreamDestroy(br[k]);
    }
  };
  bool ok = nbr == 1 || hipEventCreateWithFlags(&fork, hipEventDisableTiming) == hipSuccess;
  for (int k = 1; k < nbr && ok; ++k)
    ok = hipStreamCreateWithFlags(&br[k], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&join[k], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    cleanup();
    return fail(OCTHIP_ERR_HIP, "taylor_graph_create: stream or event creation failed");
  }
  hipStream_t used = nullptr;
  int rc = capture_graph((hipStream_t)stream,
                       [&](hipStream_t cs) {
                         used = cs;
                         int r = prepare_taylor(h, ldf, cs);
                         for (int k = 1; k < nbr && !r; ++k) r = prepare_taylor(h, ldf, br[k]);
                         return r;
                       },
                       [&](hipStream_t cs) {
                         br[0] = cs;
                         if (nbr > 1) OCTHIP_HIP(hipEventRecord(fork, cs));
                         for (int k = 1; k < nbr; ++k) OCTHIP_HIP(hipStreamWaitEvent(br[k], fork, 0));
                         int r = OCTHIP_SUCCESS;
                         for (int b = 0; b < nbatch && !r; ++b)
                           r = taylor_step(h, psi[b], ldf, deltat, order, imag_time, nullptr, br[b % nbr], nullptr);
                         for (int k = 1; k < nbr; ++k) {
                           OCTHIP_HIP(hipEventRecord(join[k], br[k]));
                           OCTHIP_HIP(hipStreamWaitEvent(cs, join[k], 0));
                         }
                         return r;
                       },
                       out);
  if (rc) {
    // the sets prepared for streams that no longer exist (the branch streams, and the
    // private capturing stream when stream == NULL) go; the caller's stream keeps its own
    for (int k = 1; k < nbr; ++k) drop_workset(h, br[k]);
    if (!stream && used) drop_workset(h, used);
    cleanup();
    return rc;
  }
  // the graph owns the buffers it was captured with (later calls on any stream get
  // others) and frees them when it is destroyed
  octhip_taylor_graph_t g = *out;
  g->h = h;
  for (int k = 0; k < nbr; ++k) {
    const hipStream_t key = pin_workset(h, k == 0 ? used : br[k]);
    if (key) g->set_keys.push_back(key);
  }
  h->graphs.push_back(g);
  cleanup();
  return OCTHIP_SUCCESS;
}

namespace octhip {
// Capture the launches `body` issues on a capturing stream into an instantiated graph.
int capture_graph(hipStream_t s, const std::function<int(hipStream_t)>& prepare,
                  const std::function<int(hipStream_t)>& body, octhip_taylor_graph_t* out) {
  hipStream_t cs = s;
  bool own = false;
  if (!cs) {  // the legacy default stream cannot be captured
    OCTHIP_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    own = true;
  }
  if (prepare) {
    int prc = prepare(cs);
    if (prc) {
      if (own) (void)hipStreamDestroy(cs);
      return prc;
    }
  }
  OCTHIP_HIP(hipStreamSynchronize(cs));
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    if (own) (void)hipStreamDestroy(cs);
    return fail(OCTHIP_ERR_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
  }
  set_capturing(true);
  int rc = body(cs);
  set_capturing(false);
  hipGraph_t graph = nullptr;
  e = hipStreamEndCapture(cs, &graph);
  if (own) (void)hipStreamDestroy(cs);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return fail(OCTHIP_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  auto* g = new octhip_taylor_graph_s();
  g->graph = graph;
  e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(graph);
    delete g;
    return fail(OCTHIP_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  }
  *out = g;
  return OCTHIP_SUCCESS;
}

hipStream_t pin_workset(octhip_hamiltonian_s* h, hipStream_t s) {
  for (auto& w : h->sets)
    if (w.s == s) {
      w.s = reinterpret_cast<hipStream_t>(static_cast<intptr_t>(-(++h->graph_sets)));
      return w.s;
    }
  return nullptr;
}

void drop_workset(octhip_hamiltonian_s* h, hipStream_t s) {
  for (size_t i = 0; i < h->sets.size(); ++i) {
    if (h->sets[i].s != s) continue;
    const auto w = h->sets[i];
    for (void* p : {(void*)w.partial, (void*)w.proj, (void*)w.ksum, (void*)w.y0, (void*)w.y1, (void*)w.y2,
                    (void*)w.dpart, (void*)w.dw, (void*)w.ephase})
      dfree(p);
    h->sets.erase(h->sets.begin() + (long)i);
    // the mirrored pointers may name the freed set: they are rebound by the next call
    h->d_partial = h->d_proj = h->d_ksum = h->d_y[0] = h->d_y[1] = h->d_y[2] = nullptr;
    h->work_ldf = h->y_ldf = -1;
    return;
  }
}
}  // namespace octhip

extern "C" int octhip_taylor_graph_launch(octhip_taylor_graph_t g, void* stream) {
  OCTHIP_RANGE("octhip_taylor_graph_launch");
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(g != nullptr, "null graph");
  ProfScope prof("taylor_graph", (hipStream_t)stream);
  OCTHIP_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_taylor_graph_destroy(octhip_taylor_graph_t g) {
  if (!g) return OCTHIP_SUCCESS;
  if (g->h) {
    // launches of the graph may still be running on any stream
    (void)hipDeviceSynchronize();
    for (hipStream_t key : g->set_keys) drop_workset(g->h, key);
    auto& gs = g->h->graphs;
    gs.erase(std::remove(gs.begin(), gs.end(), g), gs.end());
  }
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return OCTHIP_SUCCESS;
}

namespace octhip {
// One Taylor step in place.  save (optional, order pointers to np-row buffers) receives
// save[0] = psi (boundary cleared) and save[k] = H^k psi for k = 1..order-1: the terms the
// discrete adjoint contracts with the backward states.
int taylor_step(octhip_hamiltonian_s* h, double* psi, int ldf, double deltat, int order, int imag_time,
                double* const* save, hipStream_t s, const double* psi_in, const double* v_override) {
  const int L = 1 << ldf;
  const octhip_nl_operator_s* op = h->kin;
  int rc = prepare_taylor(h, ldf, s);  // the work buffers of stream s (allocated before a capture)
  if (rc) return rc;
  const double* src = psi_in ? psi_in : psi;
  // in place: psi's boundary rows are cleared by the first bra launch (or here when there
  // is no nonlocal term); every pass reads boundary rows as zero (lim = np) in any case.
  // Out of place (psi_in): only the np inner rows of psi are written.
  if (!psi_in && h->nmat == 0) {
    rc = zero_boundary(psi, op->np, op->np_part, L, s);
    if (rc) return rc;
  }
  if (save && save[0] != src)
    OCTHIP_HIP(hipMemcpyAsync(save[0], src, sizeof(double) * op->np * L, hipMemcpyDeviceToDevice, s));
  if (psi_in && order == 1)  // the order-1 axpy works in place on psi
    OCTHIP_HIP(hipMemcpyAsync(psi, psi_in, sizeof(double) * op->np * L, hipMemcpyDeviceToDevice, s));
  std::complex<double> zf[33];
  taylor_coefs(deltat, imag_time, order, zf);
  if (taylor_deferred(order)) {
    // y_k = H y_{k-1} (y_0 = psi) stored for k < order (save[k] when saving); the last
    // pass forms psi + z_1 y_1 + ... + z_order H y_{order-1}, adding the terms in order k,
    // as exponential.F90:733-744 accumulates them order by order.  No accumulator is
    // read and written at the orders in between.
    auto Y = [&](int k) -> double* { return save ? save[k] : h->d_y[k - 1]; };
    for (int k = 1; k <= order; ++k) {
      const double* yin = (k == 1) ? src : Y(k - 1);
      rc = run_bra(h, yin, L, s, (k == 1 && h->nmat > 0 && !psi_in) ? psi : nullptr);
      if (rc) return rc;
      ApplyArgs a;
      std::memset(&a, 0, sizeof(a));
      fill_terms(h, OCTHIP_TERM_ALL, a);
      if (v_override) a.v = v_override;
      a.fi = yin;
      a.L = L;
      a.lim = op->np;  // boundary rows are zero: psi's were just cleared, work buffers have none
      if (k < order) {
        a.fo = Y(k);
        rc = launch_operator(h->kin, a, EPI_STORE, s);
      } else {
        a.acc = psi;
        a.acc_in = psi_in;  // null: in place
        if (order >= 3) {
          a.ya = Y(1);
          a.car = zf[1].real();
          a.cai = zf[1].imag();
        }
        if (order >= 4) {
          a.yb = Y(2);
          a.cbr = zf[2].real();
          a.cbi = zf[2].imag();
        }
        a.has_c1 = 1;  // the pass's own input y_{order-1}
        a.c1r = zf[order - 1].real();
        a.c1i = zf[order - 1].imag();
        a.c2r = zf[order].real();
        a.c2i = zf[order].imag();
        rc = launch_operator(h->kin, a, EPI_TAYLOR, s);
      }
      if (rc) return rc;
    }
    return OCTHIP_SUCCESS;
  }
  std::complex<double> z, z1;
  for (int k = 1; k <= order; ++k) {
    z = zf[k];
    const double* yin = (k == 1) ? src : (save ? save[k - 1] : h->d_y[(k - 2) & 1]);
    double* yout = (k < order) ? (save ? save[k] : h->d_y[(k - 1) & 1]) : nullptr;
    rc = run_bra(h, yin, L, s, (k == 1 && h->nmat > 0 && !psi_in) ? psi : nullptr);
    if (rc) return rc;
    ApplyArgs a;
    std::memset(&a, 0, sizeof(a));
    fill_terms(h, OCTHIP_TERM_ALL, a);
    if (v_override) a.v = v_override;
    a.fi = yin;
    a.L = L;
    a.lim = op->np;  // boundary rows are zero: psi's were just cleared, work buffers have none
    if (order == 1) {
      a.fo = h->d_y[0];
      rc = launch_operator(h->kin, a, EPI_STORE, s);
      if (rc) return rc;
      const size_t count = (size_t)op->np * L;
      size_t grid = std::min<size_t>((count + BLOCK - 1) / BLOCK, 8192);
      ProfScope prof("k_zaxpy", s);
      hipLaunchKernelGGL(k_zaxpy, dim3((unsigned)grid), dim3(BLOCK), 0, s, psi, h->d_y[0], count, z.real(), z.imag());
      OCTHIP_HIP(hipGetLastError());
      break;
    }
    a.yout = yout;
    if (k == 1) {
      a.acc = nullptr;
      z1 = z;
    } else {
      a.acc = psi;
      a.acc_in = (k == 2) ? psi_in : nullptr;  // from the third order on psi holds the sum
      a.has_c1 = (k == 2);
      a.c1r = z1.real();
      a.c1i = z1.imag();
      a.c2r = z.real();
      a.c2i = z.imag();
    }
    rc = launch_operator(h->kin, a, EPI_TAYLOR, s);
    if (rc) return rc;
  }
  return OCTHIP_SUCCESS;
}
}  // namespace octhip

namespace octhip {
namespace {
// density buffers of stream s: nchunk x np partials and the weights of 2^(ldf-1) states
int ensure_density(octhip_hamiltonian_s* h, int nchunk, hipStream_t s, double** dpart, double** dw) {
  auto& w = workset(h, s);
  if (!w.dw) {
    int rc = dalloc((void**)&w.dw, sizeof(double) << (MAX_LDF - 1));
    if (rc) return rc;
  }
  if (nchunk > 0 && w.dchunks < nchunk) {
    if (w.dpart) h->retired.push_back(w.dpart);
    w.dpart = nullptr;
    w.dchunks = 0;
    int rc = dalloc((void**)&w.dpart, sizeof(double) * (size_t)nchunk * h->np);
    if (rc) return rc;
    w.dchunks = nchunk;
  }
  *dpart = w.dpart;
  *dw = w.dw;
  return OCTHIP_SUCCESS;
}

// zfact_k of exponential_taylor_series_batch (src/electrons/exponential.F90:716-720) for a
// complex deltat: zfact_0 (1, or deltat with an inhomogeneous term) and
// zfact_k = zfact_{k-1} (-i deltat) / (k [+1]), with the products written out as the
// oracle does.
void series_coefs(std::complex<double> dt, bool inh, int order, std::complex<double>* z) {
  const double mr = -0.0, mi = -1.0;
  const double sr = mr * dt.real() - mi * dt.imag(), si = mr * dt.imag() + mi * dt.real();
  z[0] = inh ? dt : std::complex<double>(1.0, 0.0);
  for (int k = 1; k <= order; ++k) {
    const int denom = inh ? k + 1 : k;
    const double tr = z[k - 1].real() * sr - z[k - 1].imag() * si;
    const double ti = z[k - 1].real() * si + z[k - 1].imag() * sr;
    z[k] = std::complex<double>(tr / denom, ti / denom);
  }
}

// One series of exponential_taylor_series_batch, kernel by kernel: psi2 = psi (np rows)
// first when given; with inh, psi (psi2) += real(zfact_0) inh and the chain of H-applies
// starts from inh; each order stores y_k = H y_{k-1} and adds zfact_k y_k to psi and
// zfact2_k y_k to psi2 (the reference's axpy per order, 733-744).
int series_generic(octhip_hamiltonian_s* h, double* psi, int ldf, std::complex<double> dt, int order,
                   double* psi2, std::complex<double> dt2, double* inh, hipStream_t s) {
  const int L = 1 << ldf;
  const size_t count = (size_t)h->np * L;
  int rc = prepare_taylor(h, ldf, s);
  if (rc) return rc;
  std::complex<double> z[34], z2[34];
  series_coefs(dt, inh != nullptr, order, z);
  series_coefs(dt2, inh != nullptr, order, z2);
  if (psi2) OCTHIP_HIP(hipMemcpyAsync(psi2, psi, sizeof(double) * count, hipMemcpyDeviceToDevice, s));
  if (inh) {
    rc = launch_daxpy(psi, inh, count, z[0].real(), s);
    if (!rc && psi2) rc = launch_daxpy(psi2, inh, count, z2[0].real(), s);
    if (rc) return rc;
  }
  double* src = inh ? inh : psi;
  // periodic handles: the chain needs np_part rows (its boundary rows are filled from the
  // periodic images by every apply), taken from the stream's Krylov buffers
  double* ychain[2] = {h->d_y[0], h->d_y[1]};
  if (is_periodic(h)) {
    double* base = nullptr;
    rc = krylov_vectors(h, s, 2, ldf, &base);
    if (rc) return rc;
    ychain[0] = base;
    ychain[1] = base + ((size_t)h->kin->np_part << ldf);
  }
  for (int k = 1; k <= order; ++k) {
    double* yin = (k == 1) ? src : ychain[(k - 2) & 1];
    double* yout = ychain[(k - 1) & 1];
    if (is_periodic(h)) {
      rc = propagator_apply(h, yin, yout, ldf, s);
    } else {
      // boundaries_set of the chain's input (its H-apply, set_bc default)
      rc = run_bra(h, yin, L, s, k == 1 ? src : nullptr);
      if (rc) return rc;
      if (k == 1 && h->nmat == 0) {
        rc = zero_boundary(src, h->np, h->kin->np_part, L, s);
        if (rc) return rc;
      }
      ApplyArgs a;
      std::memset(&a, 0, sizeof(a));
      fill_terms(h, OCTHIP_TERM_ALL, a);
      a.fi = yin;
      a.fo = yout;
      a.L = L;
      a.lim = h->np;
      rc = launch_operator(h->kin, a, EPI_STORE, s);
    }
    if (!rc) rc = launch_zaxpy(psi, yout, count, z[k].real(), z[k].imag(), s);
    if (!rc && psi2) rc = launch_zaxpy(psi2, yout, count, z2[k].real(), z2[k].imag(), s);
    if (rc) return rc;
  }
  return OCTHIP_SUCCESS;
}

// The homogeneous series at orders 2..4 with the deferred accumulation: plain H-apply
// passes store y_k = H^k psi (k < order); the last pass forms psi (and psi2 with the
// deltat2 factors, from the same inputs) and, when dpart is given, the density partials of
// the final states.
int series_fused(octhip_hamiltonian_s* h, double* psi, int ldf, std::complex<double> dt, int order,
                 double* psi2, std::complex<double> dt2, double* dpart, const double* dw, hipStream_t s) {
  const int L = 1 << ldf;
  int rc = prepare_taylor(h, ldf, s);
  if (rc) return rc;
  std::complex<double> z[34], z2[34];
  series_coefs(dt, false, order, z);
  series_coefs(dt2, false, order, z2);
  if (h->nmat == 0) {
    rc = zero_boundary(psi, h->np, h->kin->np_part, L, s);
    if (rc) return rc;
  }
  for (int k = 1; k <= order; ++k) {
    const double* yin = (k == 1) ? psi : h->d_y[k - 2];
    rc = run_bra(h, yin, L, s, (k == 1 && h->nmat > 0) ? psi : nullptr);
    if (rc) return rc;
    ApplyArgs a;
    std::memset(&a, 0, sizeof(a));
    fill_terms(h, OCTHIP_TERM_ALL, a);
    a.fi = yin;
    a.L = L;
    a.lim = h->np;
    if (k < order) {
      a.fo = h->d_y[k - 1];
      rc = launch_operator(h->kin, a, EPI_STORE, s);
      if (rc) return rc;
      continue;
    }
    a.acc = psi;
    if (order >= 3) {
      a.ya = h->d_y[0];
      a.car = z[1].real();
      a.cai = z[1].imag();
      a.dar = z2[1].real();
      a.dai = z2[1].imag();
    }
    if (order >= 4) {
      a.yb = h->d_y[1];
      a.cbr = z[2].real();
      a.cbi = z[2].imag();
      a.dbr = z2[2].real();
      a.dbi = z2[2].imag();
    }
    a.has_c1 = 1;
    a.c1r = z[order - 1].real();
    a.c1i = z[order - 1].imag();
    a.c2r = z[order].real();
    a.c2i = z[order].imag();
    a.d1r = z2[order - 1].real();
    a.d1i = z2[order - 1].imag();
    a.d2r = z2[order].real();
    a.d2i = z2[order].imag();
    a.acc2 = psi2;
    a.dpart = dpart;
    a.dw = dw;
    return launch_operator(h->kin, a, EPI_TAYLOR_X, s);
  }
  return OCTHIP_SUCCESS;
}
}  // namespace

int taylor_apply(octhip_hamiltonian_s* h, double* psi, int ldf, double deltat, int order, int imag_time,
                 const TaylorOptions& o, hipStream_t s) {
  const int L = 1 << ldf;
  const std::complex<double> dt = imag_time ? std::complex<double>(0.0, deltat) : std::complex<double>(deltat, 0.0);
  const std::complex<double> dt2 =
      imag_time ? std::complex<double>(0.0, o.deltat2) : std::complex<double>(o.deltat2, 0.0);
  double* dpart = nullptr;
  double* dw = nullptr;
  int rc = OCTHIP_SUCCESS;
  // periodic handles (psi carries the Bloch phase): plain passes through the periodic apply
  const bool fused = taylor_deferred(order) && !is_periodic(h);
  const bool fuse_density = o.density && fused && !o.inh;
  int nchunk = 0;
  if (o.density) {
    // density partials per lane chunk when the last pass forms them, else none
    ApplyArgs probe;
    std::memset(&probe, 0, sizeof(probe));
    fill_terms(h, OCTHIP_TERM_ALL, probe);
    probe.L = L;
    nchunk = fuse_density ? apply_lane_chunks(h->kin, probe) : 0;
    rc = ensure_density(h, nchunk, s, &dpart, &dw);
    if (rc) return rc;
    OCTHIP_HIP(hipMemcpyAsync(dw, o.weights, sizeof(double) * (L / 2), hipMemcpyHostToDevice, s));
  }
  if (fused) {
    rc = series_fused(h, psi, ldf, dt, order, o.psi2, dt2, fuse_density ? dpart : nullptr, dw, s);
  } else {
    rc = series_generic(h, psi, ldf, dt, order, o.psi2, dt2, nullptr, s);
  }
  if (!rc && o.inh) rc = series_generic(h, psi, ldf, dt, order, o.psi2, dt2, o.inh, s);
  if (rc || !o.density) return rc;
  if (fuse_density) return launch_density_reduce(o.density, dpart, nchunk, h->np, s);
  return launch_density(o.psi2 ? o.psi2 : psi, L, h->np, 1, dw, o.density, s);
}
}  // namespace octhip

extern "C" int octhip_exponential_apply_batch(octhip_hamiltonian_t h, double* psi, int ldf, double deltat,
                                              int order, int imag_time, double* psi2, double deltat2,
                                              double* inh, const double* weights, double* density,
                                              void* stream) {
  OCTHIP_RANGE("octhip_exponential_apply_batch");
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(h && psi, "null argument");
  OCTHIP_REQUIRE(ldf >= 1 && ldf <= MAX_LDF, "Taylor needs a complex batch (ldf in 1..%d)", MAX_LDF);
  OCTHIP_REQUIRE(order >= 1 && order <= 32, "order %d outside 1..32", order);
  OCTHIP_REQUIRE(psi2 != psi && inh != psi && (!psi2 || psi2 != inh), "psi, psi2 and inh must be distinct");
  OCTHIP_REQUIRE((weights == nullptr) == (density == nullptr), "weights and density go together");
  TaylorOptions o;
  o.psi2 = psi2;
  o.deltat2 = deltat2;
  o.inh = inh;
  o.weights = weights;
  o.density = density;
  return taylor_apply(h, psi, ldf, deltat, order, imag_time, o, (hipStream_t)stream);
}

extern "C" int octhip_density_accumulate(const double* psi, int ldf, int np, int is_complex, const double* weights,
                                         double* density, void* stream) {
  OCTHIP_RANGE("octhip_density_accumulate");
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(psi && weights && density, "null argument");
  OCTHIP_REQUIRE(ldf >= (is_complex ? 1 : 0) && ldf <= MAX_LDF, "ldf=%d outside the packed range", ldf);
  OCTHIP_REQUIRE(np >= 0, "np < 0");
  return launch_density(psi, 1 << ldf, np, is_complex ? 1 : 0, weights, density, (hipStream_t)stream);
}

namespace {
// density = 0 over np doubles with a fill kernel (density_calc_init)
int clear_density(double* density, int np, hipStream_t s) {
  const unsigned grid = (unsigned)std::min((np + BLOCK - 1) / BLOCK, 8192);
  hipLaunchKernelGGL(k_zero_rows, dim3(grid), dim3(BLOCK), 0, s, density, (size_t)0, (size_t)np);
  OCTHIP_HIP(hipGetLastError());
  return OCTHIP_SUCCESS;
}

// propagation_ops_elec_exp_apply / _fuse_density_exp_apply over the batches
// (src/td/propagation_ops_elec.F90:213-333): every batch propagated by dt in turn, the
// density of the results accumulated when requested
int exp_apply_batches(octhip_hamiltonian_s* h, double* const* psi, int nbatch, int ldf, double dt, int order,
                      const double* weights, double* density, hipStream_t s) {
  const int pack = 1 << (ldf - 1);
  if (density) {
    int rc = clear_density(density, h->np, s);
    if (rc) return rc;
  }
  for (int b = 0; b < nbatch; ++b) {
    TaylorOptions o;
    if (density) {
      o.weights = weights + (size_t)b * pack;
      o.density = density;
    }
    // k-points: set_phase_corr / unset_phase_corr around the apply
    // (src/td/propagation_ops_elec.F90:238-246)
    int rc = h->has_kpoint ? launch_phase(psi[b], psi[b], h->d_phase, h->np, 1 << ldf, false, s) : 0;
    if (!rc) rc = taylor_apply(h, psi[b], ldf, dt, order, 0, o, s);
    if (!rc && h->has_kpoint) rc = launch_phase(psi[b], psi[b], h->d_phase, h->np, 1 << ldf, true, s);
    if (rc) return rc;
  }
  return OCTHIP_SUCCESS;
}

int check_driver_args(octhip_hamiltonian_s* h, double* const* psi, int nbatch, int ldf, int order,
                      const double* weights, double* density) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(h && psi && nbatch > 0, "bad argument");
  OCTHIP_REQUIRE(ldf >= 1 && ldf <= MAX_LDF, "propagators need complex batches (ldf in 1..%d)", MAX_LDF);
  OCTHIP_REQUIRE(order >= 1 && order <= 32, "order %d outside 1..32", order);
  OCTHIP_REQUIRE((weights == nullptr) == (density == nullptr), "weights and density go together");
  for (int b = 0; b < nbatch; ++b) OCTHIP_REQUIRE(psi[b] != nullptr, "batch %d is null", b);
  return OCTHIP_SUCCESS;
}
}  // namespace

extern "C" int octhip_propagator_expmid(octhip_hamiltonian_t h, double* const* psi, int nbatch, int ldf, double dt,
                                        int order, const double* field_mid, const double* weights, double* density,
                                        void* stream) {
  OCTHIP_RANGE("octhip_propagator_expmid");
  int rc = check_driver_args(h, psi, nbatch, ldf, order, weights, density);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (field_mid) {
    OCTHIP_REQUIRE(h->d_xyz != nullptr, "propagator_expmid: positions never uploaded (set_dipole_field)");
    rc = set_field(h, field_mid, s);
    if (rc) return rc;
  }
  return exp_apply_batches(h, psi, nbatch, ldf, dt, order, weights, density, s);
}

extern "C" int octhip_propagator_etrs(octhip_hamiltonian_t h, double* const* psi, int nbatch, int ldf, double dt,
                                     int order, const double* field_old, const double* field_new,
                                     const double* weights, double* density, void* stream) {
  OCTHIP_RANGE("octhip_propagator_etrs");
  int rc = check_driver_args(h, psi, nbatch, ldf, order, weights, density);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  OCTHIP_REQUIRE(!(field_old || field_new) || h->d_xyz != nullptr,
                 "propagator_etrs: positions never uploaded (set_dipole_field)");
  if (field_old) {
    rc = set_field(h, field_old, s);
    if (rc) return rc;
  }
  // dt/2 with H(t - dt), then the Hamiltonian at t and dt/2 with the density
  rc = exp_apply_batches(h, psi, nbatch, ldf, 0.5 * dt, order, nullptr, nullptr, s);
  if (rc) return rc;
  if (field_new) {
    rc = set_field(h, field_new, s);
    if (rc) return rc;
  }
  return exp_apply_batches(h, psi, nbatch, ldf, 0.5 * dt, order, weights, density, s);
}

// ------------------------------------------------------- periodic / k-points --
extern "C" int octhip_hamiltonian_set_periodic(octhip_hamiltonian_t h, const int* bnd_src) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(h != nullptr, "null hamiltonian handle");
  // a captured Taylor graph keeps the zero-boundary launches it was built with
  OCTHIP_REQUIRE(h->graphs.empty(), "destroy the Taylor graphs of this Hamiltonian before changing its boundaries");
  const int np = h->np, nb = h->kin->np_part - h->np;
  (void)hipDeviceSynchronize();  // launches in flight may read the old map
  dfree(h->d_bnd_src);
  h->d_bnd_src = nullptr;
  h->nbnd = 0;
  h->periodic = false;
  if (!bnd_src) return OCTHIP_SUCCESS;
  for (int b = 0; b < nb; ++b)
    OCTHIP_REQUIRE(bnd_src[b] >= 0 && bnd_src[b] < np, "boundary row %d: image %d outside 0..np-1", np + b,
                   bnd_src[b]);
  int rc = upload(&h->d_bnd_src, bnd_src, (size_t)nb, nullptr);
  if (rc) return rc;
  h->nbnd = nb;
  h->periodic = true;
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_hamiltonian_set_phases(octhip_hamiltonian_t h, const double* phase, const double* phase_corr,
                                             const double* projector_phases) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(h != nullptr, "null hamiltonian handle");
  OCTHIP_REQUIRE(h->graphs.empty(), "destroy the Taylor graphs of this Hamiltonian before setting Bloch phases");
  const int np_part = h->kin->np_part, nb = np_part - h->np;
  (void)hipDeviceSynchronize();
  for (double** p : {&h->d_phase, &h->d_phase_corr, &h->d_pphase}) {
    dfree(*p);
    *p = nullptr;
  }
  h->has_kpoint = false;
  if (!phase) return OCTHIP_SUCCESS;
  OCTHIP_REQUIRE(nb == 0 || phase_corr, "phase_corr is required with the phase (one per boundary row)");
  OCTHIP_REQUIRE(h->nmat == 0 || projector_phases, "projector_phases are required with projectors");
  int rc = upload(&h->d_phase, phase, 2 * (size_t)np_part, nullptr);
  if (!rc && nb > 0) rc = upload(&h->d_phase_corr, phase_corr, 2 * (size_t)nb, nullptr);
  if (!rc && h->nmat > 0) rc = upload(&h->d_pphase, projector_phases, 2 * (size_t)h->total_points, nullptr);
  if (rc) return rc;
  h->has_kpoint = true;
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_hamiltonian_apply_batch_phase(octhip_hamiltonian_t h, double* psi, double* hpsi, int ldf,
                                                    int terms, int set_bc, int psi_has_phase, void* stream) {
  OCTHIP_RANGE("octhip_hamiltonian_apply_batch_phase");
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(h != nullptr, "null hamiltonian handle");
  OCTHIP_REQUIRE(psi && hpsi && psi != hpsi, "bad batch pointers");
  OCTHIP_REQUIRE(ldf >= 1 && ldf <= MAX_LDF, "a complex batch is needed (ldf in 1..%d)", MAX_LDF);
  OCTHIP_REQUIRE(!psi_has_phase || h->has_kpoint, "psi_has_phase without Bloch phases on the handle");
  OCTHIP_REQUIRE(set_bc || !psi_has_phase || !h->periodic,
                 "a batch with the phase needs set_bc (the phase correction is applied while setting the boundary)");
  hipStream_t s = (hipStream_t)stream;
  if (!is_periodic(h)) return octhip_hamiltonian_apply_batch(h, psi, hpsi, ldf, 1, terms, set_bc, stream);
  return periodic_apply(h, psi, hpsi, ldf, 1, terms, set_bc, psi_has_phase != 0, s);
}

extern "C" int octhip_hamiltonian_base_phase(octhip_hamiltonian_t h, double* psi, const double* src, int ldf, int np,
                                             int conjugate, void* stream) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(h && psi, "null argument");
  OCTHIP_REQUIRE(h->has_kpoint, "no Bloch phases on this handle");
  OCTHIP_REQUIRE(ldf >= 1 && ldf <= MAX_LDF, "a complex batch is needed (ldf in 1..%d)", MAX_LDF);
  OCTHIP_REQUIRE(np >= 0 && np <= h->kin->np_part, "np=%d outside 0..np_part", np);
  return launch_phase(psi, src ? src : psi, h->d_phase, np, 1 << ldf, conjugate != 0, (hipStream_t)stream);
}
