// Autotuner of the H-apply launch shape and of the states block size.
//
// For one Hamiltonian and one row length (ldf) it times every candidate launch shape of
// the fused stencil pass (z-march lane chunk and lanes per thread, the z segment left to
// the per-launch choice; the tiled gather with its tile depth) and both forms of the
// projector bra (FP64 MFMA, plain) on scratch batches, with HIP events on the caller's
// stream, and leaves the fastest set on the handle.  Results are kept in a process-wide
// table keyed by the problem shape (np, np_part, projector matrices, ldf), so a second
// Hamiltonian of the same shape is
// tuned without timing; the table is saved to / loaded from a text file.  Every candidate
// computes the same result (the per-lane arithmetic does not depend on the launch shape,
// except the bra's summation order between MFMA and plain, both within the H-apply
// tolerance), so tuning never changes more than rounding in the projector sums.
#include <cstdio>
#include <cstring>
#include <vector>

#include "octhip_internal.h"

using namespace octhip;

namespace {

struct TuneKey {
  int np, np_part, nmat, ldf;
  bool operator==(const TuneKey& o) const {
    return np == o.np && np_part == o.np_part && nmat == o.nmat && ldf == o.ldf;
  }
};

struct TuneVal {
  int kernel, lane_chunk, lpt, zseg, gather_tile, bra;
  double us;  // one H-apply of one batch
};

std::vector<std::pair<TuneKey, TuneVal>> g_table;

// block size (states per packed batch) per (mesh, projectors, number of states)
struct PackKey {
  int np, np_part, nmat, nst;
  bool operator==(const PackKey& o) const {
    return np == o.np && np_part == o.np_part && nmat == o.nmat && nst == o.nst;
  }
};
struct PackVal {
  int pack;
  double us;  // one Taylor-4 step of all nst states in batches of `pack`
};
std::vector<std::pair<PackKey, PackVal>> g_packs;

TuneVal* lookup(const TuneKey& k) {
  for (auto& e : g_table)
    if (e.first == k) return &e.second;
  return nullptr;
}

__global__ void k_fill_pattern(double* __restrict__ f, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    f[i] = 1.0 / (1.0 + (double)(i % 977));
}

int apply_config(octhip_hamiltonian_s* h, const TuneVal& v) {
  int rc = octhip_hamiltonian_set_tuning(h, OCTHIP_TUNE_KERNEL, v.kernel);
  if (!rc) rc = octhip_hamiltonian_set_tuning(h, OCTHIP_TUNE_LANE_CHUNK, v.lane_chunk);
  if (!rc) rc = octhip_hamiltonian_set_tuning(h, OCTHIP_TUNE_LANES_PER_THREAD, v.lpt);
  if (!rc) rc = octhip_hamiltonian_set_tuning(h, OCTHIP_TUNE_ZSEG, v.zseg);
  if (!rc) rc = octhip_hamiltonian_set_tuning(h, OCTHIP_TUNE_GATHER_TILE, v.gather_tile);
  if (!rc) rc = octhip_hamiltonian_set_tuning(h, OCTHIP_TUNE_BRA_MODE, v.bra);
  return rc;
}

// average microseconds of one H-apply with the handle's current configuration
int time_apply(octhip_hamiltonian_s* h, double* psi, double* hpsi, int ldf, int reps, hipStream_t s,
               hipEvent_t e0, hipEvent_t e1, double* us) {
  int rc = octhip_hamiltonian_apply_batch(h, psi, hpsi, ldf, 1, OCTHIP_TERM_ALL, 1, s);  // warm-up
  if (rc) return rc;
  OCTHIP_HIP(hipEventRecord(e0, s));
  for (int r = 0; r < reps && !rc; ++r) rc = octhip_hamiltonian_apply_batch(h, psi, hpsi, ldf, 1, OCTHIP_TERM_ALL, 1, s);
  if (rc) return rc;
  OCTHIP_HIP(hipEventRecord(e1, s));
  OCTHIP_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  OCTHIP_HIP(hipEventElapsedTime(&ms, e0, e1));
  *us = 1e3 * ms / reps;
  return OCTHIP_SUCCESS;
}

}  // namespace

extern "C" int octhip_hamiltonian_autotune(octhip_hamiltonian_t h, int ldf, int reps, int force, void* stream,
                                           double* best_us) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(h != nullptr, "null hamiltonian handle");
  OCTHIP_REQUIRE(ldf >= 1 && ldf <= MAX_LDF, "autotune needs a complex row length (ldf in 1..%d)", MAX_LDF);
  OCTHIP_REQUIRE(reps >= 1 && reps <= 1000, "reps %d outside 1..1000", reps);
  const octhip_nl_operator_s* op = h->kin;
  const TuneKey key{op->np, op->np_part, h->nmat, ldf};
  if (!force) {
    if (const TuneVal* v = lookup(key)) {
      if (best_us) *best_us = v->us;
      return apply_config(h, *v);
    }
  }
  hipStream_t s = (hipStream_t)stream;
  const int L = 1 << ldf;
  double* psi = nullptr;
  double* hpsi = nullptr;
  int rc = dalloc((void**)&psi, sizeof(double) * (size_t)op->np_part * L);
  if (!rc) rc = dalloc((void**)&hpsi, sizeof(double) * (size_t)op->np * L);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(OCTHIP_ERR_HIP, "autotune: hipEventCreate failed");
  if (!rc) {
    hipLaunchKernelGGL(k_fill_pattern, dim3(2048), dim3(256), 0, s, psi, (size_t)op->np_part * L);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = fail(OCTHIP_ERR_HIP, "autotune fill: %s", hipGetErrorString(e));
  }
  // candidates of the stencil pass (bra fixed to MFMA where it applies), then the bra.
  // The z segment stays automatic (0): its choice depends on the epilogue (the Taylor
  // instance keeps one workgroup per CU, the H-apply instance two), and this timing sees
  // only the H-apply.
  std::vector<TuneVal> cands;
  const int bra0 = L >= 16 ? 1 : 0;
  if (op->zmarch_ok && op->n == 25) {
    for (int lc : {8, 16, 32})
      for (int lpt : {1, 2}) {
        if (lc > L || lpt > lc || (lc == 32 && lpt == 1)) continue;
        cands.push_back(TuneVal{2, lc, lpt, 0, op->tile_iters, bra0, 0.0});
      }
  }
  for (int gt : {2, 4}) cands.push_back(TuneVal{1, 0, 0, op->zseg, gt, bra0, 0.0});
  TuneVal best{};
  best.us = 1e300;
  for (size_t i = 0; i < cands.size() && !rc; ++i) {
    rc = apply_config(h, cands[i]);
    if (!rc) rc = time_apply(h, psi, hpsi, ldf, reps, s, e0, e1, &cands[i].us);
    if (!rc && cands[i].us < best.us) best = cands[i];
  }
  if (!rc && bra0 == 1 && h->nmat > 0) {
    TuneVal plain = best;
    plain.bra = 0;
    rc = apply_config(h, plain);
    if (!rc) rc = time_apply(h, psi, hpsi, ldf, reps, s, e0, e1, &plain.us);
    if (!rc && plain.us < best.us) best = plain;
  }
  if (!rc) rc = apply_config(h, best);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(s);
  dfree(psi);
  dfree(hpsi);
  if (rc) return rc;
  if (TuneVal* v = lookup(key)) *v = best;
  else g_table.push_back({key, best});
  if (best_us) *best_us = best.us;
  return OCTHIP_SUCCESS;
}

// StatesBlockSize (src/states/states_elec.F90:447-475: 32 on GPUs, at most pad_pow2(nst))
// chosen by time: for each power-of-two block size from 8 to 64 (up to pad_pow2(nst)),
// the launch shape is tuned for that row length and one Taylor-4 step of a scratch batch
// is timed; the cost of nst states is that time times the number of batches.
extern "C" int octhip_hamiltonian_autotune_pack(octhip_hamiltonian_t h, int nst, int reps, int force, void* stream,
                                                int* pack_size, double* step_us) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(h != nullptr && pack_size != nullptr, "null argument");
  OCTHIP_REQUIRE(nst >= 1, "nst %d < 1", nst);
  OCTHIP_REQUIRE(reps >= 1 && reps <= 1000, "reps %d outside 1..1000", reps);
  const octhip_nl_operator_s* op = h->kin;
  const PackKey key{op->np, op->np_part, h->nmat, nst};
  hipStream_t s = (hipStream_t)stream;
  if (!force)
    for (const auto& e : g_packs)
      if (e.first == key) {
        *pack_size = e.second.pack;
        if (step_us) *step_us = e.second.us;
        int lg = 0;
        while ((1 << lg) < 2 * e.second.pack) ++lg;
        return octhip_hamiltonian_autotune(h, lg, reps, 0, stream, nullptr);
      }
  int top = 8;
  while (top < nst && top < 64) top *= 2;
  const int Lmax = 2 * top;
  double* psi = nullptr;
  int rc = dalloc((void**)&psi, sizeof(double) * (size_t)op->np_part * Lmax);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(OCTHIP_ERR_HIP, "autotune_pack: hipEventCreate failed");
  PackVal best{0, 1e300};
  int best_ldf = 0;
  for (int pack = 8; pack <= top && !rc; pack *= 2) {
    int ldf = 0;
    while ((1 << ldf) < 2 * pack) ++ldf;
    const int L = 1 << ldf;
    rc = octhip_hamiltonian_autotune(h, ldf, reps, 0, stream, nullptr);
    if (rc) break;
    hipLaunchKernelGGL(k_fill_pattern, dim3(2048), dim3(256), 0, s, psi, (size_t)op->np_part * L);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      rc = fail(OCTHIP_ERR_HIP, "autotune_pack fill: %s", hipGetErrorString(e));
      break;
    }
    // a short step keeps the scratch values bounded over the repetitions
    rc = octhip_exponential_taylor_batch(h, psi, ldf, 1e-4, 4, 0, stream);  // warm-up
    if (!rc && hipEventRecord(e0, s) != hipSuccess) rc = fail(OCTHIP_ERR_HIP, "autotune_pack: event record");
    for (int r = 0; r < reps && !rc; ++r) rc = octhip_exponential_taylor_batch(h, psi, ldf, 1e-4, 4, 0, stream);
    if (!rc && hipEventRecord(e1, s) != hipSuccess) rc = fail(OCTHIP_ERR_HIP, "autotune_pack: event record");
    if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = fail(OCTHIP_ERR_HIP, "autotune_pack: event sync");
    float ms = 0.f;
    if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = fail(OCTHIP_ERR_HIP, "autotune_pack: elapsed");
    if (rc) break;
    const double cost = (double)((nst + pack - 1) / pack) * 1e3 * ms / reps;
    if (cost < best.us) {
      best = PackVal{pack, cost};
      best_ldf = ldf;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(s);
  dfree(psi);
  if (rc) return rc;
  rc = octhip_hamiltonian_autotune(h, best_ldf, reps, 0, stream, nullptr);  // leave its shape configured
  if (rc) return rc;
  bool found = false;
  for (auto& e : g_packs)
    if (e.first == key) {
      e.second = best;
      found = true;
    }
  if (!found) g_packs.push_back({key, best});
  *pack_size = best.pack;
  if (step_us) *step_us = best.us;
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_tuning_save(const char* path) {
  OCTHIP_REQUIRE(path != nullptr, "null path");
  FILE* f = std::fopen(path, "w");
  if (!f) return fail(OCTHIP_ERR_INVALID, "cannot write %s", path);
  std::fprintf(f, "# np np_part nmat ldf kernel lane_chunk lanes_per_thread zseg gather_tile bra us\n");
  for (const auto& e : g_table)
    std::fprintf(f, "%d %d %d %d %d %d %d %d %d %d %.3f\n", e.first.np, e.first.np_part, e.first.nmat, e.first.ldf,
                 e.second.kernel, e.second.lane_chunk, e.second.lpt, e.second.zseg, e.second.gather_tile,
                 e.second.bra, e.second.us);
  std::fprintf(f, "# pack np np_part nmat nst block_size us_per_step\n");
  for (const auto& e : g_packs)
    std::fprintf(f, "pack %d %d %d %d %d %.3f\n", e.first.np, e.first.np_part, e.first.nmat, e.first.nst,
                 e.second.pack, e.second.us);
  std::fclose(f);
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_tuning_load(const char* path) {
  OCTHIP_REQUIRE(path != nullptr, "null path");
  FILE* f = std::fopen(path, "r");
  if (!f) return fail(OCTHIP_ERR_INVALID, "cannot read %s", path);
  char line[512];
  int n = 0;
  while (std::fgets(line, sizeof(line), f)) {
    if (line[0] == '#') continue;
    if (std::strncmp(line, "pack ", 5) == 0) {
      PackKey k;
      PackVal v;
      if (std::sscanf(line + 5, "%d %d %d %d %d %lf", &k.np, &k.np_part, &k.nmat, &k.nst, &v.pack, &v.us) != 6 ||
          v.pack < 1 || v.pack > 512 || (v.pack & (v.pack - 1)) != 0) {
        std::fclose(f);
        return fail(OCTHIP_ERR_INVALID, "%s: malformed block-size line %d", path, n + 1);
      }
      bool found = false;
      for (auto& e : g_packs)
        if (e.first == k) {
          e.second = v;
          found = true;
        }
      if (!found) g_packs.push_back({k, v});
      ++n;
      continue;
    }
    TuneKey k;
    TuneVal v;
    if (std::sscanf(line, "%d %d %d %d %d %d %d %d %d %d %lf", &k.np, &k.np_part, &k.nmat, &k.ldf, &v.kernel,
                    &v.lane_chunk, &v.lpt, &v.zseg, &v.gather_tile, &v.bra, &v.us) != 11) {
      std::fclose(f);
      return fail(OCTHIP_ERR_INVALID, "%s: malformed line %d", path, n + 1);
    }
    const bool ok = v.kernel >= 0 && v.kernel <= 2 && v.zseg >= 0 && v.zseg <= 64 && v.gather_tile >= 1 &&
                    v.gather_tile <= 8 && (v.bra == 0 || v.bra == 1) && v.lpt >= 0 && v.lpt <= 2;
    if (!ok) {
      std::fclose(f);
      return fail(OCTHIP_ERR_INVALID, "%s: invalid configuration on line %d", path, n + 1);
    }
    if (TuneVal* e = lookup(k)) *e = v;
    else g_table.push_back({k, v});
    ++n;
  }
  std::fclose(f);
  return OCTHIP_SUCCESS;
}
