// Internal declarations shared by the HIP translation units of liboctopus_hip.so.
#pragma once

#include <complex>
#include <functional>

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "octhip.h"

namespace octhip {

// largest packed row handled: 2^MAX_LDF doubles (pack_size_real(1) <= 1024)
constexpr int MAX_LDF = 10;
// largest stencil size of the device path
constexpr int MAX_STENCIL = 64;
// threads per workgroup of every streaming kernel
constexpr int BLOCK = 256;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
bool initialised();
int current_device();
// device buffer of >= 2^MAX_LDF zero doubles (stands for zero-Dirichlet boundary rows)
const double* zero_row();

#define OCTHIP_HIP(call)                                                                   \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return ::octhip::fail(OCTHIP_ERR_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                            __FILE__, __LINE__);                                           \
  } while (0)

#define OCTHIP_REQUIRE(cond, ...)                                  \
  do {                                                             \
    if (!(cond)) return ::octhip::fail(OCTHIP_ERR_INVALID, __VA_ARGS__); \
  } while (0)

#define OCTHIP_REQUIRE_INIT()                                                        \
  do {                                                                               \
    if (!::octhip::initialised())                                                    \
      return ::octhip::fail(OCTHIP_ERR_STATE, "octhip_init has not been called");    \
  } while (0)

// Device allocation and release: every allocation of the library goes through dalloc
// (RAII-free: every handle frees what it owns in its *_destroy).
int dalloc(void** p, size_t bytes);
void dfree(void* p);

template <typename T>
int upload(T** dst, const T* src, size_t count, hipStream_t s) {
  *dst = nullptr;
  if (count == 0) return OCTHIP_SUCCESS;
  const int rc = dalloc((void**)dst, count * sizeof(T));
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(*dst, src, count * sizeof(T), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(OCTHIP_ERR_HIP, "upload: %s", hipGetErrorString(e));
  return OCTHIP_SUCCESS;
}


// While a graph is being captured no profiling events are recorded.
void set_capturing(bool on);

// Device timing of one kernel launch when profiling is on (events on the launch stream);
// always a roctx range of the region's name around the launch (host side).
class ProfScope {
 public:
  ProfScope(const char* region, hipStream_t s);
  ~ProfScope();

 private:
  int slot_ = -1;
  hipStream_t s_;
};

// roctx range around one C-ABI entry point (the reference's profiling_in/out regions,
// src/basic/profiling.F90:503-625); rocprofv3 --marker-trace records them.
class ApiRange {
 public:
  explicit ApiRange(const char* name);
  ~ApiRange();
};
#define OCTHIP_RANGE(name) ::octhip::ApiRange octhip_api_range_(name)

}  // namespace octhip

// ------------------------------------------------------------------ handles --
struct octhip_nl_operator_s {
  int n = 0;             // stencil size
  int np = 0, np_part = 0;
  int nri = 0;
  int center = -1;       // stencil index of the zero offset, -1 if none
  std::vector<double> w; // const weights
  // lattice (site-map) indexing
  bool lattice = false;
  int dcube[octhip::MAX_STENCIL] = {0};  // site-index delta of stencil point j
  int* d_cidx = nullptr;                 // np: site index of each inner point
  int* d_cube = nullptr;                 // sites: point index or -1
  int* d_dcube = nullptr;                // device copy of dcube
  size_t ncube = 0;
  // run-compressed indexing (Octopus's own ri/rimap)
  int* d_ri = nullptr;                   // nri*n
  int* d_rimap = nullptr;                // np
  // runs split as nl_operator_build does for domain parallelism (src/grid/nl_operator.F90:
  // 457-514): [0] inner runs (every neighbour an inner point), [1] outer runs; ri rows and
  // half-open point ranges [imin, imax) of each part
  int part_nri[2] = {0, 0};
  int* d_part_ri[2] = {nullptr, nullptr};
  int* d_part_imin[2] = {nullptr, nullptr};
  int* d_part_imax[2] = {nullptr, nullptr};
  // host copy of lattice coordinates (3*np_part), kept for schedules
  std::vector<int> coords;
  int xcd_chunk = 32;                    // gather kernel: blockIdx -> tile remap granularity
  int tile_iters = 4;                    // gather kernel: passes over the point slots per tile
  // z-march (order-4 star on a lattice): padded march cube and (tile, z-segment) items
  int kernel = 0;                        // 0 = auto, 1 = gather, 2 = z-march
  bool zmarch_ok = false;
  std::vector<int> mcube;                // site -> point index or -1 (host; builds the item tables)
  int mx = 0, my = 0, mz = 0;
  int zseg = 0;                          // output planes per work item, 0 = chosen per launch
  int lane_chunk = 0;                    // lanes per workgroup (0 = automatic)
  int lanes_per_thread = 0;              // doubles per thread (0 = automatic)
  int zseg_built = 0;
  mutable std::vector<int4> zseg_cache;  // automatic segment lengths: {nchunk, lc, resident (+0x100 in a batched launch), zseg}
  std::vector<int> tile_zlo, tile_zhi;   // per tile: inner plane range (cube z), -1 if empty
  int ntx = 0, nty = 0;
  int4* d_items = nullptr;
  int* d_tab = nullptr;                  // per item: site -> point index table (planes zb-4 .. ze+3)
  int* d_tab_off = nullptr;              // per item: first int of its table
  int nitems = 0;
  std::vector<int4> items;
  std::vector<std::pair<int, int2*>> sched;  // lane chunks -> device schedule
  std::vector<int> sched_len;
  // item lists built so far, one per segment length (the current one is mirrored in
  // d_items / nitems / items / sched / zseg_built)
  struct ZPlan {
    int zseg;
    int4* d_items;
    int* d_tab;
    int* d_tab_off;
    std::vector<int4> items;
    std::vector<std::pair<int, int2*>> sched;
    std::vector<int> sched_len;
  };
  std::vector<ZPlan> plans;
  // device buffers replaced by a rebuild stay allocated until destroy: captured graphs and
  // batched launches may still point at them
  std::vector<void*> retired;
};

struct octhip_hamiltonian_s {
  octhip_nl_operator_t kin = nullptr;
  double mass = 1.0;
  int np = 0;
  double* d_v = nullptr;        // np local potential
  double* d_xyz = nullptr;      // 3*np positions for the dipole term (optional)
  double* d_veff = nullptr;     // np: v + field . x when a dipole field is set
  double field[3] = {0, 0, 0};
  bool has_field = false;
  // projectors
  int nmat = 0;
  int max_nprojs = 0;
  int full_projection_size = 0;
  int total_points = 0;
  int pjs = 0;                  // row stride of the transposed projector table
  std::vector<int> offsets;     // 6*nmat host copy
  double* d_matrices = nullptr; // column-major blocks
  int* d_maps = nullptr;        // 1-based
  double* d_scals = nullptr;
  double* d_pt = nullptr;       // total_points x pjs (row-major, transposed blocks)
  int* d_mat_pbase = nullptr;   // nmat: offset of the matrix's projections
  int* d_mat_npj = nullptr;     // nmat: nprojs
  // bra chunking: chunk c = (matrix, first sphere point, last+1)
  int nchunk = 0;
  int4* d_chunks = nullptr;     // {matrix, q0, q1, 0}
  int* d_mat_chunk0 = nullptr;  // nmat+1 first chunk of each matrix
  int4* d_mat_info = nullptr;   // nmat: {npoints, nprojs, matrix offset, map offset}
  int* d_proj_mat = nullptr;    // projection index -> matrix
  int nslot = 0;                // mesh points inside at least one sphere
  int nmulti = 0;               // mesh points inside several spheres
  int* d_kslot = nullptr;       // np: row of the ket buffer for each mesh point, or -1
  int* d_multi_pos = nullptr;   // nmulti+1
  int* d_multi_g = nullptr;     // sphere points of the multi-sphere mesh points
  // Work buffers, one set per stream: applies on different streams (Octopus gives every
  // batch its own stream, accel_set_stream) may run concurrently.  The set of the stream
  // of the current call is mirrored in the fields below; sets captured into a graph are
  // re-keyed so no later stream reuses them.
  struct WorkSet {
    hipStream_t s;
    int work_ldf, y_ldf;
    double *partial, *proj, *ksum, *y0, *y1, *y2;
    // density of the propagated states: per-lane-chunk partials (dchunks x np) and the
    // state weights (created on first use)
    double *dpart, *dw;
    int dchunks;
    // the phased copy of the input of a k-point apply (np_part rows)
    double* ephase;
    int ephase_ldf;
  };
  std::vector<WorkSet> sets;
  long graph_sets = 0;
  // Taylor graphs holding sets of this handle (each frees its own sets when destroyed;
  // destroying the handle first frees them all and detaches the graphs)
  std::vector<octhip_taylor_graph_t> graphs;
  int work_ldf = -1;
  double* d_partial = nullptr;  // nchunk * max_nprojs * L
  double* d_proj = nullptr;     // full_projection_size * L
  double* d_ksum = nullptr;     // (total_points + nmulti) * L ket rows
  double* d_y[3] = {nullptr, nullptr, nullptr};  // Taylor powers H^k psi, np rows each
  int y_ldf = -1;
  std::vector<void*> retired;   // work buffers outgrown by a larger ldf, freed at destroy
  int use_mfma = 1;
  // Lanczos / Chebyshev buffers per stream: Krylov vectors (np_part rows each) and a small
  // block of per-state factors and dot-product partials
  struct KrylovSet {
    hipStream_t s;
    double* vec;
    size_t bytes;
    double* small;
  };
  std::vector<KrylovSet> krylov;
  // periodic boundaries: boundary row np+b takes inner row bnd_src[b]
  int* d_bnd_src = nullptr;
  int nbnd = 0;
  bool periodic = false;
  // Bloch phases of one k-point (complex pairs): phase (np_part), phase_corr (np_part-np),
  // projector phases (total_points)
  double* d_phase = nullptr;
  double* d_phase_corr = nullptr;
  double* d_pphase = nullptr;
  bool has_kpoint = false;
};

namespace octhip {
// Taylor building blocks shared by the propagator entry points and the adjoint.
int prepare_taylor(octhip_hamiltonian_s* h, int ldf, hipStream_t s);
// Orders 2..4 keep every power H^k psi (k < order) and add them all at the last order;
// higher orders accumulate into psi order by order.
inline bool taylor_deferred(int order) { return order >= 2 && order <= 4; }
// psi_in (optional): take psi_s from psi_in and write psi_{s+1} to the np inner rows of psi
// v_override (optional): the local potential to use instead of the handle's (np doubles)
int taylor_step(octhip_hamiltonian_s* h, double* psi, int ldf, double deltat, int order, int imag_time,
                double* const* save, hipStream_t s, const double* psi_in = nullptr,
                const double* v_override = nullptr);
// out = v + field . x (np doubles), the dipole potential of one field, on stream s
int launch_veff(const octhip_hamiltonian_s* h, double* out, const double field[3], hipStream_t s);
void taylor_coefs(double deltat, int imag_time, int order, std::complex<double>* z);
// launches of body(capturing stream) as one instantiated graph (octhip_taylor_graph_launch)
// prepare(capturing stream) runs before the capture starts (allocations go there)
int capture_graph(hipStream_t s, const std::function<int(hipStream_t)>& prepare,
                  const std::function<int(hipStream_t)>& body, octhip_taylor_graph_t* out);
// the work-buffer set of stream s is taken out of the per-stream lookup (after a capture);
// returns the key it is kept under from then on (0 if stream s had no set)
hipStream_t pin_workset(octhip_hamiltonian_s* h, hipStream_t s);
// free the work-buffer set kept under key s (no synchronisation: nothing may use it)
void drop_workset(octhip_hamiltonian_s* h, hipStream_t s);
// H psi for the propagators: on a handle with periodic boundaries or Bloch phases, psi
// carries the phase (as between set_phase_corr and unset_phase_corr in the reference's
// propagation) and its boundary rows are filled from their periodic images; on any other
// handle this is octhip_hamiltonian_apply_batch with the zero boundary.  psi: np_part rows.
int propagator_apply(octhip_hamiltonian_s* h, double* psi, double* hpsi, int ldf, hipStream_t s);
inline bool is_periodic(const octhip_hamiltonian_s* h) { return h->periodic || h->has_kpoint; }
int launch_periodic_bc(double* f, const int* src, int np, int nb, int L, const double* corr, hipStream_t s);
int launch_phase(double* dst, const double* src, const double* phase, int n, int L, bool conjugate, hipStream_t s);
// nvec Krylov-style buffers of np_part rows of the stream's set (Lanczos, Chebyshev, and
// the Taylor chain of periodic handles)
int krylov_vectors(octhip_hamiltonian_s* h, hipStream_t s, int nvec, int ldf, double** base);
// free the Lanczos/Chebyshev buffers of stream s (every stream when all)
void drop_krylov(octhip_hamiltonian_s* h, hipStream_t s, bool all);
// field of the dipole term (positions must be uploaded); refreshes v_eff on stream s
int set_field(octhip_hamiltonian_s* h, const double field[3], hipStream_t s);
// exponential_apply_batch with its optional arguments: psi2/deltat2 (psib2), inh (inh_psib)
// and the density of the propagated states (weights: host, one per complex state of the
// row; density: device, np doubles, accumulated)
struct TaylorOptions {
  double* psi2 = nullptr;
  double deltat2 = 0.0;
  double* inh = nullptr;
  const double* weights = nullptr;
  double* density = nullptr;
};
int taylor_apply(octhip_hamiltonian_s* h, double* psi, int ldf, double deltat, int order, int imag_time,
                 const TaylorOptions& o, hipStream_t s);
}  // namespace octhip
