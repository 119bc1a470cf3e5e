// Discrete adjoint of the Taylor propagator with a dipole-coupled field (gfx950).
//
// Forward: psi_{s+1} = p(A_s) psi_s, A_s = -i dt H_s, H_s = H_0 + eps_s (pol . r),
// p(A) = sum_{n<=order} A^n / n!.  For an objective J(psi_N) with
// dJ = 2 Re <chi_N | d psi_N>  (<a|b> = vol sum conj(a) b), the backward states are
// chi_s = p(A_s)^dagger chi_{s+1} = p(+i dt H_s) chi_{s+1}  (H_s Hermitian), and
//   dJ/dtheta = 2 vol Re sum_s sum_{m+k<=order-1} zf_{m+k+1}
//                 sum_r conj(H_s^m chi_{s+1})(r) dH_s/dtheta(r) (H_s^k psi_s)(r)
// with zf_n = (-i dt)^n / n!.  dH_s/d eps_s = pol . r and dH_s/dv(r') = delta_{r r'}.
// This is the exact gradient of the discrete propagator; the reference's optimal-control
// gradient (src/opt_control/propagation.F90:1016-1103, d_j = 2 Im <chi|x pol_j|psi>) is
// its first-order-in-dt limit.  psi_s is restored from checkpoints every K steps
// (oct_prop_init/oct_prop_check, propagation.F90:1108-1205, keep states at fixed
// intervals) and recomputed forward inside each segment, keeping the H^k psi_s terms.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "kernels.h"
#include "octhip_internal.h"

using namespace octhip;

namespace {

constexpr int ADJ_MAX_ORDER = 8;
constexpr int ADJ_GRID = 1024;   // fixed block count: the eps partial sums reduce in a fixed order

struct AccumArgs {
  const double* yp[ADJ_MAX_ORDER];  // H^k psi_s,      k < order, np rows of L
  const double* yc[ADJ_MAX_ORDER];  // H^m chi_{s+1},  m < order
  double zr[ADJ_MAX_ORDER + 1], zi[ADJ_MAX_ORDER + 1];  // zf_n
  int order, np, L;
  const double* xp;                 // pol . r, np
  double* dv;                       // dJ/dv accumulator, np
  double* part;                     // ADJ_GRID partial sums of sum_r xp(r) Re rho(r)
  double two_vol;
};

// rho(r) = sum_states sum_{m+k<=order-1} zf_{m+k+1} conj(yc_m) yp_k; only Re rho is needed.
// TP threads per point (a power of two <= 64), each over states t, t+TP, ...; ORD > 0
// fixes the order at compile time so every (re, im) pair of all 2*ORD arrays is loaded
// up front as one 16-byte load.
template <int TP, int ORD>
__global__ __launch_bounds__(256) void k_adj_accum(const AccumArgs a) {
  __shared__ double wsum[4];
  const int pack = a.L / 2;
  const int t = threadIdx.x % TP;
  const int grp = threadIdx.x / TP;
  constexpr int GPB = 256 / TP;        // points per block iteration
  double eps_acc = 0.0;
  for (int r0 = blockIdx.x * GPB; r0 < a.np; r0 += gridDim.x * GPB) {
    const int r = r0 + grp;
    const bool okr = r < a.np;
    const size_t row = (size_t)(okr ? r : 0) * a.L;
    double re = 0.0;
    for (int j = t; j < pack; j += TP) {
      const size_t o = row + 2 * (size_t)j;
      double sj = 0.0;
      if constexpr (ORD > 0) {
        double2 p[ORD], w[ORD];
#pragma unroll
        for (int k = 0; k < ORD; ++k) p[k] = *reinterpret_cast<const double2*>(a.yp[k] + o);
#pragma unroll
        for (int m = 0; m < ORD; ++m) w[m] = *reinterpret_cast<const double2*>(a.yc[m] + o);
#pragma unroll
        for (int m = 0; m < ORD; ++m) {
          double ur = 0.0, ui = 0.0;
#pragma unroll
          for (int k = 0; k + m < ORD; ++k) {
            const double cr = a.zr[m + k + 1], ci = a.zi[m + k + 1];
            ur += cr * p[k].x - ci * p[k].y;
            ui += cr * p[k].y + ci * p[k].x;
          }
          sj += w[m].x * ur + w[m].y * ui;  // Re(conj(w) u)
        }
      } else {
        for (int m = 0; m < a.order; ++m) {
          double ur = 0.0, ui = 0.0;
          for (int k = 0; k + m < a.order; ++k) {
            const double pr = a.yp[k][o], pi = a.yp[k][o + 1];
            const double cr = a.zr[m + k + 1], ci = a.zi[m + k + 1];
            ur += cr * pr - ci * pi;
            ui += cr * pi + ci * pr;
          }
          sj += a.yc[m][o] * ur + a.yc[m][o + 1] * ui;
        }
      }
      re += sj;
    }
    re = okr ? re : 0.0;
#pragma unroll
    for (int d = TP / 2; d >= 1; d >>= 1) re += __shfl_xor(re, d, TP);
    if (t == 0 && okr) {
      a.dv[r] += a.two_vol * re;
      eps_acc += a.xp[r] * re;
    }
  }
  // block sum of eps_acc in a fixed order
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) eps_acc += __shfl_xor(eps_acc, d, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = eps_acc;
  __syncthreads();
  if (threadIdx.x == 0) a.part[blockIdx.x] = ((wsum[0] + wsum[1]) + (wsum[2] + wsum[3]));
}

template <int ORD>
int launch_accum(int tp, const AccumArgs& acc, hipStream_t s) {
  switch (tp) {
    case 1: hipLaunchKernelGGL((k_adj_accum<1, ORD>), dim3(ADJ_GRID), dim3(256), 0, s, acc); break;
    case 2: hipLaunchKernelGGL((k_adj_accum<2, ORD>), dim3(ADJ_GRID), dim3(256), 0, s, acc); break;
    case 4: hipLaunchKernelGGL((k_adj_accum<4, ORD>), dim3(ADJ_GRID), dim3(256), 0, s, acc); break;
    case 8: hipLaunchKernelGGL((k_adj_accum<8, ORD>), dim3(ADJ_GRID), dim3(256), 0, s, acc); break;
    case 16: hipLaunchKernelGGL((k_adj_accum<16, ORD>), dim3(ADJ_GRID), dim3(256), 0, s, acc); break;
    case 32: hipLaunchKernelGGL((k_adj_accum<32, ORD>), dim3(ADJ_GRID), dim3(256), 0, s, acc); break;
    default: hipLaunchKernelGGL((k_adj_accum<64, ORD>), dim3(ADJ_GRID), dim3(256), 0, s, acc); break;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(OCTHIP_ERR_HIP, "k_adj_accum: %s", hipGetErrorString(e));
  return OCTHIP_SUCCESS;
}

__global__ __launch_bounds__(256) void k_adj_reduce(const double* __restrict__ part, int n, double scale,
                                                    double* __restrict__ out) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += part[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = scale * sh[0];
}

__global__ void k_fill_zero(double* __restrict__ f, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    f[i] = 0.0;
}

}  // namespace

struct octhip_adjoint_s {
  octhip_hamiltonian_s* h = nullptr;
  int ldf = 0, L = 0, nsteps = 0, K = 0, order = 4;
  double dt = 0.0, vol = 0.0;
  double pol[3] = {0.0, 0.0, 0.0};
  double* d_xp = nullptr;       // np
  double* d_slot = nullptr;     // 2 x (K+1) x np x L: psi_s of a segment (recomputed), two sets
  double* d_ckpt = nullptr;     // nckpt x np x L: psi at every K-th step
  double* d_seg = nullptr;      // 2 x K x (order-1) x np x L: H^k psi_s, k >= 1, two sets
  double* d_ychi = nullptr;     // 2 x order x np x L: H^m chi_{s+1}, m >= 1 (slot 0 unused), two sets
  double* d_chi2 = nullptr;     // 2 x np x L: the other two chi buffers of the backward rotation
  double* d_part = nullptr;     // ADJ_GRID
  double* d_deps = nullptr;     // nsteps
  double* d_veff = nullptr;     // 2 x np: v + eps_s x of the recompute and the backward stream
  hipStream_t aux = nullptr;    // recompute stream of the backward
  hipStream_t accs = nullptr;   // gradient accumulation stream of the backward
  hipEvent_t ev_fork = nullptr, ev_rec[2] = {nullptr, nullptr}, ev_back[2] = {nullptr, nullptr};
  hipEvent_t ev_chi[2] = {nullptr, nullptr}, ev_acc[2] = {nullptr, nullptr};
  int nckpt = 0;
  std::vector<double> eps;      // field of every step (set by forward)
  bool forward_done = false;
};

namespace {

void free_adjoint(octhip_adjoint_s* a) {
  dfree(a->d_xp);
  dfree(a->d_slot);
  dfree(a->d_chi2);
  dfree(a->d_ckpt);
  dfree(a->d_seg);
  dfree(a->d_ychi);
  dfree(a->d_part);
  dfree(a->d_deps);
  dfree(a->d_veff);
  for (hipEvent_t e : {a->ev_fork, a->ev_rec[0], a->ev_rec[1], a->ev_back[0], a->ev_back[1], a->ev_chi[0],
                       a->ev_chi[1], a->ev_acc[0], a->ev_acc[1]})
    if (e) (void)hipEventDestroy(e);
  for (hipStream_t st : {a->aux, a->accs})
    if (st) {
      (void)hipStreamSynchronize(st);
      (void)hipStreamDestroy(st);
    }
  delete a;
}

int step_field(octhip_adjoint_s* a, int s, hipStream_t st) {
  const double f[3] = {a->eps[s] * a->pol[0], a->eps[s] * a->pol[1], a->eps[s] * a->pol[2]};
  return set_field(a->h, f, st);
}

size_t rows_bytes(size_t rows, int L) { return sizeof(double) * rows * (size_t)L; }

}  // namespace

extern "C" int octhip_adjoint_create(octhip_hamiltonian_t h, const double* positions, const double* pol,
                                     int ldf, int nsteps, double deltat, int order, int checkpoint_every,
                                     double vol_pp, octhip_adjoint_t* out) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(h && positions && pol && out, "null argument");
  *out = nullptr;
  OCTHIP_REQUIRE(ldf >= 1 && ldf <= MAX_LDF, "the adjoint needs a complex batch (ldf in 1..%d)", MAX_LDF);
  OCTHIP_REQUIRE(nsteps >= 1, "nsteps = %d", nsteps);
  OCTHIP_REQUIRE(order >= 1 && order <= ADJ_MAX_ORDER, "order %d outside 1..%d", order, ADJ_MAX_ORDER);
  OCTHIP_REQUIRE(checkpoint_every >= 1, "checkpoint_every = %d", checkpoint_every);
  OCTHIP_REQUIRE(vol_pp > 0.0, "vol_pp must be positive");
  if (is_periodic(h)) return fail(OCTHIP_ERR_UNSUPPORTED, "adjoint: finite systems only (zero boundary, no phases)");
  const int np = h->np;
  auto* a = new octhip_adjoint_s();
  a->h = h;
  a->ldf = ldf;
  a->L = 1 << ldf;
  a->nsteps = nsteps;
  a->K = std::min(checkpoint_every, nsteps);
  a->order = order;
  a->dt = deltat;
  a->vol = vol_pp;
  for (int d = 0; d < 3; ++d) a->pol[d] = pol[d];
  a->nckpt = (nsteps + a->K - 1) / a->K;
  a->eps.assign(nsteps, 0.0);
  const double zero[3] = {0.0, 0.0, 0.0};
  int rc = octhip_hamiltonian_set_dipole_field(h, positions, zero, nullptr);
  std::vector<double> xp(np);
  for (int p = 0; p < np; ++p)
    xp[p] = pol[0] * positions[3 * (size_t)p] + pol[1] * positions[3 * (size_t)p + 1] +
            pol[2] * positions[3 * (size_t)p + 2];
  if (!rc) rc = upload(&a->d_xp, xp.data(), xp.size(), nullptr);
  if (!rc) rc = dalloc((void**)&a->d_slot, rows_bytes((size_t)2 * (a->K + 1) * np, a->L));
  if (!rc) rc = dalloc((void**)&a->d_chi2, rows_bytes((size_t)2 * np, a->L));
  if (!rc) rc = dalloc((void**)&a->d_ckpt, rows_bytes((size_t)a->nckpt * np, a->L));
  if (!rc && order > 1) rc = dalloc((void**)&a->d_seg, rows_bytes((size_t)2 * a->K * (order - 1) * np, a->L));
  if (!rc) rc = dalloc((void**)&a->d_veff, sizeof(double) * 2 * (size_t)np);
  if (!rc && (hipStreamCreateWithFlags(&a->aux, hipStreamNonBlocking) != hipSuccess ||
              hipEventCreateWithFlags(&a->ev_fork, hipEventDisableTiming) != hipSuccess ||
              hipEventCreateWithFlags(&a->ev_rec[0], hipEventDisableTiming) != hipSuccess ||
              hipEventCreateWithFlags(&a->ev_rec[1], hipEventDisableTiming) != hipSuccess ||
              hipEventCreateWithFlags(&a->ev_back[0], hipEventDisableTiming) != hipSuccess ||
              hipEventCreateWithFlags(&a->ev_back[1], hipEventDisableTiming) != hipSuccess ||
              hipStreamCreateWithFlags(&a->accs, hipStreamNonBlocking) != hipSuccess ||
              hipEventCreateWithFlags(&a->ev_chi[0], hipEventDisableTiming) != hipSuccess ||
              hipEventCreateWithFlags(&a->ev_chi[1], hipEventDisableTiming) != hipSuccess ||
              hipEventCreateWithFlags(&a->ev_acc[0], hipEventDisableTiming) != hipSuccess ||
              hipEventCreateWithFlags(&a->ev_acc[1], hipEventDisableTiming) != hipSuccess))
    rc = fail(OCTHIP_ERR_HIP, "adjoint_create: stream or event creation failed");
  if (!rc) rc = dalloc((void**)&a->d_ychi, rows_bytes((size_t)2 * order * np, a->L));
  if (!rc) rc = dalloc((void**)&a->d_part, sizeof(double) * ADJ_GRID);
  if (!rc) rc = dalloc((void**)&a->d_deps, sizeof(double) * nsteps);
  if (!rc) rc = prepare_taylor(h, ldf, nullptr);
  if (!rc) {
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) rc = fail(OCTHIP_ERR_HIP, "adjoint_create: %s", hipGetErrorString(e));
  }
  if (rc) {
    free_adjoint(a);
    return rc;
  }
  *out = a;
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_adjoint_destroy(octhip_adjoint_t a) {
  if (!a) return OCTHIP_SUCCESS;
  (void)hipDeviceSynchronize();
  free_adjoint(a);
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_adjoint_forward(octhip_adjoint_t a, double* psi, const double* eps, void* stream) {
  OCTHIP_RANGE("octhip_adjoint_forward");
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(a && psi && eps, "null argument");
  hipStream_t s = (hipStream_t)stream;
  const int np = a->h->np;
  for (int i = 0; i < a->nsteps; ++i) a->eps[i] = eps[i];
  for (int st = 0; st < a->nsteps; ++st) {
    if (st % a->K == 0) {
      OCTHIP_HIP(hipMemcpyAsync(a->d_ckpt + (size_t)(st / a->K) * np * a->L, psi, rows_bytes(np, a->L),
                                hipMemcpyDeviceToDevice, s));
    }
    int rc = step_field(a, st, s);
    if (!rc) rc = taylor_step(a->h, psi, a->ldf, a->dt, a->order, 0, nullptr, s);
    if (rc) return rc;
  }
  a->forward_done = true;
  return OCTHIP_SUCCESS;
}

namespace {
// The backward pass proper; octhip_adjoint_backward releases the recompute stream's work
// set after it whether it succeeded or not.
int backward_run(octhip_adjoint_s* a, double* chi, double* dj_deps, double* dj_dv, void* stream) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(a && chi && dj_deps && dj_dv, "null argument");
  OCTHIP_REQUIRE(a->forward_done, "adjoint_backward before adjoint_forward");
  hipStream_t s = (hipStream_t)stream;
  const int np = a->h->np, np_part = a->h->kin->np_part, L = a->L, order = a->order;
  const size_t arr = (size_t)np * L;
  hipLaunchKernelGGL(k_fill_zero, dim3(1024), dim3(256), 0, s, dj_dv, (size_t)np);
  OCTHIP_HIP(hipGetLastError());
  // zf_n = (-i dt)^n / n!
  AccumArgs acc;
  std::memset(&acc, 0, sizeof(acc));
  {
    std::complex<double> z(1.0, 0.0);
    acc.zr[0] = 1.0;
    for (int n = 1; n <= order; ++n) {
      const std::complex<double> zz(z.imag() * a->dt, -z.real() * a->dt);  // z * (-i dt)
      z = std::complex<double>(zz.real() / n, zz.imag() / n);
      acc.zr[n] = z.real();
      acc.zi[n] = z.imag();
    }
  }
  acc.order = order;
  acc.np = np;
  acc.L = L;
  acc.xp = a->d_xp;
  acc.dv = dj_dv;
  acc.part = a->d_part;
  acc.two_vol = 2.0 * a->vol;
  double* ysave[ADJ_MAX_ORDER];
  const int tp = std::min(L / 2, 64);
  // Backward step t (t = 0, 1, ... in backward order) reads chi_{s+1} from cb[t % 3] and
  // writes chi_s into cb[(t + 1) % 3], keeping H^m chi_{s+1} in set t & 1 of d_ychi; the
  // gradient accumulation of step t runs on the accumulation stream (in step order, so
  // the sums are unchanged) while step t + 1 runs: step t waits only for the accumulation
  // of step t - 2, whose chi buffer and ychi set it overwrites.
  double* const cb[3] = {chi, a->d_chi2, a->d_chi2 + arr};
  hipStream_t s3 = a->accs;
  int t = 0;
  // Segment c is recomputed from its checkpoint into buffer set c & 1 on the aux stream
  // while the backward steps of segment c + 1 run on s: recompute(c) waits for the backward
  // of segment c + 2 (the same set) to finish reading, and the backward of c waits for
  // recompute(c).  Each stream forms its own potential v + eps_s x (d_veff[0] recompute,
  // d_veff[1] backward) and has its own work buffers.
  const size_t slot_set = (size_t)(a->K + 1) * arr, seg_set = (size_t)a->K * (order - 1) * arr;
  double* const vr = a->d_veff;
  double* const vb = a->d_veff + np;
  hipStream_t s2 = a->aux;
  auto field_of = [&](int st, double f[3]) {
    for (int d = 0; d < 3; ++d) f[d] = a->eps[st] * a->pol[d];
  };
  bool back_recorded[2] = {false, false};
  auto recompute = [&](int c) -> int {
    const int b = c & 1;
    if (back_recorded[b]) OCTHIP_HIP(hipStreamWaitEvent(s2, a->ev_back[b], 0));
    double* slot = a->d_slot + b * slot_set;
    double* seg = a->d_seg + b * seg_set;
    const int s0 = c * a->K, s1 = std::min(s0 + a->K, a->nsteps);
    OCTHIP_HIP(hipMemcpyAsync(slot, a->d_ckpt + (size_t)c * np * L, rows_bytes(np, L), hipMemcpyDeviceToDevice, s2));
    double* ys[ADJ_MAX_ORDER];
    for (int st = s0; st < s1; ++st) {
      const int i = st - s0;
      ys[0] = slot + (size_t)i * arr;
      for (int k = 1; k < order; ++k) ys[k] = seg + ((size_t)i * (order - 1) + (k - 1)) * arr;
      double f[3];
      field_of(st, f);
      int rc = launch_veff(a->h, vr, f, s2);
      if (!rc) rc = taylor_step(a->h, slot + (size_t)(i + 1) * arr, a->ldf, a->dt, order, 0, ys, s2, ys[0], vr);
      if (rc) return rc;
    }
    OCTHIP_HIP(hipEventRecord(a->ev_rec[b], s2));
    return OCTHIP_SUCCESS;
  };
  // the aux and accumulation streams start after everything the caller queued on s
  OCTHIP_HIP(hipEventRecord(a->ev_fork, s));
  OCTHIP_HIP(hipStreamWaitEvent(s2, a->ev_fork, 0));
  OCTHIP_HIP(hipStreamWaitEvent(s3, a->ev_fork, 0));
  int rc = recompute(a->nckpt - 1);
  if (rc) return rc;
  for (int c = a->nckpt - 1; c >= 0; --c) {
    const int b = c & 1;
    if (c >= 1) {
      rc = recompute(c - 1);
      if (rc) return rc;
    }
    OCTHIP_HIP(hipStreamWaitEvent(s, a->ev_rec[b], 0));
    double* slot = a->d_slot + b * slot_set;
    double* seg = a->d_seg + b * seg_set;
    const int s0 = c * a->K, s1 = std::min(s0 + a->K, a->nsteps);
    // backward through the segment
    for (int st = s1 - 1; st >= s0; --st, ++t) {
      const int i = st - s0, u = t & 1;
      double* const cur = cb[t % 3];
      double* const nxt = cb[(t + 1) % 3];
      double* const ychi = a->d_ychi + (size_t)u * order * arr;
      if (t >= 2) OCTHIP_HIP(hipStreamWaitEvent(s, a->ev_acc[u], 0));
      ysave[0] = cur;
      for (int m = 1; m < order; ++m) ysave[m] = ychi + m * arr;
      double f[3];
      field_of(st, f);
      rc = launch_veff(a->h, vb, f, s);
      // chi_s = p(+i dt H_s) chi_{s+1}, keeping H^m chi_{s+1}
      if (!rc) rc = taylor_step(a->h, nxt, a->ldf, -a->dt, order, 0, ysave, s, cur, vb);
      if (rc) return rc;
      OCTHIP_HIP(hipEventRecord(a->ev_chi[u], s));
      OCTHIP_HIP(hipStreamWaitEvent(s3, a->ev_chi[u], 0));
      acc.yp[0] = slot + (size_t)i * arr;
      for (int k = 1; k < order; ++k) acc.yp[k] = seg + ((size_t)i * (order - 1) + (k - 1)) * arr;
      acc.yc[0] = cur;
      for (int m = 1; m < order; ++m) acc.yc[m] = ychi + m * arr;
      {
        ProfScope prof("k_adj_accum", s3);
        rc = order == 4 ? launch_accum<4>(tp, acc, s3) : launch_accum<0>(tp, acc, s3);
        if (rc) return rc;
      }
      hipLaunchKernelGGL(k_adj_reduce, dim3(1), dim3(256), 0, s3, a->d_part, ADJ_GRID, 2.0 * a->vol,
                         a->d_deps + st);
      OCTHIP_HIP(hipGetLastError());
      OCTHIP_HIP(hipEventRecord(a->ev_acc[u], s3));
    }
    // the segment's psi arrays are free once its last accumulation is done
    OCTHIP_HIP(hipEventRecord(a->ev_back[b], s3));
    back_recorded[b] = true;
  }
  // everything accumulated before the results leave
  OCTHIP_HIP(hipEventRecord(a->ev_acc[0], s3));
  OCTHIP_HIP(hipStreamWaitEvent(s, a->ev_acc[0], 0));
  double* const cur = cb[t % 3];
  // chi_0 into the caller's buffer, boundary rows zero
  if (cur != chi) OCTHIP_HIP(hipMemcpyAsync(chi, cur, rows_bytes(np, L), hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_fill_zero, dim3(1024), dim3(256), 0, s, chi + arr, (size_t)(np_part - np) * L);
  OCTHIP_HIP(hipGetLastError());
  OCTHIP_HIP(hipMemcpyAsync(dj_deps, a->d_deps, sizeof(double) * a->nsteps, hipMemcpyDeviceToHost, s));
  OCTHIP_HIP(hipStreamSynchronize(s));
  return OCTHIP_SUCCESS;
}
}  // namespace

extern "C" int octhip_adjoint_backward(octhip_adjoint_t a, double* chi, double* dj_deps, double* dj_dv,
                                       void* stream) {
  OCTHIP_RANGE("octhip_adjoint_backward");
  const int rc = backward_run(a, chi, dj_deps, dj_dv, stream);
  if (!a) return rc;
  // The Hamiltonian's work set of the recompute stream is not kept between backwards, so
  // adjoints created and destroyed on one Hamiltonian leave nothing behind in it.  After
  // a failure the work already queued on the three streams drains first.
  if (rc) {
    (void)hipStreamSynchronize(a->aux);
    (void)hipStreamSynchronize(a->accs);
    (void)hipStreamSynchronize((hipStream_t)stream);
    (void)octhip_hamiltonian_release_stream(a->h, a->aux);
    return rc;
  }
  return octhip_hamiltonian_release_stream(a->h, a->aux);
}
