// Lanczos and Chebyshev exponentials of a packed complex batch
// (exponential_lanczos_batch / exponential_cheby_batch, src/electrons/exponential.F90:766-971)
// and the mesh reductions they use (mesh_batch_nrm2, mesh_batch_orthogonalization,
// src/grid/mesh_batch_inc.F90:970-1274).
//
// The Krylov vectors, the Chebyshev recurrence buffers and the small per-state buffers live
// in a set per (Hamiltonian, stream), like the H-apply's work sets.  The small projected
// problems are solved on the host: the Lanczos matrix of a Hermitian H is Hermitian
// tridiagonal (zheev reads its upper triangle, exponential.F90:829-836), which a diagonal
// unitary turns into a real symmetric tridiagonal matrix; its eigenpairs come from the
// implicit QL iteration with Wilkinson shifts (the tql2 algorithm).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "kernels.h"

namespace octhip {
size_t dotp_scratch_bytes(int L, int is_complex);
int launch_dotp(const double* a, const double* b, int L, int np, int is_complex, double vol, double* scratch,
                hipStream_t s, double** result, double* dest = nullptr);
int launch_axpy_cols(double* y, const double* x, int L, int np, int is_complex, const double* coef, hipStream_t s);
int launch_scal_cols(double* x, int L, int np, int is_complex, const double* coef, hipStream_t s);
int launch_cheby_update(double* out, const double* psi1, const double* b, const double* psi2, size_t count,
                        double mid, double span, std::complex<double> z, hipStream_t s);
int launch_cheby_final(double* psib, const double* psi0, const double* psi2, size_t count, std::complex<double> z,
                       hipStream_t s);
}  // namespace octhip

using namespace octhip;
using cplx = std::complex<double>;

namespace {

// Device buffers of one stream: nvec batches of np_part rows (Krylov vectors or the
// Chebyshev recurrence) and a small block (per-state factors + dot-product scratch).
struct KrylovBufs {
  double* vec = nullptr;
  double* small = nullptr;
  double* coef = nullptr;
  double* dots = nullptr;
};

int krylov_bufs(octhip_hamiltonian_s* h, hipStream_t s, int nvec, int ldf, KrylovBufs* out) {
  const size_t L = (size_t)1 << ldf;
  const size_t vec_bytes = sizeof(double) * (size_t)h->kin->np_part * L * nvec;
  const size_t small_bytes = sizeof(double) * 2 * (1 << MAX_LDF) + dotp_scratch_bytes(1 << MAX_LDF, 0);
  octhip_hamiltonian_s::KrylovSet* k = nullptr;
  for (auto& e : h->krylov)
    if (e.s == s) k = &e;
  if (!k) {
    h->krylov.push_back({s, nullptr, 0, nullptr});
    k = &h->krylov.back();
  }
  if (k->bytes < vec_bytes) {
    dfree(k->vec);
    k->vec = nullptr;
    k->bytes = 0;
    int rc = dalloc((void**)&k->vec, vec_bytes);
    if (rc) return rc;
    k->bytes = vec_bytes;
  }
  if (!k->small) {
    int rc = dalloc((void**)&k->small, small_bytes);
    if (rc) return rc;
  }
  out->vec = k->vec;
  out->small = k->small;
  out->coef = k->small;
  out->dots = k->small + 2 * (1 << MAX_LDF);
  return OCTHIP_SUCCESS;
}

// per-state dot products <a|b> (complex, vol included) of the first nst states, on the host
int dots(const double* a, const double* b, int L, int np, double vol, double* scratch, hipStream_t s, int nst,
         std::vector<cplx>& out) {
  double* res = nullptr;
  int rc = launch_dotp(a, b, L, np, 1, vol, scratch, s, &res);
  if (rc) return rc;
  std::vector<double> host(2 * (size_t)nst);
  OCTHIP_HIP(hipMemcpyAsync(host.data(), res, sizeof(double) * 2 * nst, hipMemcpyDeviceToHost, s));
  OCTHIP_HIP(hipStreamSynchronize(s));
  out.resize(nst);
  for (int i = 0; i < nst; ++i) out[i] = cplx(host[2 * i], host[2 * i + 1]);
  return OCTHIP_SUCCESS;
}

// per-state factors to the device: states nst.. of the row get `pad` (0 for axpy, 1 for scal)
int upload_coefs(double* coef, const std::vector<cplx>& c, int ncol, cplx pad, hipStream_t s) {
  std::vector<double> host(2 * (size_t)ncol);
  for (int i = 0; i < ncol; ++i) {
    const cplx v = i < (int)c.size() ? c[i] : pad;
    host[2 * i] = v.real();
    host[2 * i + 1] = v.imag();
  }
  OCTHIP_HIP(hipMemcpyAsync(coef, host.data(), sizeof(double) * 2 * ncol, hipMemcpyHostToDevice, s));
  OCTHIP_HIP(hipStreamSynchronize(s));  // the host vector goes out of scope
  return OCTHIP_SUCCESS;
}

// Eigenpairs of the real symmetric tridiagonal matrix (diagonal d[0..n-1], off-diagonal
// e[0..n-2]): implicit QL with Wilkinson shifts (tql2).  z (n x n, row-major, z[i*n+k] =
// component i of eigenvector k) starts as the identity.
bool tridiagonal_eigen(int n, std::vector<double>& d, std::vector<double> e, std::vector<double>& z) {
  z.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) z[(size_t)i * n + i] = 1.0;
  e.resize(n);
  e[n - 1] = 0.0;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    int m;
    do {
      for (m = l; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= 1e-16 * dd) break;
      }
      if (m != l) {
        if (++iter > 100) return false;
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
        double sn = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = sn * e[i];
          const double b = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            break;
          }
          sn = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * sn + 2.0 * c * b;
          p = sn * r;
          d[i + 1] = g + p;
          g = c * r - b;
          for (int k = 0; k < n; ++k) {
            f = z[(size_t)k * n + i + 1];
            z[(size_t)k * n + i + 1] = sn * z[(size_t)k * n + i] + c * f;
            z[(size_t)k * n + i] = c * z[(size_t)k * n + i] - sn * f;
          }
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  return true;
}

// First column of f(pp * A) for the n x n Hermitian matrix whose upper triangle is
// hm(i, j) = a[i + j*ld] (tridiagonal: the diagonal and the first superdiagonal are read),
// f = exp or phi(x) = (e^x - 1)/x (zlalg_exp / zlalg_phi, src/math/lalg_adv.F90:216-340).
bool krylov_function(int n, const cplx* a, int ld, cplx pp, bool phi, cplx* col) {
  std::vector<double> d(n), e(std::max(n - 1, 0));
  std::vector<cplx> ph(n);  // D = diag(ph): A = D T D^H with T real
  ph[0] = 1.0;
  for (int k = 0; k < n; ++k) d[k] = a[k + (size_t)k * ld].real();
  for (int k = 0; k + 1 < n; ++k) {
    const cplx b = a[k + (size_t)(k + 1) * ld];  // A(k, k+1)
    const double m = std::abs(b);
    e[k] = m;
    // T(k, k+1) = conj(ph_k) b ph_{k+1} = |b|
    ph[k + 1] = m > 0.0 ? ph[k] * std::conj(b) / m : ph[k];
  }
  std::vector<double> z;
  if (!tridiagonal_eigen(n, d, e, z)) return false;
  for (int i = 0; i < n; ++i) {
    cplx s = 0.0;
    for (int k = 0; k < n; ++k) {
      const cplx x = pp * d[k];
      const cplx f = phi ? (std::exp(x) - 1.0) / x : std::exp(x);
      s += z[(size_t)i * n + k] * f * z[k];  // z[0*n + k]: component 0 of eigenvector k
    }
    col[i] = ph[i] * s;  // (D U f U^T D^H e_1)_i with ph[0] = 1
  }
  return true;
}

int check_exp_args(octhip_hamiltonian_t h, const double* psi, int ldf, int order) {
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(h && psi, "null argument");
  OCTHIP_REQUIRE(ldf >= 1 && ldf <= MAX_LDF, "exponentials need a complex batch (ldf in 1..%d)", MAX_LDF);
  OCTHIP_REQUIRE(order >= 1 && order <= 256, "order %d outside 1..256", order);
  return OCTHIP_SUCCESS;
}

}  // namespace

namespace octhip {
int krylov_vectors(octhip_hamiltonian_s* h, hipStream_t s, int nvec, int ldf, double** base) {
  KrylovBufs kb;
  int rc = krylov_bufs(h, s, nvec, ldf, &kb);
  if (rc) return rc;
  *base = kb.vec;
  return OCTHIP_SUCCESS;
}

void drop_krylov(octhip_hamiltonian_s* h, hipStream_t s, bool all) {
  for (size_t i = 0; i < h->krylov.size();) {
    if (all || h->krylov[i].s == s) {
      dfree(h->krylov[i].vec);
      dfree(h->krylov[i].small);
      h->krylov.erase(h->krylov.begin() + (long)i);
    } else {
      ++i;
    }
  }
}
}  // namespace octhip

extern "C" int octhip_exponential_lanczos_batch(octhip_hamiltonian_t h, double* psi, int ldf, int nst, double deltat,
                                                int imag_time, int order, double lanczos_tol, const double* inh,
                                                double vol, int* iterations, double* residual, void* stream) {
  OCTHIP_RANGE("octhip_exponential_lanczos_batch");
  int rc = check_exp_args(h, psi, ldf, order);
  if (rc) return rc;
  const int L = 1 << ldf, ncol = L / 2;
  OCTHIP_REQUIRE(nst >= 1 && nst <= ncol, "nst=%d outside 1..%d", nst, ncol);
  OCTHIP_REQUIRE(lanczos_tol > 0.0 && vol > 0.0, "lanczos_tol and vol must be positive");
  hipStream_t s = (hipStream_t)stream;
  const int np = h->np, np_part = h->kin->np_part;
  const size_t rowsz = (size_t)np_part * L;
  KrylovBufs kb;
  rc = krylov_bufs(h, s, order + 1, ldf, &kb);
  if (rc) return rc;
  auto V = [&](int k) { return kb.vec + (size_t)k * rowsz; };  // vb(k + 1)
  const cplx dt = imag_time ? cplx(0.0, deltat) : cplx(deltat, 0.0);
  const cplx pp = cplx(-0.0, -1.0) * dt;  // -i deltat
  const double eps = 2.220446049250313e-16;
  // vb(1) = inh or psi (inner rows), beta = |vb(1)|
  OCTHIP_HIP(hipMemcpyAsync(V(0), inh ? inh : psi, sizeof(double) * (size_t)np * L, hipMemcpyDeviceToDevice, s));
  std::vector<cplx> d;
  rc = dots(V(0), V(0), L, np, vol, kb.dots, s, nst, d);
  if (rc) return rc;
  std::vector<double> beta(nst);
  bool all_null = true;
  for (int i = 0; i < nst; ++i) {
    beta[i] = std::sqrt(d[i].real());
    if (std::fabs(beta[i]) > 1e-12) all_null = false;
  }
  if (iterations) *iterations = 0;
  if (residual) *residual = 0.0;
  if (all_null) return OCTHIP_SUCCESS;
  std::vector<cplx> c(nst);
  for (int i = 0; i < nst; ++i) c[i] = cplx(1.0 / beta[i], 0.0);
  rc = upload_coefs(kb.coef, c, ncol, 1.0, s);
  if (!rc) rc = launch_scal_cols(V(0), L, np, 1, kb.coef, s);
  if (rc) return rc;
  const int M = order + 1;  // hamilt(1:M, 1:M) per state, column-major
  std::vector<cplx> hamilt((size_t)nst * M * M, 0.0), expo((size_t)nst * M, 0.0);
  auto H = [&](int st, int i, int j) -> cplx& { return hamilt[(size_t)st * M * M + i + (size_t)j * M]; };
  std::vector<double> res(nst, 0.0);
  int iter = 1;
  for (; iter <= order; ++iter) {
    rc = propagator_apply(h, V(iter - 1), V(iter), ldf, s);
    if (rc) return rc;
    // classical Gram-Schmidt against vb(l..iter), l = max(1, iter - 1) (Hermitian H):
    // every overlap first, then the subtractions (mesh_batch_inc.F90:1218-1231)
    const int l = std::max(1, iter - 1);
    std::vector<std::vector<cplx>> ss;
    for (int j = l; j <= iter; ++j) {
      std::vector<cplx> o;
      rc = dots(V(j - 1), V(iter), L, np, vol, kb.dots, s, nst, o);
      if (rc) return rc;
      for (int st = 0; st < nst; ++st) H(st, j - 1, iter - 1) = o[st];
      ss.push_back(o);
    }
    for (int j = l; j <= iter; ++j) {
      for (int st = 0; st < nst; ++st) c[st] = -ss[j - l][st];
      rc = upload_coefs(kb.coef, c, ncol, 0.0, s);
      if (!rc) rc = launch_axpy_cols(V(iter), V(j - 1), L, np, 1, kb.coef, s);
      if (rc) return rc;
    }
    rc = dots(V(iter), V(iter), L, np, vol, kb.dots, s, nst, d);
    if (rc) return rc;
    bool happy = true;
    for (int st = 0; st < nst; ++st) {
      const double nrm = std::sqrt(d[st].real());
      H(st, iter, iter - 1) = nrm;
      if (!krylov_function(iter, &H(st, 0, 0), M, pp, inh != nullptr, &expo[(size_t)st * M]))
        return fail(OCTHIP_ERR_HIP, "Lanczos: the projected eigenproblem did not converge");
      res[st] = std::abs(H(st, iter, iter - 1) * std::abs(expo[(size_t)st * M + iter - 1]));
      if (std::abs(H(st, iter, iter - 1)) >= 1e4 * eps) happy = false;
    }
    if (happy) break;  // happy breakdown
    for (int st = 0; st < nst; ++st) {
      const double a = std::abs(H(st, iter, iter - 1));
      c[st] = a >= 1e4 * eps ? cplx(1.0 / a, 0.0) : cplx(1.0, 0.0);
    }
    rc = upload_coefs(kb.coef, c, ncol, 1.0, s);
    if (!rc) rc = launch_scal_cols(V(iter), L, np, 1, kb.coef, s);
    if (rc) return rc;
    if (iter > 3 && *std::max_element(res.begin(), res.end()) < lanczos_tol) break;
  }
  const int used = std::min(iter, order);  // a completed loop adds vb(order+1) with factor 0
  if (iterations) *iterations = used;
  if (residual) *residual = *std::max_element(res.begin(), res.end());
  if (inh) {
    // psi += deltat nrm V expo(:, 1)  (exponential.F90:883-887)
    for (int ii = 1; ii <= used; ++ii) {
      for (int st = 0; st < nst; ++st) {
        const double rb = dt.real() * beta[st];
        const cplx ex = expo[(size_t)st * M + ii - 1];
        c[st] = cplx(rb * ex.real(), rb * ex.imag());
      }
      rc = upload_coefs(kb.coef, c, ncol, 0.0, s);
      if (!rc) rc = launch_axpy_cols(psi, V(ii - 1), L, np, 1, kb.coef, s);
      if (rc) return rc;
    }
    return OCTHIP_SUCCESS;
  }
  // psi = expo(1,1) psi + sum_{ii >= 2} beta expo(ii,1) vb(ii)  (exponential.F90:888-896)
  for (int st = 0; st < nst; ++st) c[st] = expo[(size_t)st * M];
  rc = upload_coefs(kb.coef, c, ncol, 1.0, s);
  if (!rc) rc = launch_scal_cols(psi, L, np, 1, kb.coef, s);
  if (rc) return rc;
  for (int ii = 2; ii <= used; ++ii) {
    for (int st = 0; st < nst; ++st) {
      const cplx ex = expo[(size_t)st * M + ii - 1];
      c[st] = cplx(beta[st] * ex.real(), beta[st] * ex.imag());
    }
    rc = upload_coefs(kb.coef, c, ncol, 0.0, s);
    if (!rc) rc = launch_axpy_cols(psi, V(ii - 1), L, np, 1, kb.coef, s);
    if (rc) return rc;
  }
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_exponential_cheby_batch(octhip_hamiltonian_t h, double* psi, int ldf, double deltat, int order,
                                              double spectral_middle_point, double spectral_half_span,
                                              void* stream) {
  OCTHIP_RANGE("octhip_exponential_cheby_batch");
  int rc = check_exp_args(h, psi, ldf, order);
  if (rc) return rc;
  OCTHIP_REQUIRE(spectral_half_span > 0.0, "spectral_half_span must be positive");
  hipStream_t s = (hipStream_t)stream;
  const int L = 1 << ldf;
  const int np = h->np, np_part = h->kin->np_part;
  const size_t rowsz = (size_t)np_part * L, count = (size_t)np * L;
  KrylovBufs kb;
  rc = krylov_bufs(h, s, 3, ldf, &kb);
  if (rc) return rc;
  // psi0 = psi1 = 0 (batch_set_zero); the copies of the recurrence become buffer rotations
  double* b0 = kb.vec;
  double* b1 = kb.vec + rowsz;
  double* b2 = kb.vec + 2 * rowsz;
  {
    const unsigned grid = (unsigned)std::min<size_t>((3 * rowsz + BLOCK - 1) / BLOCK, 8192);
    hipLaunchKernelGGL(k_zero_rows, dim3(grid), dim3(BLOCK), 0, s, kb.vec, (size_t)0, 3 * rowsz);
    OCTHIP_HIP(hipGetLastError());
  }
  const double x = spectral_half_span * deltat;
  for (int j = order - 1; j >= 0; --j) {
    // psi2 <- psi1, psi1 <- psi0, psi0 <- H psi1 (exponential.F90:936-940)
    double* t = b2;
    b2 = b1;
    b1 = b0;
    b0 = t;
    rc = propagator_apply(h, b1, b0, ldf, s);
    if (rc) return rc;
    static const cplx mi_pow[4] = {cplx(1.0, 0.0), cplx(0.0, -1.0), cplx(-1.0, 0.0), cplx(0.0, 1.0)};
    const cplx two_pow = cplx(2.0 * mi_pow[j & 3].real(), 2.0 * mi_pow[j & 3].imag());
    const double bj = jn(j, x);
    const cplx zfact(two_pow.real() * bj, two_pow.imag() * bj);
    rc = launch_cheby_update(b0, b1, psi, b2, count, spectral_middle_point, spectral_half_span, zfact, s);
    if (rc) return rc;
  }
  // psib = (psi0 - psi2) * exp(-i mid deltat) / 2
  const double ang = spectral_middle_point * deltat;
  const cplx z(0.5 * std::cos(ang), 0.5 * -std::sin(ang));
  return launch_cheby_final(psi, b0, b2, count, z, s);
}

extern "C" int octhip_mesh_batch_nrm2(const double* a, int ldf, int np, int is_complex, double vol, double* nrm2,
                                      void* stream) {
  OCTHIP_REQUIRE(nrm2 != nullptr, "null argument");
  const int L = 1 << (ldf < 0 ? 0 : ldf);
  const int n = is_complex ? L / 2 : L;
  std::vector<double> d(2 * (size_t)n);
  int rc = octhip_mesh_batch_dotp_vector(a, a, ldf, np, is_complex, vol, d.data(), stream);
  if (rc) return rc;
  for (int c = 0; c < n; ++c) nrm2[c] = std::sqrt(is_complex ? d[2 * c] : d[c]);
  return OCTHIP_SUCCESS;
}

extern "C" int octhip_mesh_batch_orthogonalization(int nvec, const double* const* psi, double* phi, int ldf, int nst,
                                                   int np, double vol, int normalize, double* overlap, double* norm,
                                                   void* stream) {
  OCTHIP_RANGE("octhip_mesh_batch_orthogonalization");
  OCTHIP_REQUIRE_INIT();
  OCTHIP_REQUIRE(nvec >= 0 && (nvec == 0 || psi) && phi, "bad argument");
  OCTHIP_REQUIRE(ldf >= 1 && ldf <= MAX_LDF, "complex batches only (ldf in 1..%d)", MAX_LDF);
  const int L = 1 << ldf, ncol = L / 2;
  OCTHIP_REQUIRE(nst >= 1 && nst <= ncol && np >= 0, "bad sizes nst=%d np=%d", nst, np);
  hipStream_t s = (hipStream_t)stream;
  double *scratch = nullptr, *coef = nullptr;
  int rc = dalloc((void**)&scratch, dotp_scratch_bytes(L, 1));
  if (!rc) rc = dalloc((void**)&coef, sizeof(double) * 2 * ncol);
  // classical Gram-Schmidt: every overlap <psi_k|phi> first, then phi -= sum_k <psi_k|phi> psi_k
  std::vector<std::vector<cplx>> ss(nvec);
  for (int k = 0; k < nvec && !rc; ++k) rc = dots(psi[k], phi, L, np, vol, scratch, s, nst, ss[k]);
  std::vector<cplx> c(nst);
  for (int k = 0; k < nvec && !rc; ++k) {
    for (int st = 0; st < nst; ++st) c[st] = -ss[k][st];
    rc = upload_coefs(coef, c, ncol, 0.0, s);
    if (!rc) rc = launch_axpy_cols(phi, psi[k], L, np, 1, coef, s);
  }
  if (!rc && overlap)  // overlap(k, st) = <psi_k|phi_st>, column-major (nvec x nst)
    for (int k = 0; k < nvec; ++k)
      for (int st = 0; st < nst; ++st) {
        overlap[2 * ((size_t)k + (size_t)st * nvec)] = ss[k][st].real();
        overlap[2 * ((size_t)k + (size_t)st * nvec) + 1] = ss[k][st].imag();
      }
  if (!rc && (norm || normalize)) {
    std::vector<cplx> d;
    rc = dots(phi, phi, L, np, vol, scratch, s, nst, d);
    for (int st = 0; st < nst && !rc; ++st) {
      const double nr = std::sqrt(d[st].real());
      if (norm) norm[st] = nr;
      c[st] = cplx(1.0 / nr, 0.0);
    }
    if (!rc && normalize) {
      rc = upload_coefs(coef, c, ncol, 1.0, s);
      if (!rc) rc = launch_scal_cols(phi, L, np, 1, coef, s);
      if (!rc) {
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = fail(OCTHIP_ERR_HIP, "orthogonalization: %s", hipGetErrorString(e));
      }
    }
  }
  dfree(scratch);
  dfree(coef);
  return rc;
}
