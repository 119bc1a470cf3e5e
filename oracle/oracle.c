/*
 * oracle.c -- CPU restatement of Octopus's H-apply path, used ONLY as the checker of the
 * device path (tests, smoke check) and as the timed CPU baseline.  It is test
 * infrastructure: the product library never links or calls it.
 *
 * Every function names the reference lines it restates.  Arithmetic conventions, fixed so
 * the device kernels can be compared bit for bit where the order of operations is the
 * same:
 *   - the stencil sum runs over j = 0..n-1 with fused multiply-adds from 0, one lane at
 *     a time, exactly like the AVX leaf (src/grid/operate_inc.c:76-118, VEC_FMA);
 *   - hpsi += v*psi is one fused multiply-add;
 *   - the projector contractions sum over sphere points (bra) or projectors (ket) in
 *     increasing order with fused multiply-adds (the reference calls BLAS dgemm there,
 *     whose order is unspecified -- see the header of the tests for how it is pinned);
 *   - complex axpy on interleaved lanes: re += zr*xr - zi*xi, im += zr*xi + zi*xr,
 *     products rounded separately.
 * Built with -ffp-contract=off so the compiler adds no contractions of its own.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(p, k, L) ((size_t)(p) * (size_t)(L) + (size_t)(k))

/* src/grid/operate_inc.c:29-342 (doperate_ri_vec): runs l, points i in
 * [imin[l], imax[l]), lanes k < 2^ldf: fo[i,k] = sum_j w[j] * fi[i + ri[l*n+j], k]. */
void oracle_operate_ri_vec(int n, const double* w, int nri, const int* ri, const int* imin,
                           const int* imax, const double* fi, int ldf, double* fo) {
  const int L = 1 << ldf;
#pragma omp parallel for schedule(dynamic, 64)
  for (int l = 0; l < nri; ++l) {
    const int* index = ri + (size_t)l * n;
    for (int i = imin[l]; i < imax[l]; ++i) {
      /* lanes innermost (one row of accumulators, vectorisable as the leaf's VEC_FMA
       * blocks are): every lane still sums j = 0..n-1 with fma from 0 */
      double* out = fo + IDX(i, 0, L);
      for (int k = 0; k < L; ++k) out[k] = 0.0;
      for (int j = 0; j < n; ++j) {
        const double wj = w[j];
        const double* src = fi + IDX(index[j] + i, 0, L);
        for (int k = 0; k < L; ++k) out[k] = fma(wj, src[k], out[k]);
      }
    }
  }
}

/* boundaries_set, zero Dirichlet branch (src/grid/boundaries_inc.F90:275-363):
 * rows np..np_part-1 of a packed batch are set to zero. */
void oracle_zero_boundary(int np, int np_part, int ldf, double* f) {
  const size_t L = (size_t)1 << ldf;
  memset(f + (size_t)np * L, 0, sizeof(double) * (size_t)(np_part - np) * L);
}

/* X(hamiltonian_elec_base_local_sub), packed real-potential branch
 * (src/hamiltonian/hamiltonian_elec_base_inc.F90:231-241): hpsi(k,p) += v(p) psi(k,p).
 * field != NULL adds the dipole term sum_d field[d]*x_d(p) to v (laser_potential,
 * src/hamiltonian/lasers.F90:823-853, + sign). */
void oracle_vlocal(int np, int ldf, const double* v, const double* xyz, const double* field,
                   const double* psi, double* hpsi) {
  const int L = 1 << ldf;
#pragma omp parallel for schedule(static)
  for (int p = 0; p < np; ++p) {
    double vv = v[p];
    if (field)
      vv = vv + (field[0] * xyz[3 * (size_t)p] + field[1] * xyz[3 * (size_t)p + 1] +
                 field[2] * xyz[3 * (size_t)p + 2]);
    for (int k = 0; k < L; ++k) hpsi[IDX(p, k, L)] = fma(vv, psi[IDX(p, k, L)], hpsi[IDX(p, k, L)]);
  }
}

/* X(hamiltonian_elec_base_nlocal_start), CPU packed branch
 * (src/hamiltonian/hamiltonian_elec_base_inc.F90:618-786): for each matrix, gather
 * lpsi(k, q) = psi(k, map(q)), proj(k, j) = sum_q lpsi(k,q) P(q,j) (dgemm 'N','N'),
 * then proj(k, j) *= scal(j).  offsets use build_accel's layout
 * (src/hamiltonian/hamiltonian_elec_base.F90:785-816); maps are 1-based. */
void oracle_nlocal_start(int nmat, const int* offsets, const double* matrices, const int* maps,
                         const double* scals, int ldf, const double* psi, double* proj) {
  const int L = 1 << ldf;
  for (int a = 0; a < nmat; ++a) {
    const int* o = offsets + 6 * (size_t)a;
    const int npt = o[0], npj = o[1], moff = o[2], mapoff = o[3], soff = o[4];
    /* blocks of up to 8 lanes innermost; each lane sums q = 0..npt-1 with fma from 0 */
    const int kb = L < 8 ? L : 8;
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 0; j < npj; ++j)
      for (int k0 = 0; k0 < L; k0 += kb) {
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int q = 0; q < npt; ++q) {
          const double pj = matrices[moff + (size_t)j * npt + q];
          const double* row = psi + IDX(maps[mapoff + q] - 1, k0, L);
          for (int k = 0; k < kb; ++k) acc[k] = fma(row[k], pj, acc[k]);
        }
        for (int k = 0; k < kb; ++k) proj[IDX(soff + j, k0 + k, L)] = acc[k] * scals[soff + j];
      }
  }
}

/* X(hamiltonian_elec_base_nlocal_finish) (src/hamiltonian/hamiltonian_elec_base_inc.F90:
 * 859-943): per matrix in order, psi_loc(k, q) = sum_j proj(k, j) P(q, j) (dgemm 'N','T'),
 * then hpsi(k, map(q)) += psi_loc(k, q).  Matrices are applied one after the other, so
 * points shared by several spheres accumulate in matrix order. */
void oracle_nlocal_finish(int nmat, const int* offsets, const double* matrices, const int* maps,
                          int ldf, const double* proj, double* hpsi) {
  const int L = 1 << ldf;
  for (int a = 0; a < nmat; ++a) {
    const int* o = offsets + 6 * (size_t)a;
    const int npt = o[0], npj = o[1], moff = o[2], mapoff = o[3], soff = o[4];
#pragma omp parallel for schedule(static)
    for (int q = 0; q < npt; ++q) {
      const int p = maps[mapoff + q] - 1;
      /* lanes innermost; each lane sums j = 0..npj-1 with fma from 0, then adds */
      double t[1024];
      for (int k = 0; k < L; ++k) t[k] = 0.0;
      for (int j = 0; j < npj; ++j) {
        const double pq = matrices[moff + (size_t)j * npt + q];
        const double* pr = proj + IDX(soff + j, 0, L);
        for (int k = 0; k < L; ++k) t[k] = fma(pq, pr[k], t[k]);
      }
      double* out = hpsi + IDX(p, 0, L);
      for (int k = 0; k < L; ++k) out[k] = out[k] + t[k];
    }
  }
}

static int full_projection_size(int nmat, const int* offsets) {
  int s = 0;
  for (int a = 0; a < nmat; ++a)
    if (offsets[6 * a + 4] + offsets[6 * a + 1] > s) s = offsets[6 * a + 4] + offsets[6 * a + 1];
  return s;
}

/* X(hamiltonian_elec_apply_batch) for the terms kinetic + local + KB nonlocal
 * (src/hamiltonian/hamiltonian_elec_inc.F90:77-261): set_bc -> boundaries_set; projections
 * from psi; hpsi = (-1/2m) L psi (weights pre-multiplied, nl_operator_inc.F90:72-80);
 * hpsi += v psi; hpsi += nonlocal. */
int oracle_hamiltonian_apply(const oracle_ham_t* h, double* psi, double* hpsi, int ldf, int terms,
                             int set_bc) {
  const int L = 1 << ldf;
  if (set_bc) oracle_zero_boundary(h->np, h->np_part, ldf, psi);
  double* proj = NULL;
  const int nl = (terms & ORACLE_TERM_NON_LOCAL) && h->nmat > 0;
  if (nl) {
    const int fps = full_projection_size(h->nmat, h->offsets);
    proj = (double*)malloc(sizeof(double) * (size_t)fps * L);
    if (!proj) return -1;
    oracle_nlocal_start(h->nmat, h->offsets, h->matrices, h->maps, h->scals, ldf, psi, proj);
  }
  if (terms & ORACLE_TERM_KINETIC) {
    double wre[128];
    if (h->n > 128) return -1;
    for (int j = 0; j < h->n; ++j) wre[j] = h->w[j] * (-0.5 / h->mass);
    oracle_operate_ri_vec(h->n, wre, h->nri, h->ri, h->rimap_inv, h->rimap_inv + 1, psi, ldf, hpsi);
  } else {
    memset(hpsi, 0, sizeof(double) * (size_t)h->np * L);
  }
  if ((terms & ORACLE_TERM_LOCAL) && h->v) oracle_vlocal(h->np, ldf, h->v, h->xyz, h->field, psi, hpsi);
  if (nl) {
    oracle_nlocal_finish(h->nmat, h->offsets, h->matrices, h->maps, ldf, proj, hpsi);
    free(proj);
  }
  return 0;
}

/* batch_axpy on the inner rows with a complex factor (src/grid/batch_ops_inc.F90:23-106),
 * complex lanes interleaved. */
void oracle_zaxpy(int np, int ldf, double zr, double zi, const double* x, double* y) {
  const size_t L = (size_t)1 << ldf;
  const size_t count = (size_t)np * L;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < count; i += 2) {
    const double xr = x[i], xi = x[i + 1];
    y[i] = y[i] + (zr * xr - zi * xi);
    y[i + 1] = y[i + 1] + (zr * xi + zi * xr);
  }
}

/* exponential_taylor_series_batch (src/electrons/exponential.F90:677-757):
 * zfact = 1; for k = 1..order: zfact *= (-i*deltat)/k; hpsi1 = H (k == 1 ? psi : psi1);
 * psi += zfact * hpsi1; psi1 = hpsi1.  imag_time: deltat -> i*deltat
 * (exponential.F90:609-619).  Work arrays are np_part rows like the reference's clones;
 * every H-apply sets the zero boundary of its input (set_bc default). */
int oracle_taylor(const oracle_ham_t* h, double* psi, int ldf, double deltat, int order, int imag_time) {
  const size_t L = (size_t)1 << ldf;
  const size_t rows = (size_t)h->np_part;
  double* psi1 = (double*)calloc(rows * L, sizeof(double));
  double* hpsi1 = (double*)calloc(rows * L, sizeof(double));
  if (!psi1 || !hpsi1) {
    free(psi1);
    free(hpsi1);
    return -1;
  }
  const double dtr = imag_time ? 0.0 : deltat, dti = imag_time ? deltat : 0.0;
  /* -i * deltat with complex arithmetic, then zfact = (zfact * that) / k */
  const double mr = -0.0, mi = -1.0;
  const double sr = mr * dtr - mi * dti, si = mr * dti + mi * dtr;
  double zr = 1.0, zi = 0.0;
  int rc = 0;
  for (int k = 1; k <= order && rc == 0; ++k) {
    const double tr = zr * sr - zi * si, ti = zr * si + zi * sr;
    zr = tr / k;
    zi = ti / k;
    rc = oracle_hamiltonian_apply(h, k == 1 ? psi : psi1, hpsi1, ldf, ORACLE_TERM_ALL, 1);
    oracle_zaxpy(h->np, ldf, zr, zi, hpsi1, psi);
    if (k != order) memcpy(psi1, hpsi1, sizeof(double) * (size_t)h->np * L);
  }
  free(psi1);
  free(hpsi1);
  return rc;
}

/* batch_axpy with a real factor on the inner rows (src/grid/batch_ops_inc.F90:23-106,
 * the real-constant branch used for real(zfact)): y += a * x on every double. */
void oracle_daxpy(int np, int ldf, double a, const double* x, double* y) {
  const size_t count = (size_t)np << ldf;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < count; ++i) y[i] = y[i] + a * x[i];
}

/* exponential_taylor_series_batch with its optional arguments
 * (src/electrons/exponential.F90:677-757):
 *   psi2 (optional): copy of psi's inner rows first (700), then receives the series at
 *     deltat2 with the same H^k terms (740-743);
 *   inh (optional): zfact = deltat, psi += real(zfact) inh (705-708; psi2 likewise with
 *     deltat2, 710-713), the chain starts from inh (729-731) and the denominators are
 *     k + 1 (718-719): psi += deltat sum_k (-i deltat H)^k / (k+1)! inh.
 * Complex factors multiply as in oracle_taylor; every H-apply zeroes its input's
 * boundary rows (set_bc default), inh's included. */
static int taylor_series(const oracle_ham_t* h, double* psi, int ldf, double dtr, double dti, int order,
                         double* psi2, double dt2r, double dt2i, double* inh) {
  const size_t L = (size_t)1 << ldf;
  const size_t rows = (size_t)h->np_part;
  double* psi1 = (double*)calloc(rows * L, sizeof(double));
  double* hpsi1 = (double*)calloc(rows * L, sizeof(double));
  if (!psi1 || !hpsi1) {
    free(psi1);
    free(hpsi1);
    return -1;
  }
  if (psi2) memcpy(psi2, psi, sizeof(double) * (size_t)h->np * L);
  double zr = 1.0, zi = 0.0, z2r = 1.0, z2i = 0.0;
  if (inh) {
    zr = dtr;
    zi = dti;
    oracle_daxpy(h->np, ldf, zr, inh, psi);
    if (psi2) {
      z2r = dt2r;
      z2i = dt2i;
      oracle_daxpy(h->np, ldf, z2r, inh, psi2);
    }
  }
  const double mr = -0.0, mi = -1.0;
  const double sr = mr * dtr - mi * dti, si = mr * dti + mi * dtr;
  const double s2r = mr * dt2r - mi * dt2i, s2i = mr * dt2i + mi * dt2r;
  int rc = 0;
  for (int k = 1; k <= order && rc == 0; ++k) {
    const int denom = inh ? k + 1 : k;
    const double tr = zr * sr - zi * si, ti = zr * si + zi * sr;
    zr = tr / denom;
    zi = ti / denom;
    if (psi2) {
      const double t2r = z2r * s2r - z2i * s2i, t2i = z2r * s2i + z2i * s2r;
      z2r = t2r / denom;
      z2i = t2i / denom;
    }
    double* src = k > 1 ? psi1 : (inh ? inh : psi);
    rc = oracle_hamiltonian_apply(h, src, hpsi1, ldf, ORACLE_TERM_ALL, 1);
    oracle_zaxpy(h->np, ldf, zr, zi, hpsi1, psi);
    if (psi2) oracle_zaxpy(h->np, ldf, z2r, z2i, hpsi1, psi2);
    if (k != order) memcpy(psi1, hpsi1, sizeof(double) * (size_t)h->np * L);
  }
  free(psi1);
  free(hpsi1);
  return rc;
}

/* exponential_apply_batch, Taylor branch (src/electrons/exponential.F90:581-674):
 * deltat -> i deltat (and deltat2 -> i deltat2) in imaginary time; the homogeneous series
 * (with psi2 when given), then, with inh, the inhomogeneous series (which starts by
 * copying the propagated psi into psi2 again, 700). */
int oracle_taylor_ex(const oracle_ham_t* h, double* psi, int ldf, double deltat, int order, int imag_time,
                     double* psi2, double deltat2, double* inh) {
  const double dtr = imag_time ? 0.0 : deltat, dti = imag_time ? deltat : 0.0;
  const double d2r = imag_time ? 0.0 : deltat2, d2i = imag_time ? deltat2 : 0.0;
  int rc = taylor_series(h, psi, ldf, dtr, dti, order, psi2, d2r, d2i, NULL);
  if (!rc && inh) rc = taylor_series(h, psi, ldf, dtr, dti, order, psi2, d2r, d2i, inh);
  return rc;
}

/* density_calc_accumulate on a device-packed batch, kernels density_real /
 * density_complex (src/states/density.F90:375-417, share/opencl/density.cl:24-61):
 * dd = sum_ist weight(ist) * |psi(ist, p)|^2 in state order, density(p) += dd.  Complex
 * batches: |z|^2 = re*re + im*im. */
void oracle_density_accumulate(int np, int ldf, int nst, int is_complex, const double* weights,
                               const double* psi, double* density) {
  const size_t L = (size_t)1 << ldf;
#pragma omp parallel for schedule(static)
  for (int p = 0; p < np; ++p) {
    double dd = 0.0;
    for (int ist = 0; ist < nst; ++ist) {
      double a2;
      if (is_complex) {
        const double re = psi[IDX(p, 2 * ist, L)], im = psi[IDX(p, 2 * ist + 1, L)];
        a2 = re * re + im * im;
      } else {
        const double f = psi[IDX(p, ist, L)];
        a2 = f * f;
      }
      dd = dd + weights[ist] * a2;
    }
    density[p] = density[p] + dd;
  }
}

int oracle_version(void) { return 2; }
