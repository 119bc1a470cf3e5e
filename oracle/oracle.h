/* oracle.h -- CPU restatement of the H-apply path (checker and CPU baseline only). */
#ifndef ORACLE_H
#define ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_TERM_KINETIC 1
#define ORACLE_TERM_LOCAL 2
#define ORACLE_TERM_NON_LOCAL 4
#define ORACLE_TERM_ALL 0x7fffffff

typedef struct {
  int np, np_part;
  /* kinetic operator (nl_operator_t after nl_operator_build) */
  int n, nri;
  const double* w;
  const int* ri;        /* n*nri */
  const int* rimap_inv; /* nri+1 */
  double mass;
  /* local potential and optional dipole term */
  const double* v;      /* np or NULL */
  const double* xyz;    /* 3*np or NULL */
  const double* field;  /* 3 or NULL */
  /* KB projectors in build_accel layout */
  int nmat;
  const int* offsets;
  const double* matrices;
  const int* maps;
  const double* scals;
} oracle_ham_t;

void oracle_operate_ri_vec(int n, const double* w, int nri, const int* ri, const int* imin,
                           const int* imax, const double* fi, int ldf, double* fo);
void oracle_zero_boundary(int np, int np_part, int ldf, double* f);
void oracle_vlocal(int np, int ldf, const double* v, const double* xyz, const double* field,
                   const double* psi, double* hpsi);
void oracle_nlocal_start(int nmat, const int* offsets, const double* matrices, const int* maps,
                         const double* scals, int ldf, const double* psi, double* proj);
void oracle_nlocal_finish(int nmat, const int* offsets, const double* matrices, const int* maps,
                          int ldf, const double* proj, double* hpsi);
int oracle_hamiltonian_apply(const oracle_ham_t* h, double* psi, double* hpsi, int ldf, int terms,
                             int set_bc);
void oracle_zaxpy(int np, int ldf, double zr, double zi, const double* x, double* y);
int oracle_taylor(const oracle_ham_t* h, double* psi, int ldf, double deltat, int order, int imag_time);
void oracle_daxpy(int np, int ldf, double a, const double* x, double* y);
int oracle_taylor_ex(const oracle_ham_t* h, double* psi, int ldf, double deltat, int order, int imag_time,
                     double* psi2, double deltat2, double* inh);
void oracle_density_accumulate(int np, int ldf, int nst, int is_complex, const double* weights,
                               const double* psi, double* density);
int oracle_version(void);

#ifdef __cplusplus
}
#endif

#endif
