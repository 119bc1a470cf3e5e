/*
 * octhip.h -- C ABI of the MI355X (gfx950) real-space H-apply path.
 *
 * Drop-in device backend for three Octopus routines, keeping Octopus's packed batch
 * layout (batch_t ff_pack: rows = mesh points, 2^ldf contiguous doubles per row, state
 * index fastest, complex states as interleaved re/im doubles):
 *
 *   nl_operator_operate_batch      src/grid/nl_operator_inc.F90:21-28 (device branch
 *                                  operate_opencl, 257-469; CPU leaf
 *                                  doperate_ri_vec, src/grid/operate.c:40-64)
 *   hamiltonian_elec_apply_batch   src/hamiltonian/hamiltonian_elec_inc.F90:77-261
 *   exponential_apply_batch        src/electrons/exponential.F90:581-757 (Taylor)
 *
 * Conventions shared by every entry point:
 *   - every call returns an int status: OCTHIP_SUCCESS (0) or a negative OCTHIP_ERR_*;
 *     octhip_get_last_error() copies the message of the last failure on this thread.
 *     Octopus's binding turns a nonzero status into messages_fatal, as accel.F90 does
 *     for CUDA errors.
 *   - state buffers (fi, fo, psi, hpsi) are DEVICE pointers owned by the caller
 *     (batch_t%ff_device); descriptors passed to *_create are HOST arrays that the
 *     library deep-copies to the device (as nl_operator_build / build_accel upload).
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).  Calls are
 *     asynchronous on that stream; octhip_stream_synchronize() mirrors accel_finish.
 *   - an input batch has at least np_part rows, an output batch at least np rows,
 *     both with row length 2^ldf doubles (ldf = log2(pack_size_real(1))).
 *   - call from one host thread per handle; handles are not thread safe.
 */
#ifndef OCTHIP_H
#define OCTHIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define OCTHIP_SUCCESS           0
#define OCTHIP_ERR_INVALID      -1   /* bad argument / descriptor */
#define OCTHIP_ERR_HIP          -2   /* HIP runtime error */
#define OCTHIP_ERR_ALLOC        -3   /* device or host allocation failed */
#define OCTHIP_ERR_UNSUPPORTED  -4   /* valid input the device path does not cover */
#define OCTHIP_ERR_STATE        -5   /* library not initialised / wrong handle state */

/* points selector, values of src/grid/nl_operator.F90:130 */
#define OCTHIP_OP_INNER 1
#define OCTHIP_OP_OUTER 2
#define OCTHIP_OP_ALL   3

/* Hamiltonian terms, values of src/hamiltonian/hamiltonian_elec_base.F90:150-155 */
#define OCTHIP_TERM_KINETIC            1
#define OCTHIP_TERM_LOCAL_POTENTIAL    2
#define OCTHIP_TERM_NON_LOCAL_POTENTIAL 4
#define OCTHIP_TERM_ALL                0x7fffffff

typedef struct octhip_nl_operator_s* octhip_nl_operator_t;
typedef struct octhip_hamiltonian_s* octhip_hamiltonian_t;
typedef struct octhip_taylor_graph_s* octhip_taylor_graph_t;
typedef struct octhip_adjoint_s* octhip_adjoint_t;
typedef struct octhip_batched_s* octhip_batched_t;

/* ---------------------------------------------------------------- lifecycle --
 * Replaces accel_init / accel_end (src/basic/accel.F90:264-700, 1105-1175). */
int octhip_init(int device);
int octhip_finalize(void);
int octhip_get_last_error(char* buf, size_t len);
int octhip_version(int* major, int* minor);

/* ------------------------------------------------------------------- memory --
 * Backing store of accel_mem_t%mem (src/basic/accel.F90:150-161, accel_create_buffer,
 * accel_write_buffer, accel_read_buffer, accel_release_buffer, accel_finish). */
int octhip_malloc(void** ptr, size_t bytes);
int octhip_free(void* ptr);
int octhip_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream);
int octhip_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream);
int octhip_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream);
int octhip_stream_synchronize(void* stream);

/* -------------------------------------------------------------------- batch --
 * Device pack/unpack of a batch: the transposes of batch_write_unpacked_to_device and
 * batch_read_device_to_unpacked (src/grid/batch.F90:625-760) with the kernels
 * dpack/zpack/dunpack/zunpack (share/opencl/pack.cl:23-67), which back batch_t's
 * do_pack/do_unpack on the device.  unpacked: nst states, state i starting at element
 * i*ld (ld >= np elements; an element is one double for real batches, one re/im pair of
 * doubles for complex ones), as the columns of batch_t%dff_linear / zff_linear.
 * packed: rows of 2^ldf doubles; state i is element first+i of every row (complex: the
 * doubles 2(first+i), 2(first+i)+1).  Only rows 0..np-1 and the nst states' elements are
 * written; other lanes keep their values.  All pointers are device memory. */
int octhip_batch_pack(const double* unpacked, size_t ld, int nst, int np, int is_complex,
                      double* packed, int ldf, int first, void* stream);
int octhip_batch_unpack(const double* packed, int ldf, int first, int nst, int np,
                        int is_complex, double* unpacked, size_t ld, void* stream);

/* ------------------------------------------------------------- nl_operator --
 * octhip_nl_operator_create replaces the device half of nl_operator_build
 * (src/grid/nl_operator.F90:516-582: buff_ri, buff_map, kernel build).  Inputs are the
 * host arrays nl_operator_build produced, with the C leaf's conventions
 * (src/grid/operate.c:40-45): n = stencil size, w = const weights op%w(:,1),
 * ri = op%ri as n*nri ints (offset of neighbour j of run l at ri[l*n+j]),
 * rimap_inv = op%rimap_inv (nri+1 ints: run l covers points [rimap_inv[l],
 * rimap_inv[l+1]) 0-based).  stencil_points = op%stencil%points (3*n ints, x,y,z of
 * point j at [3*j..3*j+2]) lets the library rebuild the lattice and index neighbours
 * through a compact site map instead of ri; pass NULL to index through ri.
 * coords (3*np_part ints, optional) are the lattice coordinates of every point. */
int octhip_nl_operator_create(int n, const double* w, const int* stencil_points,
                              int nri, const int* ri, const int* rimap_inv,
                              int np, int np_part, const int* coords,
                              octhip_nl_operator_t* op);
int octhip_nl_operator_destroy(octhip_nl_operator_t op);
/* 1 if neighbours are indexed through the rebuilt lattice, 0 if through ri */
int octhip_nl_operator_uses_lattice(octhip_nl_operator_t op, int* flag);

/* fo(:, 0:np-1) = sum_j (factor*w[j]) * fi(:, p + ri[rimap(p)][j]).
 * Mirrors X(nl_operator_operate_batch)(op, fi, fo, points=points, factor=factor) on a
 * BATCH_DEVICE_PACKED batch (src/grid/nl_operator_inc.F90:72-101); the per-lane sum runs
 * over j in stencil order with fused multiply-adds, like the C leaf
 * (src/grid/operate_inc.c:76-118).  Boundary rows (np..np_part-1) of fi are READ as
 * they are: the caller sets them (boundaries_set), as Octopus does before operate.
 * points: OCTHIP_OP_ALL, or the inner / outer runs only (op%inner / op%outer of
 * nl_operator_build, src/grid/nl_operator.F90:457-514: a run is inner when every
 * neighbour of its points is an inner point), through the run-compressed leaf with the
 * same per-lane arithmetic; rows of the other part are not written.  The split is built
 * for every operator, so a domain's operator can overlap its ghost exchange with the
 * inner points as the reference does. */
int octhip_nl_operator_operate_batch(octhip_nl_operator_t op, const double* fi, double* fo,
                                     int ldf, double factor, int points, void* stream);

/* Same contract as the CPU leaf doperate_ri_vec/zoperate_ri_vec
 * (src/grid/operate.c:40-76), on device memory: ri, rimap_inv, rimap_inv_max, fi, fo
 * are device pointers, w is a host array of n doubles (read during the call; it travels
 * as a kernel argument).  Runs l = 0..nri-1 cover points [rimap_inv[l],
 * rimap_inv_max[l]).  Stateless, for callers that keep Octopus's own index buffers on
 * the device: allocates nothing and returns without synchronising (asynchronous on
 * `stream`, like every other entry point). */
int octhip_operate_ri_vec(int n, const double* w, int nri, const int* ri,
                          const int* rimap_inv, const int* rimap_inv_max,
                          const double* fi, int ldf, double* fo, void* stream);

/* ------------------------------------------------------------- hamiltonian --
 * octhip_hamiltonian_create replaces hamiltonian_elec_base_build_proj's device upload
 * (src/hamiltonian/hamiltonian_elec_base.F90:755-909).  kinetic: the Laplacian operator
 * (hm%hm_base%kinetic); mass: hm%mass (kinetic factor -1/(2 mass),
 * src/hamiltonian/hamiltonian_elec_inc.F90:155); vlocal: np doubles (total local
 * potential, src/hamiltonian/hamiltonian_elec.F90:1323); projector matrices in
 * build_accel's flat layout: offsets = 6*nmat ints (POINTS, PROJS, MATRIX, MAP, SCAL,
 * MIX per matrix; MIX must be -1: KB projectors carry no mixing matrix), matrices =
 * concatenated column-major (npoints x nprojs) blocks, maps = concatenated 1-based
 * mesh indices (inner points), scals = concatenated scal(:).  nmat may be 0. */
int octhip_hamiltonian_create(octhip_nl_operator_t kinetic, double mass, const double* vlocal,
                              int nmat, const int* offsets, const double* matrices,
                              const int* maps, const double* scals,
                              octhip_hamiltonian_t* hm);
int octhip_hamiltonian_destroy(octhip_hamiltonian_t hm);
/* Release the projector and Taylor work buffers this Hamiltonian keeps for `stream` (one
 * set per stream that calls it, so batches on different streams run concurrently).  No
 * counterpart in the reference: for callers that retire streams, whose sets would
 * otherwise stay allocated until destroy.  Synchronises the stream first; buffers
 * captured into a Taylor graph are not affected. */
int octhip_hamiltonian_release_stream(octhip_hamiltonian_t hm, void* stream);
/* Device memory held by a Hamiltonian handle, in bytes: bytes[0] its data (potential,
 * projector tables, periodic tables and phases), bytes[1] the work buffers of the streams
 * that called it (projector and Taylor buffers, Krylov vectors), bytes[2] the work sets
 * handed to Taylor graphs (freed with each graph), bytes[3] retired work buffers (outgrown
 * by a longer row; kept until octhip_hamiltonian_trim or destroy).  No counterpart in the
 * reference, whose device buffers are sized by the states block size once. */
int octhip_hamiltonian_memory(octhip_hamiltonian_t hm, size_t bytes[4]);
/* Free the retired work buffers (synchronises the device first). */
int octhip_hamiltonian_trim(octhip_hamiltonian_t hm);
/* replace the local potential (np doubles, host) -- hamiltonian_elec_update */
int octhip_hamiltonian_set_potential(octhip_hamiltonian_t hm, const double* vlocal, void* stream);
/* v_eff(p) = vlocal(p) + sum_d field[d] * x_d(p): the length-gauge dipole term of
 * laser_potential (src/hamiltonian/lasers.F90:823-853, added with a + sign).
 * positions = 3*np doubles (x,y,z of point p at [3p..3p+2]); field = 3 doubles.
 * Positions are uploaded once; later calls with positions == NULL only change field. */
int octhip_hamiltonian_set_dipole_field(octhip_hamiltonian_t hm, const double* positions,
                                        const double* field, void* stream);

/* Periodic systems and k-points (the inputs hamiltonian_elec_update builds,
 * src/hamiltonian/hamiltonian_elec.F90:1178-1300):
 * set_periodic: bnd_src = np_part - np ints, the 0-based inner row whose value boundary
 *   row np+b takes (mesh_periodic_point; the periodic branch of boundaries_set,
 *   src/grid/boundaries_inc.F90:366-743); NULL returns to zero boundaries.
 * set_phases: the Bloch phases of one k-point as complex (re, im) pairs:
 *   phase = hm_base%phase(:, ik), np_part values exp(-i k.x); phase_corr =
 *   hm_base%phase_corr(:, ik), np_part - np values; projector_phases =
 *   hm_base%projector_phases(:, 1, imat, ik) for every matrix in order (total_points
 *   values, as buff_projector_phases).  NULL phase removes them.
 * With either set, octhip_hamiltonian_apply_batch applies the reference's periodic path
 * to a batch WITHOUT the phase (hamiltonian_elec_inc.F90:116-154, 243-247): boundary rows
 * from their images, epsi = phase psi (np_part rows, a work buffer), H on epsi with the
 * projector phases, hpsi = conj(phase) H epsi.  The propagators (Taylor, Lanczos,
 * Chebyshev) take batches WITH the phase, as the reference calls them between
 * set_phase_corr and unset_phase_corr (src/td/propagation_ops_elec.F90:238-246); the
 * exp-mid/ETRS drivers take batches without it and set/unset it around each batch.
 * Taylor graphs, batched launches and the adjoint cover finite systems only
 * (OCTHIP_ERR_UNSUPPORTED on periodic handles), and both calls are refused while the
 * handle has Taylor graphs (they keep the launches they were captured with).
 * Synchronise the device first. */
int octhip_hamiltonian_set_periodic(octhip_hamiltonian_t hm, const int* bnd_src);
int octhip_hamiltonian_set_phases(octhip_hamiltonian_t hm, const double* phase,
                                  const double* phase_corr, const double* projector_phases);
/* X(hamiltonian_elec_apply_batch) on a complex batch whose phase state is given:
 * psi_has_phase != 0 is psib%has_phase (boundary rows set with phase_corr, no phase
 * multiplications); 0 is octhip_hamiltonian_apply_batch. */
int octhip_hamiltonian_apply_batch_phase(octhip_hamiltonian_t hm, double* psi, double* hpsi, int ldf,
                                         int terms, int set_bc, int psi_has_phase, void* stream);
/* hamiltonian_elec_base_phase (src/hamiltonian/hamiltonian_elec_base.F90:979-1139):
 * psi(:, p) = phase(p) src(p) (conjugate: conj(phase(p)) src(p)) for p < np (np <=
 * np_part); src NULL = in place.  With np = mesh np this is set_phase_corr /
 * unset_phase_corr (933-975). */
int octhip_hamiltonian_base_phase(octhip_hamiltonian_t hm, double* psi, const double* src, int ldf,
                                  int np, int conjugate, void* stream);

/* hpsi = H psi on a packed batch: X(hamiltonian_elec_apply_batch)(hm, mesh, psib, hpsib,
 * terms, set_bc) (src/hamiltonian/hamiltonian_elec_inc.F90:77-261) for the terms
 * kinetic (-1/2m Laplacian), local potential and KB nonlocal projectors, in Octopus's
 * order: hpsi = K psi; hpsi += v psi; hpsi += sum_a P_a diag(scal_a) P_a^T psi.
 * psi: device, >= np_part rows; hpsi: device, >= np rows; 2^ldf doubles per row.
 * set_bc != 0 zeroes psi's boundary rows first (boundaries_set, zero Dirichlet,
 * src/grid/boundaries_inc.F90:275-363); with set_bc == 0 boundary rows are read as
 * they are.  is_complex: batch holds complex states (lanes are re/im pairs). */
int octhip_hamiltonian_apply_batch(octhip_hamiltonian_t hm, double* psi, double* hpsi,
                                   int ldf, int is_complex, int terms, int set_bc,
                                   void* stream);

/* ----------------------------------------------------------------- taylor --
 * psi <- sum_{k=0..order} (-i deltat H)^k / k! psi in place: exponential_apply_batch
 * with TDExponentialMethod = taylor (src/electrons/exponential.F90:581-757).
 * imag_time != 0 uses deltat -> i*deltat (exponential.F90:609-619).  The batch must be
 * complex (ldf >= 1).  psi's boundary rows are zeroed (boundaries_set).  Work buffers
 * are owned by the handle and reused across calls. */
int octhip_exponential_taylor_batch(octhip_hamiltonian_t hm, double* psi, int ldf,
                                    double deltat, int order, int imag_time, void* stream);

/* exponential_apply_batch with its optional arguments (src/electrons/exponential.F90:581-674,
 * Taylor branch; exponential_taylor_series_batch 677-757):
 *   psi2 (optional, device, np_part rows, same ldf) = psib2: receives psi's inner rows and
 *     the series at deltat2 built from the same H^k psi (only its np inner rows are
 *     written);
 *   inh (optional, device, np_part rows) = inh_psib: after the homogeneous series a second
 *     series adds deltat sum_k (-i deltat H)^k/(k+1)! inh (psi2 likewise, after receiving
 *     the propagated psi again, as the reference's second call does); inh's boundary rows
 *     are zeroed by its H-apply;
 *   weights (optional, host, 2^(ldf-1) doubles: kweight*occ per state, 0 for padding) and
 *     density (device, np doubles): density += sum_states weights |final state|^2 of psi2
 *     when given, else psi -- the density_calc_accumulate that
 *     propagation_ops_elec_fuse_density_exp_apply (src/td/propagation_ops_elec.F90:
 *     261-333) runs after each batch, here fused into the last pass for orders 2..4.
 * With all optional pointers NULL this is octhip_exponential_taylor_batch. */
int octhip_exponential_apply_batch(octhip_hamiltonian_t hm, double* psi, int ldf, double deltat,
                                   int order, int imag_time, double* psi2, double deltat2,
                                   double* inh, const double* weights, double* density,
                                   void* stream);

/* density_calc_accumulate on a BATCH_DEVICE_PACKED batch (src/states/density.F90:375-417,
 * kernels density_real/density_complex of share/opencl/density.cl): density(p) +=
 * sum_ist weights[ist] |psi(ist, p)|^2 over the np inner rows.  psi, weights
 * (2^ldf doubles for real rows, 2^(ldf-1) for complex) and density are DEVICE pointers,
 * as the reference kernel takes buff_weight and buff_density. */
int octhip_density_accumulate(const double* psi, int ldf, int np, int is_complex,
                              const double* weights, double* density, void* stream);

/* ------------------------------------------------------------ mesh_batch --
 * Batched reductions over the np inner rows of packed batches (DEVICE pointers), with the
 * mesh inner product <a|b> = vol * sum_p conj(a) b (vol = the volume element).
 * mesh_batch_dotp_vector (src/grid/mesh_batch_inc.F90:379-549): dot = <a_c|b_c> per column
 * c: complex batches give 2^(ldf-1) (re, im) pairs, real batches 2^ldf doubles (host).
 * The sums are deterministic (fixed partition and order).  Allocates and frees its
 * reduction scratch on every call and synchronises the stream. */
int octhip_mesh_batch_dotp_vector(const double* a, const double* b, int ldf, int np, int is_complex,
                                  double vol, double* dot, void* stream);
/* octhip_mesh_batch_dotp_vector for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): no counterpart in the
 * reference.  Same kernels and partition, so the same bits; it allocates nothing and
 * does not synchronise: the reduction scratch is a fixed device array of the library
 * (8 MiB + 16 KiB per device, held until the process ends, enough for every row length
 * up to 2^10 doubles), and the results go to `dot`, a DEVICE array of 2 doubles per
 * column ((re, im); im = 0 for real rows).  Every call on one device shares that array:
 * calls that overlap on one GPU (different streams, host threads) race. */
int octhip_mesh_batch_dotp_vector_noalloc(const double* a, const double* b, int ldf, int np,
                                          int is_complex, double vol, double* dot, void* stream);
/* mesh_batch_nrm2 (src/grid/mesh_batch.F90:74-95): nrm2 = sqrt(<a_c|a_c>) per column (host) */
int octhip_mesh_batch_nrm2(const double* a, int ldf, int np, int is_complex, double vol,
                           double* nrm2, void* stream);
/* mesh_batch_orthogonalization, classical Gram-Schmidt (src/grid/mesh_batch_inc.F90:
 * 1168-1274) of the complex batch phi against the nvec batches psi[k]: all overlaps
 * <psi_k|phi> first, then phi -= sum_k overlap_k psi_k; then optionally norm (host, nst
 * doubles) and normalize.  overlap (optional, host): nvec x nst complex pairs,
 * column-major (the transposed layout the reference returns).  Only the first nst
 * states of the rows are touched. */
int octhip_mesh_batch_orthogonalization(int nvec, const double* const* psi, double* phi, int ldf,
                                        int nst, int np, double vol, int normalize,
                                        double* overlap, double* norm, void* stream);

/* exponential_lanczos_batch (src/electrons/exponential.F90:766-900): psi <- exp(-i deltat
 * H) psi (imag_time: deltat -> i deltat) in the Krylov space of up to `order` Lanczos
 * steps for the first nst states of the complex batch; stops when every state's error
 * estimate (Saad's Er2) is below lanczos_tol after more than 3 steps, or on a happy
 * breakdown.  With inh (device, np_part rows, read only): psi += deltat phi(-i deltat H)
 * inh instead (zlalg_phi).  vol = the volume element of the mesh inner product.
 * iterations/residual (optional, host): the Krylov dimension used and the largest error
 * estimate (the reference warns when it exceeds lanczos_tol).  Synchronises the stream
 * (the projected problems are solved on the host). */
int octhip_exponential_lanczos_batch(octhip_hamiltonian_t hm, double* psi, int ldf, int nst,
                                     double deltat, int imag_time, int order, double lanczos_tol,
                                     const double* inh, double vol, int* iterations,
                                     double* residual, void* stream);
/* exponential_cheby_batch (src/electrons/exponential.F90:903-971): Chebyshev expansion of
 * exp(-i deltat H) of `order` terms with Clenshaw's recurrence, for H's spectrum inside
 * spectral_middle_point +- spectral_half_span (hamiltonian_elec_span,
 * src/hamiltonian/hamiltonian_elec.F90:898-909). */
int octhip_exponential_cheby_batch(octhip_hamiltonian_t hm, double* psi, int ldf, double deltat,
                                   int order, double spectral_middle_point,
                                   double spectral_half_span, void* stream);

/* Exponential-midpoint step for independent particles (exponential_midpoint,
 * src/td/propagator_expmid.F90:52-100): the Hamiltonian is updated to t - dt/2
 * (field_mid = the dipole field there, 3 doubles; NULL keeps the current potential; the
 * positions must have been uploaded with octhip_hamiltonian_set_dipole_field), then every
 * batch psi[b] (device, np_part rows, one ldf) is propagated by dt with Taylor of `order`
 * and the density of the results accumulated (propagation_ops_elec_fuse_density_exp_apply).
 * weights: host, nbatch x 2^(ldf-1); density: device, np doubles, zeroed first
 * (density_calc_init); both may be NULL. */
int octhip_propagator_expmid(octhip_hamiltonian_t hm, double* const* psi, int nbatch, int ldf,
                             double dt, int order, const double* field_mid,
                             const double* weights, double* density, void* stream);
/* Enforced time-reversal-symmetry step for independent particles (td_etrs,
 * src/td/propagator_etrs.F90:68-150): every batch propagated by dt/2 with H(t - dt)
 * (field_old), the Hamiltonian updated to t (field_new), every batch propagated by dt/2
 * with the density accumulated.  Arguments as octhip_propagator_expmid. */
int octhip_propagator_etrs(octhip_hamiltonian_t hm, double* const* psi, int nbatch, int ldf,
                           double dt, int order, const double* field_old, const double* field_new,
                           const double* weights, double* density, void* stream);

/* One Taylor time step over nbatch batches (psi[b], device, same ldf) captured once into
 * a HIP graph and replayed with a single launch: the loop over batches of
 * propagation_ops_elec_exp_apply (src/td/propagation_ops_elec.F90:232-249), each
 * te%apply_batch being octhip_exponential_taylor_batch, without per-kernel launch cost.  The graph holds the batch pointers,
 * deltat, order and the handle's potential buffers as they are at capture;
 * set_potential and set_dipole_field update those buffers in place, so a field that
 * changes between launches is seen when one was set before capture.  The graph owns the
 * work buffers it was captured with (one set per branch; destroy frees them, after
 * waiting for the device); only the potential buffers are shared with the handle, so
 * launches overlapping a set_potential/set_dipole_field race with it.  Destroy the
 * graphs of a Hamiltonian before or after it: octhip_hamiltonian_destroy frees the sets
 * of graphs still alive, which then own nothing but may no longer be launched.
 * stream == NULL captures on a private stream; the graph may be launched on any
 * stream. */
int octhip_taylor_graph_create(octhip_hamiltonian_t hm, double* const* psi, int nbatch, int ldf,
                               double deltat, int order, int imag_time, void* stream,
                               octhip_taylor_graph_t* graph);
int octhip_taylor_graph_launch(octhip_taylor_graph_t graph, void* stream);
int octhip_taylor_graph_destroy(octhip_taylor_graph_t graph);

/* ---------------------------------------------------------------- batched --
 * Several independent (Hamiltonian, batch) pairs -- systems of uneven size, several
 * batches per system -- propagated together: every kernel kind of a Taylor order runs
 * once for all pairs (a work list of (pair, tile segment, lane chunk) for the fused
 * stencil pass, block-offset tables for the projector kernels), instead of one launch
 * sequence per batch and system as in the reference's loops (propagation_ops_elec_exp_apply,
 * src/td/propagation_ops_elec.F90:232-249, per system).  Results are bit-identical to
 * octhip_exponential_taylor_batch / octhip_hamiltonian_apply_batch per pair.
 * hms[i], psi[i] (device, np_part rows), ldf[i] (complex batches), hpsi[i] (device, np
 * rows; the whole array may be NULL if octhip_batched_apply is not used).  Every pair
 * needs the order-4 star on a lattice; pairs may share a Hamiltonian (each pair has its
 * own work buffers).  order in 2..32.  The pointers are fixed at create. */
int octhip_batched_create(int npairs, const octhip_hamiltonian_t* hms, double* const* psi,
                          double* const* hpsi, const int* ldf, double deltat, int order,
                          octhip_batched_t* batched);
/* hpsi[i] = H_i psi[i] for every pair (zero boundary, every term) */
int octhip_batched_apply(octhip_batched_t batched, void* stream);
/* psi[i] <- sum_k (-i deltat H_i)^k / k! psi[i] for every pair */
int octhip_batched_taylor_step(octhip_batched_t batched, void* stream);
/* one batched Taylor step captured as a graph (launch and destroy with the
 * octhip_taylor_graph_* entry points) */
int octhip_batched_graph_create(octhip_batched_t batched, void* stream, octhip_taylor_graph_t* graph);
int octhip_batched_destroy(octhip_batched_t batched);

/* ---------------------------------------------------------------- adjoint --
 * Differentiable Taylor propagation of one complex batch under a dipole field
 * H_s = H + eps_s (pol . r) (the length-gauge laser term, src/hamiltonian/lasers.F90:823-853),
 * psi_{s+1} = sum_{n<=order} (-i deltat H_s)^n / n! psi_s for s = 0..nsteps-1.
 * Counterpart of the optimal-control forward/backward pair (src/opt_control/propagation.F90:
 * 579-783 bwd_step_2, 1016-1103 update_field, 1108-1205 checkpoints), but exact for the
 * discrete propagator: given chi_N with dJ = 2 Re <chi_N|d psi_N> (<a|b> = vol_pp sum
 * conj(a) b), backward returns dJ/d eps_s (host, nsteps) and dJ/dv(r) (device, np) and
 * leaves chi_0 = dJ/d psi_0* in chi.  Forward stores psi every checkpoint_every steps;
 * backward recomputes each segment from its checkpoint, the next segment on an internal
 * second stream while the current one runs backward, and accumulates the gradients on a
 * third (memory: ceil(nsteps/K) checkpoints of np rows plus 2*(K+1) + 2*K*(order-1) +
 * 2*order + 2 batches of np rows; the Hamiltonian's work set for the second stream exists
 * only while backward runs).
 * create uploads positions (3*np doubles, host) into the Hamiltonian; forward leaves the
 * Hamiltonian's dipole field at the field of its last step (backward forms its own
 * potentials and leaves the field as it is). */
int octhip_adjoint_create(octhip_hamiltonian_t hm, const double* positions, const double* pol,
                          int ldf, int nsteps, double deltat, int order, int checkpoint_every,
                          double vol_pp, octhip_adjoint_t* adj);
/* psi (device, np_part rows): in psi_0, out psi_N.  eps: nsteps doubles (host). */
int octhip_adjoint_forward(octhip_adjoint_t adj, double* psi, const double* eps, void* stream);
/* chi (device, np_part rows): in chi_N, out chi_0.  dj_deps: host, nsteps.
 * dj_dv: device, np (overwritten).  Synchronises the stream before returning. */
int octhip_adjoint_backward(octhip_adjoint_t adj, double* chi, double* dj_deps, double* dj_dv,
                            void* stream);
int octhip_adjoint_destroy(octhip_adjoint_t adj);

/* --------------------------------------------------------------- profiling --
 * Per-kernel device timing with HIP events recorded on the launch stream, the device-side
 * counterpart of profiling_in/profiling_out (src/basic/profiling.F90:503-625).  Regions
 * are kernel names ("k_zmarch", "k_apply", "k_bra", "k_bra_reduce", "k_ket", "k_ket_sum", "k_zero_rows", "k_zaxpy", "k_density").
 * start clears earlier records; get synchronises the recorded events and returns the
 * summed milliseconds and the number of launches of one region. */
int octhip_profiling_start(void);
int octhip_profiling_stop(void);
int octhip_profiling_get(const char* region, double* total_ms, int* count);
/* record only one region (each record is two HIP events on the launch stream, which
 * costs device time between kernels); NULL or "" records every region again */
int octhip_profiling_select(const char* region);

/* ------------------------------------------------------------------ tuning --
 * No counterpart in the reference (Octopus fixes launch shapes in accel.F90).  Key/value
 * knobs for the autotuner; results never depend on them beyond floating-point summation
 * order (the stencil kernels are bit-identical to each other under every setting).
 *   KERNEL           0 automatic (z-march for the order-4 star on a lattice, gather
 *                    otherwise), 1 gather, 2 z-march
 *   XCD_CHUNK        gather: consecutive point tiles given to one XCD (1 = round-robin)
 *   GATHER_TILE      gather: passes over a workgroup's point slots per tile (1..8)
 *   ZSEG             z-march: output planes per work item (1..64)
 *   LANE_CHUNK       z-march: lanes of a row handled by one workgroup (1..32, power of 2;
 *                    0 = min(row length, 16))
 *   LANES_PER_THREAD z-march: 1 or 2 doubles per thread (0 = automatic)
 *   BRA_MODE         projector bra: 0 plain FMA reduction, 1 FP64 MFMA
 * set returns OCTHIP_ERR_INVALID for an unknown key or a value out of range. */
#define OCTHIP_TUNE_KERNEL            1
#define OCTHIP_TUNE_XCD_CHUNK         2
#define OCTHIP_TUNE_GATHER_TILE       3
#define OCTHIP_TUNE_ZSEG              4
#define OCTHIP_TUNE_LANE_CHUNK        5
#define OCTHIP_TUNE_LANES_PER_THREAD  6
#define OCTHIP_TUNE_BRA_MODE          16
int octhip_nl_operator_set_tuning(octhip_nl_operator_t op, int key, int value);
int octhip_nl_operator_get_tuning(octhip_nl_operator_t op, int key, int* value);
int octhip_hamiltonian_set_tuning(octhip_hamiltonian_t hm, int key, int value);

/* Autotuner: times every candidate launch shape of the H-apply for rows of 2^ldf lanes
 * (z-march lane chunk x lanes per thread x z segment, the tiled gather, MFMA or plain
 * bra) on scratch batches (allocated here), reps applies each, on `stream`, and leaves
 * the fastest on the handle; *best_us (optional) = its time per H-apply.  Results are
 * kept in a process-wide table keyed by (np, np_part, number of projector matrices,
 * ldf): unless force != 0 a known shape is configured without timing.  No counterpart in
 * the reference (its device path has fixed launch shapes). */
int octhip_hamiltonian_autotune(octhip_hamiltonian_t hm, int ldf, int reps, int force, void* stream,
                                double* best_us);
/* StatesBlockSize by time (src/states/states_elec.F90:447-475, where it is an input
 * variable, 32 on GPUs): for block sizes 8..64 (up to pad_pow2(nst)) tunes the launch shape
 * of that row length and times one Taylor-4 step of a scratch batch; returns the block
 * size with the smallest time for nst states and its time per step (microseconds, all
 * ceil(nst/pack) batches) and leaves that row length's shape configured.  Kept in the
 * tuning table keyed by (np, np_part, matrices, nst).  Synchronises the stream. */
int octhip_hamiltonian_autotune_pack(octhip_hamiltonian_t hm, int nst, int reps, int force, void* stream,
                                     int* pack_size, double* step_us);
/* the tuning table as text (one line per shape and per block-size choice) */
int octhip_tuning_save(const char* path);
int octhip_tuning_load(const char* path);

/* Test support, no counterpart in the reference: the allocator hook of the failure-path
 * tests.  After `after` more successful device allocations every allocation of the
 * library fails with OCTHIP_ERR_ALLOC ("injected allocation failure"), as a
 * hipErrorOutOfMemory would; after < 0 turns the hook off (the default). */
int octhip_inject_alloc_failure(long after);

#ifdef __cplusplus
}
#endif

#endif /* OCTHIP_H */
