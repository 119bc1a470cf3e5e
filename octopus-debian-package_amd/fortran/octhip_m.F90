! iso_c_binding interface to liboctopus_hip.so (include/octhip.h).
!
! This is the module Octopus's Fortran would `use` in the BATCH_DEVICE_PACKED branches of
! X(nl_operator_operate_batch) (src/grid/nl_operator_inc.F90:257-469),
! X(hamiltonian_elec_apply_batch) (src/hamiltonian/hamiltonian_elec_inc.F90:77-261) and
! exponential_taylor_series_batch (src/electrons/exponential.F90:677-757).  Device
! buffers are type(c_ptr), as accel_mem_t%mem is on the CUDA build (src/basic/accel.F90:
! 150-161); handles are type(c_ptr); every function returns the library's status code
! (0 = success), which the caller turns into messages_fatal as the reference does with
! its device errors.  Integer arguments are c_int and passed by value, except arrays.
module octhip_m
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: &
    OCTHIP_SUCCESS = 0, OCTHIP_ERR_INVALID = -1, OCTHIP_ERR_HIP = -2, OCTHIP_ERR_ALLOC = -3, &
    OCTHIP_ERR_UNSUPPORTED = -4, OCTHIP_ERR_STATE = -5
  integer(c_int), parameter, public :: OCTHIP_OP_INNER = 1, OCTHIP_OP_OUTER = 2, OCTHIP_OP_ALL = 3
  integer(c_int), parameter, public :: &
    OCTHIP_TERM_KINETIC = 1, OCTHIP_TERM_LOCAL_POTENTIAL = 2, OCTHIP_TERM_NON_LOCAL_POTENTIAL = 4, &
    OCTHIP_TERM_ALL = int(z'7fffffff', c_int)

  public :: octhip_init, octhip_finalize, octhip_get_last_error, octhip_version, &
    octhip_malloc, octhip_free, octhip_memcpy_htod, octhip_memcpy_dtoh, octhip_stream_synchronize, &
    octhip_nl_operator_create, octhip_nl_operator_destroy, octhip_nl_operator_operate_batch, &
    octhip_operate_ri_vec, octhip_hamiltonian_create, octhip_hamiltonian_destroy, &
    octhip_hamiltonian_release_stream, &
    octhip_hamiltonian_set_potential, octhip_hamiltonian_set_dipole_field, &
    octhip_hamiltonian_apply_batch, octhip_exponential_taylor_batch, &
    octhip_taylor_graph_create, octhip_taylor_graph_launch, octhip_taylor_graph_destroy, &
    octhip_adjoint_create, octhip_adjoint_forward, octhip_adjoint_backward, octhip_adjoint_destroy, &
    octhip_batch_pack, octhip_batch_unpack, octhip_exponential_apply_batch, octhip_density_accumulate, &
    octhip_propagator_expmid, octhip_propagator_etrs, &
    octhip_memcpy_dtod, octhip_nl_operator_set_tuning, octhip_nl_operator_get_tuning, &
    octhip_nl_operator_uses_lattice, octhip_hamiltonian_set_tuning, octhip_hamiltonian_autotune, &
    octhip_hamiltonian_autotune_pack, octhip_tuning_save, octhip_tuning_load, &
    octhip_hamiltonian_memory, octhip_hamiltonian_trim, &
    octhip_hamiltonian_set_periodic, octhip_hamiltonian_set_phases, octhip_hamiltonian_apply_batch_phase, &
    octhip_hamiltonian_base_phase, octhip_exponential_lanczos_batch, octhip_exponential_cheby_batch, &
    octhip_mesh_batch_dotp_vector, octhip_mesh_batch_dotp_vector_noalloc, octhip_mesh_batch_nrm2, &
    octhip_mesh_batch_orthogonalization, octhip_batched_create, octhip_batched_apply, &
    octhip_batched_taylor_step, octhip_batched_graph_create, octhip_batched_destroy, &
    octhip_last_error_message

  interface
    integer(c_int) function octhip_init(device) bind(C, name="octhip_init")
      import :: c_int
      integer(c_int), value :: device
    end function
    integer(c_int) function octhip_finalize() bind(C, name="octhip_finalize")
      import :: c_int
    end function
    integer(c_int) function octhip_get_last_error(buf, len) bind(C, name="octhip_get_last_error")
      import :: c_int, c_char, c_size_t
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_size_t), value :: len
    end function
    integer(c_int) function octhip_version(major, minor) bind(C, name="octhip_version")
      import :: c_int
      integer(c_int), intent(out) :: major, minor
    end function
    integer(c_int) function octhip_malloc(ptr, bytes) bind(C, name="octhip_malloc")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), intent(out) :: ptr
      integer(c_size_t), value :: bytes
    end function
    integer(c_int) function octhip_free(ptr) bind(C, name="octhip_free")
      import :: c_int, c_ptr
      type(c_ptr), value :: ptr
    end function
    integer(c_int) function octhip_memcpy_htod(dst, src, bytes, stream) bind(C, name="octhip_memcpy_htod")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: dst, src, stream
      integer(c_size_t), value :: bytes
    end function
    integer(c_int) function octhip_memcpy_dtoh(dst, src, bytes, stream) bind(C, name="octhip_memcpy_dtoh")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: dst, src, stream
      integer(c_size_t), value :: bytes
    end function
    integer(c_int) function octhip_stream_synchronize(stream) bind(C, name="octhip_stream_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: stream
    end function

    ! nl_operator_build's device upload (src/grid/nl_operator.F90:516-582): w = op%w(:,1),
    ! ri = op%ri (n x nri), rimap_inv = op%rimap_inv (0-based half-open runs), points =
    ! op%stencil%points (3 x n) or c_null_ptr, coords = lattice coordinates or c_null_ptr
    integer(c_int) function octhip_nl_operator_create(n, w, stencil_points, nri, ri, rimap_inv, &
        np, np_part, coords, op) bind(C, name="octhip_nl_operator_create")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: n, nri, np, np_part
      real(c_double), intent(in) :: w(*)
      type(c_ptr), value :: stencil_points, coords
      integer(c_int), intent(in) :: ri(*), rimap_inv(*)
      type(c_ptr), intent(out) :: op
    end function
    integer(c_int) function octhip_nl_operator_destroy(op) bind(C, name="octhip_nl_operator_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: op
    end function
    ! X(nl_operator_operate_batch)(op, fi, fo, points=points, factor=factor) on packed
    ! device batches; ldf = log2(pack_size_real(1))
    integer(c_int) function octhip_nl_operator_operate_batch(op, fi, fo, ldf, factor, points, stream) &
        bind(C, name="octhip_nl_operator_operate_batch")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: op, fi, fo, stream
      integer(c_int), value :: ldf, points
      real(c_double), value :: factor
    end function
    ! doperate_ri_vec/zoperate_ri_vec (src/grid/operate.c:40-76) on device arrays
    integer(c_int) function octhip_operate_ri_vec(n, w, nri, ri, rimap_inv, rimap_inv_max, fi, ldf, fo, &
        stream) bind(C, name="octhip_operate_ri_vec")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: n, nri, ldf
      real(c_double), intent(in) :: w(*)
      type(c_ptr), value :: ri, rimap_inv, rimap_inv_max, fi, fo, stream
    end function

    ! build_accel's upload (src/hamiltonian/hamiltonian_elec_base.F90:755-909)
    integer(c_int) function octhip_hamiltonian_create(kinetic, mass, vlocal, nmat, offsets, matrices, &
        maps, scals, hm) bind(C, name="octhip_hamiltonian_create")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: kinetic, vlocal, offsets, matrices, maps, scals
      real(c_double), value :: mass
      integer(c_int), value :: nmat
      type(c_ptr), intent(out) :: hm
    end function
    integer(c_int) function octhip_hamiltonian_destroy(hm) bind(C, name="octhip_hamiltonian_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: hm
    end function
    integer(c_int) function octhip_hamiltonian_release_stream(hm, stream) &
        bind(C, name="octhip_hamiltonian_release_stream")
      import :: c_int, c_ptr
      type(c_ptr), value :: hm, stream
    end function
    integer(c_int) function octhip_hamiltonian_set_potential(hm, vlocal, stream) &
        bind(C, name="octhip_hamiltonian_set_potential")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: hm, stream
      real(c_double), intent(in) :: vlocal(*)
    end function
    integer(c_int) function octhip_hamiltonian_set_dipole_field(hm, positions, field, stream) &
        bind(C, name="octhip_hamiltonian_set_dipole_field")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: hm, positions, stream
      real(c_double), intent(in) :: field(3)
    end function
    ! X(hamiltonian_elec_apply_batch)(hm, mesh, psib, hpsib, terms, set_bc)
    integer(c_int) function octhip_hamiltonian_apply_batch(hm, psi, hpsi, ldf, is_complex, terms, set_bc, &
        stream) bind(C, name="octhip_hamiltonian_apply_batch")
      import :: c_int, c_ptr
      type(c_ptr), value :: hm, psi, hpsi, stream
      integer(c_int), value :: ldf, is_complex, terms, set_bc
    end function
    ! exponential_apply_batch(te, ..., psib, deltat, imag_time=imag_time), Taylor
    integer(c_int) function octhip_exponential_taylor_batch(hm, psi, ldf, deltat, order, imag_time, stream) &
        bind(C, name="octhip_exponential_taylor_batch")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: hm, psi, stream
      integer(c_int), value :: ldf, order, imag_time
      real(c_double), value :: deltat
    end function
    ! exponential_apply_batch(te, ..., psib, deltat, psib2, deltat2, imag_time=, inh_psib=) with
    ! the density of the result (propagation_ops_elec_fuse_density_exp_apply); pass
    ! c_null_ptr for absent optional batches and for weights/density
    integer(c_int) function octhip_exponential_apply_batch(hm, psi, ldf, deltat, order, imag_time, psi2, &
        deltat2, inh, weights, density, stream) bind(C, name="octhip_exponential_apply_batch")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: hm, psi, psi2, inh, weights, density, stream
      integer(c_int), value :: ldf, order, imag_time
      real(c_double), value :: deltat, deltat2
    end function
    ! density_calc_accumulate on a device-packed batch (weights and density on the device)
    integer(c_int) function octhip_density_accumulate(psi, ldf, np, is_complex, weights, density, stream) &
        bind(C, name="octhip_density_accumulate")
      import :: c_int, c_ptr
      type(c_ptr), value :: psi, weights, density, stream
      integer(c_int), value :: ldf, np, is_complex
    end function
    ! exponential_midpoint / td_etrs for independent particles over the batches psi(1:nbatch)
    integer(c_int) function octhip_propagator_expmid(hm, psi, nbatch, ldf, dt, order, field_mid, weights, &
        density, stream) bind(C, name="octhip_propagator_expmid")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: hm, field_mid, weights, density, stream
      type(c_ptr), intent(in) :: psi(*)
      integer(c_int), value :: nbatch, ldf, order
      real(c_double), value :: dt
    end function
    integer(c_int) function octhip_propagator_etrs(hm, psi, nbatch, ldf, dt, order, field_old, field_new, &
        weights, density, stream) bind(C, name="octhip_propagator_etrs")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: hm, field_old, field_new, weights, density, stream
      type(c_ptr), intent(in) :: psi(*)
      integer(c_int), value :: nbatch, ldf, order
      real(c_double), value :: dt
    end function
    ! batch_write_unpacked_to_device / batch_read_device_to_unpacked (batch.F90:625-760)
    integer(c_int) function octhip_batch_pack(unpacked, ld, nst, np, is_complex, packed, ldf, first, stream) &
        bind(C, name="octhip_batch_pack")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: unpacked, packed, stream
      integer(c_size_t), value :: ld
      integer(c_int), value :: nst, np, is_complex, ldf, first
    end function
    integer(c_int) function octhip_batch_unpack(packed, ldf, first, nst, np, is_complex, unpacked, ld, stream) &
        bind(C, name="octhip_batch_unpack")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: packed, unpacked, stream
      integer(c_size_t), value :: ld
      integer(c_int), value :: ldf, first, nst, np, is_complex
    end function
    integer(c_int) function octhip_taylor_graph_create(hm, psi, nbatch, ldf, deltat, order, imag_time, &
        stream, graph) bind(C, name="octhip_taylor_graph_create")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: hm, stream
      type(c_ptr), intent(in) :: psi(*)
      integer(c_int), value :: nbatch, ldf, order, imag_time
      real(c_double), value :: deltat
      type(c_ptr), intent(out) :: graph
    end function
    integer(c_int) function octhip_taylor_graph_launch(graph, stream) bind(C, name="octhip_taylor_graph_launch")
      import :: c_int, c_ptr
      type(c_ptr), value :: graph, stream
    end function
    integer(c_int) function octhip_taylor_graph_destroy(graph) bind(C, name="octhip_taylor_graph_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: graph
    end function

    integer(c_int) function octhip_adjoint_create(hm, positions, pol, ldf, nsteps, deltat, order, &
        checkpoint_every, vol_pp, adj) bind(C, name="octhip_adjoint_create")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: hm
      real(c_double), intent(in) :: positions(*), pol(3)
      integer(c_int), value :: ldf, nsteps, order, checkpoint_every
      real(c_double), value :: deltat, vol_pp
      type(c_ptr), intent(out) :: adj
    end function
    integer(c_int) function octhip_adjoint_forward(adj, psi, eps, stream) bind(C, name="octhip_adjoint_forward")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: adj, psi, stream
      real(c_double), intent(in) :: eps(*)
    end function
    integer(c_int) function octhip_adjoint_backward(adj, chi, dj_deps, dj_dv, stream) &
        bind(C, name="octhip_adjoint_backward")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: adj, chi, dj_dv, stream
      real(c_double), intent(out) :: dj_deps(*)
    end function
    integer(c_int) function octhip_adjoint_destroy(adj) bind(C, name="octhip_adjoint_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: adj
    end function

    integer(c_int) function octhip_memcpy_dtod(dst, src, bytes, stream) bind(C, name="octhip_memcpy_dtod")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: dst, src, stream
      integer(c_size_t), value :: bytes
    end function

    ! launch-shape knobs, the autotuners and their persisted table
    integer(c_int) function octhip_nl_operator_set_tuning(op, key, val) bind(C, name="octhip_nl_operator_set_tuning")
      import :: c_int, c_ptr
      type(c_ptr), value :: op
      integer(c_int), value :: key, val
    end function
    integer(c_int) function octhip_nl_operator_get_tuning(op, key, val) bind(C, name="octhip_nl_operator_get_tuning")
      import :: c_int, c_ptr
      type(c_ptr), value :: op
      integer(c_int), value :: key
      integer(c_int), intent(out) :: val
    end function
    integer(c_int) function octhip_nl_operator_uses_lattice(op, flag) bind(C, name="octhip_nl_operator_uses_lattice")
      import :: c_int, c_ptr
      type(c_ptr), value :: op
      integer(c_int), intent(out) :: flag
    end function
    integer(c_int) function octhip_hamiltonian_set_tuning(hm, key, val) bind(C, name="octhip_hamiltonian_set_tuning")
      import :: c_int, c_ptr
      type(c_ptr), value :: hm
      integer(c_int), value :: key, val
    end function
    integer(c_int) function octhip_hamiltonian_autotune(hm, ldf, reps, force, stream, best_us) &
        bind(C, name="octhip_hamiltonian_autotune")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: hm, stream
      integer(c_int), value :: ldf, reps, force
      real(c_double), intent(out) :: best_us
    end function
    ! StatesBlockSize by time (src/states/states_elec.F90:447-475)
    integer(c_int) function octhip_hamiltonian_autotune_pack(hm, nst, reps, force, stream, block_size, step_us) &
        bind(C, name="octhip_hamiltonian_autotune_pack")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: hm, stream
      integer(c_int), value :: nst, reps, force
      integer(c_int), intent(out) :: block_size
      real(c_double), intent(out) :: step_us
    end function
    ! path: a NUL-terminated string (trim(name)//c_null_char)
    integer(c_int) function octhip_tuning_save(path) bind(C, name="octhip_tuning_save")
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: path(*)
    end function
    integer(c_int) function octhip_tuning_load(path) bind(C, name="octhip_tuning_load")
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: path(*)
    end function
    ! device bytes held by a handle: data, stream work sets, graph sets, retired buffers
    integer(c_int) function octhip_hamiltonian_memory(hm, bytes) bind(C, name="octhip_hamiltonian_memory")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: hm
      integer(c_size_t), intent(out) :: bytes(4)
    end function
    integer(c_int) function octhip_hamiltonian_trim(hm) bind(C, name="octhip_hamiltonian_trim")
      import :: c_int, c_ptr
      type(c_ptr), value :: hm
    end function

    ! periodic boundaries and the Bloch phases of one k-point (src/grid/boundaries_inc.F90:
    ! 366-743, src/hamiltonian/hamiltonian_elec_base.F90:979-1139); host arrays or c_null_ptr
    integer(c_int) function octhip_hamiltonian_set_periodic(hm, bnd_src) bind(C, name="octhip_hamiltonian_set_periodic")
      import :: c_int, c_ptr
      type(c_ptr), value :: hm, bnd_src
    end function
    integer(c_int) function octhip_hamiltonian_set_phases(hm, phase, phase_corr, projector_phases) &
        bind(C, name="octhip_hamiltonian_set_phases")
      import :: c_int, c_ptr
      type(c_ptr), value :: hm, phase, phase_corr, projector_phases
    end function
    integer(c_int) function octhip_hamiltonian_apply_batch_phase(hm, psi, hpsi, ldf, terms, set_bc, has_phase, &
        stream) bind(C, name="octhip_hamiltonian_apply_batch_phase")
      import :: c_int, c_ptr
      type(c_ptr), value :: hm, psi, hpsi, stream
      integer(c_int), value :: ldf, terms, set_bc, has_phase
    end function
    integer(c_int) function octhip_hamiltonian_base_phase(hm, psi, src, ldf, np, conjugate, stream) &
        bind(C, name="octhip_hamiltonian_base_phase")
      import :: c_int, c_ptr
      type(c_ptr), value :: hm, psi, src, stream
      integer(c_int), value :: ldf, np, conjugate
    end function

    ! exponential_lanczos_batch / exponential_cheby_batch (src/electrons/exponential.F90:766-971)
    integer(c_int) function octhip_exponential_lanczos_batch(hm, psi, ldf, nst, deltat, imag_time, order, tol, &
        inh, vol, iterations, residual, stream) bind(C, name="octhip_exponential_lanczos_batch")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: hm, psi, inh, stream
      integer(c_int), value :: ldf, nst, imag_time, order
      real(c_double), value :: deltat, tol, vol
      integer(c_int), intent(out) :: iterations
      real(c_double), intent(out) :: residual
    end function
    integer(c_int) function octhip_exponential_cheby_batch(hm, psi, ldf, deltat, order, mid, half_span, stream) &
        bind(C, name="octhip_exponential_cheby_batch")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: hm, psi, stream
      integer(c_int), value :: ldf, order
      real(c_double), value :: deltat, mid, half_span
    end function

    ! mesh_batch_dotp_vector, mesh_batch_nrm2, mesh_batch_orthogonalization
    ! (src/grid/mesh_batch_inc.F90); dot/nrm2/overlap/norm are host arrays
    integer(c_int) function octhip_mesh_batch_dotp_vector(a, b, ldf, np, is_complex, vol, dot, stream) &
        bind(C, name="octhip_mesh_batch_dotp_vector")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: a, b, stream
      integer(c_int), value :: ldf, np, is_complex
      real(c_double), value :: vol
      real(c_double), intent(out) :: dot(*)
    end function
    ! the same without allocating; dot is a device array of 2 doubles per column
    integer(c_int) function octhip_mesh_batch_dotp_vector_noalloc(a, b, ldf, np, is_complex, vol, dot, stream) &
        bind(C, name="octhip_mesh_batch_dotp_vector_noalloc")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: a, b, dot, stream
      integer(c_int), value :: ldf, np, is_complex
      real(c_double), value :: vol
    end function
    integer(c_int) function octhip_mesh_batch_nrm2(a, ldf, np, is_complex, vol, nrm2, stream) &
        bind(C, name="octhip_mesh_batch_nrm2")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: a, stream
      integer(c_int), value :: ldf, np, is_complex
      real(c_double), value :: vol
      real(c_double), intent(out) :: nrm2(*)
    end function
    integer(c_int) function octhip_mesh_batch_orthogonalization(nvec, psi, phi, ldf, nst, np, vol, normalize, &
        overlap, norm, stream) bind(C, name="octhip_mesh_batch_orthogonalization")
      import :: c_int, c_double, c_ptr
      type(c_ptr), intent(in) :: psi(*)
      type(c_ptr), value :: phi, overlap, norm, stream
      integer(c_int), value :: nvec, ldf, nst, np, normalize
      real(c_double), value :: vol
    end function

    ! several (Hamiltonian, batch) pairs in one launch per kernel kind; psi, hpsi: c_loc of
    ! arrays of device pointers (hpsi may be c_null_ptr)
    integer(c_int) function octhip_batched_create(npairs, hms, psi, hpsi, ldf, deltat, order, batched) &
        bind(C, name="octhip_batched_create")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: npairs, order
      type(c_ptr), intent(in) :: hms(*)
      type(c_ptr), value :: psi, hpsi
      integer(c_int), intent(in) :: ldf(*)
      real(c_double), value :: deltat
      type(c_ptr), intent(out) :: batched
    end function
    integer(c_int) function octhip_batched_apply(batched, stream) bind(C, name="octhip_batched_apply")
      import :: c_int, c_ptr
      type(c_ptr), value :: batched, stream
    end function
    integer(c_int) function octhip_batched_taylor_step(batched, stream) bind(C, name="octhip_batched_taylor_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: batched, stream
    end function
    integer(c_int) function octhip_batched_graph_create(batched, stream, graph) &
        bind(C, name="octhip_batched_graph_create")
      import :: c_int, c_ptr
      type(c_ptr), value :: batched, stream
      type(c_ptr), intent(out) :: graph
    end function
    integer(c_int) function octhip_batched_destroy(batched) bind(C, name="octhip_batched_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: batched
    end function
  end interface

contains

  ! The last error message as a Fortran string (for messages_fatal).
  function octhip_last_error_message() result(msg)
    character(len=1024) :: msg
    character(kind=c_char) :: buf(1024)
    integer :: i, ierr
    buf = c_null_char
    ierr = octhip_get_last_error(buf, int(size(buf), c_size_t))
    msg = ''
    do i = 1, size(buf)
      if (buf(i) == c_null_char) exit
      msg(i:i) = buf(i)
    end do
  end function

end module octhip_m
