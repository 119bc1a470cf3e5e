! Fortran host driver of the device path through the iso_c_binding module: what an
! Octopus build would do around its own batches, on arrays read from a file.
!
! usage: octhip_driver <input> <output>
! input (stream, native endian): int32 n, nri, np, np_part, ldf, nmat, matrix_size,
!   total_points, nproj; float64 w(n); int32 points(3n); int32 ri(n*nri);
!   int32 rimap_inv(nri+1); float64 v(np); int32 offsets(6*nmat); float64 matrices;
!   int32 maps; float64 scals; float64 psi(L*np_part); float64 deltat; int32 order.
! output: float64 lap(L*np) (factor -1/2), hpsi(L*np), psi after one Taylor step
!   (L*np_part), that psi unpacked on the device (L/2 complex states of np points, state
!   after state: batch_read_device_to_unpacked) and packed again into a fresh batch (L*np),
!   the mesh inner products <psi_c|psi_c> of that psi with volume element 1 (L doubles: L/2
!   (re, im) pairs), and the handle's device bytes (4 values: data, work, graphs, retired).
!   Exit status 0 on success; on a library error the message of octhip_get_last_error is
!   printed and the status is 2.
program octhip_driver
  use, intrinsic :: iso_c_binding
  use octhip_m
  implicit none

  integer(c_int) :: hdr(9), n, nri, np, np_part, ldf, nmat, msize, tpts, nproj, L, order
  real(c_double), allocatable, target :: w(:), v(:), matrices(:), scals(:), psi(:), lap(:), hpsi(:), &
    unp(:), repacked(:)
  integer(c_int), allocatable, target :: points(:), ri(:), rimap_inv(:), offsets(:), maps(:)
  real(c_double) :: deltat
  real(c_double), allocatable :: dots(:)
  integer(c_size_t) :: mem(4)
  type(c_ptr) :: op, hm, d_psi, d_out, d_unp, s
  character(len=4096) :: fin, fout
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  s = c_null_ptr

  call check(octhip_init(0_c_int), 'octhip_init')

  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read(u) hdr
  n = hdr(1); nri = hdr(2); np = hdr(3); np_part = hdr(4); ldf = hdr(5)
  nmat = hdr(6); msize = hdr(7); tpts = hdr(8); nproj = hdr(9)
  L = ishft(1, ldf)
  allocate(w(n), points(3 * n), ri(n * nri), rimap_inv(nri + 1), v(np), offsets(max(1, 6 * nmat)), &
    matrices(max(1, msize)), maps(max(1, tpts)), scals(max(1, nproj)), psi(L * np_part), lap(L * np), &
    hpsi(L * np), unp(L * np), repacked(L * np))
  read(u) w, points, ri, rimap_inv, v
  if (nmat > 0) read(u) offsets(1:6 * nmat), matrices(1:msize), maps(1:tpts), scals(1:nproj)
  read(u) psi, deltat, order
  close(u)

  call check(octhip_malloc(d_psi, int(8, c_size_t) * L * np_part), 'malloc psi')
  call check(octhip_malloc(d_out, int(8, c_size_t) * L * np), 'malloc out')
  call check(octhip_memcpy_htod(d_psi, c_loc(psi), int(8, c_size_t) * L * np_part, s), 'htod')

  ! nl_operator_build -> device upload, then the kinetic operator on the batch
  call check(octhip_nl_operator_create(n, w, c_loc(points), nri, ri, rimap_inv, np, np_part, c_null_ptr, op), &
    'nl_operator_create')
  call check(octhip_nl_operator_operate_batch(op, d_psi, d_out, ldf, -0.5_c_double, OCTHIP_OP_ALL, s), &
    'operate_batch')
  call check(octhip_memcpy_dtoh(c_loc(lap), d_out, int(8, c_size_t) * L * np, s), 'dtoh lap')

  ! hamiltonian_elec_apply_batch with every term and the zero boundary
  if (nmat > 0) then
    call check(octhip_hamiltonian_create(op, 1.0_c_double, c_loc(v), nmat, c_loc(offsets), c_loc(matrices), &
      c_loc(maps), c_loc(scals), hm), 'hamiltonian_create')
  else
    call check(octhip_hamiltonian_create(op, 1.0_c_double, c_loc(v), 0_c_int, c_null_ptr, c_null_ptr, &
      c_null_ptr, c_null_ptr, hm), 'hamiltonian_create')
  end if
  call check(octhip_hamiltonian_apply_batch(hm, d_psi, d_out, ldf, 1_c_int, OCTHIP_TERM_ALL, 1_c_int, s), &
    'hamiltonian_apply_batch')
  call check(octhip_memcpy_dtoh(c_loc(hpsi), d_out, int(8, c_size_t) * L * np, s), 'dtoh hpsi')

  ! one Taylor step in place
  call check(octhip_exponential_taylor_batch(hm, d_psi, ldf, deltat, order, 0_c_int, s), 'taylor')
  call check(octhip_memcpy_dtoh(c_loc(psi), d_psi, int(8, c_size_t) * L * np_part, s), 'dtoh psi')

  ! do_unpack / do_pack on the device: the propagated batch to one column per state and back
  call check(octhip_malloc(d_unp, int(8, c_size_t) * L * np), 'malloc unpacked')
  call check(octhip_batch_unpack(d_psi, ldf, 0_c_int, L / 2, np, 1_c_int, d_unp, int(np, c_size_t), s), &
    'batch_unpack')
  call check(octhip_memcpy_dtoh(c_loc(unp), d_unp, int(8, c_size_t) * L * np, s), 'dtoh unpacked')
  call check(octhip_batch_pack(d_unp, int(np, c_size_t), L / 2, np, 1_c_int, d_out, ldf, 0_c_int, s), &
    'batch_pack')
  call check(octhip_memcpy_dtoh(c_loc(repacked), d_out, int(8, c_size_t) * L * np, s), 'dtoh repacked')
  ! mesh_batch_dotp_vector of the propagated batch with itself, and the handle's memory
  allocate(dots(L))
  call check(octhip_mesh_batch_dotp_vector(d_psi, d_psi, ldf, np, 1_c_int, 1.0_c_double, dots, s), 'dotp')
  call check(octhip_hamiltonian_memory(hm, mem), 'memory')
  call check(octhip_stream_synchronize(s), 'synchronize')

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  write(u) lap, hpsi, psi, unp, repacked, dots, real(mem, c_double)
  close(u)

  call check(octhip_hamiltonian_destroy(hm), 'hamiltonian_destroy')
  call check(octhip_nl_operator_destroy(op), 'nl_operator_destroy')
  call check(octhip_free(d_psi), 'free')
  call check(octhip_free(d_out), 'free')
  call check(octhip_free(d_unp), 'free')
  call check(octhip_finalize(), 'finalize')

contains

  subroutine check(ierr, what)
    integer(c_int), intent(in) :: ierr
    character(len=*), intent(in) :: what
    if (ierr /= OCTHIP_SUCCESS) then
      write(*, '(a, a, i0, a, a)') what, ': status ', ierr, ': ', trim(octhip_last_error_message())
      stop 2
    end if
  end subroutine

end program octhip_driver
