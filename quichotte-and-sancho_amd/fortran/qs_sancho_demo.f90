! qs_sancho_demo.f90 -- SANCHO's self-energies from Fortran through qs_hip_binding.
! Reads na, nw, order, delta, U(na), G_in%r and G_in%L (na,na,nw) from the
! stream file given as the first argument, calls qs_sancho_self_energies as
! SANCHO's scf_residual would (SANCHO/src/scf.f90:114-118), and writes
! Sig2R, Sig2L, Sig2G (and Sig3R, Sig3L, Sig3G at order 3) to the second.
program qs_sancho_demo
   use iso_c_binding
   use qs_hip_binding
   implicit none
   integer, parameter :: dp = kind(1.d0)
   integer(c_int32_t) :: na, nw, order
   real(dp) :: delta
   real(dp), allocatable :: U(:)
   complex(dp), allocatable :: GR(:, :, :), GL(:, :, :), S(:, :, :, :)
   character(len=512) :: fin, fout
   type(c_ptr) :: h
   integer :: rc, nout

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open (10, file=trim(fin), access='stream', form='unformatted', status='old')
   read (10) na, nw, order
   read (10) delta
   allocate (U(na), GR(na, na, nw), GL(na, na, nw))
   read (10) U, GR, GL
   close (10)
   nout = merge(6, 3, order == 3)
   allocate (S(na, na, nw, 6))
   S = (0.d0, 0.d0)
   rc = qs_sancho_create(na, nw, order, 0, h)
   if (rc /= 0) then
      print *, 'qs_sancho_create failed', rc
      stop 1
   end if
   rc = qs_sancho_self_energies(h, delta, U, GR, GL, S(:, :, :, 1), S(:, :, :, 2), S(:, :, :, 3), &
                                S(:, :, :, 4), S(:, :, :, 5), S(:, :, :, 6), 0)
   call qs_sancho_destroy(h)
   if (rc /= 0) then
      print *, 'qs_sancho_self_energies failed', rc
      stop 1
   end if
   open (11, file=trim(fout), access='stream', form='unformatted', status='replace')
   write (11) S(:, :, :, 1:nout)
   close (11)
end program qs_sancho_demo
