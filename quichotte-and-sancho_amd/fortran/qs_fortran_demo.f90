! qs_fortran_demo.f90 -- a Fortran voltage loop on the C ABI: helix of 4 sites,
! order 2, prints V, J_up, J_down, converged per line (list-directed, as
! Volt_Current_2.dat).  Exit status 1 if the device engine cannot be created.
program qs_fortran_demo
   use iso_c_binding
   use qs_hip_binding
   implicit none
   integer, parameter :: nsite = 4, norb = 2*nsite, nw = 256
   complex(c_double_complex), target :: H(norb, norb), GL(norb, norb), GR(norb, norb)
   real(c_double), target :: U(norb)
   type(qs_problem) :: p
   type(qs_scf_result) :: res
   type(c_ptr) :: h_
   real(c_double) :: lf(200), lr(200), ju, jd, V
   integer :: i, rc, k
   logical :: fresh

   H = (0.d0, 0.d0); GL = (0.d0, 0.d0); GR = (0.d0, 0.d0); U = 0.d0
   do i = 1, nsite
      H(2*i - 1, 2*i - 1) = 2.d0; H(2*i, 2*i) = 2.d0
      if (i < nsite) then
         H(2*i - 1, 2*i + 1) = -0.4d0; H(2*i + 1, 2*i - 1) = -0.4d0
         H(2*i, 2*i + 2) = -0.4d0; H(2*i + 2, 2*i) = -0.4d0
      end if
      U(2*i - 1) = 0.8d0
   end do
   GL(1, 1) = 1.5d0; GL(2, 2) = 1.d0; GR(norb - 1, norb - 1) = 1.d0; GR(norb, norb) = 1.d0

   p%norb = norb; p%n_omega = nw; p%w_init = 0.d0; p%delta = 4.d0/nw
   p%H = c_loc(H); p%gammaL = c_loc(GL); p%gammaR = c_loc(GR); p%U_site = c_loc(U)
   p%order = 2; p%beta = 1.d0/(8.6173303d-5*300.d0); p%mu = 2.d0
   p%n_atoms = 0; p%atom_norb = c_null_ptr
   p%alpha_mix = 0.5d0; p%nhist = 8; p%maxscf = 200; p%tolf = 1.d-10; p%tolf_order1 = 1.d-8
   p%single_history = 1; p%device = 0

   rc = qs_create(p, h_)
   if (rc /= 0) then
      write (*, *) 'qs_create failed, status', rc
      stop 1
   end if
   fresh = .true.
   do k = 0, 2
      V = 0.25d0*k
      rc = qs_solve_scf(h_, V, merge(1, 0, fresh), res, lf, lr)
      if (rc /= 0) stop 1
      rc = qs_current(h_, ju, jd)
      write (*, *) V, ju, jd, res%converged, res%n_iter
      fresh = res%converged == 0
   end do
   call qs_destroy(h_)
end program qs_fortran_demo
