! qs_hip_binding.f90 -- iso_c_binding interface to libqs_hip.so (include/qs_hip.h)
! for Fortran callers of the SCF path (see the voltage loop of src/quichotte.f90:125-143).
module qs_hip_binding
   use iso_c_binding
   implicit none
   type, bind(C) :: qs_problem
      integer(c_int32_t) :: norb, n_omega
      real(c_double)     :: w_init, delta
      type(c_ptr)        :: H, gammaL, gammaR, U_site      ! complex(dp)(Norb,Norb), real(dp)(Norb)
      integer(c_int32_t) :: order
      real(c_double)     :: beta, mu
      integer(c_int32_t) :: n_atoms
      type(c_ptr)        :: atom_norb
      real(c_double)     :: alpha_mix
      integer(c_int32_t) :: nhist, maxscf
      real(c_double)     :: tolf, tolf_order1
      integer(c_int32_t) :: single_history, device
   end type
   type, bind(C) :: qs_scf_result
      integer(c_int32_t) :: converged, n_iter
      real(c_double)     :: fmax
      integer(c_int32_t) :: runaway, unphysical, i_worst
      real(c_double)     :: n_worst, w_worst
   end type
   interface
      integer(c_int) function qs_create(p, h) bind(C, name='qs_create')
         import; type(qs_problem), intent(in) :: p; type(c_ptr), intent(out) :: h
      end function
      subroutine qs_destroy(h) bind(C, name='qs_destroy')
         import; type(c_ptr), value :: h
      end subroutine
      integer(c_int) function qs_solve_scf(h, V, fresh, res, log_fmax, log_rms) bind(C, name='qs_solve_scf')
         import; type(c_ptr), value :: h; real(c_double), value :: V; integer(c_int32_t), value :: fresh
         type(qs_scf_result), intent(out) :: res; real(c_double) :: log_fmax(*), log_rms(*)
      end function
      integer(c_int) function qs_current(h, J_up, J_down) bind(C, name='qs_current')
         import; type(c_ptr), value :: h; real(c_double), intent(out) :: J_up, J_down
      end function
      integer(c_int) function qs_g_less_obs(h, G) bind(C, name='qs_g_less_obs')
         import; type(c_ptr), value :: h; complex(c_double_complex) :: G(*)
      end function
      integer(c_int) function qs_dosd(h, d) bind(C, name='qs_dosd')
         import; type(c_ptr), value :: h; real(c_double) :: d(*)
      end function
      integer(c_int) function qs_residual2(h, V, n, x, F, abort) bind(C, name='qs_residual2')
         import; type(c_ptr), value :: h; real(c_double), value :: V; integer(c_int64_t), value :: n
         real(c_double), intent(in) :: x(*); real(c_double), intent(out) :: F(*)
         integer(c_int32_t), intent(out) :: abort
      end function
      ! SANCHO's order-2/3 self-energies: complete_input_gf + build_self_energies
      ! (SANCHO/src/green_functions.f90:163-173, SANCHO/src/self_energies.f90:38-46)
      integer(c_int) function qs_sancho_create(na, nw, order, device, h) bind(C, name='qs_sancho_create')
         import; integer(c_int32_t), value :: na, nw, order, device; type(c_ptr), intent(out) :: h
      end function
      subroutine qs_sancho_destroy(h) bind(C, name='qs_sancho_destroy')
         import; type(c_ptr), value :: h
      end subroutine
      integer(c_int) function qs_sancho_self_energies(h, delta, U, GR, GL, S2R, S2L, S2G, S3R, S3L, S3G, &
                                                      on_device) bind(C, name='qs_sancho_self_energies')
         import; type(c_ptr), value :: h; real(c_double), value :: delta; real(c_double), intent(in) :: U(*)
         complex(c_double_complex), intent(in) :: GR(*), GL(*)
         complex(c_double_complex) :: S2R(*), S2L(*), S2G(*), S3R(*), S3L(*), S3G(*)
         integer(c_int32_t), value :: on_device
      end function
   end interface
end module qs_hip_binding
