!  Step recorder for the reference's own main program (test infrastructure).
!
!  oracle/build_ref.sh links PROGRAM SPH_FEM_Driver_2019
!  ("8 SPH_FEM_Driver_2019.f90") unchanged except for the compile fix at
!  8:1993, with the linker option --wrap for the module procedure
!  stop_case_SPH ("7 SPH_MAIN_2019.f90", called by Driver_Flow_23 after every
!  step, 8:1527).  Calls from the main program then reach this routine, which
!  runs the reference's stop_case_SPH and appends one record to
!  case_trace.bin (stream, native endian) -- so the case loop of
!  Driver_Flow_23 (8:1444-1548: dt_sph cut to the next plot time and restored
!  after it, Stop_case_SW, time_end) runs exactly as the reference wrote it
!  and its effect is observed step by step:
!     int32  magic(=1195725648) nproblems_index itimestep ic_stop_cases npoin
!     real64 time_sph dt_sph time t_plot_reset
!     real64 x(2,npoin) vx(2,npoin) rho(npoin) u(npoin)
!     int32  If_Out_Domain(npoin)
!  It never ships with the product.

subroutine gfsph_probe_stop_case() bind(C, name='__wrap__QMsph_main_2019Pstop_case_sph')
   use SPH_FEM_variable_types_2019
   use SPH_FEM_Driver_time_vars_2019
   use SPH_Main_Vars_2019
   use SPH_SW_Vars_2019
   implicit none
   interface
      subroutine reference_stop_case() bind(C, name='__real__QMsph_main_2019Pstop_case_sph')
      end subroutine reference_stop_case
   end interface
   integer, save :: iu = 0
   logical, save :: opened = .false.
   integer ip

   call reference_stop_case ()
   if (.not.opened) then
      open (newunit=iu, file='case_trace.bin', access='stream', form='unformatted', status='replace')
      opened = .true.
   endif
   write (iu) int(1195725648,4), int(nproblems_index,4), int(itimestep,4), int(ic_stop_cases,4), &
              int(npoin,4)
   write (iu) real(time_sph,8), real(dt_sph,8), real(time,8), real(t_plot_reset,8)
   write (iu) (real(x(:,ip),8), ip=1,npoin)
   write (iu) (real(vx(:,ip),8), ip=1,npoin)
   write (iu) (real(rho(ip),8), ip=1,npoin)
   write (iu) (real(u(ip),8), ip=1,npoin)
   write (iu) (int(If_Out_Domain(ip),4), ip=1,npoin)
   flush (iu)
end subroutine gfsph_probe_stop_case
