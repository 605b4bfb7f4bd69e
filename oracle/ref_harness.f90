!  Test-infrastructure driver for the reference Geoflow-SPH modules.
!
!  It replaces PROGRAM SPH_FEM_Driver_2019 ("8 SPH_FEM_Driver_2019.f90":6)
!  for one plain SPH run (ic_cases_mat=0): it sets the driver-level variables
!  that Driver_Input (8:165-470) would read from <name>.master.dat, calls
!  the reference's own Init_sph (7:74) -- which reads <name>.dat, the .top
!  named on stdin and restart.dat exactly as the reference does -- and then
!  repeats the Driver_Flow step (8:1580-1620): adaptive_dt_sph,
!  itimestep_sph += 1, time_integration_sph, time_sph += dt_sph.
!
!  Control file harness.ctl (cwd):
!     line 1: problem name (the .dat stem)
!     line 2: nsteps  dt_sph  ic_adapt_dt_sph
!     line 3: ndump   followed by ndump step numbers (0 = after Init_sph)
!     line 4 (optional): nprob  plot_every  -- Monte Carlo pixel mode: the
!             problems named on line 5 (one .dat stem each, sharing the
!             release and terrain) run one after another, each from its own
!             Init_sph, with Get_MC_pixel_Vars_SW ("6 SPH_SW_2019.f90":
!             20717-20833) every plot_every steps and Norm_MC_pixel_Vars_SW(1)
!             (6:20838-20908) after its last step, as Driver_Flow_23 calls
!             them through Manage_pixel_Hazard_SPH (11) and (12)
!             ("8 SPH_FEM_Driver_2019.f90":1504-1541); then mc_pixel.bin:
!                int32  magic npoig nstat
!                per pixel with MC_Pr allocated: int32 ipoig nr_problem_pix,
!                real64 MC_Pr(1:4, 1:4)
!             Dumps of line 3 are written for the last problem only.
!
!  After Init_sph it writes topo.bin (the terrain, see write_topo) and at
!  each listed step dump_<step>.bin (stream, native endian):
!     int32  magic(=1195725648) version(=1) npoin ndimn has_uaux npairs
!     real64 time_sph
!     real64 x(ndimn,npoin) vx(ndimn,npoin) rho mass hsml u  [uaux1]
!     int32  If_Out_Domain(npoin)
!     real64 x0(ndimn,npoin)        positions the pair list was built from
!     int32  pair_i(npairs) pair_j(npairs)
!     real64 w(npairs) dwdx(ndimn,npairs)
!  with a pore-pressure column (nUaux > 1), uaux_<step>.bin:
!     real64 (Uaux(1:nUaux,ip), ip = 1, npoin)
!  and, when erosion is on (ic_erosion = 1), eros_<step>.bin:
!     real64 (topol(1,ip), topol(14,ip), topol(15,ip), ip = 1, npoig)
!  If gid_final.ctl exists in the run directory, the reference's GiD result
!  writer (OutputRes_SW) also runs once after the last step.
!  If zones.ctl exists (line 1: npts, then npts lines x y), the harness calls
!  the reference's Topo_main itself before Init_sph (whose own call then
!  returns at once) and writes zones.bin -- the terrain zones as Topo_main
!  leaves them, before Input_SW and restart.top change topol:
!     int32  magic npoig npts
!     real64 (topol(1,ip), topol(15,ip), ip = 1, npoig)
!     int32  (Which_Topo_Zone(x,y,1), Which_Topo_Zone(x,y,2), point = 1, npts)
!  (-999: in no zone of that type).
!  It is built only by oracle/build_ref.sh and never ships with the product.
!  Built with -DGFSPH_GPU it calls time_integration_sph_gfsph (module
!  SPH_MAIN_GFSPH of the product's Fortran layer) instead, i.e. the
!  reference's own inputs and Init_sph driving the GPU step, which also
!  performs adaptive_dt_sph on the device.

program gfsph_ref_harness

use SPH_FEM_variable_types_2019
use SPH_FEM_Driver_time_vars_2019
use SPH_MAIN_2019, only: Init_SPH, time_integration_sph, adaptive_dt_sph
use SPH_Main_Vars_2019
use SPH_SW_Vars_2019
use SPH_SW_2019, only: OutputRes_SW, Get_MC_pixel_Vars_SW, Norm_MC_pixel_Vars_SW
use SPH_FEM_topo_2019, only: npoigx, npoigy, Xming, Xmaxg, Yming, Ymaxg, deltxg, deltyg, &
                             xshiftg, yshiftg, topol, poig_in_model, Topo_main, Which_Topo_Zone
#ifdef GFSPH_GPU
use SPH_MAIN_GFSPH, only: time_integration_sph_gfsph, gfsph_release, gfsph_sync_state, gfsph_pulls, &
                          gfsph_pull_every_step
#endif

implicit none

character(16)  name
integer        nsteps, ndump, istep, k, ictl, iadapt
logical        gid_final
integer, allocatable :: dump_steps(:)
real(8)        dtin
integer        nprob, plot_every, iprob, ios
character(16), allocatable :: pnames(:)
integer(8)     clk0, clk1, clkrate, cd0, cd1, clkdump
character(8)   envv
logical        zones_ctl

open (newunit=ictl, file='harness.ctl', status='old')
read (ictl,'(a16)') name
read (ictl,*) nsteps, dtin, iadapt
read (ictl,*) ndump
allocate (dump_steps(max(ndump,1)))
if (ndump.gt.0) then
   backspace (ictl)
   read (ictl,*) ndump, (dump_steps(k), k=1,ndump)
endif
nprob = 0; plot_every = 0
read (ictl,*,iostat=ios) nprob, plot_every
if (ios.ne.0) nprob = 0
allocate (pnames(max(nprob,1)))
if (nprob.gt.0) read (ictl,*) (pnames(k), k=1,nprob)
close (ictl)

!      ------  what Driver_Input would have read (8:229-470)

problem_name_Global = name
SPH_SW_problem_name = name
if_Read_TopoDat  = 0
if_sph = 1; if_gfl = 0; if_tgf = 0
ic_cases_mat     = 0
SPH_ProblemType  = 1
SPH_t_Integ_Alg  = 4
dt               = dtin
initial_dt       = dtin
time_end         = 1.e30
maxtimestep      = nsteps
print_step = 1000000000; save_step = 1000000000; plot_step = 1000000000
nsearch_step     = 1
isearch_step     = 0
dt_sph           = dtin
initial_dt_sph   = dtin
ic_adapt_dt_sph  = iadapt
dt_gfl           = -999.
ntcurves = 0; mptstcurves = 0
time_print = print_step*dt; time_save = save_step*dt; time_plot = plot_step*dt
t_print_reset = 0.; t_save_reset = 0.; t_plot_reset = 0.
dt_limit_inf     = 1.e-7*dt
ic_Trigger = 0; ic_Prob_Propag = 0; ic_pixel_Hazard = 0
ic_MCarlo = 0; ic_FOSM = 0; N_Gzones = 1; npixel_HVars = 10
ic_cases_win     = 0
nproblems        = 1
nproblems_index  = 1
time_sph_restart = 0.0

do iprob = 1, max(nprob,1)
if (nprob.gt.0) then
   problem_name_Global = pnames(iprob)
   SPH_SW_problem_name = pnames(iprob)
   nproblems_index     = iprob
   dt = dtin; dt_sph = dtin
   t_print_reset = 0.; t_save_reset = 0.; t_plot_reset = 0.
endif
inquire (file='zones.ctl', exist=zones_ctl)
if (iprob.eq.1.and.zones_ctl) call write_zones
call Init_sph

!      ------  Driver_Flow (8:1565-1620) for one SPH substep per step

time = 0.0; time_sph = 0.0; itimestep_sph = 0
if (time_sph_restart.gt.0.0) then
   time     = time_sph_restart
   time_sph = time_sph_restart
endif

if (iprob.eq.1) then
   call maybe_dump (0)
   call write_topo
endif
#ifdef GFSPH_GPU
call get_environment_variable('GFSPH_PULL_EVERY_STEP', envv)
gfsph_pull_every_step = envv(1:1).eq.'1'
#endif
call system_clock (clk0, clkrate)
clkdump = 0
do istep = 1, nsteps
   t_print_reset = t_print_reset + dt
#ifndef GFSPH_GPU
   call adaptive_dt_sph
#endif
   itimestep_sph = itimestep_sph + 1
#ifdef GFSPH_GPU
   call time_integration_sph_gfsph      ! adapts dt_sph on the device
#else
   call time_integration_sph
#endif
   time_sph = time_sph + dt_sph
   time     = time + dt
   call system_clock (cd0)
   if (nprob.eq.0.or.iprob.eq.nprob) call maybe_dump (istep)
   call system_clock (cd1)
   clkdump = clkdump + (cd1 - cd0)
   if (nprob.gt.0) then
      if (mod(istep, plot_every).eq.0) call Get_MC_pixel_Vars_SW
   endif
enddo
if (nprob.gt.0) call Norm_MC_pixel_Vars_SW (1)
#ifdef GFSPH_GPU
call gfsph_sync_state
#endif
call system_clock (clk1)
write (*,'(a,i0,a,es12.5)') ' harness: ', nsteps, ' steps (dumps excluded), seconds ', &
      real(clk1-clk0-clkdump,8)/real(clkrate,8)
enddo
if (nprob.gt.0) call write_mc_pixel
!      ------  with gid_final.ctl present: the GiD result blocks of the final
!              state (OutputRes_SW, appended to <name>.post.res)
inquire (file='gid_final.ctl', exist=gid_final)
if (gid_final) call OutputRes_SW
#ifdef GFSPH_GPU
write (*,'(a,i0)') ' gfsph: state copies back to the host: ', gfsph_pulls
call gfsph_release
#endif

contains

subroutine write_mc_pixel
   integer iu, ip, nst, np9
   np9 = npoigx*npoigy
   nst = 0
   do ip = 1, np9
      if (MC_pixel(ip)%if_MC_Pr_alloc.eq.1) nst = nst + 1
   enddo
   open (newunit=iu, file='mc_pixel.bin', access='stream', form='unformatted', status='replace')
   write (iu) int(1195725648,4), int(np9,4), int(nst,4)
   do ip = 1, np9
      if (MC_pixel(ip)%if_MC_Pr_alloc.eq.1) then
         write (iu) int(ip,4), int(MC_pixel(ip)%nr_problem_pix,4)
         write (iu) real(MC_pixel(ip)%MC_Pr(1:4,1:4),8)
      endif
   enddo
   close (iu)
end subroutine write_mc_pixel

subroutine write_zones
   integer iu, ip, np9, npts, izc, z1, z2
   real(8), allocatable :: px(:), py(:)
   real(8) xx, yy
   call Topo_main
   open (newunit=izc, file='zones.ctl', status='old')
   read (izc,*) npts
   allocate (px(max(npts,1)), py(max(npts,1)))
   do ip = 1, npts
      read (izc,*) px(ip), py(ip)
   enddo
   close (izc)
   np9 = npoigx*npoigy
   open (newunit=iu, file='zones.bin', access='stream', form='unformatted', status='replace')
   write (iu) int(1195725648,4), int(np9,4), int(npts,4)
   write (iu) (real(topol(1,ip),8), real(topol(15,ip),8), ip=1,np9)
   do ip = 1, npts
      xx = px(ip); yy = py(ip)
      call Which_Topo_Zone (xx, yy, 1, z1)
      call Which_Topo_Zone (xx, yy, 2, z2)
      write (iu) int(z1,4), int(z2,4)
   enddo
   close (iu)
   deallocate (px, py)
end subroutine write_zones

! topo.bin: the terrain Init_sph built (Topo_main's module state)
!    int32  magic(=1195725648) version(=2) npoigx npoigy
!    real64 Xming Xmaxg Yming Ymaxg deltxg deltyg xshiftg yshiftg
!    real64 topol(1:9, npoig)   int32 poig_in_model(npoig)
subroutine write_topo
   integer iu, ip, np9
   np9 = npoigx*npoigy
   open (newunit=iu, file='topo.bin', access='stream', form='unformatted', status='replace')
   write (iu) int(1195725648,4), int(2,4), int(npoigx,4), int(npoigy,4)
   write (iu) real(Xming,8), real(Xmaxg,8), real(Yming,8), real(Ymaxg,8), real(deltxg,8), real(deltyg,8), &
              real(xshiftg,8), real(yshiftg,8)
   do ip = 1, np9, 8192   ! bounded records: large single records are slow to build
      write (iu) real(topol(1:9,ip:min(ip+8191,np9)),8)
   enddo
   do ip = 1, np9, 8192
      write (iu) int(poig_in_model(ip:min(ip+8191,np9)),4)
   enddo
   close (iu)
end subroutine write_topo

subroutine maybe_dump (is)
   integer, intent(in) :: is
   integer k2
   do k2 = 1, ndump
      if (dump_steps(k2).eq.is) then
#ifdef GFSPH_GPU
         call gfsph_sync_state
#endif
         call write_dump (is)
      endif
   enddo
end subroutine maybe_dump

subroutine write_dump (is)
   integer, intent(in) :: is
   integer iu, np, has_uaux, ipair, ip
   character(64) fname
   type(pairs), pointer :: node
   integer(4), allocatable :: pi_(:), pj_(:)
   real(8),    allocatable :: pw(:), pdw(:,:)

   np = 0
   node => last
   do while (associated(node))
      np = np + 1
      node => node%next
   enddo
   allocate (pi_(max(np,1)), pj_(max(np,1)), pw(max(np,1)), pdw(ndimn,max(np,1)))
   ipair = 0
   node => last
   do while (associated(node))
      ipair = ipair + 1
      pi_(ipair) = node%pair_i
      pj_(ipair) = node%pair_j
      pw(ipair)  = node%w
      pdw(:,ipair) = node%dwdx(:)
      node => node%next
   enddo

   has_uaux = 0
   if (allocated(Uaux)) has_uaux = 1
   write (fname,'(a,i0,a)') 'dump_', is, '.bin'
   open (newunit=iu, file=trim(fname), access='stream', form='unformatted', status='replace')
   write (iu) int(1195725648,4), int(1,4), int(npoin,4), int(ndimn,4), int(has_uaux,4), int(np,4)
   write (iu) real(time_sph,8)
   call wr_d (iu, x, ndimn*npoin)
   call wr_d (iu, vx, ndimn*npoin)
   call wr_d (iu, rho, npoin)
   call wr_d (iu, mass, npoin)
   call wr_d (iu, hsml, npoin)
   call wr_d (iu, u, npoin)
   if (has_uaux.eq.1) write (iu) (real(Uaux(1,ip),8), ip=1,npoin)
   call wr_i (iu, If_Out_Domain, npoin)
   call wr_d (iu, x0, ndimn*npoin)
   if (np.gt.0) then
      call wr_i (iu, pi_, np)
      call wr_i (iu, pj_, np)
      call wr_d (iu, pw, np)
      call wr_d (iu, pdw, ndimn*np)
   endif
   close (iu)
   deallocate (pi_, pj_, pw, pdw)
   if (allocated(Uaux)) then
      if (size(Uaux, 1).gt.1) then
         write (fname,'(a,i0,a)') 'uaux_', is, '.bin'
         open (newunit=iu, file=trim(fname), access='stream', form='unformatted', status='replace')
         write (iu) (real(Uaux(:,ip),8), ip=1,npoin)
         close (iu)
      endif
   endif
   if (ic_erosion.eq.1) then
      write (fname,'(a,i0,a)') 'eros_', is, '.bin'
      open (newunit=iu, file=trim(fname), access='stream', form='unformatted', status='replace')
      write (iu) (real(topol(1,ip),8), real(topol(14,ip),8), real(topol(15,ip),8), ip=1,npoigx*npoigy)
      close (iu)
   endif
end subroutine write_dump

! wr_d / wr_i: n values of an array (any shape, contiguous) as consecutive
! stream records of at most 8192 values (large single records are slow to
! build in the runtime); the bytes are those of one record of all n
subroutine wr_d (iu, a, n)
   integer, intent(in) :: iu, n
   real(8), intent(in) :: a(*)
   integer k
   do k = 1, n, 8192
      write (iu) a(k:min(k+8191,n))
   enddo
end subroutine wr_d

subroutine wr_i (iu, a, n)
   integer, intent(in) :: iu, n
   integer(4), intent(in) :: a(*)
   integer k
   do k = 1, n, 8192
      write (iu) a(k:min(k+8191,n))
   enddo
end subroutine wr_i

end program gfsph_ref_harness
