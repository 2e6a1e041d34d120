!=======================================================================
! spicule_main -- the UTRIP spicule main program on the GPU drop-in library.
!
! The time loop is the reference's (cans1d/md_spicule/hdmlw/main.f:169-227):
!     ns = ns+1; cfl_m; stop on merr; t = t+dt; mlw_m_bg; trqalfven;
!     bnd (mirror at the foot point, vx odd; free at the top); chkdav(ro);
!     chkdav(pr); loop while ns < nstop and t < tend
! with mlw_m_bg, cfl_m, chkdav, trqalfven and bd* from libcans1d_gpu.so.
! The state and the geometry come from DAC files (as main.f's restart,
! main.f:87-111, and the static arrays main.f:115-151 write), and the final
! state is written as DAC files.  The torque's noise is the Park-Miller
! stream from SEED.
!
! usage: spicule_main INDIR OUTDIR NSTOP TEND GM SEED RESIDENT
!=======================================================================
program spicule_main
  use dac_io
  implicit none
  character(len=512) :: indir, outdir, arg
  integer :: nstop, seed, resident, ix, ns, merr, dims(4), nd, margin, nf
  double precision :: tend, gm, t, dt, safety, dtmin, qav, floor
  double precision, allocatable :: ro(:), pr(:), vx(:), vy(:), by(:), x(:)
  double precision, allocatable :: bx(:), bxm(:), gx(:), gxm(:), sc(:), dsc(:), scm(:), dscm(:)
  double precision, allocatable :: rr(:), rrm(:), drr(:), drrm(:), dx(:), dxm(:)

  call get_command_argument(1, indir)
  call get_command_argument(2, outdir)
  call get_command_argument(3, arg); read(arg, *) nstop
  call get_command_argument(4, arg); read(arg, *) tend
  call get_command_argument(5, arg); read(arg, *) gm
  call get_command_argument(6, arg); read(arg, *) seed
  call get_command_argument(7, arg); read(arg, *) resident
  margin = 1
  safety = 0.8d0
  dtmin = 1.d-10
  qav = 1.d0
  floor = 1.d-9

  call dac_dims(11, trim(indir)//'/ro.dac', dims, nd)
  ix = dims(1)
  allocate(ro(ix), pr(ix), vx(ix), vy(ix), by(ix), x(ix))
  allocate(bx(ix), bxm(ix), gx(ix), gxm(ix), sc(ix), dsc(ix), scm(ix), dscm(ix))
  allocate(rr(ix), rrm(ix), drr(ix), drrm(ix), dx(ix), dxm(ix))
  call dac_read_last(11, trim(indir)//'/ro.dac', ro, ix)
  call dac_read_last(11, trim(indir)//'/pr.dac', pr, ix)
  call dac_read_last(11, trim(indir)//'/vx.dac', vx, ix)
  call dac_read_last(11, trim(indir)//'/vy.dac', vy, ix)
  call dac_read_last(11, trim(indir)//'/by.dac', by, ix)
  call dac_read_last(11, trim(indir)//'/x.dac', x, ix)
  call dac_read_last(11, trim(indir)//'/bx.dac', bx, ix)
  call dac_read_last(11, trim(indir)//'/bxm.dac', bxm, ix)
  call dac_read_last(11, trim(indir)//'/gx.dac', gx, ix)
  call dac_read_last(11, trim(indir)//'/gxm.dac', gxm, ix)
  call dac_read_last(11, trim(indir)//'/sc.dac', sc, ix)
  call dac_read_last(11, trim(indir)//'/dsc.dac', dsc, ix)
  call dac_read_last(11, trim(indir)//'/scm.dac', scm, ix)
  call dac_read_last(11, trim(indir)//'/dscm.dac', dscm, ix)
  call dac_read_last(11, trim(indir)//'/rr.dac', rr, ix)
  call dac_read_last(11, trim(indir)//'/rrm.dac', rrm, ix)
  call dac_read_last(11, trim(indir)//'/drr.dac', drr, ix)
  call dac_read_last(11, trim(indir)//'/drrm.dac', drrm, ix)
  call dac_read_last(11, trim(indir)//'/dx.dac', dx, ix)
  call dac_read_last(11, trim(indir)//'/dxm.dac', dxm, ix)

  call cans_gpu_rangen_seed(seed)
  if (resident /= 0) call cans_gpu_resident(1)
  ns = 0
  t = 0.0d0
  merr = 0
  do
     ns = ns + 1
     call cfl_m(dt, safety, dtmin, merr, gm, bx, ro, pr, vx, vy, by, dx, ix)
     if (merr /= 0) exit
     t = t + dt
     call mlw_m_bg(ro, pr, vx, vy, by, bx, bxm, dt, qav, gm, gx, gxm, sc, dsc, scm, dscm, rr, rrm, drr, drrm, &
          dx, dxm, ix)
     call trqalfven(vy, dt, t, x, ix, rr)
     call spicule_bnd(margin, ro, pr, vx, vy, by, ix)
     call chkdav(nf, ro, vx, floor, ix)
     call chkdav(nf, pr, vx, floor, ix)
     if (.not. (ns < nstop .and. t < tend)) exit
  end do
  if (resident /= 0) call cans_gpu_sync()

  call dac_write_series(12, trim(outdir)//'/ro.dac', ro, [ix], 1)
  call dac_write_series(12, trim(outdir)//'/pr.dac', pr, [ix], 1)
  call dac_write_series(12, trim(outdir)//'/vx.dac', vx, [ix], 1)
  call dac_write_series(12, trim(outdir)//'/vy.dac', vy, [ix], 1)
  call dac_write_series(12, trim(outdir)//'/by.dac', by, [ix], 1)
  call dac_write_scalar_series(12, trim(outdir)//'/t.dac', t)
  write(6, '(1x,a,a,i8,a,e10.3,a,i5)') ' stop     ', 'step=', ns, ' time=', t, ' merr=', merr
end program spicule_main

! md_spicule/hdmlw/bnd.f:10-20
subroutine spicule_bnd(margin, ro, pr, vx, vy, by, ix)
  implicit none
  integer margin, ix
  double precision ro(ix), pr(ix), vx(ix), vy(ix), by(ix)
  call bdsppx(0, margin, pr, ix)
  call bdsppx(0, margin, ro, ix)
  call bdspnx(0, margin, vx, ix)
  call bdsppx(0, margin, vy, ix)
  call bdsppx(0, margin, by, ix)
  call bdfrex(1, margin, pr, ix)
  call bdfrex(1, margin, ro, ix)
  call bdfrex(1, margin, vx, ix)
  call bdfrex(1, margin, vy, ix)
  call bdfrex(1, margin, by, ix)
end subroutine spicule_bnd
