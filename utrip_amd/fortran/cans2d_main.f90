!=======================================================================
! cans2d_main -- a CANS 2D main program on the GPU drop-in library.
!
! The time loop is the reference's (cans2d/md_mhdshktb/hdmlw/main.f:150-206):
!     ns = ns+1; cfl_m; stop on merr; t = t+dt; engine; bnd; loop test
! with the engine mlw_m (cans2d/hdmlw) or roe_m (cans2d/hdroe) and periodic
! boundaries (bdperx then bdpery on every field, cans2d/md_mhdkh/hdmlw/bnd.f).
! The initial state is restarted from DAC files (the mcont=1 path of
! main.f:89-120: the last record of INDIR/<field>.dac) and the final state
! is written as DAC files to OUTDIR.
!
! usage: cans2d_main ENGINE INDIR OUTDIR NSTOP TEND MARGIN SAFETY QAV GM RESIDENT
!   RESIDENT = 1 keeps the arrays on the GPU between calls (one copy back at
!   the end); 0 is the strict drop-in (every call copies its arrays).
!=======================================================================
program cans2d_main
  use dac_io
  implicit none
  character(len=512) :: engine, indir, outdir, arg
  integer :: nstop, margin, resident, ix, jx, ns, merr, dims(4), nd, n
  double precision :: tend, safety, qav, gm, dtmin, t, dt
  double precision, allocatable :: ro(:, :), pr(:, :), vx(:, :), vy(:, :), bx(:, :), by(:, :), az(:, :)
  double precision, allocatable :: dx(:), dxm(:), dy(:), dym(:)

  call get_command_argument(1, engine)
  call get_command_argument(2, indir)
  call get_command_argument(3, outdir)
  call get_command_argument(4, arg); read(arg, *) nstop
  call get_command_argument(5, arg); read(arg, *) tend
  call get_command_argument(6, arg); read(arg, *) margin
  call get_command_argument(7, arg); read(arg, *) safety
  call get_command_argument(8, arg); read(arg, *) qav
  call get_command_argument(9, arg); read(arg, *) gm
  call get_command_argument(10, arg); read(arg, *) resident
  dtmin = 1.d-10

  call dac_dims(11, trim(indir)//'/ro.dac', dims, nd)
  ix = dims(1); jx = dims(2); n = ix * jx
  allocate(ro(ix, jx), pr(ix, jx), vx(ix, jx), vy(ix, jx), bx(ix, jx), by(ix, jx), az(ix, jx))
  allocate(dx(ix), dxm(ix), dy(jx), dym(jx))
  call dac_read_last(11, trim(indir)//'/ro.dac', ro, n)
  call dac_read_last(11, trim(indir)//'/pr.dac', pr, n)
  call dac_read_last(11, trim(indir)//'/vx.dac', vx, n)
  call dac_read_last(11, trim(indir)//'/vy.dac', vy, n)
  call dac_read_last(11, trim(indir)//'/bx.dac', bx, n)
  call dac_read_last(11, trim(indir)//'/by.dac', by, n)
  call dac_read_last(11, trim(indir)//'/az.dac', az, n)
  call dac_read_last(11, trim(indir)//'/dx.dac', dx, ix)
  call dac_read_last(11, trim(indir)//'/dxm.dac', dxm, ix)
  call dac_read_last(11, trim(indir)//'/dy.dac', dy, jx)
  call dac_read_last(11, trim(indir)//'/dym.dac', dym, jx)

  if (resident /= 0) call cans_gpu_resident(1)
  ns = 0
  t = 0.0d0
  merr = 0
  do
     ns = ns + 1
     call cfl_m(dt, safety, dtmin, merr, gm, ro, pr, vx, vy, bx, by, dx, ix, dy, jx)
     if (merr /= 0) exit
     t = t + dt
     if (trim(engine) == 'roe') then
        call roe_m(ro, pr, vx, vy, bx, by, az, dt, gm, dx, ix, dy, jx)
     else
        call mlw_m(ro, pr, vx, vy, bx, by, az, dt, qav, gm, dx, dxm, ix, dy, dym, jx)
     end if
     call bnd(margin, ro, pr, vx, vy, bx, by, az, ix, jx)
     if (.not. (ns < nstop .and. t < tend)) exit
  end do
  if (resident /= 0) call cans_gpu_sync()

  call dac_write_series(12, trim(outdir)//'/ro.dac', ro, [ix, jx], 2)
  call dac_write_series(12, trim(outdir)//'/pr.dac', pr, [ix, jx], 2)
  call dac_write_series(12, trim(outdir)//'/vx.dac', vx, [ix, jx], 2)
  call dac_write_series(12, trim(outdir)//'/vy.dac', vy, [ix, jx], 2)
  call dac_write_series(12, trim(outdir)//'/bx.dac', bx, [ix, jx], 2)
  call dac_write_series(12, trim(outdir)//'/by.dac', by, [ix, jx], 2)
  call dac_write_series(12, trim(outdir)//'/az.dac', az, [ix, jx], 2)
  call dac_write_scalar_series(12, trim(outdir)//'/t.dac', t)
  write(6, '(1x,a,a,i8,a,e10.3,a,i5)') ' stop     ', 'step=', ns, ' time=', t, ' merr=', merr
end program cans2d_main

subroutine bnd(margin, ro, pr, vx, vy, bx, by, az, ix, jx)
  implicit none
  integer margin, ix, jx
  double precision ro(ix, jx), pr(ix, jx), vx(ix, jx), vy(ix, jx), bx(ix, jx), by(ix, jx), az(ix, jx)
  call bdperx(margin, margin, ro, ix, jx)
  call bdperx(margin, margin, pr, ix, jx)
  call bdperx(margin, margin, vx, ix, jx)
  call bdperx(margin, margin, vy, ix, jx)
  call bdperx(margin, margin, bx, ix, jx)
  call bdperx(margin, margin, by, ix, jx)
  call bdperx(margin, margin, az, ix, jx)
  call bdpery(margin, margin, ro, ix, jx)
  call bdpery(margin, margin, pr, ix, jx)
  call bdpery(margin, margin, vx, ix, jx)
  call bdpery(margin, margin, vy, ix, jx)
  call bdpery(margin, margin, bx, ix, jx)
  call bdpery(margin, margin, by, ix, jx)
  call bdpery(margin, margin, az, ix, jx)
end subroutine bnd
