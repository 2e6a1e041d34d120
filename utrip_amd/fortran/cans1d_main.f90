!=======================================================================
! cans1d_main -- a CANS 1D main program on the GPU drop-in library.
!
! The time loop is the reference's (cans1d/md_mhdshktb/hdmlw/main.f:132-186):
!     ns = ns+1; cfl_m; stop on merr; t = t+dt; engine; bnd; loop test
! with the engine mlw_m (cans1d/hdmlw) or roe_m (cans1d/hdroe) and free
! (bdfrex) or periodic (bdperx) boundaries on all five fields.  The state
! restarts from DAC files (main.f:77-102, mcont=1) and the final state is
! written as DAC files.
!
! usage: cans1d_main ENGINE BC INDIR OUTDIR NSTOP TEND MARGIN SAFETY QAV GM RESIDENT
!=======================================================================
program cans1d_main
  use dac_io
  implicit none
  character(len=512) :: engine, bc, indir, outdir, arg
  integer :: nstop, margin, resident, ix, ns, merr, dims(4), nd
  double precision :: tend, safety, qav, gm, dtmin, t, dt
  double precision, allocatable :: ro(:), pr(:), vx(:), vy(:), by(:), bx(:), bxm(:), dx(:), dxm(:)

  call get_command_argument(1, engine)
  call get_command_argument(2, bc)
  call get_command_argument(3, indir)
  call get_command_argument(4, outdir)
  call get_command_argument(5, arg); read(arg, *) nstop
  call get_command_argument(6, arg); read(arg, *) tend
  call get_command_argument(7, arg); read(arg, *) margin
  call get_command_argument(8, arg); read(arg, *) safety
  call get_command_argument(9, arg); read(arg, *) qav
  call get_command_argument(10, arg); read(arg, *) gm
  call get_command_argument(11, arg); read(arg, *) resident
  dtmin = 1.d-10

  call dac_dims(11, trim(indir)//'/ro.dac', dims, nd)
  ix = dims(1)
  allocate(ro(ix), pr(ix), vx(ix), vy(ix), by(ix), bx(ix), bxm(ix), dx(ix), dxm(ix))
  call dac_read_last(11, trim(indir)//'/ro.dac', ro, ix)
  call dac_read_last(11, trim(indir)//'/pr.dac', pr, ix)
  call dac_read_last(11, trim(indir)//'/vx.dac', vx, ix)
  call dac_read_last(11, trim(indir)//'/vy.dac', vy, ix)
  call dac_read_last(11, trim(indir)//'/by.dac', by, ix)
  call dac_read_last(11, trim(indir)//'/bx.dac', bx, ix)
  call dac_read_last(11, trim(indir)//'/bxm.dac', bxm, ix)
  call dac_read_last(11, trim(indir)//'/dx.dac', dx, ix)
  call dac_read_last(11, trim(indir)//'/dxm.dac', dxm, ix)

  if (resident /= 0) call cans_gpu_resident(1)
  ns = 0
  t = 0.0d0
  merr = 0
  do
     ns = ns + 1
     call cfl_m(dt, safety, dtmin, merr, gm, bx, ro, pr, vx, vy, by, dx, ix)
     if (merr /= 0) exit
     t = t + dt
     if (trim(engine) == 'roe') then
        call roe_m(ro, pr, vx, vy, by, bx, bxm, dt, gm, dx, ix)
     else
        call mlw_m(ro, pr, vx, vy, by, bx, bxm, dt, qav, gm, dx, dxm, ix)
     end if
     call bnd1(bc, margin, ro, pr, vx, vy, by, ix)
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
end program cans1d_main

subroutine bnd1(bc, margin, ro, pr, vx, vy, by, ix)
  implicit none
  character(len=*) bc
  integer margin, ix
  double precision ro(ix), pr(ix), vx(ix), vy(ix), by(ix)
  if (trim(bc) == 'per') then
     call bdperx(margin, margin, ro, ix)
     call bdperx(margin, margin, pr, ix)
     call bdperx(margin, margin, vx, ix)
     call bdperx(margin, margin, vy, ix)
     call bdperx(margin, margin, by, ix)
  else
     call bdfrex(0, margin, ro, ix)
     call bdfrex(0, margin, pr, ix)
     call bdfrex(0, margin, vx, ix)
     call bdfrex(0, margin, vy, ix)
     call bdfrex(0, margin, by, ix)
     call bdfrex(1, margin, ro, ix)
     call bdfrex(1, margin, pr, ix)
     call bdfrex(1, margin, vx, ix)
     call bdfrex(1, margin, vy, ix)
     call bdfrex(1, margin, by, ix)
  end if
end subroutine bnd1
