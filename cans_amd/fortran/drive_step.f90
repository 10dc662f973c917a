! A Fortran host for the device step: the call sequence of the reference's
! src/main.f90 (initialisation 241-395, RK3 substeps 437-473, checks 489-511)
! written against the C ABI through module cans_hip.  It reads a &dns-style
! namelist (subset of src/param.f90:117-140) and restart files in the
! reference's checkpoint format (src/load.f90:461-589), advances nstep steps,
! and writes the fields back in the same format.
!
! usage: drive_step <dir>   (reads <dir>/case.nml and <dir>/fld_{u,v,w,p}.bin)
! The namelist's optional is_mixed = .true. selects the mixed-precision context
! (pressure-like arrays in FP32, velocity in FP64); the call sequence is the same.
program drive_step
  use, intrinsic :: iso_c_binding
  use cans_hip
  implicit none
  integer, parameter :: dp = c_double
  real(dp), parameter :: rkcoeff(2,3) = reshape([32._dp/60._dp,  0._dp        , &
                                                 25._dp/60._dp, -17._dp/60._dp, &
                                                 45._dp/60._dp, -25._dp/60._dp], [2,3])
  integer :: ng(3), gtype, nstep, istep, irk, k, idir, ib
  real(dp) :: l(3), gr, visci, dt, time, dtrk, dtrki, alpha
  character(len=1) :: cbcvel(0:1,3,3), cbcpre(0:1,3)
  real(dp) :: bcvel(0:1,3,3), bcpre(0:1,3), bforce(3), velf(3), gacc(3) = 0.
  logical :: is_forced(3), is_mixed
  integer(c_int) :: es_p
  namelist /dns/ ng, l, gtype, gr, visci, dt, nstep, cbcvel, cbcpre, bcvel, bcpre, bforce, is_forced, velf, is_mixed
  character(len=512) :: dir
  integer(c_int) :: n(3), nb(0:1,3), is_bound(0:1,3), isf(3)
  real(dp) :: dl(3), dli(3), visc, normfft, f(3), divtot, divmax
  real(dp), allocatable, target :: dzc(:), dzf(:), zc(:), zf(:), dzci(:), dzfi(:), gvr_c(:), gvr_f(:)
  real(dp), allocatable, target :: lamx(:), lamy(:), lam(:,:), a(:), b(:), c(:)
  real(dp), allocatable, target :: host(:,:,:)
  type(c_ptr) :: ctx, plan, u, v, w, p, pp, ru, rv, rw
  type(c_ptr) :: dzci_d, dzfi_d, gvr_c_d, gvr_f_d, lam_d, a_d, b_d, c_d
  character(len=1) :: names(4) = ['u','v','w','p']
  type(c_ptr) :: flds(4)
  integer :: iunit, i, j
  !
  call get_command_argument(1, dir)
  cbcvel = 'P'; cbcpre = 'P'; bcvel = 0.; bcpre = 0.; bforce = 0.; is_forced = .false.; velf = 0.
  gtype = 1; gr = 0.; nstep = 1; is_mixed = .false.
  open(newunit=iunit, file=trim(dir)//'/case.nml', status='old')
  read(iunit, nml=dns)
  close(iunit)
  n(:) = ng(:)
  ! param.f90:196-198
  dl(:) = l(:)/(1.*ng(:))
  dli(:) = dl(:)**(-1)
  visc = visci**(-1)
  ! single GPU = one x-pencil: x bounded; y/z periodic -> own neighbour, else none (initmpi.f90:226-267)
  nb(:,1) = -1; is_bound(:,1) = 1
  do idir=2,3
    if(cbcpre(0,idir)//cbcpre(1,idir) == 'PP') then
      nb(:,idir) = 0; is_bound(:,idir) = 0
    else
      nb(:,idir) = -1; is_bound(:,idir) = 1
    end if
  end do
  ! main.f90:241-285
  allocate(dzc(0:ng(3)+1), dzf(0:ng(3)+1), zc(0:ng(3)+1), zf(0:ng(3)+1), dzci(0:ng(3)+1), dzfi(0:ng(3)+1), &
           gvr_c(0:ng(3)+1), gvr_f(0:ng(3)+1))
  call cans_check(cans_initgrid(gtype, ng(3), gr, l(3), merge(1, 0, cbcpre(0,3)//cbcpre(1,3) == 'PP'), &
                                dzc, dzf, zc, zf), 'cans_initgrid')
  dzci(:) = dzc(:)**(-1)
  dzfi(:) = dzf(:)**(-1)
  gvr_c(:) = dl(1)*dl(2)*dzc(:)/(l(1)*l(2)*l(3))
  gvr_f(:) = dl(1)*dl(2)*dzf(:)/(l(1)*l(2)*l(3))
  ! initsolver.f90:15-83 (device ordering of the eigenvalues)
  allocate(lamx(ng(1)), lamy(ng(2)), lam(ng(1),ng(2)), a(ng(3)), b(ng(3)), c(ng(3)))
  call cans_check(cans_eigenvalues(ng(1), cbcpre(:,1), 'c', lamx), 'cans_eigenvalues')
  call cans_check(cans_eigenvalues(ng(2), cbcpre(:,2), 'c', lamy), 'cans_eigenvalues')
  lamx(:) = lamx(:)*dli(1)**2
  lamy(:) = lamy(:)*dli(2)**2
  do j=1,ng(2)
    do i=1,ng(1)
      lam(i,j) = lamx(i)+lamy(j)
    end do
  end do
  call cans_check(cans_tridmatrix(cbcpre(:,3), ng(3), dzci, dzfi, 'c', a, b, c), 'cans_tridmatrix')
  ! device context, fields and small arrays (what `!$acc enter data` provides, main.f90:388)
  call cans_check(cans_ctx_create(n, 0, merge(CANS_MIXED, CANS_FP64, is_mixed), c_null_ptr, ctx), 'cans_ctx_create')
  call cans_check(cans_fftini(ctx, ng, cbcpre(:,1:2), ['c','c'], plan, normfft), 'cans_fftini')
  call cans_check(cans_field_alloc(ctx, u), 'alloc'); call cans_check(cans_field_alloc(ctx, v), 'alloc')
  call cans_check(cans_field_alloc(ctx, w), 'alloc')
  call cans_check(cans_field_alloc_kind(ctx, CANS_FIELD_PRESSURE, p), 'alloc')
  call cans_check(cans_field_alloc_kind(ctx, CANS_FIELD_PRESSURE, pp), 'alloc')
  call cans_check(cans_ctx_elem_size(ctx, CANS_FIELD_PRESSURE, es_p), 'cans_ctx_elem_size')
  call dev_copy(dzci, size(dzci), dzci_d); call dev_copy(dzfi, size(dzfi), dzfi_d)
  call dev_copy(gvr_c, size(gvr_c), gvr_c_d); call dev_copy(gvr_f, size(gvr_f), gvr_f_d)
  call dev_copy(lam, size(lam), lam_d)
  call dev_copy(a, ng(3), a_d); call dev_copy(b, ng(3), b_d); call dev_copy(c, ng(3), c_d)
  ! dudtrko & co. hold the pressure-kind element type
  call cans_check(cans_device_alloc(ctx, int(es_p, c_long)*product(int(ng, c_long)), ru), 'alloc')
  call cans_check(cans_device_alloc(ctx, int(es_p, c_long)*product(int(ng, c_long)), rv), 'alloc')
  call cans_check(cans_device_alloc(ctx, int(es_p, c_long)*product(int(ng, c_long)), rw), 'alloc')
  ! restart (main.f90:382-385) into halo'd host arrays, then upload
  allocate(host(0:ng(1)+1,0:ng(2)+1,0:ng(3)+1))
  flds = [u, v, w, p]
  do k=1,4
    host(:,:,:) = 0.
    call load_one(trim(dir)//'/fld_'//names(k)//'.bin', host, time, istep)
    call cans_check(cans_field_upload(ctx, flds(k), host), 'cans_field_upload')
  end do
  isf(:) = merge(1, 0, is_forced(:))
  ! main.f90:389-390
  call cans_check(cans_bounduvw(ctx, cbcvel, n, bcvel, nb, is_bound, dl, dzc, dzf, u, v, w, 0, &
                                CANS_HALO_ABSENT, CANS_HALO_ABSENT), 'cans_bounduvw')
  call cans_check(cans_boundp(ctx, cbcpre, n, bcpre, nb, is_bound, dl, dzc, p, CANS_HALO_UPPER), 'cans_boundp')
  alpha = 0.
  do istep=1,nstep
    time = time + dt
    do irk=1,3
      ! main.f90:438-472
      dtrk = sum(rkcoeff(:,irk))*dt
      dtrki = dtrk**(-1)
      call cans_check(cans_rk(ctx, rkcoeff(:,irk), n, dli, dzci_d, dzfi_d, gvr_c_d, gvr_f_d, dt, visc, p, &
                              isf, velf, bforce, gacc, 0._c_double, c_null_ptr, ru, rv, rw, u, v, w, f), 'cans_rk')
      call cans_check(cans_bounduvw(ctx, cbcvel, n, bcvel, nb, is_bound, dl, dzc, dzf, u, v, w, 0, &
                                    CANS_HALO_LOWER, CANS_HALO_ABSENT), 'cans_bounduvw')
      call cans_check(cans_fillps(ctx, n, dli, dzfi_d, dtrki, u, v, w, pp), 'cans_fillps')
      call cans_check(cans_solver(ctx, n, ng, plan, normfft, lam_d, a_d, b_d, c_d, cbcpre, ['c','c','c'], pp), &
                      'cans_solver')
      call cans_check(cans_boundp(ctx, cbcpre, n, bcpre, nb, is_bound, dl, dzc, pp, CANS_HALO_ABSENT), 'cans_boundp')
      call cans_check(cans_correc(ctx, n, dli, dzci_d, dtrk, pp, u, v, w), 'cans_correc')
      call cans_check(cans_bounduvw(ctx, cbcvel, n, bcvel, nb, is_bound, dl, dzc, dzf, u, v, w, 1, &
                                    CANS_HALO_UPPER, CANS_HALO_NONE), 'cans_bounduvw')
      call cans_check(cans_updatep(ctx, n, dli, dzci_d, dzfi_d, alpha, pp, p, 0), 'cans_updatep')
      call cans_check(cans_boundp(ctx, cbcpre, n, bcpre, nb, is_bound, dl, dzc, p, CANS_HALO_UPPER), 'cans_boundp')
    end do
  end do
  call cans_check(cans_chkdiv(ctx, [1,1,1], ng, l, dli, dzfi_d, u, v, w, divtot, divmax), 'cans_chkdiv')
  print '(a,es12.4,a,es12.4)', 'divergence: mean = ', divtot, ' max = ', divmax
  do k=1,4
    call cans_check(cans_field_download(ctx, flds(k), host), 'cans_field_download')
    call write_one(trim(dir)//'/out_'//names(k)//'.bin', host, time, nstep)
  end do
  call cans_check(cans_fftend(plan), 'cans_fftend')
  call cans_check(cans_ctx_destroy(ctx), 'cans_ctx_destroy')
contains
  subroutine dev_copy(arr, nel, ptr)
    real(dp), intent(in), target :: arr(*)
    integer, intent(in) :: nel
    type(c_ptr), intent(out) :: ptr
    call cans_check(cans_device_alloc(ctx, int(8*nel, c_long), ptr), 'cans_device_alloc')
    call cans_check(cans_memcpy_h2d(ctx, ptr, c_loc(arr), int(8*nel, c_long)), 'cans_memcpy_h2d')
  end subroutine dev_copy
  subroutine load_one(fname, q, time, istep)
    ! load.f90:461-589: ng(1)*ng(2)*ng(3) interior reals + [time, istep]
    character(len=*), intent(in) :: fname
    real(dp), intent(inout) :: q(0:,0:,0:)
    real(dp), intent(out) :: time
    integer, intent(out) :: istep
    real(dp) :: info(2)
    integer :: iu
    open(newunit=iu, file=fname, access='stream', form='unformatted', status='old')
    read(iu) q(1:ng(1),1:ng(2),1:ng(3))
    read(iu) info
    close(iu)
    time = info(1); istep = nint(info(2))
  end subroutine load_one
  subroutine write_one(fname, q, time, istep)
    character(len=*), intent(in) :: fname
    real(dp), intent(in) :: q(0:,0:,0:)
    real(dp), intent(in) :: time
    integer, intent(in) :: istep
    integer :: iu
    open(newunit=iu, file=fname, access='stream', form='unformatted', status='replace')
    write(iu) q(1:ng(1),1:ng(2),1:ng(3))
    write(iu) [time, 1._dp*istep]
    close(iu)
  end subroutine write_one
end program drive_step
