! Fortran interface to libcans_hip.so (include/cans_hip.h).
!
! A CaNS host keeps its own Fortran (src/main.f90, src/initsolver.f90, the
! namelist input) and replaces the bodies of the device routines by calls to
! these interfaces; argument order and meaning follow the reference routines,
! with the context handle first.  Every function returns 0 on success; the
! helper cans_check stops with the library's message otherwise, mirroring the
! reference's `error stop` on fatal conditions (src/fft.f90:525).
module cans_hip
  use, intrinsic :: iso_c_binding
  implicit none
  integer(c_int), parameter :: CANS_FP64 = 0, CANS_MIXED = 1
  integer(c_int), parameter :: CANS_FIELD_VELOCITY = 0, CANS_FIELD_PRESSURE = 1
  integer(c_int), parameter :: CANS_HALO_ABSENT = -99, CANS_HALO_LOWER = -1, &
                               CANS_HALO_NONE = 0, CANS_HALO_UPPER = 1
  !
  ! cans_step_params (the argument bundle of one fused RK3 substep)
  !
  ! one implicit-diffusion system (what initsolver returns for a variable)
  type, bind(C) :: cans_helmholtz
    type(c_ptr)            :: plan = c_null_ptr
    real(c_double)         :: normfft = 1.
    type(c_ptr)            :: lambdaxy_d = c_null_ptr, a_d = c_null_ptr, b_d = c_null_ptr, c_d = c_null_ptr
    type(c_ptr)            :: rhsbx_d = c_null_ptr, rhsby_d = c_null_ptr, rhsbz_d = c_null_ptr
    character(kind=c_char) :: cbc(6) = 'P'
    character(kind=c_char) :: c_or_f(3) = 'c'
  end type cans_helmholtz
  type, bind(C) :: cans_step_params
    character(kind=c_char) :: cbcvel(18)
    real(c_double)         :: bcvel(18)
    character(kind=c_char) :: cbcpre(6)
    real(c_double)         :: bcpre(6)
    integer(c_int)         :: nb(6), is_bound(6)
    real(c_double)         :: dl(3), dli(3)
    real(c_double)         :: visc
    integer(c_int)         :: is_forced(3)
    real(c_double)         :: velf(3), bforce(3)
    type(c_ptr)            :: dzc_h, dzf_h, dzci_d, dzfi_d, grid_vol_ratio_c_d, grid_vol_ratio_f_d, plan
    real(c_double)         :: normfft
    type(c_ptr)            :: lambdaxy_d, a_d, b_d, c_d, rhsbx_d, rhsby_d, rhsbz_d
    type(cans_helmholtz)   :: hvel(3)
    integer(c_int)         :: nscal = 0
    type(c_ptr)            :: scal = c_null_ptr      ! nscal entries of type(cans_scalar)
    integer(c_int)         :: is_buoyancy = 0
    real(c_double)         :: gacc(3) = 0., beta = 0.
  end type cans_step_params
  ! the members of src/scal.f90's scalar type that the step uses
  type, bind(C) :: cans_scalar
    type(c_ptr)            :: val, dsdtrko
    real(c_double)         :: alpha
    character(kind=c_char) :: cbc(6)
    real(c_double)         :: bc(6)
    real(c_double)         :: source
    integer(c_int)         :: is_forced
    real(c_double)         :: scalf
    type(cans_helmholtz)   :: h
  end type cans_scalar
  type, bind(C) :: cans_state
    type(c_ptr) :: u, v, w, p, pp, dudtrko, dvdtrko, dwdtrko
  end type cans_state
  ! one independent Poisson problem of a batch: the arguments of one solver call
  type, bind(C) :: cans_poisson_box
    integer(c_int)         :: n(3)
    integer(c_long)        :: sy, sz
    character(kind=c_char) :: cbc(6)
    type(c_ptr)            :: lambdaxy_h, a_h, b_h, c_h, p_d
  end type cans_poisson_box
  !
  interface
    ! ---- context and memory (replace initmpi/workspaces device setup) ----
    integer(c_int) function cans_ctx_create(n, ld, precision, stream, ctx) bind(C, name='cans_ctx_create')
      import :: c_int, c_ptr
      integer(c_int), intent(in) :: n(3)
      integer(c_int), value :: ld, precision
      type(c_ptr), value :: stream
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function cans_ctx_destroy(ctx) bind(C, name='cans_ctx_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function cans_ctx_sync(ctx) bind(C, name='cans_ctx_sync')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function cans_last_error() bind(C, name='cans_last_error')
      import :: c_ptr
    end function
    integer(c_int) function cans_ctx_elem_size(ctx, kind, bytes) bind(C, name='cans_ctx_elem_size')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: kind
      integer(c_int), intent(out) :: bytes
    end function
    integer(c_int) function cans_field_alloc(ctx, field) bind(C, name='cans_field_alloc')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), intent(out) :: field
    end function
    integer(c_int) function cans_field_alloc_kind(ctx, kind, field) bind(C, name='cans_field_alloc_kind')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: kind
      type(c_ptr), intent(out) :: field
    end function
    integer(c_int) function cans_field_free(ctx, field) bind(C, name='cans_field_free')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, field
    end function
    integer(c_int) function cans_ctx_set_numerics(ctx, is_impdiff, is_impdiff_1d) &
                   bind(C, name='cans_ctx_set_numerics')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: is_impdiff, is_impdiff_1d
    end function
    integer(c_int) function cans_ctx_set_dtdma(ctx, nslab) bind(C, name='cans_ctx_set_dtdma')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nslab
    end function
    ! FP64 refinement sweeps of the mixed fused substep's pressure solve (substep_mask: bit irk-1)
    integer(c_int) function cans_ctx_set_refinement(ctx, sweeps, substep_mask) bind(C, name='cans_ctx_set_refinement')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: sweeps, substep_mask
    end function
    ! mixed mode: the RK intermediate velocity as FP32 increments (1) or an FP64 field (0, default)
    integer(c_int) function cans_ctx_set_mixed_increments(ctx, on) bind(C, name='cans_ctx_set_mixed_increments')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function
    integer(c_int) function cans_field_register(ctx, field, kind) bind(C, name='cans_field_register')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, field
      integer(c_int), value :: kind
    end function
    integer(c_int) function cans_field_upload(ctx, field, host) bind(C, name='cans_field_upload')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, field
      real(c_double), intent(in) :: host(*)
    end function
    integer(c_int) function cans_field_download(ctx, field, host) bind(C, name='cans_field_download')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, field
      real(c_double), intent(out) :: host(*)
    end function
    integer(c_int) function cans_device_alloc(ctx, bytes, ptr) bind(C, name='cans_device_alloc')
      import :: c_int, c_ptr, c_long
      type(c_ptr), value :: ctx
      integer(c_long), value :: bytes
      type(c_ptr), intent(out) :: ptr
    end function
    integer(c_int) function cans_device_free(ctx, ptr) bind(C, name='cans_device_free')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, ptr
    end function
    integer(c_int) function cans_memcpy_h2d(ctx, dst, src, bytes) bind(C, name='cans_memcpy_h2d')
      import :: c_int, c_ptr, c_long
      type(c_ptr), value :: ctx, dst, src
      integer(c_long), value :: bytes
    end function
    integer(c_int) function cans_memcpy_d2h(ctx, dst, src, bytes) bind(C, name='cans_memcpy_d2h')
      import :: c_int, c_ptr, c_long
      type(c_ptr), value :: ctx, dst, src
      integer(c_long), value :: bytes
    end function
    ! ---- host set-up helpers (initgrid.f90, initsolver.f90) ----
    integer(c_int) function cans_initgrid(gtype, n, gr, lz, is_periodic, dzc, dzf, zc, zf) &
                   bind(C, name='cans_initgrid')
      import :: c_int, c_double
      integer(c_int), value :: gtype, n, is_periodic
      real(c_double), value :: gr, lz
      real(c_double), intent(out) :: dzc(*), dzf(*), zc(*), zf(*)
    end function
    integer(c_int) function cans_initgrid_natural(n, lz, is_periodic, dzc, dzf, zc, zf) &
                   bind(C, name='cans_initgrid_natural')
      import :: c_int, c_double
      integer(c_int), value :: n, is_periodic
      real(c_double), value :: lz
      real(c_double), intent(out) :: dzc(*), dzf(*), zc(*), zf(*)
    end function
    integer(c_int) function cans_eigenvalues(n, bc, c_or_f, lambda) bind(C, name='cans_eigenvalues')
      import :: c_int, c_double, c_char
      integer(c_int), value :: n
      character(kind=c_char), intent(in) :: bc(2)
      character(kind=c_char), value :: c_or_f
      real(c_double), intent(out) :: lambda(*)
    end function
    integer(c_int) function cans_transform_matrix(n, bc, c_or_f, dir, m) bind(C, name='cans_transform_matrix')
      import :: c_int, c_double, c_char
      integer(c_int), value :: n, dir
      character(kind=c_char), intent(in) :: bc(2)
      character(kind=c_char), value :: c_or_f
      real(c_double), intent(out) :: m(*)
    end function
    integer(c_int) function cans_bc_rhs(cbc, bc, dlc, dlf, c_or_f, rhs) bind(C, name='cans_bc_rhs')
      import :: c_int, c_double, c_char
      character(kind=c_char), intent(in) :: cbc(2)
      real(c_double), intent(in) :: bc(2), dlc(2), dlf(2)
      character(kind=c_char), value :: c_or_f
      real(c_double), intent(out) :: rhs(2)
    end function
    integer(c_int) function cans_tridmatrix(bc, n, dzci, dzfi, c_or_f, a, b, c) bind(C, name='cans_tridmatrix')
      import :: c_int, c_double, c_char
      character(kind=c_char), intent(in) :: bc(2)
      integer(c_int), value :: n
      real(c_double), intent(in) :: dzci(*), dzfi(*)
      character(kind=c_char), value :: c_or_f
      real(c_double), intent(out) :: a(*), b(*), c(*)
    end function
    ! ---- fftini / fftend (fft.f90:25-245) ----
    integer(c_int) function cans_fftini(ctx, ng, bcxy, c_or_f, arrplan, normfft) bind(C, name='cans_fftini')
      import :: c_int, c_ptr, c_char, c_double
      type(c_ptr), value :: ctx
      integer(c_int), intent(in) :: ng(3)
      character(kind=c_char), intent(in) :: bcxy(4), c_or_f(2)
      type(c_ptr), intent(out) :: arrplan
      real(c_double), intent(out) :: normfft
    end function
    integer(c_int) function cans_fftend(arrplan) bind(C, name='cans_fftend')
      import :: c_int, c_ptr
      type(c_ptr), value :: arrplan
    end function
    ! ---- operators (same argument lists as the reference, context first) ----
    integer(c_int) function cans_solver(ctx, n, ng, arrplan, normfft, lambdaxy, a, b, c, bc, c_or_f, p) &
                   bind(C, name='cans_solver')
      import :: c_int, c_ptr, c_double, c_char
      type(c_ptr), value :: ctx, arrplan, lambdaxy, a, b, c, p
      integer(c_int), intent(in) :: n(3), ng(3)
      real(c_double), value :: normfft
      character(kind=c_char), intent(in) :: bc(6), c_or_f(3)
    end function
    ! cans_solver on a fixed device scratch: no allocation in the call (for callers that cannot allocate)
    integer(c_int) function cans_solver_noalloc(ctx, n, ng, arrplan, normfft, lambdaxy, a, b, c, bc, c_or_f, p) &
                   bind(C, name='cans_solver_noalloc')
      import :: c_int, c_ptr, c_double, c_char
      type(c_ptr), value :: ctx, arrplan, lambdaxy, a, b, c, p
      integer(c_int), intent(in) :: n(3), ng(3)
      real(c_double), value :: normfft
      character(kind=c_char), intent(in) :: bc(6), c_or_f(3)
    end function
    ! many independent Poisson problems of different sizes, five launches per solve for all of them
    integer(c_int) function cans_poisson_batch_create(ctx, nbox, boxes, normfft, batch) &
                   bind(C, name='cans_poisson_batch_create')
      import :: c_int, c_ptr, c_double, cans_poisson_box
      type(c_ptr), value :: ctx
      integer(c_int), value :: nbox
      type(cans_poisson_box), intent(in) :: boxes(*)
      real(c_double), intent(out) :: normfft(*)
      type(c_ptr), intent(out) :: batch
    end function
    integer(c_int) function cans_poisson_batch_solve(batch) bind(C, name='cans_poisson_batch_solve')
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
    end function
    integer(c_int) function cans_poisson_batch_destroy(batch) bind(C, name='cans_poisson_batch_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
    end function
    integer(c_int) function cans_solve_helmholtz(ctx, n, ng, arrplan, normfft, alpha, lambdaxy, a, b, c, &
                                                 rhsbx, rhsby, rhsbz, is_bound, cbc, c_or_f, p) &
                   bind(C, name='cans_solve_helmholtz')
      import :: c_int, c_ptr, c_double, c_char
      type(c_ptr), value :: ctx, arrplan, lambdaxy, a, b, c, rhsbx, rhsby, rhsbz, p
      integer(c_int), intent(in) :: n(3), ng(3), is_bound(6)
      real(c_double), value :: normfft, alpha
      character(kind=c_char), intent(in) :: cbc(6), c_or_f(3)
    end function
    ! several solve_helmholtz calls (main.f90:453-462) as one batch: one launch per transform pass
    integer(c_int) function cans_solve_helmholtz_batch(ctx, nsys, h, alpha, is_bound, f) &
                   bind(C, name='cans_solve_helmholtz_batch')
      import :: c_int, c_ptr, c_double, cans_helmholtz
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsys
      type(cans_helmholtz), intent(in) :: h(*)
      real(c_double), intent(in) :: alpha(*)
      integer(c_int), intent(in) :: is_bound(6)
      type(c_ptr), intent(in) :: f(*)
    end function
    ! s: the first scalar's field with Boussinesq buoyancy (rk.f90:69-72), else c_null_ptr
    integer(c_int) function cans_rk(ctx, rkpar, n, dli, dzci, dzfi, grid_vol_ratio_c, grid_vol_ratio_f, dt, visc, &
                                    p, is_forced, velf, bforce, gacc, beta, s, dudtrko, dvdtrko, dwdtrko, &
                                    u, v, w, f) bind(C, name='cans_rk')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, dzci, dzfi, grid_vol_ratio_c, grid_vol_ratio_f, p, s, dudtrko, dvdtrko, dwdtrko, &
                            u, v, w
      real(c_double), intent(in) :: rkpar(2), dli(3), velf(3), bforce(3), gacc(3)
      integer(c_int), intent(in) :: n(3), is_forced(3)
      real(c_double), value :: dt, visc, beta
      real(c_double), intent(out) :: f(3)
    end function
    integer(c_int) function cans_rk_scal(ctx, rkpar, n, dli, l, dzci, dzfi, grid_vol_ratio_f, alpha, dt, is_bound, &
                                         u, v, w, is_forced, scalf, ssource, fluxo, dsdtrko, s, f) &
                   bind(C, name='cans_rk_scal')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, dzci, dzfi, grid_vol_ratio_f, u, v, w, dsdtrko, s
      real(c_double), intent(in) :: rkpar(2), dli(3), l(3)
      integer(c_int), intent(in) :: n(3), is_bound(6)
      real(c_double), value :: alpha, dt, scalf, ssource
      integer(c_int), value :: is_forced
      real(c_double), intent(inout) :: fluxo(6)
      real(c_double), intent(out) :: f
    end function
    integer(c_int) function cans_bulk_forcing_s(ctx, n, is_forced, ff, s) bind(C, name='cans_bulk_forcing_s')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, s
      integer(c_int), intent(in) :: n(3)
      integer(c_int), value :: is_forced
      real(c_double), value :: ff
    end function
    integer(c_int) function cans_fillps(ctx, n, dli, dzfi, dti, u, v, w, p) bind(C, name='cans_fillps')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, dzfi, u, v, w, p
      integer(c_int), intent(in) :: n(3)
      real(c_double), intent(in) :: dli(3)
      real(c_double), value :: dti
    end function
    integer(c_int) function cans_correc(ctx, n, dli, dzci, dt, p, u, v, w) bind(C, name='cans_correc')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, dzci, p, u, v, w
      integer(c_int), intent(in) :: n(3)
      real(c_double), intent(in) :: dli(3)
      real(c_double), value :: dt
    end function
    integer(c_int) function cans_updatep(ctx, n, dli, dzci, dzfi, alpha, pp, p, is_impdiff) &
                   bind(C, name='cans_updatep')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, dzci, dzfi, pp, p
      integer(c_int), intent(in) :: n(3)
      real(c_double), intent(in) :: dli(3)
      real(c_double), value :: alpha
      integer(c_int), value :: is_impdiff
    end function
    integer(c_int) function cans_bounduvw(ctx, cbc, n, bc, nb, is_bound, dl, dzc, dzf, u, v, w, &
                                          keep_norm, halos_norm, halos_tang) bind(C, name='cans_bounduvw')
      import :: c_int, c_ptr, c_double, c_char
      type(c_ptr), value :: ctx, u, v, w
      character(kind=c_char), intent(in) :: cbc(18)
      integer(c_int), intent(in) :: n(3), nb(6), is_bound(6)
      real(c_double), intent(in) :: bc(18), dl(3), dzc(*), dzf(*)
      integer(c_int), value :: keep_norm, halos_norm, halos_tang
    end function
    integer(c_int) function cans_boundp(ctx, cbc, n, bc, nb, is_bound, dl, dzc, p, halo_side) &
                   bind(C, name='cans_boundp')
      import :: c_int, c_ptr, c_double, c_char
      type(c_ptr), value :: ctx, p
      character(kind=c_char), intent(in) :: cbc(6)
      integer(c_int), intent(in) :: n(3), nb(6), is_bound(6)
      real(c_double), intent(in) :: bc(6), dl(3), dzc(*)
      integer(c_int), value :: halo_side
    end function
    integer(c_int) function cans_updt_rhs_b(ctx, c_or_f, cbc, n, is_bound, rhsbx, rhsby, rhsbz, p, alpha) &
                   bind(C, name='cans_updt_rhs_b')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx, rhsbx, rhsby, rhsbz, p, alpha
      character(kind=c_char), intent(in) :: c_or_f(3), cbc(6)
      integer(c_int), intent(in) :: n(3), is_bound(6)
    end function
    integer(c_int) function cans_chkdiv(ctx, lo, hi, l, dli, dzfi, u, v, w, divtot, divmax) &
                   bind(C, name='cans_chkdiv')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, dzfi, u, v, w
      integer(c_int), intent(in) :: lo(3), hi(3)
      real(c_double), intent(in) :: l(3), dli(3)
      real(c_double), intent(out) :: divtot, divmax
    end function
    integer(c_int) function cans_chkdt(ctx, n, dl, dzci, dzfi, visc, alpha, u, v, w, is_impdiff, is_impdiff_1d, &
                                       dtmax) bind(C, name='cans_chkdt')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, dzci, dzfi, u, v, w
      integer(c_int), intent(in) :: n(3)
      real(c_double), intent(in) :: dl(3)
      real(c_double), value :: visc, alpha
      integer(c_int), value :: is_impdiff, is_impdiff_1d
      real(c_double), intent(out) :: dtmax
    end function
    integer(c_int) function cans_bulk_mean(ctx, n, grid_vol_ratio, p, mean) bind(C, name='cans_bulk_mean')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, grid_vol_ratio, p
      integer(c_int), intent(in) :: n(3)
      real(c_double), intent(out) :: mean
    end function
    integer(c_int) function cans_rk3_substep(ctx, prm, irk, dt, st) bind(C, name='cans_rk3_substep')
      import :: c_int, c_ptr, c_double, cans_step_params, cans_state
      type(c_ptr), value :: ctx
      type(cans_step_params), intent(in) :: prm
      integer(c_int), value :: irk
      real(c_double), value :: dt
      type(cans_state), intent(inout) :: st
    end function
    ! the whole do irk = 1,3 loop body of main.f90 (projections folded into the next momentum kernel)
    integer(c_int) function cans_rk3_step(ctx, prm, dt, st) bind(C, name='cans_rk3_step')
      import :: c_int, c_ptr, c_double, cans_step_params, cans_state
      type(c_ptr), value :: ctx
      type(cans_step_params), intent(in) :: prm
      real(c_double), value :: dt
      type(cans_state), intent(inout) :: st
    end function
    ! device address of the bulk forcing f(1:3) of the last rk (no counterpart: rk.f90 keeps it on the host)
    integer(c_int) function cans_ctx_forcing(ctx, f_d) bind(C, name='cans_ctx_forcing')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), intent(out) :: f_d
    end function
    integer(c_int) function cans_ctx_layout(ctx, sy, sz, field_elems) bind(C, name='cans_ctx_layout')
      import :: c_int, c_ptr, c_long
      type(c_ptr), value :: ctx
      integer(c_long), intent(out) :: sy, sz, field_elems
    end function
    ! launch shapes the plan's tuner chose: x threads (FP32, FP64), y lines (FP32, FP64)
    integer(c_int) function cans_fft_tuning(arrplan, shapes) bind(C, name='cans_fft_tuning')
      import :: c_int, c_ptr
      type(c_ptr), value :: arrplan
      integer(c_int), intent(out) :: shapes(4)
    end function
    ! radix of the power-of-two FFT stages: x (FP32, FP64), y (FP32, FP64); 8, 4, or 0 (dense)
    integer(c_int) function cans_fft_radix(arrplan, radix) bind(C, name='cans_fft_radix')
      import :: c_int, c_ptr
      type(c_ptr), value :: arrplan
      integer(c_int), intent(out) :: radix(4)
    end function
    integer(c_int) function cans_fft_set_radix(arrplan, radix) bind(C, name='cans_fft_set_radix')
      import :: c_int, c_ptr
      type(c_ptr), value :: arrplan
      integer(c_int), intent(in) :: radix(4)
    end function
    integer(c_int) function cans_poisson_batch_tuning(batch, threads) bind(C, name='cans_poisson_batch_tuning')
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
      integer(c_int), intent(out) :: threads
    end function
    ! solver with FP64 iterative refinement (p FP64; inner solves in the context's pressure storage)
    integer(c_int) function cans_solver_refined(ctx, n, ng, arrplan, normfft, lambdaxy, a, b, c, bc, c_or_f, &
                                                dli, dzci, dzfi, sweeps, p) bind(C, name='cans_solver_refined')
      import :: c_int, c_ptr, c_double, c_char
      type(c_ptr), value :: ctx, arrplan, lambdaxy, a, b, c, dzci, dzfi, p
      integer(c_int), intent(in) :: n(3), ng(3)
      real(c_double), value :: normfft
      character(kind=c_char), intent(in) :: bc(6), c_or_f(3)
      real(c_double), intent(in) :: dli(3)
      integer(c_int), value :: sweeps
    end function
    ! adjoint (tri-periodic uniform grids, FP64): o = J(u)^T g + b of the momentum right-hand side
    integer(c_int) function cans_mom_jt(ctx, n, dli, visc, u, v, w, gu, gv, gw, bu, bv, bw, ou, ov, ow) &
                   bind(C, name='cans_mom_jt')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, u, v, w, gu, gv, gw, bu, bv, bw, ou, ov, ow
      integer(c_int), intent(in) :: n(3)
      real(c_double), intent(in) :: dli(3)
      real(c_double), value :: visc
    end function
    ! one RK3 substep run backwards: adjoints (ub, fb) of its outputs -> (ubo, fbo) of its inputs
    ! (FP64; each direction periodic or walled, a periodic z uniform; forcing of u and v)
    integer(c_int) function cans_rk3_substep_adjoint(ctx, prm, irk, dt, u, v, w, ub, vb, wb, fu, fv, fw, &
                                                     ubo, vbo, wbo, fuo, fvo, fwo, pp) &
                   bind(C, name='cans_rk3_substep_adjoint')
      import :: c_int, c_ptr, c_double, cans_step_params
      type(c_ptr), value :: ctx, u, v, w, ub, vb, wb, fu, fv, fw, ubo, vbo, wbo, fuo, fvo, fwo, pp
      type(cans_step_params), intent(in) :: prm
      integer(c_int), value :: irk
      real(c_double), value :: dt
    end function
    ! the same with scalars: s(i) the scalars the forward substep started from, (sb, fsb) -> (sbo, fsbo)
    ! their adjoints and those of their stored right-hand sides (arrays of nscal device addresses)
    integer(c_int) function cans_rk3_substep_adjoint_scal(ctx, prm, irk, dt, u, v, w, ub, vb, wb, fu, fv, fw, &
                                                          ubo, vbo, wbo, fuo, fvo, fwo, pp, s, sb, fsb, sbo, fsbo) &
                   bind(C, name='cans_rk3_substep_adjoint_scal')
      import :: c_int, c_ptr, c_double, cans_step_params
      type(c_ptr), value :: ctx, u, v, w, ub, vb, wb, fu, fv, fw, ubo, vbo, wbo, fuo, fvo, fwo, pp
      type(cans_step_params), intent(in) :: prm
      integer(c_int), value :: irk
      real(c_double), value :: dt
      type(c_ptr), intent(in) :: s(*), sb(*), fsb(*), sbo(*), fsbo(*)
    end function
  end interface
contains
  subroutine cans_check(istat, what)
    integer(c_int), intent(in) :: istat
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: i
    if(istat == 0) return
    call c_f_pointer(cans_last_error(), msg, [1024])
    write(*,'(a)', advance='no') 'ERROR: '//what//' failed: '
    do i=1,1024
      if(msg(i) == c_null_char) exit
      write(*,'(a)', advance='no') msg(i)
    end do
    write(*,*)
    error stop
  end subroutine cans_check
end module cans_hip
