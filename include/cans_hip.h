/*
 * cans_hip.h -- C ABI of the MI355X (gfx950) pressure-correction time step.
 *
 * Drop-in boundary for CaNS's per-routine Fortran interfaces on one GPU.  Every
 * entry point takes plain pointers and sizes, is ordered on the context's HIP
 * stream, returns 0 on success or a negative CANS_E* code (message in
 * cans_last_error()), and synchronises only where the reference reads a result
 * back on the host (reductions that return a value, downloads).
 *
 * Arrays follow the reference's Fortran conventions:
 *   - fields are halo'd (0:n1+1, 0:n2+1, 0:n3+1) with i fastest; the context
 *     fixes the leading dimension ld >= n1+2 and plane stride ld*(n2+2); a field
 *     pointer is the address of element (0,0,0);
 *   - z arrays are (0:n3+1); a,b,c are (1:n3); lambdaxy is (1:n1,1:n2);
 *   - character arrays are the flattened Fortran arrays (cbcvel(0:1,3,3) -> 18
 *     chars, cbcpre(0:1,3) -> 6, bcxy(0:1,2) -> 4), logicals are int 0/1;
 *   - nb(0:1,3): on one GPU a value >= 0 marks a periodic neighbour (this rank),
 *     a negative value (MPI_PROC_NULL) marks none.
 * Device pointers are marked _d, host pointers _h; unmarked scalars are by value.
 *
 * Precision modes (cans_ctx_create):
 *   CANS_FP64  - every array is double.
 *   CANS_MIXED - velocity fields (u, v, w) stay double; pressure-like arrays
 *                (p, pp, the field of cans_solver / cans_solve_helmholtz /
 *                cans_updt_rhs_b / cans_fillps's output, dudtrko, dvdtrko,
 *                dwdtrko) are float, as are the transform intermediates.  The
 *                FFT butterflies of the transforms compute in float; every
 *                other kernel (stencils, RK update, projection, the DCT/DST
 *                pre/post-processing and the z elimination) computes in
 *                double, and only stores to float arrays round.  Small arrays (grid, eigenvalues, a/b/c, rhsb, bc values) stay
 *                double.  cans_ctx_elem_size reports the size per kind.
 */
#ifndef CANS_HIP_H
#define CANS_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cans_ctx cans_ctx;
typedef struct cans_plan cans_plan;

enum { CANS_OK = 0, CANS_EINVAL = -1, CANS_EHIP = -2, CANS_ENOMEM = -3, CANS_EUNSUPPORTED = -4 };
enum { CANS_FP64 = 0, CANS_MIXED = 1 };
enum { CANS_FIELD_VELOCITY = 0, CANS_FIELD_PRESSURE = 1 };
enum { CANS_HALO_ABSENT = -99, CANS_HALO_LOWER = -1, CANS_HALO_NONE = 0, CANS_HALO_UPPER = 1 };

/* ---------------------------------------------------------------------------
 * Context, memory (replaces the device setup of src/initmpi.f90:40-279 and the
 * work buffers of src/workspaces.f90:16-147; the host keeps ownership of its
 * Fortran arrays and moves them with upload/download, as `!$acc enter data` /
 * `!$acc update self` do in src/main.f90:388,423).
 * ------------------------------------------------------------------------- */
/* n: interior sizes; ld: leading dimension (0 selects n1+2); stream: hipStream_t (NULL = default) */
int cans_ctx_create(const int *n, int ld, int precision, void *stream, cans_ctx **ctx);
int cans_ctx_destroy(cans_ctx *ctx);
/* strides of the halo'd layout, in elements */
int cans_ctx_layout(const cans_ctx *ctx, long *sy, long *sz, long *field_elems);
int cans_ctx_sync(cans_ctx *ctx);
const char *cans_last_error(void);

/* bytes per element of a field kind (CANS_FIELD_*) in this context */
int cans_ctx_elem_size(const cans_ctx *ctx, int kind, int *bytes);

/* field = device pointer to element (0,0,0) of a halo'd field owned by the context;
 * cans_field_alloc allocates a velocity field, cans_field_alloc_kind either kind */
int cans_field_alloc(cans_ctx *ctx, void **field_d);
int cans_field_alloc_kind(cans_ctx *ctx, int kind, void **field_d);
int cans_field_free(cans_ctx *ctx, void *field_d);
/* Declares the kind of a field the caller allocated itself (origin = element (0,0,0)),
 * so that entry points whose element type depends on the field (cans_boundp on a
 * scalar in a CANS_MIXED context, cans_field_upload/download) read it right.
 * Registering an address again replaces its kind; cans_field_free only forgets it.
 * No counterpart in the reference (its fields are typed Fortran arrays). */
int cans_field_register(cans_ctx *ctx, void *field_d, int kind);
/* the &numerics switches the reference's routines read from mod_param (param.f90:96,331):
 * is_impdiff (implicit diffusion, solve_helmholtz of u, v, w and the scalars) and
 * is_impdiff_1d (implicit in z only, which implies is_impdiff).  They change cans_rk,
 * cans_rk_scal, cans_updatep, cans_solve_helmholtz and cans_rk3_substep as in the
 * reference.  Implicit diffusion runs with FP64 velocity and scalar fields and FP64
 * Helmholtz solves; the fused substep supports it in both modes (under CANS_MIXED the
 * pressure side -- p, pp, RHS_old and the Poisson solve -- stays FP32). */
int cans_ctx_set_numerics(cans_ctx *ctx, int is_impdiff, int is_impdiff_1d);
/* is_poisson_dtdma (param.f90, solver.f90:309-517): the z elimination of the pressure
 * solve (and of the Helmholtz solves' 3-D form) as the distributed tridiagonal algorithm
 * with the z rows split into nslab slabs, the ranks of a z decomposition (each slab at
 * least 3 rows).  0 or 1: the plain sweep.  A periodic z closes the reduced system
 * cyclically, as the reference does.  Unlike the reference's gaussel_dtdma, a line whose
 * reduced system is singular (the constant mode of a Neumann/periodic operator) has its
 * last unknown pinned to zero, the gauge of the plain sweep. */
int cans_ctx_set_dtdma(cans_ctx *ctx, int nslab);
/* Iterative refinement of the pressure solve inside cans_rk3_substep (no counterpart in the
 * reference, whose solves run in one precision; CANS_MIXED contexts only).  Each substep irk
 * whose bit (irk-1) is set in substep_mask follows its FP32-storage solve with `sweeps` (0-2)
 * sweeps: the residual r = b - L(pp + x_1 + ...) of the pressure equation is formed in FP64
 * from the intermediate velocity (b, the rhs of fillps) and the corrections so far, L the
 * solver's 7-point operator, and solved again with FP32 storage into x_q; correc and the
 * pressure update then use pp + x_1 + x_2 summed in FP64 (pp itself keeps the first solve's
 * part).  After the substep div(u) = dtrk * (b - L(pp + x_1 + x_2)), the FP64 residual, so a
 * refined last substep (mask 4) leaves every step's velocity divergence-free to FP64 level.
 * Homogeneous pressure boundary values only.  Default: 0 sweeps (the unrefined mixed step). */
int cans_ctx_set_refinement(cans_ctx *ctx, int sweeps, int substep_mask);
/* How the fused mixed substep (explicit diffusion) keeps the RK intermediate velocity u*
 * (no counterpart in the reference, which stores it in place of u): on = 0 (default) as an
 * FP64 field, as in CANS_FP64; on = 1 as FP32 increments e = u* - u over the step's velocity,
 * half the scratch memory, with div(u) kept in FP32 and every consumer rebuilding u + e + f.
 * Both meet the mixed mode's 1e-6 bar; the FP64 field moves fewer bytes per step (the folded
 * momentum kernel and correc read u* instead of u and e, and re-read no u for the increment)
 * and is faster on MI355X.  CANS_MIXED contexts only. */
int cans_ctx_set_mixed_increments(cans_ctx *ctx, int on);
/* host arrays are the reference's contiguous (0:n1+1,0:n2+1,0:n3+1) Fortran arrays
 * of double; float fields round on upload.  In a CANS_MIXED context the field
 * must come from cans_field_alloc_kind (its element type is looked up) */
int cans_field_upload(cans_ctx *ctx, void *field_d, const double *host_h);
int cans_field_download(cans_ctx *ctx, const void *field_d, double *host_h);

/* plain device buffers for the small arrays (dzci, lambdaxy, a, b, c, ...): what
 * `!$acc enter data copyin(...)` provides to the reference (src/main.f90:233-337) */
int cans_device_alloc(cans_ctx *ctx, long bytes, void **ptr_d);
int cans_device_free(cans_ctx *ctx, void *ptr_d);
int cans_memcpy_h2d(cans_ctx *ctx, void *dst_d, const void *src_h, long bytes);
int cans_memcpy_d2h(cans_ctx *ctx, void *dst_h, const void *src_d, long bytes);

/* ---------------------------------------------------------------------------
 * Host-side set-up helpers, for hosts that do not keep the Fortran versions.
 * ------------------------------------------------------------------------- */
/* src/initgrid.f90:14-118 (gtype 1-4); arrays (0:n+1) */
int cans_initgrid(int gtype, int n, double gr, double lz, int is_periodic,
                  double *dzc_h, double *dzf_h, double *zc_h, double *zf_h);
/* src/initgrid.f90:14-99 with is_gridpoint_natural_channel = .true. (param.f90:99): the
 * faces of gridpoint_natural (initgrid.f90:162-203, Pirozzoli & Orlandi 2021), which takes
 * neither gtype nor gr; arrays (0:n+1) */
int cans_initgrid_natural(int n, double lz, int is_periodic, double *dzc_h, double *dzf_h, double *zc_h,
                          double *zf_h);
/* src/initsolver.f90:85-144 in the device ordering of this library (the
 * reference's GPU ordering, initsolver.f90:98-117, for periodic directions) */
int cans_eigenvalues(int n, const char *bc, char c_or_f, double *lambda_h);
/* The n x n matrix (row-major, m[k*n + j] multiplies input j into output k) of one
 * direction's forward (dir = 1) or backward (dir = -1) transform: the kind of
 * fft.f90:260-313 for the boundary pair bc and grid c_or_f, with FFTW's unnormalised
 * definition, its output in the order cans_eigenvalues uses.  A plan whose length the
 * FFT passes do not take (cans_fftini) applies exactly this matrix; n <= 8192. */
int cans_transform_matrix(int n, const char *bc, char c_or_f, int dir, double *m);
/* src/initsolver.f90:189-232: the two plane values of rhs(:,:,0:1) (every entry of
 * plane ib equals rhs_h[ib]); dlc, dlf: spacings at the two boundaries */
int cans_bc_rhs(const char *cbc, const double *bc, const double *dlc, const double *dlf, char c_or_f,
                double *rhs_h);
/* src/initsolver.f90:146-187; a,b,c (1:n) */
int cans_tridmatrix(const char *bc, int n, const double *dzci_h, const double *dzfi_h,
                    char c_or_f, double *a_h, double *b_h, double *c_h);

/* ---------------------------------------------------------------------------
 * Transforms: twins of fftini/fftend (src/fft.f90:25-245).  arrplan holds the
 * four batched transforms ((fwd,bwd) x (x,y)); normfft is returned as the
 * reference computes it (fft.f90:75,140,202,208).  Any length up to 8192 along x and y:
 * 2^m, or 2^m (m >= 3) times at most three factors 3 and 5 (e.g. 24, 40, 72, 120), at most
 * 4096, run the FFT passes; other lengths (and both directions when n1 is odd) apply the
 * dense matrix of cans_transform_matrix, O(n) work per element.
 * Boundary pairs along x and y (fft.f90:260-313): PP; cell-centred NN, DD, ND and DN (ND/DN
 * run on a line of twice the length, which must be such a length); face-centred ND and DN
 * (REDFT10/01, RODFT01/10) and DD (RODFT00 on the n-1 interior faces, fft.f90:80-82, through a
 * line of 2n points).  Face-centred NN (REDFT00) is refused: the reference forbids it with
 * implicit diffusion (sanity.f90:151-162), the only user of face-centred transforms.
 * ------------------------------------------------------------------------- */
int cans_fftini(cans_ctx *ctx, const int *ng, const char *bcxy, const char *c_or_f,
                cans_plan **arrplan, double *normfft);
int cans_fftend(cans_plan *arrplan);
/* The launch shapes the plan's tuner chose at cans_fftini (no counterpart in the
 * reference, whose FFT libraries tune themselves): shapes[0..1] x-pass threads per
 * work-group for FP32/FP64 storage, shapes[2..3] y-pass lines per work-group (0: not
 * tuned for that type; a heuristic applies). */
int cans_fft_tuning(const cans_plan *arrplan, int *shapes);
/* The radix of the power-of-two FFT stages the tuner chose (no counterpart in the reference):
 * radix[0..1] x pass FP32/FP64, radix[2..3] y pass FP32/FP64; 8 (radix-8 stages), 4 (radix-4
 * stages: twice the LDS round trips, half the values per butterfly) or 0 (a dense direction). */
int cans_fft_radix(const cans_plan *arrplan, int *radix);
/* Replaces the tuner's radix choice (4 or 8 each, same order; ignored for dense directions). */
int cans_fft_set_radix(cans_plan *arrplan, const int *radix);

/* ---------------------------------------------------------------------------
 * Operators.  Each replaces the reference routine cited; argument meaning and
 * order follow the Fortran interface, with the context first.
 * ------------------------------------------------------------------------- */
/* src/solver.f90:17 / src/solver_gpu.f90:34 -- p(0:,0:,0:) overwritten with the solution */
int cans_solver(cans_ctx *ctx, const int *n, const int *ng, const cans_plan *arrplan,
                double normfft, const double *lambdaxy_d, const double *a_d, const double *b_d,
                const double *c_d, const char *bc, const char *c_or_f, void *p_d);
/* No counterpart in the reference.  cans_solver with the same arguments and results, for
 * callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): the transform work buffer and the z
 * pass's checkpoints come from a fixed device array of CANS_NOALLOC_SCRATCH_BYTES that
 * the library holds on every GPU it runs on until the process ends, and the call makes
 * no allocation (cans_solver grows the context's scratch on its first call).  A grid
 * whose solve needs more scratch is refused with CANS_EINVAL (64 MiB covers, e.g., a
 * 128^3 tri-periodic FP64 solve, 38.8 MB).  The distributed z solve (cans_ctx_set_dtdma)
 * is refused.  All calls on one GPU share its array, so calls that overlap on one GPU
 * (two streams, or two host threads) race; calls ordered on one stream do not. */
#define CANS_NOALLOC_SCRATCH_BYTES (64L << 20)
int cans_solver_noalloc(cans_ctx *ctx, const int *n, const int *ng, const cans_plan *arrplan,
                        double normfft, const double *lambdaxy_d, const double *a_d, const double *b_d,
                        const double *c_d, const char *bc, const char *c_or_f, void *p_d);
/* Many independent Poisson problems of different sizes, solved together: each box
 * carries the arguments of one src/solver.f90:17 call, and cans_poisson_batch_create
 * does each box's fftini (src/fft.f90:25) as well.  A solve is five launches for
 * all boxes (x, y, z, y, x passes) in place on the boxes' fields, on the context's stream,
 * with no allocation or host synchronisation (it can be captured in a HIP graph).
 * The boxes are cell-centred (c_or_f = 'ccc'); x and y pairs PP, NN, DD, ND or DN with
 * lengths 2^m, or 2^m (m >= 3) times at most three factors 3 and 5, up to 2048 (ND, DN:
 * the doubled length);
 * n1 even; z non-periodic.  The fields
 * are in the context's pressure storage type (float under CANS_MIXED) with element
 * (1,0,0) on an element-pair boundary.  normfft (nbox, may be NULL) receives each
 * box's normfft.  The batch keeps the field addresses given at creation. */
typedef struct cans_poisson_batch cans_poisson_batch;
typedef struct cans_poisson_box {
  int n[3];                   /* cells */
  long sy, sz;                /* strides of p in elements; 0: n1+2 and sy*(n2+2) */
  char cbc[6];                /* cbcpre(0:1,1:3) of the box */
  const double *lambdaxy_h;   /* (n1,n2), device ordering (cans_eigenvalues) */
  const double *a_h, *b_h, *c_h; /* (n3) (cans_tridmatrix) */
  void *p_d;                  /* &p(0,0,0): rhs in, solution out */
} cans_poisson_box;
int cans_poisson_batch_create(cans_ctx *ctx, int nbox, const cans_poisson_box *boxes,
                              double *normfft, cans_poisson_batch **batch);
int cans_poisson_batch_solve(cans_poisson_batch *batch);
/* lanes per work-group of the batch's x and y passes (128, 256 or 512), chosen at
 * creation by timing the forward passes (no counterpart in the reference) */
int cans_poisson_batch_tuning(const cans_poisson_batch *batch, int *threads);
int cans_poisson_batch_destroy(cans_poisson_batch *batch);
/* Poisson solve with iterative refinement (no counterpart in the reference): p_d is an
 * FP64 halo'd field holding b on entry and x with L x = b on return, whatever the
 * context's precision.  The inner solves (src/solver.f90:17, same arguments) run in the
 * context's pressure storage type -- FP32 transforms and storage under CANS_MIXED --
 * and the residuals b - L x and the sum x in FP64: x = S(b), then `sweeps` times
 * x += S(b - L x), L the solver's 7-point operator with homogeneous cell-centred
 * ghost values (c_or_f "ccc").  dli = (1/dx, 1/dy); dzci_d, dzfi_d (0:n3+1). */
int cans_solver_refined(cans_ctx *ctx, const int *n, const int *ng, const cans_plan *arrplan,
                        double normfft, const double *lambdaxy_d, const double *a_d, const double *b_d,
                        const double *c_d, const char *bc, const char *c_or_f, const double *dli,
                        const double *dzci_d, const double *dzfi_d, int sweeps, void *p_d);
/* src/solve_helmholtz.f90:28 (3-D branch) -- rhsb*_d may be NULL */
int cans_solve_helmholtz(cans_ctx *ctx, const int *n, const int *ng, const cans_plan *arrplan,
                         double normfft, double alpha, const double *lambdaxy_d,
                         const double *a_d, const double *b_d, const double *c_d,
                         const double *rhsbx_d, const double *rhsby_d, const double *rhsbz_d,
                         const int *is_bound, const char *cbc, const char *c_or_f, void *p_d);
/* src/rk.f90:24 (explicit diffusion) -- s_d is the first scalar's halo'd FP64 field
 * when Boussinesq buoyancy is on (rk.f90:69-72, 163-177; the reference passes the
 * scalars array), NULL otherwise (gacc and beta are then not read).  f_h (3)
 * receives the bulk forcing like the reference's intent(out) f; pass NULL to keep
 * it on the device (no host synchronisation) */
int cans_rk(cans_ctx *ctx, const double *rkpar, const int *n, const double *dli,
            const double *dzci_d, const double *dzfi_d, const double *grid_vol_ratio_c_d,
            const double *grid_vol_ratio_f_d, double dt, double visc, const void *p_d,
            const int *is_forced, const double *velf, const double *bforce,
            const double *gacc, double beta, const void *s_d,
            void *dudtrko_d, void *dvdtrko_d, void *dwdtrko_d,
            void *u_d, void *v_d, void *w_d, double *f_h);
/* src/rk.f90:335 (rk_scal), explicit diffusion.  l, is_bound and fluxo serve the
 * wall-flux branch, which the reference compiles out (is_cmpt_wallflux is a .false.
 * parameter there); they are accepted and not read.  The scalar and dsdtrko are
 * FP64 fields in both precision modes (cans_field_alloc_kind with
 * CANS_FIELD_VELOCITY; dsdtrko holds n1*n2*n3 doubles).  *f_h (if not NULL)
 * receives the bulk forcing f = scalf - <s> when is_forced, else 0. */
int cans_rk_scal(cans_ctx *ctx, const double *rkpar, const int *n, const double *dli,
                 const double *l, const double *dzci_d, const double *dzfi_d,
                 const double *grid_vol_ratio_f_d, double alpha, double dt, const int *is_bound,
                 const void *u_d, const void *v_d, const void *w_d, int is_forced, double scalf,
                 double ssource, double *fluxo, void *dsdtrko_d, void *s_d, double *f_h);
/* src/scal.f90:313 (bulk_forcing_s): s(1:n) += ff when is_forced */
int cans_bulk_forcing_s(cans_ctx *ctx, const int *n, int is_forced, double ff, void *s_d);
/* src/fillps.f90:13 */
int cans_fillps(cans_ctx *ctx, const int *n, const double *dli, const double *dzfi_d, double dti,
                const void *u_d, const void *v_d, const void *w_d, void *p_d);
/* src/correc.f90:13 */
int cans_correc(cans_ctx *ctx, const int *n, const double *dli, const double *dzci_d, double dt,
                const void *p_d, void *u_d, void *v_d, void *w_d);
/* src/updatep.f90:14 (is_impdiff adds alpha*lap(pp)) */
int cans_updatep(cans_ctx *ctx, const int *n, const double *dli, const double *dzci_d,
                 const double *dzfi_d, double alpha, const void *pp_d, void *p_d, int is_impdiff);
/* src/bound.f90:17 -- keep_norm < 0 and halos_* == CANS_HALO_ABSENT mean "argument absent" */
int cans_bounduvw(cans_ctx *ctx, const char *cbc, const int *n, const double *bc, const int *nb,
                  const int *is_bound, const double *dl, const double *dzc_h, const double *dzf_h,
                  void *u_d, void *v_d, void *w_d, int keep_norm, int halos_norm, int halos_tang);
/* src/bound.f90:124 */
int cans_boundp(cans_ctx *ctx, const char *cbc, const int *n, const double *bc, const int *nb,
                const int *is_bound, const double *dl, const double *dzc_h, void *p_d, int halo_side);
/* src/bound.f90:514 -- rhsb*_d are (n2,n3,0:1), (n1,n3,0:1), (n1,n2,0:1) or NULL; alpha_h NULL = absent */
int cans_updt_rhs_b(cans_ctx *ctx, const char *c_or_f, const char *cbc, const int *n,
                    const int *is_bound, const double *rhsbx_d, const double *rhsby_d,
                    const double *rhsbz_d, void *p_d, const double *alpha_h);
/* src/chkdiv.f90:15 (synchronises) */
int cans_chkdiv(cans_ctx *ctx, const int *lo, const int *hi, const double *l, const double *dli,
                const double *dzfi_d, const void *u_d, const void *v_d, const void *w_d,
                double *divtot_h, double *divmax_h);
/* src/chkdt.f90:16 (synchronises) */
int cans_chkdt(cans_ctx *ctx, const int *n, const double *dl, const double *dzci_d,
               const double *dzfi_d, double visc, double alpha, const void *u_d, const void *v_d,
               const void *w_d, int is_impdiff, int is_impdiff_1d, double *dtmax_h);
/* src/utils.f90:15 (synchronises) */
int cans_bulk_mean(cans_ctx *ctx, const int *n, const double *grid_vol_ratio_d, const void *p_d,
                   double *mean_h);

/* ---------------------------------------------------------------------------
 * Fused RK3 substep: the whole call sequence of src/main.f90:452-472 (no
 * scalars, explicit diffusion) in one stream-ordered call with no host
 * synchronisation.  Same results as calling the operators above in order.
 * ------------------------------------------------------------------------- */
/* one implicit-diffusion system: what initsolver returns for a variable (initsolver.f90:15-83) */
typedef struct cans_helmholtz {
  const cans_plan *plan;                       /* NULL with is_impdiff_1d */
  double normfft;
  const double *lambdaxy_d, *a_d, *b_d, *c_d;
  const double *rhsbx_d, *rhsby_d, *rhsbz_d;   /* boundary rhs planes, NULL if absent */
  char cbc[6];                                 /* the variable's cbc(0:1,1:3) */
  char c_or_f[3];
} cans_helmholtz;

/* No counterpart in the reference, whose main.f90:453-462 (and 440-450 for the scalars) calls
 * solve_helmholtz once per system: the nsys systems h[q], each with its alpha[q], on the FP64
 * fields f_d[q] (halo'd, rhs in, solution out), each as cans_solve_helmholtz would solve it, as
 * one batch -- every system's boundary rhs, then one launch per transform pass for up to four
 * systems together (more run in groups of four), also when their x and y transforms differ (the
 * velocity components next to walls); with is_impdiff_1d one z launch for all.  Systems with a
 * periodic z, or under the distributed z solve, are solved one after another. */
int cans_solve_helmholtz_batch(cans_ctx *ctx, int nsys, const cans_helmholtz *h, const double *alpha,
                               const int *is_bound, void *const *f_d);

typedef struct cans_step_params {
  char cbcvel[18];
  double bcvel[18];
  char cbcpre[6];
  double bcpre[6];
  int nb[6];
  int is_bound[6];
  double dl[3], dli[3];
  double visc;
  int is_forced[3];
  double velf[3];
  double bforce[3];
  const double *dzc_h, *dzf_h;                 /* (0:n3+1) host copies */
  const double *dzci_d, *dzfi_d;               /* (0:n3+1) device copies */
  const double *grid_vol_ratio_c_d, *grid_vol_ratio_f_d;
  const cans_plan *plan;                       /* pressure plan from cans_fftini */
  double normfft;
  const double *lambdaxy_d, *a_d, *b_d, *c_d;  /* pressure solver arrays */
  const double *rhsbx_d, *rhsby_d, *rhsbz_d;   /* pressure BC rhs, NULL if homogeneous */
  /* implicit diffusion (cans_ctx_set_numerics): the Helmholtz systems of u, v, w
   * (initsolver with c_or_f 'fcc', 'cfc', 'ccf', main.f90:304-309) */
  cans_helmholtz hvel[3];
  /* scalars (main.f90:440-450, at most CANS_MAX_SCALARS) and the Boussinesq buoyancy
   * of scal[0] (rk.f90:69-72); nscal = 0 and is_buoyancy = 0 give the velocity-only step */
  int nscal;
  const struct cans_scalar *scal;
  int is_buoyancy;
  double gacc[3], beta;
} cans_step_params;

/* the members of src/scal.f90:17-41's scalar type that the step uses */
typedef struct cans_scalar {
  void *val, *dsdtrko;                         /* FP64: halo'd field; (1:n1,1:n2,1:n3) */
  double alpha;                                /* diffusivity 1/alphai */
  char cbc[6];                                 /* cbcscal(0:1,1:3) */
  double bc[6];                                /* bcscal(0:1,1:3) */
  double source;                               /* ssource */
  int is_forced;                               /* is_sforced */
  double scalf;                                /* scalf */
  cans_helmholtz h;                            /* implicit diffusion: initsolver(..., 'ccc', ...) */
} cans_scalar;
#define CANS_MAX_SCALARS 32

typedef struct cans_state {
  void *u, *v, *w, *p, *pp;                    /* halo'd fields */
  void *dudtrko, *dvdtrko, *dwdtrko;           /* (1:n1,1:n2,1:n3) */
} cans_state;

/* irk in 1..3; rkcoeff as src/param.f90:20-22.  The fused path skips work whose
 * result is never read: substep 1 does not read dudtrko (its weight rkcoeff(2,1) is
 * 0) and substep 3 does not write it (the next step's substep 1 overwrites it before
 * any use), so after substep 3 dudtrko still holds substep 2's right-hand side. */
int cans_rk3_substep(cans_ctx *ctx, const cans_step_params *prm, int irk, double dt, cans_state *st);
/* One whole RK3 step, the loop body of src/main.f90:437-473 (do irk = 1,3): the three
 * substeps of cans_rk3_substep with the same results, bit for bit.  Where diffusion is
 * explicit, there are no scalars, the pressure boundary values are homogeneous and
 * substeps 1 and 2 are not refined, each substep's projection (correc, bounduvw, updatep,
 * boundp) is folded into the next substep's momentum kernel, which forms the projected
 * velocity and p + pp while it stages them, so substeps 1 and 2 write no corrected
 * velocity; the context keeps that state between the substeps in its own scratch. */
int cans_rk3_step(cans_ctx *ctx, const cans_step_params *prm, double dt, cans_state *st);
/* forcing of the last substep, device pointer to 3 doubles (f(1:3) of rk.f90) */
int cans_ctx_forcing(cans_ctx *ctx, const double **f_d);

/* ---------------------------------------------------------------------------
 * Adjoint of the RK3 substep (no counterpart in the reference; for gradients of a
 * loss on the velocity).  Scope: CANS_FP64 and CANS_MIXED contexts (under CANS_MIXED
 * the adjoint's projection runs in FP32 pressure storage, as the forward's does; pp_d
 * is then an FP32 field), no scalars or buoyancy; explicit
 * diffusion, or 3-D implicit diffusion on tri-periodic uniform grids (where the
 * substep's Helmholtz operator H = I - visc dtrk/2 L commutes with the gradient and is
 * symmetric: the adjoint applies H^-1 after P^T, the advection's transposed Jacobian,
 * and the diffusion's explicit half dtrk/2 visc L); each direction periodic or bounded by walls (Neumann pressure,
 * Dirichlet normal velocity, Dirichlet or Neumann tangential velocity, any wall
 * values), a periodic z on a uniform grid and walls in z on any grid; bulk forcing of
 * u and v along periodic x and y.  There substep irk maps the degrees of freedom
 * (every interior value except a normal velocity's wall face: u at i = n1, v at
 * j = n2, w at k = n3) as
 * (u, Fo) -> (P Fr(u + f1 F(B u) + f2 Fo), F(B u)), F the right-hand side of
 * mom_xyz_ad (src/mom.f90:771-1019) at the velocity whose ghost cells bound.f90 sets
 * (B), f1, f2 = rkcoeff(1:2,irk)*dt, Fr the bulk forcing (rk.f90:493-518) and P the
 * projection of fillps/solver/correc; the pressure drops out (P G p = 0).  P is
 * self-adjoint in the inner product weighted by the cells' dz (dzf for u, v; dzc for
 * w), so its transpose is P itself conjugated by those weights, whatever transforms
 * (FFT, DCT) the solve uses along x and y.
 * ------------------------------------------------------------------------- */
/* out = base + J(q)^T g: J the Jacobian of mom_xyz_ad's right-hand side at the
 * velocity q = (u,v,w) on a tri-periodic uniform grid; dli = (1/dx, 1/dy, 1/dz).  q
 * and g are halo'd with their periodic ghost cells set; out may be base but no q or g
 * field. */
int cans_mom_jt(cans_ctx *ctx, const int *n, const double *dli, double visc, const void *u_d,
                const void *v_d, const void *w_d, const void *gu_d, const void *gv_d, const void *gw_d,
                const void *bu_d, const void *bv_d, const void *bw_d, void *ou_d, void *ov_d, void *ow_d);
/* One substep backwards: u_d, v_d, w_d is the velocity the forward substep irk started
 * from (halo'd, ghosts set); (ub, vb, wb) holds the adjoint of the substep's output
 * velocity, (fu, fv, fw) that of the right-hand side this substep stored for the next
 * one (zero after substep 3).  (ubo, vbo, wbo) and (fuo, fvo, fwo) receive the adjoints of
 * the substep's input velocity and of the right-hand side it read; ub..wb are
 * overwritten (they receive Fr'^T P^T ub).  Values on the normal wall faces are
 * ignored on input and zero on output.  pp_d is a halo'd scratch field; all sixteen fields are
 * distinct.  prm as for cans_rk3_substep (plan, lambdaxy, grid arrays, is_forced,
 * grid_vol_ratio_f). */
int cans_rk3_substep_adjoint(cans_ctx *ctx, const cans_step_params *prm, int irk, double dt,
                             const void *u_d, const void *v_d, const void *w_d, void *ub_d,
                             void *vb_d, void *wb_d, void *fu_d, void *fv_d, void *fw_d,
                             void *ubo_d, void *vbo_d, void *wbo_d, void *fuo_d, void *fvo_d,
                             void *fwo_d, void *pp_d);
/* The same with scalars (prm->nscal > 0; cans_rk3_substep_adjoint refuses them): explicit
 * diffusion; scalars bulk-forced or not, periodic where the box is and Dirichlet or
 * Neumann at its walls.  s_d[i] is scalar i as the forward substep started from (halo'd, ghosts
 * set); sb_d[i] and fsb_d[i] hold the adjoints of its output and of the right-hand side
 * it stored (sb_d[i] is overwritten); sbo_d[i] and fsbo_d[i] receive those of its input
 * and of the right-hand side it read.  The first scalar's adjoint also takes the
 * Boussinesq term's transpose (rk.f90:163-177) when prm->is_buoyancy.  Every field is
 * distinct. */
int cans_rk3_substep_adjoint_scal(cans_ctx *ctx, const cans_step_params *prm, int irk, double dt,
                                  const void *u_d, const void *v_d, const void *w_d, void *ub_d,
                                  void *vb_d, void *wb_d, void *fu_d, void *fv_d, void *fw_d,
                                  void *ubo_d, void *vbo_d, void *wbo_d, void *fuo_d, void *fvo_d,
                                  void *fwo_d, void *pp_d, const void *const *s_d, void *const *sb_d,
                                  void *const *fsb_d, void *const *sbo_d, void *const *fsbo_d);

#ifdef __cplusplus
}
#endif
#endif
