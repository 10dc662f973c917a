/*
 * CPU restatement of CaNS's FP64 pressure-correction time step (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path and
 * the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links or calls it.
 *
 * Conventions (mirroring the reference's Fortran arrays):
 *   - halo'd fields are (0:n1+1, 0:n2+1, 0:n3+1), column-major, contiguous;
 *   - z-spacing arrays are (0:n3+1);
 *   - boundary-condition characters are the flattened Fortran arrays, e.g.
 *     cbcpre(0:1,3) -> 6 chars [x0,x1,y0,y1,z0,z1], cbcvel(0:1,3,3) -> 18
 *     chars (component is the slowest index);
 *   - logical arrays are passed as int (0/1).
 * Every function cites the reference routine it restates.
 */
#ifndef CANS_ORACLE_H
#define CANS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* FFTW r2r kinds (numbering is FFTW's own public enum fftw_r2r_kind). */
enum {
  ORC_R2HC = 0, ORC_HC2R = 1, ORC_DHT = 2,
  ORC_REDFT00 = 3, ORC_REDFT01 = 4, ORC_REDFT10 = 5, ORC_REDFT11 = 6,
  ORC_RODFT00 = 7, ORC_RODFT01 = 8, ORC_RODFT10 = 9, ORC_RODFT11 = 10
};

/* grid: src/initgrid.f90:14-99 (+ stretching functions 105-200) */
void orc_initgrid(int gtype, int n, double gr, double lz, int is_periodic,
                  double *dzc, double *dzf, double *zc, double *zf);
void orc_initgrid_natural(int n, double lz, int is_periodic, double *dzc, double *dzf, double *zc, double *zf);
/* src/initsolver.f90:85-144 (CPU ordering, i.e. FFTW halfcomplex for PP) */
void orc_eigenvalues(int n, const char *bc, char c_or_f, double *lambda);
/* src/initsolver.f90:146-187 */
void orc_tridmatrix(const char *bc, int n, const double *dzci, const double *dzfi,
                    char c_or_f, double *a, double *b, double *c);
/* src/initsolver.f90:189-232; returns the two (constant) plane values */
void orc_bc_rhs(const char *cbc, const double *bc, const double *dlc,
                const double *dlf, char c_or_f, double *factor_out);
/* src/fft.f90:260-313 */
int orc_find_fft(const char *bc, char c_or_f, int *kind_fwd, int *kind_bwd, double *norm);
/* src/fft.f90:25-209 (kinds, transform lengths and normfft only) */
int orc_fftini(const int *ng, const char *bcxy, const char *c_or_f,
               int *kinds, int *lens, double *normfft);

/* batched in-place r2r transform with FFTW's unnormalised definitions:
 * howmany lines of length n; element e of line l at data[l*dist + e*stride] */
/* 0 (default) or 1: an alternative factorisation / real-FFT route with the same
 * definitions and different rounding, used to measure the algorithm's own spread */
void orc_set_fft_variant(int v);
int orc_r2r(int kind, int n, long howmany, long stride, long dist, double *data);

/* src/solver.f90:114-307 on a compact (nx,ny,>=n) array */
void orc_gaussel(int nx, int ny, int n, const double *a, const double *b, const double *c,
                 int is_periodic, double norm, double *p, const double *lambdaxy);
/* src/solver.f90:17-112 (single rank: transposes are identities) */
int orc_solver(const int *n, const int *kinds, const int *lens, double normfft,
               const double *lambdaxy, const double *a, const double *b, const double *c,
               const char *bc, const char *c_or_f, double *p);
/* src/solve_helmholtz.f90:28-75 (3-D branch) */
int orc_solve_helmholtz(const int *n, const int *kinds, const int *lens, double normfft,
                        double alpha, const double *lambdaxy, const double *a,
                        const double *b, const double *c, const double *rhsbx,
                        const double *rhsby, const double *rhsbz, const int *is_bound,
                        const char *cbc, const char *c_or_f, double *p);

/* src/solver.f90:309-517 (gaussel_dtdma): z rows split into nslab slabs; periodic z closes
   the reduced system cyclically (nslab >= 2) */
void orc_slab_range(int n, int nslab, int s, int *lo, int *cnt);
int orc_gaussel_dtdma(int nx, int ny, int n, int nslab, const double *a, const double *b,
                      const double *c, double norm, double *p, const double *lambdaxy, int is_periodic);
/* src/solver.f90:547-616 (one rank): z-only elimination of the halo'd field p */
int orc_solver_gaussel_z(const int *n, const double *a, const double *b, const double *c,
                         const char *bcz, char c_or_f_z, double norm, double *p);
/* src/solve_helmholtz.f90:28-75, is_impdiff_1d branch */
int orc_solve_helmholtz_1d(const int *n, double alpha, const double *a, const double *b,
                           const double *c, const double *rhsbz, const int *is_bound,
                           const char *cbc, const char *c_or_f, double *p);
/* src/mom.f90:771-1019 */
void orc_mom_xyz_ad(const int *n, double dxi, double dyi, const double *dzci,
                    const double *dzfi, double visc, const double *u, const double *v,
                    const double *w, double *dudt, double *dvdt, double *dwdt,
                    double *dudtd, double *dvdtd, double *dwdtd,
                    int is_impdiff, int is_impdiff_1d);
/* src/rk.f90:24-333 (+ cmpt_bulk_forcing 493-518, bulk_forcing mom.f90:725-769);
 * s = NULL: no buoyancy, else the Boussinesq term -gacc*beta*s of the first scalar;
 * work holds 3*n1*n2*n3 doubles (6* with implicit diffusion) */
void orc_rk(const double *rkpar, const int *n, const double *dli, const double *dzci,
            const double *dzfi, const double *grid_vol_ratio_c,
            const double *grid_vol_ratio_f, double dt, double visc, const double *p,
            const int *is_forced, const double *velf, const double *bforce,
            const double *gacc, double beta, const double *s,
            double *dudtrko, double *dvdtrko, double *dwdtrko,
            double *u, double *v, double *w, double *f, double *work,
            int is_impdiff, int is_impdiff_1d);
/* src/scal.f90:41-98: dsdt (and dsdtd with implicit diffusion) of scalar s */
void orc_scal(const int *n, double dxi, double dyi, const double *dzci, const double *dzfi,
              double visc, const double *u, const double *v, const double *w, const double *s,
              double *dsdt, double *dsdtd, int is_impdiff, int is_impdiff_1d);
/* src/rk.f90:335-491; work holds n1*n2*n3 doubles (2* with implicit diffusion) */
void orc_rk_scal(const double *rkpar, const int *n, const double *dli, const double *dzci,
                 const double *dzfi, const double *grid_vol_ratio_f, double alpha, double dt,
                 const double *u, const double *v, const double *w, int is_forced, double scalf,
                 double ssource, double *dsdtrko, double *s, double *f, double *work,
                 int is_impdiff, int is_impdiff_1d);
/* src/scal.f90:313-329 */
void orc_bulk_forcing_s(const int *n, int is_forced, double ff, double *p);
/* src/utils.f90:15-42 */
double orc_bulk_mean(const int *n, const double *grid_vol_ratio, const double *p);
/* src/fillps.f90:13-51 */
void orc_fillps(const int *n, const double *dli, const double *dzfi, double dti,
                const double *u, const double *v, const double *w, double *p);
/* src/correc.f90:13-61 */
void orc_correc(const int *n, const double *dli, const double *dzci, double dt,
                const double *p, double *u, double *v, double *w);
/* src/updatep.f90:14-88 */
void orc_updatep(const int *n, const double *dli, const double *dzci, const double *dzfi,
                 double alpha, const double *pp, double *p, int is_impdiff, int is_impdiff_1d);
/* src/bound.f90:17-122; keep_norm < 0 / halos_* == -99 mean "argument absent" */
void orc_bounduvw(const char *cbc, const int *n, const double *bc, const int *is_bound,
                  const int *is_periodic, const double *dl, const double *dzc,
                  const double *dzf, double *u, double *v, double *w,
                  int keep_norm, int halos_norm, int halos_tang);
/* src/bound.f90:124-180 */
void orc_boundp(const char *cbc, const int *n, const double *bc, const int *is_bound,
                const int *is_periodic, const double *dl, const double *dzc, double *p,
                int halo_side);
/* src/bound.f90:514-598; rhsb* are (n2,n3,0:1), (n1,n3,0:1), (n1,n2,0:1) or NULL */
void orc_updt_rhs_b(const char *c_or_f, const char *cbc, const int *n, const int *is_bound,
                    const double *rhsbx, const double *rhsby, const double *rhsbz,
                    double *p, double alpha, int has_alpha);
/* src/chkdiv.f90:15-54 */
void orc_chkdiv(const int *n, const double *l, const double *dli, const double *dzfi,
                const double *u, const double *v, const double *w,
                double *divtot, double *divmax);
/* src/chkdt.f90:16-78 */
double orc_chkdt(const int *n, const double *dl, const double *dzci, const double *dzfi,
                 double visc, double alpha, const double *u, const double *v,
                 const double *w, int is_impdiff, int is_impdiff_1d);

int orc_num_threads(void);
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
