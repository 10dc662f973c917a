/*
 * Oracle: CPU restatement of CaNS's FP64 pressure-correction time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP path and the
 * timed CPU baseline; the product library never links or calls it.
 *
 * Each routine follows the reference routine cited in its comment, with the
 * same loop order and the same floating-point association.  Fortran default
 * REAL literals in the reference are single precision (no -fdefault-real-8 in
 * configs/defaults/flags-default.mk); where such a literal or a default-real
 * expression is not exactly representable its float rounding is reproduced.
 */
#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "cans_oracle.h"

#define PI 3.14159265358979323846
#define EPS DBL_EPSILON

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* halo'd (0:n1+1,0:n2+1,0:n3+1) and interior (1:n1,1:n2,1:n3) index maps */
#define HX(i, j, k) ((size_t)(i) + s1 * ((size_t)(j) + s2 * (size_t)(k)))
#define IX(i, j, k) ((size_t)((i) - 1) + (size_t)n1 * ((size_t)((j) - 1) + (size_t)n2 * (size_t)((k) - 1)))
#define DIMS                                                         \
  const int n1 = n[0], n2 = n[1], n3 = n[2];                         \
  const size_t s1 = (size_t)n1 + 2, s2 = (size_t)n2 + 2;             \
  (void)n1; (void)n2; (void)n3; (void)s1; (void)s2;

/* ------------------------------------------------------------------------ */
/* grid: src/initgrid.f90                                                    */
/* ------------------------------------------------------------------------ */
static double gp_two_end(double alpha, double z0) {   /* initgrid.f90:105-118 */
  if (alpha > EPS) return 0.5 * (1. + tanh((z0 - 0.5) * alpha) / tanh(alpha / 2.));
  return z0;
}
static double gp_one_end(double alpha, double z0) {   /* initgrid.f90:119-131 */
  if (alpha > EPS) return 1.0 * (1. + tanh((z0 - 1.0) * alpha) / tanh(alpha / 1.));
  return z0;
}
static double gp_one_end_r(double alpha, double r0) { /* initgrid.f90:132-144 */
  if (alpha > EPS) return 1. - 1.0 * (1. + tanh((1. - r0 - 1.0) * alpha) / tanh(alpha / 1.));
  return r0;
}
static double gp_middle(double alpha, double z0) {    /* initgrid.f90:145-161 */
  if (alpha > EPS) {
    if (z0 <= 0.5) return 0.5 * (1. - 1. + tanh(2. * alpha * (z0 - 0.)) / tanh(alpha));
    return 0.5 * (1. + 1. + tanh(2. * alpha * (z0 - 1.)) / tanh(alpha));
  }
  return z0;
}

/* initgrid.f90:162-203 (gridpoint_natural, Pirozzoli & Orlandi, JCP 439, 2021): kb = 32,
   alpha = pi/1.5, c_eta = 0.8, dyp = 0.05; everything in double precision (_rp literals) */
static double gp_natural(int kg, int nzg) {
  const double pi = acos(-1.);
  const double kb = 32., alpha = pi / 1.5, c_eta = 0.8, dyp = 0.05;
  const double n = nzg / 2.;
  const double retau = 1. / (1. + (n / kb) * (n / kb)) *
                       (dyp * n + pow(3. / 4. * alpha * c_eta * n, 4. / 3.) * ((n / kb) * (n / kb)));
  const double k = 1. * (kg < nzg - kg ? kg : nzg - kg);
  double z = 1. / (1. + (k / kb) * (k / kb)) * (dyp * k + pow(3. / 4. * alpha * c_eta * k, 4. / 3.) * ((k / kb) * (k / kb))) /
             (2. * retau);
  if (kg > nzg - kg) z = 1. - z;
  return z;
}

static void initgrid(int gtype, int n, double gr, double lz, int is_periodic, int natural,
                     double *dzc, double *dzf, double *zc, double *zf) {
  /* initgrid.f90:14-99 */
  double (*gp)(double, double) = gp_two_end;
  if (gtype == 2) gp = gp_one_end;
  else if (gtype == 3) gp = gp_one_end_r;
  else if (gtype == 4) gp = gp_middle;
  zf[0] = 0.;
  for (int k = 1; k <= n; k++) {
    if (natural) { zf[k] = gp_natural(k, n); continue; }
    /* z0 = (k-0.)/(1.*n) is a default-real (single precision) expression */
    float z0f = ((float)k - 0.f) / (1.f * (float)n);
    zf[k] = gp(gr, (double)z0f);
  }
  for (int k = 1; k <= n; k++) zf[k] = zf[k] * lz;
  if (fabs(gr) < EPS && !natural) {
    for (int k = 0; k <= n + 1; k++) { dzf[k] = lz / (double)(1.f * (float)n); dzc[k] = dzf[k]; }
  } else {
    for (int k = 1; k <= n; k++) dzf[k] = zf[k] - zf[k - 1];
    if (!is_periodic) { dzf[0] = dzf[1]; dzf[n + 1] = dzf[n]; }
    else { dzf[0] = dzf[n]; dzf[n + 1] = dzf[1]; }
    for (int k = 0; k <= n; k++) dzc[k] = .5 * (dzf[k] + dzf[k + 1]);
    if (!is_periodic) dzc[n + 1] = dzc[n]; else dzc[n + 1] = dzc[1];
  }
  zc[0] = -dzc[0] / 2.;
  zf[0] = 0.;
  for (int k = 1; k <= n + 1; k++) { zc[k] = zc[k - 1] + dzc[k - 1]; zf[k] = zf[k - 1] + dzf[k]; }
}

void orc_initgrid(int gtype, int n, double gr, double lz, int is_periodic,
                  double *dzc, double *dzf, double *zc, double *zf) {
  initgrid(gtype, n, gr, lz, is_periodic, 0, dzc, dzf, zc, zf);
}

/* is_gridpoint_natural_channel = .true. (param.f90:99, initgrid.f90:46-58) */
void orc_initgrid_natural(int n, double lz, int is_periodic, double *dzc, double *dzf, double *zc, double *zf) {
  initgrid(1, n, 0., lz, is_periodic, 1, dzc, dzf, zc, zf);
}

/* ------------------------------------------------------------------------ */
/* solver setup: src/initsolver.f90                                          */
/* ------------------------------------------------------------------------ */
void orc_eigenvalues(int n, const char *bc, char c_or_f, double *lambda) {
  /* initsolver.f90:85-144, CPU (FFTW halfcomplex) ordering */
  const double pi = PI;
  float fn = 1.f * (float)n;  /* (1.*n) is default real */
  if (bc[0] == 'P' && bc[1] == 'P') {
    for (int l = 1; l <= n; l++) lambda[l - 1] = -2. * (1. - cos((2 * (l - 1)) * pi / (double)fn));
  } else if (bc[0] == 'N' && bc[1] == 'N') {
    if (c_or_f == 'c') for (int l = 1; l <= n; l++) lambda[l - 1] = -2. * (1. - cos((l - 1) * pi / (double)fn));
    else {
      float fm = 1.f * (float)(n - 1 + 1);
      for (int l = 1; l <= n; l++) lambda[l - 1] = -2. * (1. - cos((l - 1) * pi / (double)fm));
    }
  } else if (bc[0] == 'D' && bc[1] == 'D') {
    if (c_or_f == 'c') for (int l = 1; l <= n; l++) lambda[l - 1] = -2. * (1. - cos((l)*pi / (double)fn));
    else {
      float fm = 1.f * (float)(n + 1 - 1);
      for (int l = 1; l <= n - 1; l++) lambda[l - 1] = -2. * (1. - cos((l)*pi / (double)fm));
      lambda[n - 1] = 0.;
    }
  } else {
    float f2n = 2.f * (float)n;
    for (int l = 1; l <= n; l++) lambda[l - 1] = -2. * (1. - cos((2 * l - 1) * pi / (double)f2n));
  }
}

void orc_tridmatrix(const char *bc, int n, const double *dzci, const double *dzfi,
                    char c_or_f, double *a, double *b, double *c) {
  /* initsolver.f90:146-187; a,b,c are 1-based in the reference, 0-based here */
  if (c_or_f == 'c') {
    for (int k = 1; k <= n; k++) { a[k - 1] = dzfi[k] * dzci[k - 1]; c[k - 1] = dzfi[k] * dzci[k]; }
  } else {
    for (int k = 1; k <= n; k++) { a[k - 1] = dzfi[k] * dzci[k]; c[k - 1] = dzfi[k + 1] * dzci[k]; }
  }
  for (int k = 0; k < n; k++) b[k] = -(a[k] + c[k]);
  double factor[2];
  for (int ib = 0; ib < 2; ib++) factor[ib] = bc[ib] == 'P' ? 0. : bc[ib] == 'D' ? -1. : 1.;
  if (c_or_f == 'c') {
    b[0] = b[0] + factor[0] * a[0];
    b[n - 1] = b[n - 1] + factor[1] * c[n - 1];
  } else {
    if (bc[0] == 'N') b[0] = b[0] + factor[0] * a[0];
    if (bc[1] == 'N') b[n - 1] = b[n - 1] + factor[1] * c[n - 1];
  }
}

void orc_bc_rhs(const char *cbc, const double *bc, const double *dlc, const double *dlf,
                char c_or_f, double *out) {
  /* initsolver.f90:189-232: every plane value of rhs(:,:,ib) is factor/dlc/dlf */
  double factor[2], sgn = 0;
  for (int ib = 0; ib < 2; ib++) {
    if (cbc[ib] == 'P') factor[ib] = 0.;
    else if (cbc[ib] == 'D') factor[ib] = (c_or_f == 'c') ? -2. * bc[ib] : -bc[ib];
    else {
      sgn = ib == 0 ? 1. : -1.;
      factor[ib] = sgn * ((c_or_f == 'c') ? dlc[ib] : dlf[ib]) * bc[ib];
    }
    out[ib] = factor[ib] / dlc[ib] / dlf[ib];
  }
}

int orc_find_fft(const char *bc, char c_or_f, int *kind_fwd, int *kind_bwd, double *norm) {
  /* fft.f90:260-313 */
  char a = bc[0], b = bc[1];
  int pp = a == 'P' && b == 'P', nn = a == 'N' && b == 'N', dd = a == 'D' && b == 'D';
  int nd = a == 'N' && b == 'D', dn = a == 'D' && b == 'N';
  if (c_or_f == 'c') {
    if (pp) { *kind_fwd = ORC_R2HC; *kind_bwd = ORC_HC2R; norm[0] = 1; norm[1] = 0; }
    else if (nn) { *kind_fwd = ORC_REDFT10; *kind_bwd = ORC_REDFT01; norm[0] = 2; norm[1] = 0; }
    else if (dd) { *kind_fwd = ORC_RODFT10; *kind_bwd = ORC_RODFT01; norm[0] = 2; norm[1] = 0; }
    else if (nd) { *kind_fwd = ORC_REDFT11; *kind_bwd = ORC_REDFT11; norm[0] = 2; norm[1] = 0; }
    else if (dn) { *kind_fwd = ORC_RODFT11; *kind_bwd = ORC_RODFT11; norm[0] = 2; norm[1] = 0; }
    else return -1;
  } else {
    if (pp) { *kind_fwd = ORC_R2HC; *kind_bwd = ORC_HC2R; norm[0] = 1; norm[1] = 0; }
    else if (nn) { *kind_fwd = ORC_REDFT00; *kind_bwd = ORC_REDFT00; norm[0] = 2; norm[1] = -1; }
    else if (dd) { *kind_fwd = ORC_RODFT00; *kind_bwd = ORC_RODFT00; norm[0] = 2; norm[1] = 1; }
    else if (nd) { *kind_fwd = ORC_REDFT10; *kind_bwd = ORC_REDFT01; norm[0] = 2; norm[1] = 0; }
    else if (dn) { *kind_fwd = ORC_RODFT01; *kind_bwd = ORC_RODFT10; norm[0] = 2; norm[1] = 0; }
    else return -1;
  }
  return 0;
}

int orc_fftini(const int *ng, const char *bcxy, const char *c_or_f,
               int *kinds, int *lens, double *normfft) {
  /* fft.f90:25-209: kinds = [fwd x, bwd x, fwd y, bwd y]; lens = transform lengths */
  double nf = 1.;
  for (int d = 0; d < 2; d++) {
    double norm[2];
    if (orc_find_fft(bcxy + 2 * d, c_or_f[d], &kinds[2 * d], &kinds[2 * d + 1], norm)) return -1;
    int ix = (bcxy[2 * d] == 'D' && bcxy[2 * d + 1] == 'D' && c_or_f[d] == 'f') ? 1 : 0;
    lens[d] = ng[d] - ix;
    nf = nf * norm[0] * (ng[d] + norm[1] - ix);
  }
  *normfft = 1. / nf;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* tridiagonal solve: src/solver.f90:114-307                                 */
/* ------------------------------------------------------------------------ */
void orc_gaussel(int nx, int ny, int n, const double *a, const double *b, const double *c,
                 int is_periodic, double norm, double *p, const double *lambdaxy) {
  const size_t pl = (size_t)nx * ny;
  int nn = is_periodic ? n - 1 : n;
  /* lines are independent: line-major loop order gives identical arithmetic */
  #pragma omp parallel
  {
    double *d = (double *)malloc(sizeof(double) * (size_t)(nn > 0 ? nn : 1));
    double *p2 = (double *)malloc(sizeof(double) * (size_t)(nn > 0 ? nn : 1));
    #pragma omp for collapse(2) schedule(static)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        double *pl_ = p + (size_t)i + (size_t)nx * j;
        #define P(k) pl_[(size_t)((k) - 1) * pl]
        double lxy = lambdaxy ? lambdaxy[(size_t)i + (size_t)nx * j] : 0.;
        double z, den, tol;
        if (lambdaxy) {
          z = 1. / (b[0] + lxy);
          d[0] = c[0] * z;
          P(1) = P(1) * norm * z;
          for (int k = 2; k <= nn; k++) {
            den = b[k - 1] + lxy - a[k - 1] * d[k - 2];
            tol = EPS * fmax(fabs(b[k - 1] + lxy), fabs(a[k - 1] * d[k - 2]));
            if (k == nn && fabs(den) <= tol) {
              d[k - 1] = 0.; P(k) = 0.;
            } else {
              z = 1. / den;
              d[k - 1] = c[k - 1] * z;
              P(k) = (P(k) * norm - a[k - 1] * P(k - 1)) * z;
            }
          }
        } else {
          z = 1. / b[0];
          d[0] = c[0] * z;
          P(1) = P(1) * norm * z;
          for (int k = 2; k <= nn; k++) {
            z = 1. / (b[k - 1] - a[k - 1] * d[k - 2]);
            d[k - 1] = c[k - 1] * z;
            P(k) = (P(k) * norm - a[k - 1] * P(k - 1)) * z;
          }
        }
        for (int k = nn - 1; k >= 1; k--) P(k) = P(k) - d[k - 1] * P(k + 1);
        if (is_periodic) {
          for (int k = 0; k < nn; k++) p2[k] = 0.;
          p2[0] = -a[0];
          p2[nn - 1] = p2[nn - 1] - c[nn - 1];
          z = 1. / (b[0] + lxy);
          d[0] = c[0] * z;
          p2[0] = p2[0] * z;
          for (int k = 2; k <= nn; k++) {
            z = 1. / (b[k - 1] + lxy - a[k - 1] * d[k - 2]);
            d[k - 1] = c[k - 1] * z;
            p2[k - 1] = (p2[k - 1] - a[k - 1] * p2[k - 2]) * z;
          }
          for (int k = nn - 1; k >= 1; k--) p2[k - 1] = p2[k - 1] - d[k - 1] * p2[k];
          if (lambdaxy) {
            den = b[nn] + lxy + c[nn] * p2[0] + a[nn] * p2[nn - 1];
            tol = EPS * fmax(fabs(b[nn] + lxy), fabs(c[nn] * p2[0] + a[nn] * p2[nn - 1]));
            if (fabs(den) <= tol) P(nn + 1) = 0.;
            else P(nn + 1) = (P(nn + 1) * norm - c[nn] * P(1) - a[nn] * P(nn)) / den;
          } else {
            P(nn + 1) = (P(nn + 1) * norm - c[nn] * P(1) - a[nn] * P(nn)) /
                        (b[nn] + c[nn] * p2[0] + a[nn] * p2[nn - 1]);
          }
          for (int k = 1; k <= nn; k++) P(k) = P(k) + p2[k - 1] * P(nn + 1);
        }
        #undef P
      }
    free(d); free(p2);
  }
}

/* ------------------------------------------------------------------------ */
/* distributed tridiagonal solve: src/solver.f90:309-517 (gaussel_dtdma),     */
/* with the z rows split into nslab consecutive slabs (the ranks of the z     */
/* decomposition; slab s has n/nslab rows, the first n%nslab slabs one        */
/* more).  Every slab eliminates its rows down to its first and last unknown, */
/* the 2*nslab boundary rows form the reduced system (unit diagonal) solved   */
/* per line, and the slab interiors are back-substituted.  Periodic z: a(1)   */
/* couples row 1 to row n and c(n) row n to row 1, so the reduced system is   */
/* cyclic; it is closed as solver.f90:469-498 does (first 2*nslab-1 rows,     */
/* the auxiliary solution for a unit last unknown, then the last unknown).    */
/* ------------------------------------------------------------------------ */
void orc_slab_range(int n, int nslab, int s, int *lo, int *cnt) {
  const int m = n / nslab, r = n % nslab;
  *cnt = m + (s < r ? 1 : 0);
  *lo = s * m + (s < r ? s : r);
}

int orc_gaussel_dtdma(int nx, int ny, int n, int nslab, const double *a, const double *b,
                      const double *c, double norm, double *p, const double *lambdaxy, int is_periodic) {
  const size_t pl = (size_t)nx * ny;
  if (nslab < 1 || n / nslab < 3) return -1;  /* the boundary-row step needs three rows per slab */
  if (is_periodic && nslab < 2) return -1;    /* one slab's first and last rows would be one pair */
  int ok = 0;
  #pragma omp parallel
  {
    double *aa = (double *)malloc(sizeof(double) * (size_t)n);
    double *cc = (double *)malloc(sizeof(double) * (size_t)n);
    double *bb = (double *)malloc(sizeof(double) * (size_t)n);
    double *ra = (double *)malloc(sizeof(double) * 2 * (size_t)nslab);
    double *rc = (double *)malloc(sizeof(double) * 2 * (size_t)nslab);
    double *rp = (double *)malloc(sizeof(double) * 2 * (size_t)nslab);
    double *rc0 = (double *)malloc(sizeof(double) * 2 * (size_t)nslab);
    double *r2 = (double *)malloc(sizeof(double) * 2 * (size_t)nslab);
    #pragma omp for collapse(2) schedule(static)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        double *pl_ = p + (size_t)i + (size_t)nx * j;
        #define PK(k) pl_[(size_t)(k) * pl]
        const double lxy = lambdaxy ? lambdaxy[(size_t)i + (size_t)nx * j] : 0.;
        for (int s = 0; s < nslab; s++) {
          int lo, ns;
          orc_slab_range(n, nslab, s, &lo, &ns);
          /* local rows 0..ns-1 are global rows lo..lo+ns-1 (0-based) */
          #define A(k) a[lo + (k)]
          #define C(k) c[lo + (k)]
          for (int k = 0; k < ns; k++) bb[k] = lambdaxy ? b[lo + k] + lxy : b[lo + k];
          for (int k = 0; k < 2; k++) {
            const double zz = 1. / bb[k];
            aa[lo + k] = A(k) * zz;
            cc[lo + k] = C(k) * zz;
            PK(lo + k) = PK(lo + k) * norm * zz;
          }
          for (int k = 2; k < ns; k++) {  /* elimination of lower diagonals */
            const double z = 1. / (bb[k] - A(k) * cc[lo + k - 1]);
            PK(lo + k) = (PK(lo + k) * norm - A(k) * PK(lo + k - 1)) * z;
            aa[lo + k] = -A(k) * aa[lo + k - 1] * z;
            cc[lo + k] = C(k) * z;
          }
          for (int k = ns - 3; k >= 1; k--) {  /* elimination of upper diagonals */
            PK(lo + k) = PK(lo + k) - cc[lo + k] * PK(lo + k + 1);
            aa[lo + k] = aa[lo + k] - cc[lo + k] * aa[lo + k + 1];
            cc[lo + k] = -cc[lo + k] * cc[lo + k + 1];
          }
          {
            const double z = 1. / (1. - aa[lo + 1] * cc[lo]);
            PK(lo) = (PK(lo) - cc[lo] * PK(lo + 1)) * z;
            aa[lo] = aa[lo] * z;
            cc[lo] = -cc[lo] * cc[lo + 1] * z;
          }
          ra[2 * s] = aa[lo];          rc[2 * s] = cc[lo];          rp[2 * s] = PK(lo);
          ra[2 * s + 1] = aa[lo + ns - 1]; rc[2 * s + 1] = cc[lo + ns - 1]; rp[2 * s + 1] = PK(lo + ns - 1);
          #undef A
          #undef C
        }
        /* reduced system (solver.f90:446-468); periodic: rows 0..nn-1 of 0..nn */
        const int nn = is_periodic ? 2 * nslab - 1 : 2 * nslab;
        for (int k = 0; k < 2 * nslab; k++) rc0[k] = rc[k];
        for (int k = 1; k < nn; k++) {
          const double z = 1. / (1. - ra[k] * rc[k - 1]);
          rp[k] = (rp[k] - ra[k] * rp[k - 1]) * z;
          rc[k] = rc[k] * z;
        }
        for (int k = nn - 2; k >= 0; k--) rp[k] = rp[k] - rc[k] * rp[k + 1];
        if (is_periodic) { /* solver.f90:469-498, with the unmodified coefficients rc0 */
          for (int k = 0; k < nn; k++) r2[k] = 0.;
          r2[0] = -ra[0];
          r2[nn - 1] = r2[nn - 1] - rc0[nn - 1];
          for (int k = 1; k < nn; k++) {
            const double z = 1. / (1. - ra[k] * rc0[k - 1]);
            r2[k] = (r2[k] - ra[k] * r2[k - 1]) * z;
            rc0[k] = rc0[k] * z;
          }
          for (int k = nn - 2; k >= 0; k--) r2[k] = r2[k] - rc0[k] * r2[k + 1];
          rp[nn] = (rp[nn] - rc0[nn] * rp[0] - ra[nn] * rp[nn - 1]) /
                   (1. + rc0[nn] * r2[0] + ra[nn] * r2[nn - 1]);
          for (int k = 0; k < nn; k++) rp[k] = rp[k] + r2[k] * rp[nn];
        }
        /* back to the slabs (solver.f90:503-514) */
        for (int s = 0; s < nslab; s++) {
          int lo, ns;
          orc_slab_range(n, nslab, s, &lo, &ns);
          const double p1 = rp[2 * s], pn = rp[2 * s + 1];
          PK(lo) = p1;
          PK(lo + ns - 1) = pn;
          for (int k = 1; k < ns - 1; k++) PK(lo + k) = PK(lo + k) - aa[lo + k] * p1 - cc[lo + k] * pn;
        }
        #undef PK
      }
    free(aa); free(cc); free(bb); free(ra); free(rc); free(rp); free(rc0); free(r2);
  }
  return ok;
}

/* ------------------------------------------------------------------------ */
/* Poisson/Helmholtz solver: src/solver.f90:17-112                           */
/* ------------------------------------------------------------------------ */
int orc_solver(const int *n, const int *kinds, const int *lens, double normfft,
               const double *lambdaxy, const double *a, const double *b, const double *c,
               const char *bc, const char *c_or_f, double *p) {
  DIMS
  const size_t N = (size_t)n1 * n2 * n3;
  double *px = (double *)malloc(sizeof(double) * (N ? N : 1));
  if (!px) return -1;
  #pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) px[IX(i, j, k)] = p[HX(i, j, k)];
  /* x transform: stride 1, one line per (j,k); y: stride n1, one line per (i,k) */
  orc_r2r(kinds[0], lens[0], (long)n2 * n3, 1, n1, px);
  for (int k = 0; k < n3; k++)
    orc_r2r(kinds[2], lens[1], n1, n1, 1, px + (size_t)k * n1 * n2);
  int q = (c_or_f[2] == 'f' && bc[5] == 'D') ? 1 : 0;
  int is_periodic_z = bc[4] == 'P' && bc[5] == 'P';
  orc_gaussel(n1, n2, n3 - q, a, b, c, is_periodic_z, normfft, px, lambdaxy);
  for (int k = 0; k < n3; k++)
    orc_r2r(kinds[3], lens[1], n1, n1, 1, px + (size_t)k * n1 * n2);
  orc_r2r(kinds[1], lens[0], (long)n2 * n3, 1, n1, px);
  #pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) p[HX(i, j, k)] = px[IX(i, j, k)];
  free(px);
  return 0;
}

int orc_solve_helmholtz(const int *n, const int *kinds, const int *lens, double normfft,
                        double alpha, const double *lambdaxy, const double *a,
                        const double *b, const double *c, const double *rhsbx,
                        const double *rhsby, const double *rhsbz, const int *is_bound,
                        const char *cbc, const char *c_or_f, double *p) {
  /* solve_helmholtz.f90:28-75 */
  int nz = n[2];
  orc_updt_rhs_b(c_or_f, cbc, n, is_bound, rhsbx, rhsby, rhsbz, p, alpha, 1);
  double alphai = 1. / alpha;
  double *bb = (double *)malloc(sizeof(double) * (size_t)nz);
  for (int k = 0; k < nz; k++) bb[k] = b[k] + alphai;
  int rc = orc_solver(n, kinds, lens, normfft * alphai, lambdaxy, a, bb, c, cbc, c_or_f, p);
  free(bb);
  return rc;
}

/* src/solver.f90:547-616 on one rank (z not decomposed): gaussel along z of the
 * halo'd field, without eigenvalues (the lambda-absent branch of gaussel) */
int orc_solver_gaussel_z(const int *n, const double *a, const double *b, const double *c,
                         const char *bcz, char c_or_f_z, double norm, double *p) {
  DIMS
  const int q = (c_or_f_z == 'f' && bcz[1] == 'D') ? 1 : 0;
  const int is_periodic_z = bcz[0] == 'P' && bcz[1] == 'P';
  const size_t N = (size_t)n1 * n2 * n3;
  double *px = (double *)malloc(sizeof(double) * (N ? N : 1));
  if (!px) return -1;
  #pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) px[IX(i, j, k)] = p[HX(i, j, k)];
  orc_gaussel(n1, n2, n3 - q, a, b, c, is_periodic_z, norm, px, NULL);
  #pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= n3 - q; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) p[HX(i, j, k)] = px[IX(i, j, k)];
  free(px);
  return 0;
}

/* solve_helmholtz.f90:28-75, is_impdiff_1d branch: z boundary terms only (the x/y
 * ones are deallocated for the 1-D form, main.f90:317-329) */
int orc_solve_helmholtz_1d(const int *n, double alpha, const double *a, const double *b,
                           const double *c, const double *rhsbz, const int *is_bound,
                           const char *cbc, const char *c_or_f, double *p) {
  int nz = n[2];
  orc_updt_rhs_b(c_or_f, cbc, n, is_bound, NULL, NULL, rhsbz, p, alpha, 1);
  double alphai = 1. / alpha;
  double *bb = (double *)malloc(sizeof(double) * (size_t)nz);
  for (int k = 0; k < nz; k++) bb[k] = b[k] + alphai;
  int rc = orc_solver_gaussel_z(n, a, bb, c, cbc + 4, c_or_f[2], alphai, p);
  free(bb);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* momentum right-hand side: src/mom.f90:771-1019                            */
/* ------------------------------------------------------------------------ */
void orc_mom_xyz_ad(const int *n, double dxi, double dyi, const double *dzci,
                    const double *dzfi, double visc, const double *u, const double *v,
                    const double *w, double *dudt, double *dvdt, double *dwdt,
                    double *dudtd, double *dvdtd, double *dwdtd,
                    int is_impdiff, int is_impdiff_1d) {
  DIMS
  double *rdzm = (double *)malloc(sizeof(double) * (n3 + 2));
  double *rdzp = (double *)malloc(sizeof(double) * (n3 + 2));
  for (int k = 1; k <= n3; k++) { rdzm[k] = dzci[k] / dzfi[k]; rdzp[k] = dzci[k] / dzfi[k + 1]; }
  #pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
#define LD(f, di, dj, dk) f[HX(i + (di), j + (dj), k + (dk))]
        double u_ccm = LD(u, 0, 0, -1), u_pcm = LD(u, 1, 0, -1), u_cpm = LD(u, 0, 1, -1),
               u_cmc = LD(u, 0, -1, 0), u_pmc = LD(u, 1, -1, 0), u_mcc = LD(u, -1, 0, 0),
               u_ccc = LD(u, 0, 0, 0), u_pcc = LD(u, 1, 0, 0), u_mpc = LD(u, -1, 1, 0),
               u_cpc = LD(u, 0, 1, 0), u_cmp = LD(u, 0, -1, 1), u_mcp = LD(u, -1, 0, 1),
               u_ccp = LD(u, 0, 0, 1);
        double v_ccm = LD(v, 0, 0, -1), v_pcm = LD(v, 1, 0, -1), v_cpm = LD(v, 0, 1, -1),
               v_cmc = LD(v, 0, -1, 0), v_pmc = LD(v, 1, -1, 0), v_mcc = LD(v, -1, 0, 0),
               v_ccc = LD(v, 0, 0, 0), v_pcc = LD(v, 1, 0, 0), v_mpc = LD(v, -1, 1, 0),
               v_cpc = LD(v, 0, 1, 0), v_cmp = LD(v, 0, -1, 1), v_mcp = LD(v, -1, 0, 1),
               v_ccp = LD(v, 0, 0, 1);
        double w_ccm = LD(w, 0, 0, -1), w_pcm = LD(w, 1, 0, -1), w_cpm = LD(w, 0, 1, -1),
               w_cmc = LD(w, 0, -1, 0), w_pmc = LD(w, 1, -1, 0), w_mcc = LD(w, -1, 0, 0),
               w_ccc = LD(w, 0, 0, 0), w_pcc = LD(w, 1, 0, 0), w_mpc = LD(w, -1, 1, 0),
               w_cpc = LD(w, 0, 1, 0), w_cmp = LD(w, 0, -1, 1), w_mcp = LD(w, -1, 0, 1),
               w_ccp = LD(w, 0, 0, 1);
#undef LD


        /* x */
        double dudxp = (u_pcc - u_ccc) * dxi, dudxm = (u_ccc - u_mcc) * dxi;
        double dudyp = (u_cpc - u_ccc) * dyi, dudym = (u_ccc - u_cmc) * dyi;
        double dudzp = (u_ccp - u_ccc) * dzci[k], dudzm = (u_ccc - u_ccm) * dzci[k - 1];
        double uuip = 0.25 * (u_ccc + u_pcc) * u_pcc, uuim = 0.25 * (u_ccc + u_mcc) * u_mcc;
        double vujp = 0.25 * (v_ccc + v_pcc) * u_cpc, vujm = 0.25 * (v_cmc + v_pmc) * u_cmc;
        double wukp = 0.25 * (w_ccc + w_pcc) * u_ccp, wukm = 0.25 * (w_ccm + w_pcm) * u_ccm;
        double dudtd_xy_s = visc * (dudxp - dudxm) * dxi + visc * (dudyp - dudym) * dyi;
        double dudtd_z_s = visc * (dudzp - dudzm) * dzfi[k];
        double dudt_s = -(uuip - uuim) * dxi - (vujp - vujm) * dyi - (wukp - wukm) * dzfi[k];
        /* y */
        double dvdxp = (v_pcc - v_ccc) * dxi, dvdxm = (v_ccc - v_mcc) * dxi;
        double dvdyp = (v_cpc - v_ccc) * dyi, dvdym = (v_ccc - v_cmc) * dyi;
        double dvdzp = (v_ccp - v_ccc) * dzci[k], dvdzm = (v_ccc - v_ccm) * dzci[k - 1];
        double uvip = 0.25 * (u_ccc + u_cpc) * v_pcc, uvim = 0.25 * (u_mcc + u_mpc) * v_mcc;
        double vvjp = 0.25 * (v_ccc + v_cpc) * v_cpc, vvjm = 0.25 * (v_ccc + v_cmc) * v_cmc;
        double wvkp = 0.25 * (w_ccc + w_cpc) * v_ccp, wvkm = 0.25 * (w_ccm + w_cpm) * v_ccm;
        double dvdtd_xy_s = visc * (dvdxp - dvdxm) * dxi + visc * (dvdyp - dvdym) * dyi;
        double dvdtd_z_s = visc * (dvdzp - dvdzm) * dzfi[k];
        double dvdt_s = -(uvip - uvim) * dxi - (vvjp - vvjm) * dyi - (wvkp - wvkm) * dzfi[k];
        /* z */
        double dwdxp = (w_pcc - w_ccc) * dxi, dwdxm = (w_ccc - w_mcc) * dxi;
        double dwdyp = (w_cpc - w_ccc) * dyi, dwdym = (w_ccc - w_cmc) * dyi;
        double dwdzp = (w_ccp - w_ccc) * dzfi[k + 1], dwdzm = (w_ccc - w_ccm) * dzfi[k];
        double uwip = 0.25 * (rdzm[k] * u_ccc + rdzp[k] * u_ccp) * w_pcc;
        double uwim = 0.25 * (rdzm[k] * u_mcc + rdzp[k] * u_mcp) * w_mcc;
        double vwjp = 0.25 * (rdzm[k] * v_ccc + rdzp[k] * v_ccp) * w_cpc;
        double vwjm = 0.25 * (rdzm[k] * v_cmc + rdzp[k] * v_cmp) * w_cmc;
        double wwkp = 0.25 * (w_ccc + w_ccp) * w_ccp, wwkm = 0.25 * (w_ccc + w_ccm) * w_ccm;
        double dwdtd_xy_s = visc * (dwdxp - dwdxm) * dxi + visc * (dwdyp - dwdym) * dyi;
        double dwdtd_z_s = visc * (dwdzp - dwdzm) * dzci[k];
        double dwdt_s = -(uwip - uwim) * dxi - (vwjp - vwjm) * dyi - (wwkp - wwkm) * dzci[k];
        size_t o = IX(i, j, k);
        if (is_impdiff) {
          double a1, a2, a3;
          if (is_impdiff_1d) {
            dudt_s = dudt_s + dudtd_xy_s; dvdt_s = dvdt_s + dvdtd_xy_s; dwdt_s = dwdt_s + dwdtd_xy_s;
            a1 = dudtd_z_s; a2 = dvdtd_z_s; a3 = dwdtd_z_s;
          } else {
            a1 = dudtd_xy_s + dudtd_z_s; a2 = dvdtd_xy_s + dvdtd_z_s; a3 = dwdtd_xy_s + dwdtd_z_s;
          }
          dudt[o] = dudt_s; dvdt[o] = dvdt_s; dwdt[o] = dwdt_s;
          dudtd[o] = a1; dvdtd[o] = a2; dwdtd[o] = a3;
        } else {
          dudt_s = dudt_s + dudtd_xy_s + dudtd_z_s;
          dvdt_s = dvdt_s + dvdtd_xy_s + dvdtd_z_s;
          dwdt_s = dwdt_s + dwdtd_xy_s + dwdtd_z_s;
          dudt[o] = dudt_s; dvdt[o] = dvdt_s; dwdt[o] = dwdt_s;
        }
      }
  free(rdzm); free(rdzp);
}

/* ------------------------------------------------------------------------ */
/* src/utils.f90:15-42                                                       */
/* ------------------------------------------------------------------------ */
double orc_bulk_mean(const int *n, const double *grid_vol_ratio, const double *p) {
  DIMS
  double mean = 0.;
  #pragma omp parallel for collapse(2) reduction(+ : mean) schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) mean = mean + p[HX(i, j, k)] * grid_vol_ratio[k];
  return mean;
}

/* ------------------------------------------------------------------------ */
/* src/rk.f90:24-333, cmpt_bulk_forcing rk.f90:493-518                        */
/* ------------------------------------------------------------------------ */
void orc_rk(const double *rkpar, const int *n, const double *dli, const double *dzci,
            const double *dzfi, const double *grid_vol_ratio_c,
            const double *grid_vol_ratio_f, double dt, double visc, const double *p,
            const int *is_forced, const double *velf, const double *bforce,
            const double *gacc, double beta, const double *s,
            double *dudtrko, double *dvdtrko, double *dwdtrko,
            double *u, double *v, double *w, double *f, double *work,
            int is_impdiff, int is_impdiff_1d) {
  DIMS
  const size_t N = (size_t)n1 * n2 * n3;
  double factor1 = rkpar[0] * dt, factor2 = rkpar[1] * dt, factor12 = factor1 + factor2;
  double *dudtrk = work, *dvdtrk = work + N, *dwdtrk = work + 2 * N;
  double *dudtrkd = is_impdiff ? work + 3 * N : NULL;
  double *dvdtrkd = is_impdiff ? work + 4 * N : NULL;
  double *dwdtrkd = is_impdiff ? work + 5 * N : NULL;
  orc_mom_xyz_ad(n, dli[0], dli[1], dzci, dzfi, visc, u, v, w, dudtrk, dvdtrk, dwdtrk,
                 dudtrkd, dvdtrkd, dwdtrkd, is_impdiff, is_impdiff_1d);
  /* rk.f90:156-186 (the default, non-unswitched loop) */
  #pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
        size_t o = IX(i, j, k), h = HX(i, j, k);
        u[h] = u[h] + factor1 * dudtrk[o] + factor2 * dudtrko[o] +
               factor12 * (bforce[0] - dli[0] * (p[HX(i + 1, j, k)] - p[h]));
        v[h] = v[h] + factor1 * dvdtrk[o] + factor2 * dvdtrko[o] +
               factor12 * (bforce[1] - dli[1] * (p[HX(i, j + 1, k)] - p[h]));
        w[h] = w[h] + factor1 * dwdtrk[o] + factor2 * dwdtrko[o] +
               factor12 * (bforce[2] - dzci[k] * (p[HX(i, j, k + 1)] - p[h]));
        if (s) { /* Boussinesq buoyancy of the first scalar, rk.f90:163-177 */
          u[h] = u[h] - factor12 * gacc[0] * beta * 0.5 * (s[HX(i + 1, j, k)] + s[h]);
          v[h] = v[h] - factor12 * gacc[1] * beta * 0.5 * (s[HX(i, j + 1, k)] + s[h]);
          w[h] = w[h] - factor12 * gacc[2] * beta * 0.5 * (s[HX(i, j, k + 1)] + s[h]);
        }
        if (is_impdiff) {
          u[h] = u[h] + factor12 * dudtrkd[o];
          v[h] = v[h] + factor12 * dvdtrkd[o];
          w[h] = w[h] + factor12 * dwdtrkd[o];
        }
      }
  /* rk.f90:266-276 */
  #pragma omp parallel for schedule(static)
  for (size_t o = 0; o < N; o++) { dudtrko[o] = dudtrk[o]; dvdtrko[o] = dvdtrk[o]; dwdtrko[o] = dwdtrk[o]; }
  /* rk.f90:493-518 */
  f[0] = f[1] = f[2] = 0.;
  if (is_forced[0]) f[0] = velf[0] - orc_bulk_mean(n, grid_vol_ratio_f, u);
  if (is_forced[1]) f[1] = velf[1] - orc_bulk_mean(n, grid_vol_ratio_f, v);
  if (is_forced[2]) f[2] = velf[2] - orc_bulk_mean(n, grid_vol_ratio_c, w);
  if (is_impdiff) {
    /* rk.f90:312-326 */
    #pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) {
          size_t o = IX(i, j, k), h = HX(i, j, k);
          u[h] = (u[h] - .5 * factor12 * dudtrkd[o]) + f[0];
          v[h] = (v[h] - .5 * factor12 * dvdtrkd[o]) + f[1];
          w[h] = (w[h] - .5 * factor12 * dwdtrkd[o]) + f[2];
        }
  } else if (is_forced[0] || is_forced[1] || is_forced[2]) {
    /* mom.f90:725-769 */
    double *fld[3] = {u, v, w};
    for (int c = 0; c < 3; c++) {
      if (!is_forced[c]) continue;
      double ff = f[c];
      double *q = fld[c];
      #pragma omp parallel for collapse(2) schedule(static)
      for (int k = 1; k <= n3; k++)
        for (int j = 1; j <= n2; j++)
          for (int i = 1; i <= n1; i++) q[HX(i, j, k)] = q[HX(i, j, k)] + ff;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* scalar transport: src/scal.f90:41-98 (the default, non-unswitched loop)    */
/* ------------------------------------------------------------------------ */
void orc_scal(const int *n, double dxi, double dyi, const double *dzci, const double *dzfi,
              double visc, const double *u, const double *v, const double *w, const double *s,
              double *dsdt, double *dsdtd, int is_impdiff, int is_impdiff_1d) {
  DIMS
  #pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
        const size_t h = HX(i, j, k), o = IX(i, j, k);
        const double sc = s[h];
        const double usim = 0.5 * (s[HX(i - 1, j, k)] + sc) * u[HX(i - 1, j, k)];
        const double usip = 0.5 * (s[HX(i + 1, j, k)] + sc) * u[h];
        const double vsjm = 0.5 * (s[HX(i, j - 1, k)] + sc) * v[HX(i, j - 1, k)];
        const double vsjp = 0.5 * (s[HX(i, j + 1, k)] + sc) * v[h];
        const double wskm = 0.5 * (s[HX(i, j, k - 1)] + sc) * w[HX(i, j, k - 1)];
        const double wskp = 0.5 * (s[HX(i, j, k + 1)] + sc) * w[h];
        const double dsdxp = (s[HX(i + 1, j, k)] - sc) * dxi;
        const double dsdxm = (sc - s[HX(i - 1, j, k)]) * dxi;
        const double dsdyp = (s[HX(i, j + 1, k)] - sc) * dyi;
        const double dsdym = (sc - s[HX(i, j - 1, k)]) * dyi;
        const double dsdzp = (s[HX(i, j, k + 1)] - sc) * dzci[k];
        const double dsdzm = (sc - s[HX(i, j, k - 1)]) * dzci[k - 1];
        double r = dxi * (-usip + usim) + dyi * (-vsjp + vsjm) + dzfi[k] * (-wskp + wskm);
        const double dxy = (dsdxp - dsdxm) * visc * dxi + (dsdyp - dsdym) * visc * dyi;
        const double dz = (dsdzp - dsdzm) * visc * dzfi[k];
        if (is_impdiff) {
          if (is_impdiff_1d) {
            r = r + dxy;
            dsdtd[o] = dz;
          } else {
            dsdtd[o] = dxy + dz;
          }
        } else {
          r = r + dxy + dz;
        }
        dsdt[o] = r;
      }
}

/* src/rk.f90:335-491 (is_cmpt_wallflux is a .false. parameter there: no wall flux) */
void orc_rk_scal(const double *rkpar, const int *n, const double *dli, const double *dzci,
                 const double *dzfi, const double *grid_vol_ratio_f, double alpha, double dt,
                 const double *u, const double *v, const double *w, int is_forced, double scalf,
                 double ssource, double *dsdtrko, double *s, double *f, double *work,
                 int is_impdiff, int is_impdiff_1d) {
  DIMS
  const size_t N = (size_t)n1 * n2 * n3;
  const double factor1 = rkpar[0] * dt, factor2 = rkpar[1] * dt, factor12 = factor1 + factor2;
  double *dsdtrk = work, *dsdtrkd = is_impdiff ? work + N : NULL;
  orc_scal(n, dli[0], dli[1], dzci, dzfi, alpha, u, v, w, s, dsdtrk, dsdtrkd, is_impdiff, is_impdiff_1d);
  #pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
        const size_t o = IX(i, j, k), h = HX(i, j, k);
        s[h] = s[h] + factor1 * dsdtrk[o] + factor2 * dsdtrko[o] + factor12 * ssource;
        if (is_impdiff) s[h] = s[h] + factor12 * dsdtrkd[o];
      }
  *f = 0.;
  if (is_forced) *f = scalf - orc_bulk_mean(n, grid_vol_ratio_f, s);
  if (is_impdiff) {
    #pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) {
          const size_t o = IX(i, j, k), h = HX(i, j, k);
          s[h] = s[h] - .5 * factor12 * dsdtrkd[o];
        }
  }
  #pragma omp parallel for schedule(static)
  for (size_t o = 0; o < N; o++) dsdtrko[o] = dsdtrk[o];
}

/* src/scal.f90:313-329 */
void orc_bulk_forcing_s(const int *n, int is_forced, double ff, double *p) {
  DIMS
  if (!is_forced) return;
  #pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) p[HX(i, j, k)] = p[HX(i, j, k)] + ff;
}

/* ------------------------------------------------------------------------ */
/* src/fillps.f90, src/correc.f90, src/updatep.f90                           */
/* ------------------------------------------------------------------------ */
void orc_fillps(const int *n, const double *dli, const double *dzfi, double dti,
                const double *u, const double *v, const double *w, double *p) {
  DIMS
  double dtidxi = dti * dli[0], dtidyi = dti * dli[1];
  #pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
        size_t h = HX(i, j, k);
        p[h] = ((w[h] - w[HX(i, j, k - 1)]) * dti * dzfi[k] +
                (v[h] - v[HX(i, j - 1, k)]) * dtidyi +
                (u[h] - u[HX(i - 1, j, k)]) * dtidxi);
      }
}

void orc_correc(const int *n, const double *dli, const double *dzci, double dt,
                const double *p, double *u, double *v, double *w) {
  DIMS
  double factori = dt * dli[0], factorj = dt * dli[1];
  #pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k <= n3 + 1; k++)
    for (int j = 0; j <= n2 + 1; j++)
      for (int i = 0; i <= n1; i++) {
        size_t h = HX(i, j, k);
        u[h] = u[h] - factori * (p[HX(i + 1, j, k)] - p[h]);
      }
  #pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k <= n3 + 1; k++)
    for (int j = 0; j <= n2; j++)
      for (int i = 0; i <= n1 + 1; i++) {
        size_t h = HX(i, j, k);
        v[h] = v[h] - factorj * (p[HX(i, j + 1, k)] - p[h]);
      }
  #pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k <= n3; k++)
    for (int j = 0; j <= n2 + 1; j++)
      for (int i = 0; i <= n1 + 1; i++) {
        size_t h = HX(i, j, k);
        w[h] = w[h] - dt * dzci[k] * (p[HX(i, j, k + 1)] - p[h]);
      }
}

void orc_updatep(const int *n, const double *dli, const double *dzci, const double *dzfi,
                 double alpha, const double *pp, double *p, int is_impdiff, int is_impdiff_1d) {
  DIMS
  if (is_impdiff) {
    double dxi = dli[0], dyi = dli[1];
    #pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) {
          size_t h = HX(i, j, k);
          double lap_pp = 0.;
          if (!is_impdiff_1d)
            lap_pp = lap_pp + (pp[HX(i + 1, j, k)] - 2. * pp[h] + pp[HX(i - 1, j, k)]) * (dxi * dxi) +
                              (pp[HX(i, j + 1, k)] - 2. * pp[h] + pp[HX(i, j - 1, k)]) * (dyi * dyi);
          lap_pp = lap_pp + ((pp[HX(i, j, k + 1)] - pp[h]) * dzci[k] -
                             (pp[h] - pp[HX(i, j, k - 1)]) * dzci[k - 1]) * dzfi[k];
          p[h] = p[h] + pp[h] + alpha * lap_pp;
        }
  } else {
    #pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) { size_t h = HX(i, j, k); p[h] = p[h] + pp[h]; }
  }
}

/* ------------------------------------------------------------------------ */
/* boundary conditions: src/bound.f90                                        */
/* ------------------------------------------------------------------------ */
enum { HALO_ABSENT = -99, HALO_LOWER = -1, HALO_NONE = 0, HALO_UPPER = 1 };

/* one face of one field, bound.f90:182-463 (nh = 1) */
static void set_bc(char ctype, int ibound, int idir, int centered, double rvalue, double dr,
                   const int *n, double *p) {
  DIMS
  int nd = n[idir - 1];
  double factor = rvalue, sgn = 0.;
  if (ctype == 'D' && centered) { factor = 2. * factor; sgn = -1.; }
  if (ctype == 'N') {
    if (ibound == 0) factor = -dr * factor; else factor = dr * factor;
    sgn = 1.;
  }
  /* plane index ranges over the full halo'd extent of the two other directions */
  int e1 = idir == 1 ? n2 + 1 : n1 + 1, e2 = idir == 3 ? n2 + 1 : n3 + 1;
#define AT(a, b, c) p[idir == 1 ? HX(c, a, b) : idir == 2 ? HX(a, c, b) : HX(a, b, c)]
  #pragma omp parallel for collapse(2) schedule(static)
  for (int b = 0; b <= e2; b++)
    for (int a = 0; a <= e1; a++) {
      if (ctype == 'P') {
        if (ibound == 0) AT(a, b, 0) = AT(a, b, nd);
        else AT(a, b, nd + 1) = AT(a, b, 1);
      } else if (centered) {
        if (ibound == 0) AT(a, b, 0) = factor + sgn * AT(a, b, 1);
        else AT(a, b, nd + 1) = factor + sgn * AT(a, b, nd);
      } else if (ctype == 'D') {
        if (ibound == 0) AT(a, b, 0) = factor;
        else { AT(a, b, nd + 1) = AT(a, b, nd - 1); AT(a, b, nd) = factor; }
      } else { /* staggered N */
        if (ibound == 0) AT(a, b, 0) = 1. * factor + AT(a, b, 1);
        else { AT(a, b, nd + 1) = AT(a, b, nd); AT(a, b, nd) = 1. * factor + AT(a, b, nd - 1); }
      }
    }
#undef AT
}

/* single-rank halo exchange, bound.f90:600-713: along a periodic decomposed
 * direction the neighbour is this rank itself, otherwise MPI_PROC_NULL */
static void updthalo(int idir, int periodic, const int *n, double *p, int halo_side) {
  DIMS
  if (idir == 1 || !periodic) return;  /* x is the pencil axis (ipencil_axis = 1) */
  int up_lower = 1, up_upper = 1;
  if (halo_side == HALO_LOWER) up_upper = 0;
  else if (halo_side == HALO_NONE) { up_lower = 0; up_upper = 0; }
  else if (halo_side == HALO_UPPER) up_lower = 0;
  int nd = n[idir - 1];
  int e1 = n1 + 1, e2 = idir == 3 ? n2 + 1 : n3 + 1;
#define AT(a, b, c) p[idir == 2 ? HX(a, c, b) : HX(a, b, c)]
  if (up_upper) {
    #pragma omp parallel for collapse(2) schedule(static)
    for (int b = 0; b <= e2; b++)
      for (int a = 0; a <= e1; a++) AT(a, b, nd + 1) = AT(a, b, 1);
  }
  if (up_lower) {
    #pragma omp parallel for collapse(2) schedule(static)
    for (int b = 0; b <= e2; b++)
      for (int a = 0; a <= e1; a++) AT(a, b, 0) = AT(a, b, nd);
  }
#undef AT
}

void orc_bounduvw(const char *cbc, const int *n, const double *bc, const int *is_bound,
                  const int *is_periodic, const double *dl, const double *dzc,
                  const double *dzf, double *u, double *v, double *w,
                  int keep_norm_values, int halos_norm, int halos_tang) {
  /* bound.f90:17-122; cbc/bc index (ib, idir, ivel) -> ib + 2*(idir-1) + 6*(ivel-1) */
#define CBC(ib, d, c) cbc[(ib) + 2 * ((d) - 1) + 6 * ((c) - 1)]
#define BCV(ib, d, c) bc[(ib) + 2 * ((d) - 1) + 6 * ((c) - 1)]
  int keep_norm = keep_norm_values > 0;
  int updt_norm[2] = {1, 1};
  if (halos_norm == HALO_LOWER) updt_norm[1] = 0;
  else if (halos_norm == HALO_NONE) { updt_norm[0] = 0; updt_norm[1] = 0; }
  else if (halos_norm == HALO_UPPER) updt_norm[0] = 0;
  int apply_norm_bc[2][4];
  for (int d = 1; d <= 3; d++) {
    int pp = CBC(0, d, d) == 'P' && CBC(1, d, d) == 'P';
    apply_norm_bc[0][d] = !keep_norm || (pp && updt_norm[0]);
    apply_norm_bc[1][d] = !keep_norm || (pp && updt_norm[1]);
  }
  int hn = halos_norm, ht = halos_tang;
  updthalo(1, is_periodic[0], n, u, hn); updthalo(1, is_periodic[0], n, v, ht); updthalo(1, is_periodic[0], n, w, ht);
  updthalo(2, is_periodic[1], n, u, ht); updthalo(2, is_periodic[1], n, v, hn); updthalo(2, is_periodic[1], n, w, ht);
  updthalo(3, is_periodic[2], n, u, ht); updthalo(3, is_periodic[2], n, v, ht); updthalo(3, is_periodic[2], n, w, hn);
  double *f[4] = {NULL, u, v, w};
  for (int d = 1; d <= 3; d++)
    for (int ib = 0; ib <= 1; ib++) {
      if (!is_bound[ib + 2 * (d - 1)]) continue;
      double dr = d == 1 ? dl[0] : d == 2 ? dl[1] : 0.;
      for (int c = 1; c <= 3; c++) {
        int normal = c == d;
        if (normal && !apply_norm_bc[ib][d]) continue;
        if (d == 3) dr = normal ? (ib == 0 ? dzf[0] : dzf[n[2]]) : (ib == 0 ? dzc[0] : dzc[n[2]]);
        set_bc(CBC(ib, d, c), ib, d, !normal, BCV(ib, d, c), dr, n, f[c]);
      }
    }
#undef CBC
#undef BCV
}

void orc_boundp(const char *cbc, const int *n, const double *bc, const int *is_bound,
                const int *is_periodic, const double *dl, const double *dzc, double *p,
                int halo_side) {
  /* bound.f90:124-180 */
  int update_halo[2] = {1, 1};
  if (halo_side == HALO_LOWER) update_halo[1] = 0;
  else if (halo_side == HALO_NONE) { update_halo[0] = 0; update_halo[1] = 0; }
  else if (halo_side == HALO_UPPER) update_halo[0] = 0;
  for (int d = 1; d <= 3; d++) updthalo(d, is_periodic[d - 1], n, p, halo_side);
  for (int d = 1; d <= 3; d++)
    for (int ib = 0; ib <= 1; ib++) {
      if (!(is_bound[ib + 2 * (d - 1)] && update_halo[ib])) continue;
      double dr = d == 1 ? dl[0] : d == 2 ? dl[1] : (ib == 0 ? dzc[0] : dzc[n[2]]);
      set_bc(cbc[ib + 2 * (d - 1)], ib, d, 1, bc[ib + 2 * (d - 1)], dr, n, p);
    }
}

void orc_updt_rhs_b(const char *c_or_f, const char *cbc, const int *n, const int *is_bound,
                    const double *rhsbx, const double *rhsby, const double *rhsbz,
                    double *p, double alpha, int has_alpha) {
  /* bound.f90:514-598 */
  DIMS
  int q[3] = {0, 0, 0};
  for (int d = 0; d < 3; d++) if (c_or_f[d] == 'f' && cbc[1 + 2 * d] == 'D') q[d] = 1;
  double norm = has_alpha ? alpha : 1.;
  if (rhsbx) {
    if (is_bound[0])
      for (int k = 1; k <= n3; k++) for (int j = 1; j <= n2; j++)
        p[HX(1, j, k)] = p[HX(1, j, k)] + rhsbx[(j - 1) + (size_t)n2 * (k - 1)] * norm;
    if (is_bound[1]) {
      int nn = n1 - q[0];
      for (int k = 1; k <= n3; k++) for (int j = 1; j <= n2; j++)
        p[HX(nn, j, k)] = p[HX(nn, j, k)] + rhsbx[(j - 1) + (size_t)n2 * (k - 1) + (size_t)n2 * n3] * norm;
    }
  }
  if (rhsby) {
    if (is_bound[2])
      for (int k = 1; k <= n3; k++) for (int i = 1; i <= n1; i++)
        p[HX(i, 1, k)] = p[HX(i, 1, k)] + rhsby[(i - 1) + (size_t)n1 * (k - 1)] * norm;
    if (is_bound[3]) {
      int nn = n2 - q[1];
      for (int k = 1; k <= n3; k++) for (int i = 1; i <= n1; i++)
        p[HX(i, nn, k)] = p[HX(i, nn, k)] + rhsby[(i - 1) + (size_t)n1 * (k - 1) + (size_t)n1 * n3] * norm;
    }
  }
  if (rhsbz) {
    if (is_bound[4])
      for (int j = 1; j <= n2; j++) for (int i = 1; i <= n1; i++)
        p[HX(i, j, 1)] = p[HX(i, j, 1)] + rhsbz[(i - 1) + (size_t)n1 * (j - 1)] * norm;
    if (is_bound[5]) {
      int nn = n3 - q[2];
      for (int j = 1; j <= n2; j++) for (int i = 1; i <= n1; i++)
        p[HX(i, j, nn)] = p[HX(i, j, nn)] + rhsbz[(i - 1) + (size_t)n1 * (j - 1) + (size_t)n1 * n2] * norm;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* checks: src/chkdiv.f90:15-54, src/chkdt.f90:16-78                         */
/* ------------------------------------------------------------------------ */
void orc_chkdiv(const int *n, const double *l, const double *dli, const double *dzfi,
                const double *u, const double *v, const double *w,
                double *divtot, double *divmax) {
  DIMS
  double dxi = dli[0], dyi = dli[1];
  double tot = 0., mx = 0.;
  #pragma omp parallel for collapse(2) reduction(+ : tot) reduction(max : mx) schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
        size_t h = HX(i, j, k);
        double div = (w[h] - w[HX(i, j, k - 1)]) * dzfi[k] + (v[h] - v[HX(i, j - 1, k)]) * dyi +
                     (u[h] - u[HX(i - 1, j, k)]) * dxi;
        mx = fmax(mx, fabs(div));
        tot = tot + fabs(div) / (dxi * dyi * dzfi[k]);
      }
  *divtot = tot / (l[0] * l[1] * l[2]);
  *divmax = mx;
}

double orc_chkdt(const int *n, const double *dl, const double *dzci, const double *dzfi,
                 double visc, double alpha, const double *u, const double *v,
                 const double *w, int is_impdiff, int is_impdiff_1d) {
  DIMS
  double dxi = 1. / dl[0], dyi = 1. / dl[1];
  double dlmin = fmin(dl[0], dl[1]);
  if (!is_impdiff_1d)
    for (int k = 0; k <= n3 + 1; k++) dlmin = fmin(dlmin, 1. / dzfi[k]);
  double dti = 0.;
  #pragma omp parallel for collapse(2) reduction(max : dti) schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
#define F(q, a, b, c) q[HX(i + (a), j + (b), k + (c))]
        double ux = fabs(F(u, 0, 0, 0));
        double vx = 0.25 * fabs(F(v, 0, 0, 0) + F(v, 0, -1, 0) + F(v, 1, 0, 0) + F(v, 1, -1, 0));
        double wx = 0.25 * fabs(F(w, 0, 0, 0) + F(w, 0, 0, -1) + F(w, 1, 0, 0) + F(w, 1, 0, -1));
        double dtix = ux * dxi + vx * dyi + wx * dzfi[k];
        double uy = 0.25 * fabs(F(u, 0, 0, 0) + F(u, 0, 1, 0) + F(u, -1, 1, 0) + F(u, -1, 0, 0));
        double vy = fabs(F(v, 0, 0, 0));
        double wy = 0.25 * fabs(F(w, 0, 0, 0) + F(w, 0, 1, 0) + F(w, 0, 1, -1) + F(w, 0, 0, -1));
        double dtiy = uy * dxi + vy * dyi + wy * dzfi[k];
        double uz = 0.25 * fabs(F(u, 0, 0, 0) + F(u, -1, 0, 0) + F(u, -1, 0, 1) + F(u, 0, 0, 1));
        double vz = 0.25 * fabs(F(v, 0, 0, 0) + F(v, 0, -1, 0) + F(v, 0, -1, 1) + F(v, 0, 0, 1));
        double wz = fabs(F(w, 0, 0, 0));
        double dtiz = uz * dxi + vz * dyi + wz * dzci[k];
#undef F
        dti = fmax(dti, fmax(dtix, fmax(dtiy, dtiz)));
      }
  if (dti < EPS) dti = 1.;
  /* sqrt(3.) and 1.65/12. are default-real (single precision) expressions */
  double sqrt3 = (double)sqrtf(3.f);
  double c165 = (double)(1.65f / 12.f);
  if (is_impdiff && !is_impdiff_1d) return sqrt3 / dti;
  return fmin(c165 / fmax(visc, alpha) * (dlmin * dlmin), sqrt3 / dti);
}
