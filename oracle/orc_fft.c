/*
 * Oracle transforms: batched in-place real-to-real transforms with FFTW's
 * unnormalised definitions (the arithmetic the reference's CPU build obtains
 * from fftw_plan_guru_r2r, src/fft.f90:87-97,152-162,247-258).
 *
 * TEST INFRASTRUCTURE ONLY (checker and CPU baseline, never the product path).
 *
 * FFTW itself is an external, un-vendored dependency of the reference (system
 * libfftw3, version unpinned).  Its published kind definitions are restated
 * here on top of our own complex FFT:
 *   - complex DFT: Stockham autosort, mixed radix 4/2/3/5/generic, following
 *     the formulation of Govindaraju et al., "High performance discrete Fourier
 *     transforms on graphics processors" (SC'08), with exact-argument twiddles;
 *   - R2HC/HC2R (even n): the half-length complex FFT with the standard
 *     split-radix post/pre-processing; odd n: full-length complex FFT;
 *   - REDFT10/01 and RODFT10/01: Makhoul's reordering (J. Makhoul, "A fast
 *     cosine transform in one and two dimensions", IEEE TASSP 28, 1980), the
 *     DST kinds through the sign/reversal identities;
 *   - REDFT11/RODFT11: a length-2n complex DFT with pre/post twiddles;
 *   - REDFT00/RODFT00: the DFT of the even/odd extension (lengths 2(n-1), 2(n+1)).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "cans_oracle.h"

typedef struct { double re, im; } cx;

static inline cx cx_mul(cx a, cx b) { cx r = {a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re}; return r; }
static inline cx cx_add(cx a, cx b) { cx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cx cx_sub(cx a, cx b) { cx r = {a.re - b.re, a.im - b.im}; return r; }
static inline cx cx_conj(cx a) { cx r = {a.re, -a.im}; return r; }

/* e^{-2 pi i t/n}, computed in long double from an octant-reduced argument */
static cx twiddle(long t, long n) {
  t %= n; if (t < 0) t += n;
  long double x = (long double)t / (long double)n; /* in [0,1) */
  long double s, c;
  const long double twopi = 6.283185307179586476925286766559005768L;
  c = cosl(twopi * x); s = sinl(twopi * x);
  /* exact values at multiples of a quarter turn */
  if (4 * t % n == 0) {
    long q = 4 * t / n;
    c = (q == 0) ? 1 : (q == 2) ? -1 : 0;
    s = (q == 1) ? 1 : (q == 3) ? -1 : 0;
  }
  cx r = {(double)c, -(double)s};
  return r;
}

/* Rounding variant: 0 = radix-4-first factorisation and half-length real FFTs;
 * 1 = radix-2/odd factorisation and full-length complex FFTs for R2HC/HC2R.  Both
 * compute FFTW's definitions; comparing them measures how far two valid FP64
 * implementations of the same transforms drift apart through the solver. */
static int g_variant = 0;
void orc_set_fft_variant(int v) { g_variant = v; }

typedef struct {
  int n, nfac, fac[40];
  cx *tw;     /* e^{-2 pi i t/n}, t = 0..n-1 */
} cplan;

static void cplan_init(cplan *p, int n) {
  p->n = n; p->nfac = 0;
  int m = n;
  while (!g_variant && m % 4 == 0) { p->fac[p->nfac++] = 4; m /= 4; }
  while (m % 2 == 0) { p->fac[p->nfac++] = 2; m /= 2; }
  for (int f = 3; m > 1; f += 2)
    while (m % f == 0) { p->fac[p->nfac++] = f; m /= f; }
  p->tw = (cx *)malloc(sizeof(cx) * (size_t)(n > 0 ? n : 1));
  for (int t = 0; t < n; t++) p->tw[t] = twiddle(t, n);
}
static void cplan_free(cplan *p) { free(p->tw); p->tw = NULL; }

/* small DFT of radix R on v (sign -1 forward, +1 backward) */
static void dft_small(cx *v, int R, int sign, const cplan *p) {
  if (R == 2) {
    cx a = v[0], b = v[1];
    v[0] = cx_add(a, b); v[1] = cx_sub(a, b);
  } else if (R == 4) {
    cx a0 = cx_add(v[0], v[2]), a1 = cx_sub(v[0], v[2]);
    cx b0 = cx_add(v[1], v[3]), b1 = cx_sub(v[1], v[3]);
    /* b1 * (-i*sign... ) : forward multiplies by -i */
    cx jb1 = (sign < 0) ? (cx){b1.im, -b1.re} : (cx){-b1.im, b1.re};
    v[0] = cx_add(a0, b0); v[2] = cx_sub(a0, b0);
    v[1] = cx_add(a1, jb1); v[3] = cx_sub(a1, jb1);
  } else {
    cx tmp[64];
    cx *out = tmp;
    cx *heap = NULL;
    if (R > 64) { heap = (cx *)malloc(sizeof(cx) * R); out = heap; }
    long step = p->n / R;
    for (int q = 0; q < R; q++) {
      cx s = {0, 0};
      for (int r = 0; r < R; r++) {
        cx w = p->tw[(long)((long)r * q % R) * step];
        if (sign > 0) w = cx_conj(w);
        s = cx_add(s, cx_mul(v[r], w));
      }
      out[q] = s;
    }
    memcpy(v, out, sizeof(cx) * R);
    free(heap);
  }
}

/* in-place complex DFT of length p->n; work holds n elements */
static void cfft(const cplan *p, cx *x, cx *work, int sign) {
  int n = p->n;
  if (n <= 1) return;
  cx *in = x, *out = work;
  long Ns = 1;
  cx v[64];
  for (int s = 0; s < p->nfac; s++) {
    int R = p->fac[s];
    long m = n / R;
    cx *vv = v;
    cx *heap = NULL;
    if (R > 64) { heap = (cx *)malloc(sizeof(cx) * R); vv = heap; }
    for (long j = 0; j < m; j++) {
      long k = j % Ns;
      long tstep = (long)n / (Ns * R);
      for (int r = 0; r < R; r++) {
        cx a = in[j + r * m];
        if (r && k) {
          cx w = p->tw[(k * r * tstep) % n];
          if (sign > 0) w = cx_conj(w);
          a = cx_mul(a, w);
        }
        vv[r] = a;
      }
      dft_small(vv, R, sign, p);
      long idxD = (j / Ns) * Ns * R + k;
      for (int r = 0; r < R; r++) out[idxD + r * Ns] = vv[r];
    }
    free(heap);
    cx *t = in; in = out; out = t;
    Ns *= R;
  }
  if (in != x) memcpy(x, in, sizeof(cx) * n);
}

/* ---- per-length transform context -------------------------------------- */
typedef struct {
  int kind, n;
  cplan cp;        /* main complex plan */
  cx *post;        /* auxiliary twiddles */
} rplan;

static void rplan_init(rplan *rp, int kind, int n) {
  memset(rp, 0, sizeof(*rp));
  rp->kind = kind; rp->n = n;
  switch (kind) {
  case ORC_R2HC: case ORC_HC2R:
    if (n % 2 == 0 && !g_variant) {
      cplan_init(&rp->cp, n / 2);
      rp->post = (cx *)malloc(sizeof(cx) * (n / 2 + 1));
      for (int k = 0; k <= n / 2; k++) rp->post[k] = twiddle(k, n);
    } else {
      cplan_init(&rp->cp, n);
    }
    break;
  case ORC_REDFT10: case ORC_REDFT01: case ORC_RODFT10: case ORC_RODFT01:
    cplan_init(&rp->cp, n);
    rp->post = (cx *)malloc(sizeof(cx) * (n + 1));
    for (int k = 0; k <= n; k++) rp->post[k] = twiddle(k, 4L * n); /* e^{-i pi k/(2n)} */
    break;
  case ORC_REDFT11: case ORC_RODFT11:
    cplan_init(&rp->cp, 2 * n);
    rp->post = (cx *)malloc(sizeof(cx) * (2 * n + 1));
    for (int k = 0; k < n; k++) rp->post[k] = twiddle(k, 4L * n);            /* e^{-i pi j/(2n)} */
    for (int k = 0; k < n; k++) rp->post[n + k] = twiddle(2L * k + 1, 8L * n); /* e^{-i pi (2k+1)/(4n)} */
    break;
  case ORC_REDFT00:
    if (n > 1) cplan_init(&rp->cp, 2 * (n - 1));
    break;
  case ORC_RODFT00:
    cplan_init(&rp->cp, 2 * (n + 1));
    break;
  default: break;
  }
}
static void rplan_free(rplan *rp) { cplan_free(&rp->cp); free(rp->post); rp->post = NULL; }

/* one line; x has n contiguous values; w1,w2 scratch of >= 2n+4 complex */
static void r2r_line(const rplan *rp, double *x, cx *w1, cx *w2) {
  int n = rp->n;
  switch (rp->kind) {
  case ORC_R2HC: {
    if (n == 1) return;
    if (n % 2 == 0 && rp->post) {
      int h = n / 2;
      for (int m = 0; m < h; m++) { w1[m].re = x[2 * m]; w1[m].im = x[2 * m + 1]; }
      cfft(&rp->cp, w1, w2, -1);
      /* X_k = (Z_k + conj Z_{h-k})/2 - i W^k (Z_k - conj Z_{h-k})/2 */
      for (int k = 0; k <= h; k++) {
        cx zk = w1[k % h], zc = cx_conj(w1[(h - k) % h]);
        cx e = {0.5 * (zk.re + zc.re), 0.5 * (zk.im + zc.im)};
        cx o = {0.5 * (zk.re - zc.re), 0.5 * (zk.im - zc.im)};
        cx wo = cx_mul(rp->post[k], o);
        cx X = {e.re + wo.im, e.im - wo.re};
        w2[k] = X;
      }
      x[0] = w2[0].re;
      for (int k = 1; k < h; k++) { x[k] = w2[k].re; x[n - k] = w2[k].im; }
      x[h] = w2[h].re;
    } else {
      for (int j = 0; j < n; j++) { w1[j].re = x[j]; w1[j].im = 0; }
      cfft(&rp->cp, w1, w2, -1);
      x[0] = w1[0].re;
      for (int k = 1; k <= (n - 1) / 2; k++) { x[k] = w1[k].re; x[n - k] = w1[k].im; }
      if (n % 2 == 0) x[n / 2] = w1[n / 2].re;
    }
    return;
  }
  case ORC_HC2R: {
    if (n == 1) return;
    if (n % 2 == 0 && rp->post) {
      int h = n / 2;
      /* unpack halfcomplex */
      w2[0].re = x[0]; w2[0].im = 0;
      for (int k = 1; k < h; k++) { w2[k].re = x[k]; w2[k].im = x[n - k]; }
      w2[h].re = x[h]; w2[h].im = 0;
      /* Z_k = (X_k + conj X_{h-k}) + i conj(W^k) (X_k - conj X_{h-k}) */
      for (int k = 0; k < h; k++) {
        cx xk = w2[k], xc = cx_conj(w2[h - k]);
        cx e = cx_add(xk, xc), d = cx_sub(xk, xc);
        cx od = cx_mul(cx_conj(rp->post[k]), d);
        w1[k].re = e.re - od.im; w1[k].im = e.im + od.re;
      }
      cfft(&rp->cp, w1, w2, +1);
      for (int m = 0; m < h; m++) { x[2 * m] = w1[m].re; x[2 * m + 1] = w1[m].im; }
    } else {
      w1[0].re = x[0]; w1[0].im = 0;
      for (int k = 1; k <= (n - 1) / 2; k++) {
        w1[k].re = x[k]; w1[k].im = x[n - k];
        w1[n - k].re = x[k]; w1[n - k].im = -x[n - k];
      }
      if (n % 2 == 0) { w1[n / 2].re = x[n / 2]; w1[n / 2].im = 0; }
      cfft(&rp->cp, w1, w2, +1);
      for (int j = 0; j < n; j++) x[j] = w1[j].re;
    }
    return;
  }
  case ORC_REDFT10: case ORC_RODFT10: {
    /* RODFT10(x)_k = REDFT10((-1)^j x_j)_{n-1-k} */
    int odd = rp->kind == ORC_RODFT10;
    for (int m = 0; 2 * m < n; m++) {
      double a = x[2 * m];
      w1[m].re = a; w1[m].im = 0;
    }
    for (int m = 0; 2 * m + 1 < n; m++) {
      double b = x[2 * m + 1];
      if (odd) b = -b;
      w1[n - 1 - m].re = b; w1[n - 1 - m].im = 0;
    }
    cfft(&rp->cp, w1, w2, -1);
    for (int k = 0; k < n; k++) {
      cx t = cx_mul(rp->post[k], w1[k]);
      double y = 2.0 * t.re;
      if (odd) x[n - 1 - k] = y; else x[k] = y;
    }
    return;
  }
  case ORC_REDFT01: case ORC_RODFT01: {
    /* inverse Makhoul: REDFT01(X)_j from V_k = e^{i pi k/(2n)} (X_k - i X_{n-k}), X_n = 0,
     * v = IDFT(V), y_{2m} = v_m, y_{2m+1} = v_{n-1-m}.
     * RODFT01(x)_k = (-1)^k REDFT01(reverse x)_k */
    int odd = rp->kind == ORC_RODFT01;
    for (int k = 0; k < n; k++) {
      double xk = odd ? x[n - 1 - k] : x[k];
      double xnk = (k == 0) ? 0.0 : (odd ? x[k - 1] : x[n - k]);
      cx t = {xk, -xnk};
      w1[k] = cx_mul(cx_conj(rp->post[k]), t);
    }
    cfft(&rp->cp, w1, w2, +1);
    for (int m = 0; 2 * m < n; m++) x[2 * m] = w1[m].re;
    for (int m = 0; 2 * m + 1 < n; m++) x[2 * m + 1] = w1[n - 1 - m].re;
    if (odd) for (int k = 1; k < n; k += 2) x[k] = -x[k];
    return;
  }
  case ORC_REDFT11: case ORC_RODFT11: {
    int n2 = 2 * n;
    for (int j = 0; j < n; j++) { cx t = {x[j], 0}; w1[j] = cx_mul(rp->post[j], t); }
    for (int j = n; j < n2; j++) { w1[j].re = 0; w1[j].im = 0; }
    cfft(&rp->cp, w1, w2, -1);
    for (int k = 0; k < n; k++) {
      cx t = cx_mul(rp->post[n + k], w1[k]);
      x[k] = (rp->kind == ORC_REDFT11) ? 2.0 * t.re : -2.0 * t.im;
    }
    return;
  }
  case ORC_REDFT00: {
    if (n == 1) return;
    int M = 2 * (n - 1);
    for (int j = 0; j < n; j++) { w1[j].re = x[j]; w1[j].im = 0; }
    for (int j = n; j < M; j++) { w1[j].re = x[M - j]; w1[j].im = 0; }
    cfft(&rp->cp, w1, w2, -1);
    for (int k = 0; k < n; k++) x[k] = w1[k].re;
    return;
  }
  case ORC_RODFT00: {
    int M = 2 * (n + 1);
    w1[0].re = 0; w1[0].im = 0;
    for (int j = 0; j < n; j++) { w1[j + 1].re = x[j]; w1[j + 1].im = 0; }
    w1[n + 1].re = 0; w1[n + 1].im = 0;
    for (int j = 0; j < n; j++) { w1[n + 2 + j].re = -x[n - 1 - j]; w1[n + 2 + j].im = 0; }
    (void)M;
    cfft(&rp->cp, w1, w2, -1);
    for (int k = 0; k < n; k++) x[k] = -w1[k + 1].im;
    return;
  }
  default: return;
  }
}

int orc_r2r(int kind, int n, long howmany, long stride, long dist, double *data) {
  if (n <= 0 || howmany < 0) return -1;
  if (kind == ORC_DHT || kind < 0 || kind > ORC_RODFT11) return -2;
  rplan rp;
  rplan_init(&rp, kind, n);
  #pragma omp parallel
  {
    cx *w1 = (cx *)malloc(sizeof(cx) * (4 * (size_t)n + 8));
    cx *w2 = (cx *)malloc(sizeof(cx) * (4 * (size_t)n + 8));
    double *line = (double *)malloc(sizeof(double) * ((size_t)n + 1));
    #pragma omp for schedule(static)
    for (long l = 0; l < howmany; l++) {
      double *base = data + l * dist;
      if (stride == 1) {
        r2r_line(&rp, base, w1, w2);
      } else {
        for (int e = 0; e < n; e++) line[e] = base[e * stride];
        r2r_line(&rp, line, w1, w2);
        for (int e = 0; e < n; e++) base[e * stride] = line[e];
      }
    }
    free(w1); free(w2); free(line);
  }
  rplan_free(&rp);
  return 0;
}
