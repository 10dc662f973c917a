// Poisson/Helmholtz solver (src/solver.f90:17-112 and src/solver_gpu.f90:34-276):
// batched r2r transforms along x and y, Thomas elimination along z, back.
//
// Data stay x-contiguous for the whole solve, so no pencil transposes exist:
//   pass X  : rows along x are contiguous; one complex line carries two rows.
//   pass Y  : a work-group holds a tile of x-columns x all y of one z-plane in LDS.
//   pass Z  : one thread per (kx,ky) line, lines adjacent in x -> coalesced rows.
// Spectral ordering matches the reference's GPU build: periodic directions are
// stored packed as (r0, r_{n/2}, r1, i1, ..., r_{n/2-1}, i_{n/2-1}) (fft.f90:402-451,
// initsolver.f90:98-117); DCT/DST directions in natural order.
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "fft.hpp"
#include "kernels.hpp"

// transform family per direction
enum { K_PP = 0, K_NN = 1, K_DD = 2 };

struct cans_plan {
  cans_ctx* ctx;
  int ng[3];
  int len[2];        // transform lengths (x, y)
  int logn[2];
  int kind[2];
  double normfft;
  cans::cplx<double>* tw[2];  // W_N^t = exp(-2 pi i t/N), t < N
  cans::cplx<double>* mk[2];  // exp(-i pi k/(2N)), k <= N (DCT/DST pre/post twiddles)
};

namespace cans {

static int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) l++;
  return (1 << l) == n ? l : -1;
}

// position of input element e in Makhoul's even/odd reordering (v_m = x_{2m}, v_{N-1-m} = x_{2m+1})
__device__ __forceinline__ int mk_pos(int e, int N) { return (e & 1) ? N - 1 - (e >> 1) : (e >> 1); }

// ---------------------------------------------------------------------------
// forward: real lines A, B -> complex line z in LDS (pre-processing)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void fwd_pre(cplx<T>* buf, int S, int N, int lt, int Tn, int kind, const T* xa,
                                        const T* xb, long est) {
  for (int e = lt; e < N; e += Tn) {
    T a = xa[e * est];
    T b = xb ? xb[e * est] : (T)0;
    if (kind == K_PP) {
      buf[e * S] = {a, b};
    } else {
      if (kind == K_DD && (e & 1)) {  // RODFT10 = REDFT10 of (-1)^j x_j, reversed output
        a = -a;
        b = -b;
      }
      buf[mk_pos(e, N) * S] = {a, b};
    }
  }
}

// forward post-processing: complex spectrum Z of (A + iB) -> spectra of A and B.
// Output element index o of line A is written to ya[o*est] (B likewise).
template <typename T>
__device__ __forceinline__ void fwd_post(const cplx<T>* buf, int S, int N, int lt, int Tn, int kind,
                                         const cplx<T>* __restrict__ mk, T* ya, T* yb, long est) {
  if (kind == K_PP) {
    // A_k = (Z_k + conj Z_{N-k})/2, B_k = -i (Z_k - conj Z_{N-k})/2; packed output
    const int H = N >> 1;
    if (N == 1) {
      if (lt == 0) {
        ya[0] = buf[0].re;
        if (yb) yb[0] = buf[0].im;
      }
      return;
    }
    for (int k = lt; k < H; k += Tn) {
      if (k == 0) {
        const cplx<T> z0 = buf[0], zh = buf[H * S];
        ya[0] = z0.re;
        ya[est] = zh.re;
        if (yb) {
          yb[0] = z0.im;
          yb[est] = zh.im;
        }
      } else {
        const cplx<T> zk = buf[k * S], zc = buf[(N - k) * S];
        const T are = (T)0.5 * (zk.re + zc.re), aim = (T)0.5 * (zk.im - zc.im);
        const T bre = (T)0.5 * (zk.im + zc.im), bim = -(T)0.5 * (zk.re - zc.re);
        ya[(2 * k) * est] = are;
        ya[(2 * k + 1) * est] = aim;
        if (yb) {
          yb[(2 * k) * est] = bre;
          yb[(2 * k + 1) * est] = bim;
        }
      }
    }
    return;
  }
  // REDFT10: Y_k = 2 Re(M_k V_k), V_A = (Z_k + conj Z_{N-k})/2 -> Y_k = Re(M_k (Z_k + conj Z_{N-k}))
  for (int k = lt; k < N; k += Tn) {
    const cplx<T> zk = buf[k * S];
    const cplx<T> zc = buf[((N - k) & (N - 1)) * S];
    const cplx<T> m = mk[k];
    // A: Z_k + conj Z_{N-k};  B: -i (Z_k - conj Z_{N-k})
    const T sre = zk.re + zc.re, sim = zk.im - zc.im;
    const T dre = zk.im + zc.im, dim = -(zk.re - zc.re);
    const T ya_ = m.re * sre - m.im * sim;
    const T yb_ = m.re * dre - m.im * dim;
    const int o = kind == K_DD ? N - 1 - k : k;
    ya[o * est] = ya_;
    if (yb) yb[o * est] = yb_;
  }
}

// ---------------------------------------------------------------------------
// backward: spectra of A, B -> complex Z in LDS (pre-processing)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void bwd_pre(cplx<T>* buf, int S, int N, int lt, int Tn, int kind,
                                        const cplx<T>* __restrict__ mk, const T* xa, const T* xb, long est) {
  if (kind == K_PP) {
    const int H = N >> 1;
    for (int k = lt; k < N; k += Tn) {
      // A_k from the packed halfcomplex line, with A_{N-k} = conj(A_k)
      T are, aim, bre = 0, bim = 0;
      if (N == 1 || k == 0) {
        are = xa[0];
        aim = 0;
        if (xb) bre = xb[0];
      } else if (k == H) {
        are = xa[est];
        aim = 0;
        if (xb) bre = xb[est];
      } else {
        const int kk = k < H ? k : N - k;
        const T sg = k < H ? (T)1 : (T)-1;
        are = xa[(2 * kk) * est];
        aim = sg * xa[(2 * kk + 1) * est];
        if (xb) {
          bre = xb[(2 * kk) * est];
          bim = sg * xb[(2 * kk + 1) * est];
        }
      }
      buf[k * S] = {are - bim, aim + bre};  // Z = A + iB
    }
    return;
  }
  // REDFT01: V_k = conj(M_k) (X_k - i X_{N-k}), X_N = 0; RODFT01 reads the reversed input
  for (int k = lt; k < N; k += Tn) {
    T xk_a, xnk_a, xk_b = 0, xnk_b = 0;
    if (kind == K_DD) {
      xk_a = xa[(N - 1 - k) * est];
      xnk_a = k ? xa[(k - 1) * est] : (T)0;
      if (xb) {
        xk_b = xb[(N - 1 - k) * est];
        xnk_b = k ? xb[(k - 1) * est] : (T)0;
      }
    } else {
      xk_a = xa[k * est];
      xnk_a = k ? xa[(N - k) * est] : (T)0;
      if (xb) {
        xk_b = xb[k * est];
        xnk_b = k ? xb[(N - k) * est] : (T)0;
      }
    }
    const cplx<T> m = mk[k];
    // conj(m) * (x, -y) = (m.re x - m.im y, -m.re y - m.im x)
    const T va_re = m.re * xk_a - m.im * xnk_a, va_im = -m.re * xnk_a - m.im * xk_a;
    const T vb_re = m.re * xk_b - m.im * xnk_b, vb_im = -m.re * xnk_b - m.im * xk_b;
    buf[k * S] = {va_re - vb_im, va_im + vb_re};
  }
}

template <typename T>
__device__ __forceinline__ void bwd_post(const cplx<T>* buf, int S, int N, int lt, int Tn, int kind, T* ya, T* yb,
                                         long est) {
  for (int e = lt; e < N; e += Tn) {
    cplx<T> z = kind == K_PP ? buf[e * S] : buf[mk_pos(e, N) * S];
    if (kind == K_DD && (e & 1)) {
      z.re = -z.re;
      z.im = -z.im;
    }
    ya[e * est] = z.re;
    if (yb) yb[e * est] = z.im;
  }
}

// ---------------------------------------------------------------------------
// pass X: rows of the halo'd field (or the work buffer) along x.
// rows r = j + n2*k (0-based); a complex line carries rows 2q and 2q+1.
// ---------------------------------------------------------------------------
struct XArgs {
  Layout L;
  double* fld;     // halo'd field (origin)
  double* W;       // compact work buffer, row r at W + r*n1
  int N, logN, Tn, pairs, kind;
  const cplx<double>* tw;
  const cplx<double>* mk;
};

template <int DIR>  // +1 forward (fld -> W), -1 backward (W -> fld)
__global__ void k_xpass(XArgs a) {
  extern __shared__ unsigned char smem_raw[];
  cplx<double>* smem = (cplx<double>*)smem_raw;
  const int team = threadIdx.x / a.Tn, lt = threadIdx.x % a.Tn;
  const long nrows = (long)a.L.n2 * a.L.n3;
  const long q = (long)blockIdx.x * a.pairs + team;
  const long ra = 2 * q, rb = 2 * q + 1;
  const bool va = ra < nrows, vb = rb < nrows;
  cplx<double>* buf = smem + (long)team * a.N;
  const Layout& L = a.L;
  auto frow = [&](long r) { return a.fld + L.at(1, (int)(r % L.n2) + 1, (int)(r / L.n2) + 1); };
  auto wrow = [&](long r) { return a.W + r * L.n1; };
  if (DIR > 0) {
    if (va) fwd_pre<double>(buf, 1, a.N, lt, a.Tn, a.kind, frow(ra), vb ? frow(rb) : nullptr, 1);
    __syncthreads();
    fft_line<double, -1>(buf, 1, a.N, a.logN, lt, a.Tn, a.tw);
    if (va) fwd_post<double>(buf, 1, a.N, lt, a.Tn, a.kind, a.mk, wrow(ra), vb ? wrow(rb) : nullptr, 1);
  } else {
    if (va) bwd_pre<double>(buf, 1, a.N, lt, a.Tn, a.kind, a.mk, wrow(ra), vb ? wrow(rb) : nullptr, 1);
    __syncthreads();
    fft_line<double, +1>(buf, 1, a.N, a.logN, lt, a.Tn, a.tw);
    if (va) bwd_post<double>(buf, 1, a.N, lt, a.Tn, a.kind, frow(ra), vb ? frow(rb) : nullptr, 1);
  }
}

// ---------------------------------------------------------------------------
// pass Y: in place on W; a work-group owns 2*LINES adjacent x-columns of one z-plane.
// thread t -> line c = t % LINES, team thread lt = t / LINES; LDS element (e, c) at [e*LINES + c].
// ---------------------------------------------------------------------------
struct YArgs {
  int n1, n2, n3;
  double* W;
  int N, logN, Tn, lines, kind;
  const cplx<double>* tw;
  const cplx<double>* mk;
};

template <int DIR>
__global__ void k_ypass(YArgs a) {
  extern __shared__ unsigned char smem_raw[];
  cplx<double>* smem = (cplx<double>*)smem_raw;
  const int LN = a.lines;
  const int c = threadIdx.x % LN, lt = threadIdx.x / LN;
  const int ncolblk = (a.n1 + 2 * LN - 1) / (2 * LN);
  const int k = blockIdx.x / ncolblk;
  const int i0 = (blockIdx.x % ncolblk) * 2 * LN;
  const int ia = i0 + 2 * c, ib = ia + 1;
  const bool va = ia < a.n1, vb = ib < a.n1;
  double* plane = a.W + (long)k * a.n1 * a.n2;
  const long est = a.n1;  // element stride along y
  cplx<double>* buf = smem + c;
  if (DIR > 0) {
    if (va) fwd_pre<double>(buf, LN, a.N, lt, a.Tn, a.kind, plane + ia, vb ? plane + ib : nullptr, est);
    __syncthreads();
    fft_line<double, -1>(buf, LN, a.N, a.logN, lt, a.Tn, a.tw);
    if (va) fwd_post<double>(buf, LN, a.N, lt, a.Tn, a.kind, a.mk, plane + ia, vb ? plane + ib : nullptr, est);
  } else {
    if (va) bwd_pre<double>(buf, LN, a.N, lt, a.Tn, a.kind, a.mk, plane + ia, vb ? plane + ib : nullptr, est);
    __syncthreads();
    fft_line<double, +1>(buf, LN, a.N, a.logN, lt, a.Tn, a.tw);
    if (va) bwd_post<double>(buf, LN, a.N, lt, a.Tn, a.kind, plane + ia, vb ? plane + ib : nullptr, est);
  }
}

// ---------------------------------------------------------------------------
// pass Z: Thomas elimination per (kx,ky) line, gaussel (solver.f90:114-307)
// with identical arithmetic.  The backward sweep needs d(k) and the forward-
// modified rhs in reverse order; instead of storing them for every k
// (3 extra field passes), the forward sweep checkpoints them every CK rows and
// the backward sweep re-reads the rhs of one block, recomputes the block's
// d/rhs from its checkpoint (bit-identical: same operations) and substitutes.
// ---------------------------------------------------------------------------
constexpr int CK = 16;

struct ZArgs {
  double* W;
  long nl;         // number of lines (n1*n2) = stride between consecutive k
  int n;           // rows of the system (n3 - q)
  int periodic;
  double norm;
  const double *a, *b, *c, *lam;
  double *ckd, *ckp, *ckq;  // checkpoints [nblk][nl]
  double* X2;               // periodic: auxiliary solution p2 [nn][nl]
};

__global__ void __launch_bounds__(256) k_gaussel(ZArgs z) {
  const long line = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= z.nl) return;
  const double eps = 2.220446049250313e-16;
  const double l = z.lam[line];
  double* P = z.W + line;
  const long ps = z.nl;
  const int nn = z.periodic ? z.n - 1 : z.n;
  const int nblk = (nn + CK - 1) / CK;
  const double* a = z.a;
  const double* b = z.b;
  const double* c = z.c;
  const double norm = z.norm;
  // ---- forward sweep: checkpoints only ----
  double d, pp, q2 = 0.;
  {
    const double zz = 1. / (b[0] + l);
    d = c[0] * zz;
    pp = P[0] * norm * zz;
    if (z.periodic) q2 = ((nn == 1) ? (-a[0] - c[nn - 1]) : -a[0]) * zz;
  }
  if (CK == 1 || nn == 1) {
    z.ckd[line] = d;
    z.ckp[line] = pp;
    if (z.periodic) z.ckq[line] = q2;
  }
  for (int k = 2; k <= nn; k++) {
    const double r = P[(long)(k - 1) * ps];
    const double bl = b[k - 1] + l;
    const double ad = a[k - 1] * d;
    const double den = bl - ad;
    if (z.periodic) {
      const double zq = 1. / den;
      const double rq = (k == nn) ? -c[nn - 1] : 0.;
      q2 = (rq - a[k - 1] * q2) * zq;
    }
    const double tol = eps * fmax(fabs(bl), fabs(ad));
    if (k == nn && fabs(den) <= tol) {
      d = 0.;
      pp = 0.;
    } else {
      const double zz = 1. / den;
      d = c[k - 1] * zz;
      pp = (r * norm - a[k - 1] * pp) * zz;
    }
    if ((k % CK) == 0 || k == nn) {
      const long o = (long)((k - 1) / CK) * ps + line;
      z.ckd[o] = d;
      z.ckp[o] = pp;
      if (z.periodic) z.ckq[o] = q2;
    }
  }
  // ---- backward sweep, one block of CK rows at a time ----
  double xn = 0., xq = 0.;
  double x1 = 0., xq1 = 0., xnn = 0., xqnn = 0.;
  for (int bb = nblk - 1; bb >= 0; bb--) {
    const int k0 = bb * CK + 1;
    const int k1 = min(k0 + CK - 1, nn);
    double rr[CK], dd[CK], pv[CK], qv[CK];
#pragma unroll
    for (int s = 0; s < CK; s++) rr[s] = (k0 + s <= k1) ? P[(long)(k0 + s - 1) * ps] : 0.;
    double dp, pq, qq;
    if (bb > 0) {
      const long o = (long)(bb - 1) * ps + line;
      dp = z.ckd[o];
      pq = z.ckp[o];
      qq = z.periodic ? z.ckq[o] : 0.;
    } else {
      dp = pq = qq = 0.;
    }
#pragma unroll
    for (int s = 0; s < CK; s++) {
      const int k = k0 + s;
      if (k <= k1) {
        if (k == 1) {
          const double zz = 1. / (b[0] + l);
          dp = c[0] * zz;
          pq = rr[s] * norm * zz;
          if (z.periodic) qq = ((nn == 1) ? (-a[0] - c[nn - 1]) : -a[0]) * zz;
        } else {
          const double bl = b[k - 1] + l;
          const double ad = a[k - 1] * dp;
          const double den = bl - ad;
          if (z.periodic) {
            const double zq = 1. / den;
            const double rq = (k == nn) ? -c[nn - 1] : 0.;
            qq = (rq - a[k - 1] * qq) * zq;
          }
          const double tol = eps * fmax(fabs(bl), fabs(ad));
          if (k == nn && fabs(den) <= tol) {
            dp = 0.;
            pq = 0.;
          } else {
            const double zz = 1. / den;
            dp = c[k - 1] * zz;
            pq = (rr[s] * norm - a[k - 1] * pq) * zz;
          }
        }
        dd[s] = dp;
        pv[s] = pq;
        qv[s] = qq;
      }
    }
#pragma unroll
    for (int s = CK - 1; s >= 0; s--) {
      const int k = k0 + s;
      if (k <= k1) {
        double x, x2 = 0.;
        if (k == nn) {
          x = pv[s];
          x2 = qv[s];
          xnn = x;
          xqnn = x2;
        } else {
          x = pv[s] - dd[s] * xn;
          if (z.periodic) x2 = qv[s] - dd[s] * xq;
        }
        P[(long)(k - 1) * ps] = x;
        if (z.periodic) z.X2[(long)(k - 1) * ps + line] = x2;
        xn = x;
        xq = x2;
        if (k == 1) {
          x1 = x;
          xq1 = x2;
        }
      }
    }
  }
  if (!z.periodic) return;
  // ---- periodic closure (solver.f90:268-296) and Sherman-Morrison correction ----
  const int m = nn + 1;  // closure row, 1-based
  const double den = b[m - 1] + l + c[m - 1] * xq1 + a[m - 1] * xqnn;
  const double tol = eps * fmax(fabs(b[m - 1] + l), fabs(c[m - 1] * xq1 + a[m - 1] * xqnn));
  double xm;
  if (fabs(den) <= tol) xm = 0.;
  else xm = (P[(long)(m - 1) * ps] * norm - c[m - 1] * x1 - a[m - 1] * xnn) / den;
  P[(long)(m - 1) * ps] = xm;
  for (int k = 1; k <= nn; k++) {
    const long o = (long)(k - 1) * ps;
    P[o] = P[o] + z.X2[o + line] * xm;
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int upload_tables(cans_plan* p, int d) {
  const int N = p->len[d];
  std::vector<cplx<double>> tw(N > 0 ? N : 1), mk(N + 1);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int t = 0; t < N; t++) {
    long double ang = 2.0L * pi * (long double)t / (long double)N;
    tw[t] = {(double)cosl(ang), -(double)sinl(ang)};
    if (4 * t % N == 0) {  // exact quarter turns
      const int qq = 4 * t / N;
      tw[t] = {qq == 0 ? 1. : qq == 2 ? -1. : 0., qq == 1 ? -1. : qq == 3 ? 1. : 0.};
    }
  }
  for (int k = 0; k <= N; k++) {
    long double ang = pi * (long double)k / (2.0L * (long double)N);
    mk[k] = {(double)cosl(ang), -(double)sinl(ang)};
  }
  CANS_HIP_TRY(hipMalloc((void**)&p->tw[d], tw.size() * sizeof(cplx<double>)));
  CANS_HIP_TRY(hipMalloc((void**)&p->mk[d], mk.size() * sizeof(cplx<double>)));
  CANS_HIP_TRY(hipMemcpy(p->tw[d], tw.data(), tw.size() * sizeof(cplx<double>), hipMemcpyHostToDevice));
  CANS_HIP_TRY(hipMemcpy(p->mk[d], mk.data(), mk.size() * sizeof(cplx<double>), hipMemcpyHostToDevice));
  return CANS_OK;
}

static int xpass(cans_ctx* c, const cans_plan* p, double* fld, double* W, int dir) {
  XArgs a;
  a.L = c->L;
  a.fld = fld;
  a.W = W;
  a.N = p->len[0];
  a.logN = p->logn[0];
  a.Tn = a.N >= 8 ? a.N / 8 : 1;
  a.pairs = a.Tn >= 256 ? 1 : 256 / a.Tn;
  a.kind = p->kind[0];
  a.tw = p->tw[0];
  a.mk = p->mk[0];
  const long nrows = (long)c->L.n2 * c->L.n3;
  const long npairs = (nrows + 1) / 2;
  dim3 grid(cdiv(npairs, a.pairs));
  const size_t shm = (size_t)a.pairs * a.N * sizeof(cplx<double>);
  if (dir > 0) k_xpass<1><<<grid, a.pairs * a.Tn, shm, c->stream>>>(a);
  else k_xpass<-1><<<grid, a.pairs * a.Tn, shm, c->stream>>>(a);
  CANS_LAUNCH_CHECK("k_xpass");
  return CANS_OK;
}

static int ypass(cans_ctx* c, const cans_plan* p, double* W, int dir) {
  YArgs a;
  a.n1 = c->L.n1;
  a.n2 = c->L.n2;
  a.n3 = c->L.n3;
  a.W = W;
  a.N = p->len[1];
  a.logN = p->logn[1];
  a.Tn = a.N >= 8 ? a.N / 8 : 1;
  int lines = 256 / a.Tn;
  if (lines < 8) lines = 8;
  while (lines > 1 && (size_t)lines * a.N * sizeof(cplx<double>) > 65536) lines /= 2;
  while ((long)lines * a.Tn > 1024) lines /= 2;
  a.lines = lines;
  a.kind = p->kind[1];
  a.tw = p->tw[1];
  a.mk = p->mk[1];
  const int ncolblk = (a.n1 + 2 * lines - 1) / (2 * lines);
  dim3 grid((unsigned)ncolblk * a.n3);
  const size_t shm = (size_t)lines * a.N * sizeof(cplx<double>);
  if (dir > 0) k_ypass<1><<<grid, lines * a.Tn, shm, c->stream>>>(a);
  else k_ypass<-1><<<grid, lines * a.Tn, shm, c->stream>>>(a);
  CANS_LAUNCH_CHECK("k_ypass");
  return CANS_OK;
}

int run_solver(cans_ctx* c, const cans_plan* p, double normfft, const double* lam, const double* a,
               const double* b, const double* cc, const char* bc, const char* c_or_f, double* fld) {
  const Layout& L = c->L;
  const long N = L.cells();
  const long nl = (long)L.n1 * L.n2;
  void* W = nullptr;
  int rc = ensure_work(c, 1, (size_t)N * sizeof(double), &W);
  if (rc) return rc;
  const int q = (c_or_f[2] == 'f' && bc[5] == 'D') ? 1 : 0;
  const int periodic = bc[4] == 'P' && bc[5] == 'P';
  const int n = L.n3 - q;
  const int nn = periodic ? n - 1 : n;
  if (nn < 1) return fail(CANS_EUNSUPPORTED, "cans_solver: fewer than 2 z rows");
  const long nblk = (nn + CK - 1) / CK;
  void* ck = nullptr;
  size_t ckbytes = (size_t)3 * nblk * nl * sizeof(double) + (periodic ? (size_t)N * sizeof(double) : 0);
  rc = ensure_work(c, 2, ckbytes, &ck);
  if (rc) return rc;
  if ((rc = xpass(c, p, fld, (double*)W, +1))) return rc;
  if ((rc = ypass(c, p, (double*)W, +1))) return rc;
  ZArgs z;
  z.W = (double*)W;
  z.nl = nl;
  z.n = n;
  z.periodic = periodic;
  z.norm = normfft;
  z.a = a;
  z.b = b;
  z.c = cc;
  z.lam = lam;
  z.ckd = (double*)ck;
  z.ckp = z.ckd + nblk * nl;
  z.ckq = z.ckp + nblk * nl;
  z.X2 = periodic ? z.ckq + nblk * nl : nullptr;
  k_gaussel<<<cdiv(nl, 256), 256, 0, c->stream>>>(z);
  CANS_LAUNCH_CHECK("k_gaussel");
  if ((rc = ypass(c, p, (double*)W, -1))) return rc;
  if ((rc = xpass(c, p, fld, (double*)W, -1))) return rc;
  return CANS_OK;
}

}  // namespace cans

using namespace cans;

extern "C" int cans_fftini(cans_ctx* c, const int* ng, const char* bcxy, const char* c_or_f, cans_plan** out,
                           double* normfft) {
  if (!c || !ng || !bcxy || !c_or_f || !out) return fail(CANS_EINVAL, "cans_fftini: null argument");
  if (ng[0] != c->L.n1 || ng[1] != c->L.n2 || ng[2] != c->L.n3)
    return fail(CANS_EINVAL, "cans_fftini: ng differs from the context grid (one GPU holds the whole domain)");
  cans_plan* p = new cans_plan();
  p->ctx = c;
  for (int d = 0; d < 3; d++) p->ng[d] = ng[d];
  double nf = 1.;
  for (int d = 0; d < 2; d++) {
    const char b0 = bcxy[2 * d], b1 = bcxy[2 * d + 1];
    if (c_or_f[d] != 'c') {
      delete p;
      return fail(CANS_EUNSUPPORTED, "cans_fftini: face-centred transforms are not implemented on the device");
    }
    if (b0 == 'P' && b1 == 'P') p->kind[d] = K_PP;
    else if (b0 == 'N' && b1 == 'N') p->kind[d] = K_NN;
    else if (b0 == 'D' && b1 == 'D') p->kind[d] = K_DD;
    else {
      delete p;
      return fail(CANS_EUNSUPPORTED, "cans_fftini: mixed N/D transforms are not implemented on the device");
    }
    p->len[d] = ng[d];
    p->logn[d] = ilog2(ng[d]);
    if (p->logn[d] < 0 || ng[d] > 4096) {
      delete p;
      return fail(CANS_EUNSUPPORTED, "cans_fftini: transform lengths must be powers of two <= 4096");
    }
    // fft.f90:75,140: normfft = 1/prod(norm(1)*(ng+norm(2)-i)), norm = [1,0] (PP) or [2,0] (NN/DD, c-grid)
    nf = nf * (p->kind[d] == K_PP ? 1. : 2.) * (ng[d] + 0.);
    int rc = upload_tables(p, d);
    if (rc) {
      cans_fftend(p);
      return rc;
    }
  }
  p->normfft = 1. / nf;
  if (normfft) *normfft = p->normfft;
  *out = p;
  return CANS_OK;
}

extern "C" int cans_fftend(cans_plan* p) {
  if (!p) return CANS_OK;
  for (int d = 0; d < 2; d++) {
    hipFree(p->tw[d]);
    hipFree(p->mk[d]);
  }
  delete p;
  return CANS_OK;
}
