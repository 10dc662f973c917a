// Poisson/Helmholtz solver (src/solver.f90:17-112 and src/solver_gpu.f90:34-276):
// batched r2r transforms along x and y, Thomas elimination along z, back.
//
// Data stay x-contiguous for the whole solve, so no pencil transposes exist:
//   pass X  : rows along x are contiguous; one complex line carries two rows.
//   pass Y  : a work-group holds a tile of x-columns x all y of a z-plane in LDS.
//   pass Z  : one thread per (kx,ky) line, lines adjacent in x -> coalesced rows.
// Spectral ordering matches the reference's GPU build: periodic directions are
// stored packed as (r0, r_{n/2}, r1, i1, ..., r_{n/2-1}, i_{n/2-1}) (fft.f90:402-451,
// initsolver.f90:98-117); DCT/DST directions in natural order.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "fft.hpp"
#include "kernels.hpp"

// transform family per direction
// K_ND, K_DN: REDFT11 and RODFT11 (cell-centred N-D and D-N pairs, fft.f90:280-287), computed on a
// line of twice the data length (below).  Face-centred pairs (fft.f90:289-311): PP as K_PP, ND as
// K_NN (REDFT10 forward, REDFT01 backward), DN as K_DNF (RODFT01 forward, RODFT10 backward: the
// K_DD pair with the directions swapped), DD as K_DDF (RODFT00 on the n-1 interior faces,
// fft.f90:80-82, by the odd extension of the line to 2n points; the boundary face n passes through).
enum { K_PP = 0, K_NN = 1, K_DD = 2, K_ND = 3, K_DN = 4, K_DNF = 5, K_DDF = 6 };
// kinds whose lines hold twice the data length
__host__ __device__ __forceinline__ bool k_double(int kind) { return kind == K_ND || kind == K_DN || kind == K_DDF; }

struct cans_plan {
  cans_ctx* ctx;
  int ng[3];
  int len[2];        // transform lengths (x, y)
  int nfft[2];       // complex line lengths: len, or 2 len for K_ND / K_DN / K_DDF
  int logn[2];       // nfft = 2^logn * odd
  int odd[2];        // odd part of len: 1, or a product of at most three factors 3 and 5
  int kind[2];
  double normfft;
  cans::cplx<double>* tw[2];  // W_N^t = exp(-2 pi i t/N), t < N; then the same N values rounded to float
  cans::cplx<double>* mk[2];  // exp(-i pi k/(2N)), k <= N (DCT/DST pre/post twiddles)
  // launch shapes chosen by the tuner at cans_fftini, [0] FP32 storage, [1] FP64 (0: heuristic)
  int xthreads[2] = {0, 0};   // x pass: threads per work-group (lines = threads / (N/8))
  int ylines[2] = {0, 0};     // y pass: complex lines (2 columns each) per work-group
  int radix4[2][2] = {{0, 0}, {0, 0}};  // [direction][storage]: radix-4 instead of radix-8 stages
  // directions whose length the FFT passes do not take: the transform as a dense n x n
  // matrix (forward dm[d][0], backward dm[d][1], row-major, FP64) applied by k_dense
  int dense[2] = {0, 0};
  double* dm[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
};

namespace cans {

static int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) l++;
  return (1 << l) == n ? l : -1;
}

// n = 2^m * odd with odd a product of at most three factors 3 and 5 (1 for a power of two);
// with odd > 1 the line needs m >= 3 (its team of n/8 threads)
static bool fft_factor(int n, int* m, int* odd) {
  if (n < 1 || n > 4096) return false;
  int o = n, e = 0;
  while (o % 2 == 0) o /= 2, e++;
  int q = o, nf = 0;
  while (q % 3 == 0) q /= 3, nf++;
  while (q % 5 == 0) q /= 5, nf++;
  *m = e;
  *odd = o;
  return q == 1 && nf <= 3 && (o == 1 || e >= 3);
}

// position of input element e in Makhoul's even/odd reordering (v_m = x_{2m}, v_{N-1-m} = x_{2m+1})
__device__ __forceinline__ int mk_pos(int e, int N) { return (e & 1) ? N - 1 - (e >> 1) : (e >> 1); }

// ---------------------------------------------------------------------------
// the twiddle table of an N-point line in the line's type (the float copy follows the N doubles)
template <typename T>
__device__ __forceinline__ const cplx<T>* line_tw(const cplx<double>* tw, int N) {
  if constexpr (sizeof(T) == sizeof(double)) return (const cplx<T>*)tw;
  else return (const cplx<T>*)(tw + N);
}

// Pre/post-processing around the complex FFT.  Two real lines A, B travel as
// one complex line z = A + iB:
//   R2HC   : A_k = (Z_k + conj Z_{N-k})/2, B_k = -i (Z_k - conj Z_{N-k})/2,
//            stored packed (r0, r_{N/2}, r1, i1, ...);
//   REDFT10: Makhoul's reorder v = (x0, x2, ..., x3, x1), then
//            Y_k = Re(M_k (Z_k + conj Z_{N-k})) with M_k = exp(-i pi k/(2N));
//   RODFT10: REDFT10 of (-1)^j x_j, output reversed;
//   HC2R, REDFT01, RODFT01: the inverses (unnormalised, as FFTW);
//   REDFT11 (length n): Y_k = 2 sum_j x_j cos(pi (2j+1)(2k+1)/(4n)) is output 2k+1 of the
//            REDFT10 of length N = 2n of x padded with n zeros, so the line has N = 2n
//            elements and Makhoul positions (n+1)/2 .. N-1-n/2 hold zeros;
//   RODFT11: (-1)^k REDFT11 of the reversed input.  Both are their own inverses (up to the
//            factor 2n), so the backward passes run the forward code for them.
// Global accesses move two elements per lane (16 bytes FP64, 8 bytes FP32 storage).
// The pre/post-processing widens to double and rounds once on the store to T; the
// FFT stages in between compute in the storage type (fft.hpp).
// ---------------------------------------------------------------------------
typedef HIP_vector_type<double, 2> d2;

template <typename T>
__device__ __forceinline__ d2 ld2(const T* p) {
  const HIP_vector_type<T, 2> v = *(const HIP_vector_type<T, 2>*)p;
  return d2((double)v.x, (double)v.y);
}
template <typename T>
__device__ __forceinline__ void st2(T* p, double x, double y) {
  *(HIP_vector_type<T, 2>*)p = HIP_vector_type<T, 2>((T)x, (T)y);
}

// spectra of A and B at index k from the complex transform of A + iB
__device__ __forceinline__ void split_pp(cplx<double> zk, cplx<double> zc, double& are, double& aim, double& bre,
                                         double& bim) {
  are = 0.5 * (zk.re + zc.re);
  aim = 0.5 * (zk.im - zc.im);
  bre = 0.5 * (zk.im + zc.im);
  bim = -0.5 * (zk.re - zc.re);
}
__device__ __forceinline__ void split_dct(cplx<double> zk, cplx<double> zc, cplx<double> m, double& ya, double& yb) {
  const double sre = zk.re + zc.re, sim = zk.im - zc.im;
  const double dre = zk.im + zc.im, dim = -(zk.re - zc.re);
  ya = m.re * sre - m.im * sim;
  yb = m.re * dre - m.im * dim;
}
// conj(M_k) (X_k - i X_{N-k}) for lines A and B, combined as V_A + i V_B
__device__ __forceinline__ cplx<double> merge_dct(cplx<double> m, double xka, double xnka, double xkb, double xnkb) {
  const double va_re = m.re * xka - m.im * xnka, va_im = -m.re * xnka - m.im * xka;
  const double vb_re = m.re * xkb - m.im * xnkb, vb_im = -m.re * xnkb - m.im * xkb;
  return {va_re - vb_im, va_im + vb_re};
}
// Z = A + iB from packed halfcomplex A_k, B_k (A_{N-k} = conj A_k)
__device__ __forceinline__ cplx<double> merge_pp(double are, double aim, double bre, double bim) {
  return {are - bim, aim + bre};
}

// ---------------------------------------------------------------------------
// Work buffer layout (z-tiled).  Spectral line (i,j) (0-based) belongs to tile
// column (i/ZT, j), slot i%ZT.  Rows are grouped in blocks of ZK: tile block
// (kb, i/ZT, j) stores its ZT lines for rows kb*ZK .. kb*ZK+ZK-1 contiguously, row k
// at offset (k%ZK)*ZT, and the blocks of one row block kb are ordered j fastest.
// The z pass streams each wave's lines as contiguous ZK-row blocks; the x and y
// passes move contiguous chunks of ZT (x) or 2*LINES (y) elements, and a y-pass
// work-group (consecutive column pairs of one plane, all j) stays inside n2 adjacent
// blocks unless its lines run on into the next plane (narrow x).
// nic = number of ZT-line chunks along x; PS = elements per row block (all tiles).
// ---------------------------------------------------------------------------
constexpr int ZT = 64;
constexpr int ZK = 16;
constexpr int NBATCH = 4;  // systems one batched solve carries (blockIdx.y of the passes)
constexpr int CK = 16;  // z pass: rows per checkpoint block (ZK is a multiple)
static_assert(ZK % CK == 0, "tile row blocks hold whole checkpoint blocks");
__host__ __device__ __forceinline__ int wnic(int n1) { return (n1 + ZT - 1) / ZT; }
__host__ __device__ __forceinline__ long wplane(int n1, int n2) { return (long)wnic(n1) * n2 * ZT * ZK; }
__host__ __device__ __forceinline__ long woff(int i, int j, int k, int n2, int nic) {
  return (((long)(k / ZK) * nic + i / ZT) * n2 + j) * (ZT * ZK) + (k % ZK) * ZT + (i % ZT);
}
// offset of row r of a line from that line's row 0
__host__ __device__ __forceinline__ long zrow(int r, long PS) { return (long)(r / ZK) * PS + (r % ZK) * ZT; }
// elements of the buffer: the z pass reads whole CK-row blocks, up to CK - 2 rows past n3
__host__ __device__ inline long wbuf_elems(int n1, int n2, int n3) {
  return (long)((n3 + CK + ZK - 1) / ZK) * wplane(n1, n2);
}

// element e of one row: a contiguous field row, or row (j,k) of the tiled work buffer.
// `r + e` maps the whole element index e through the tiling and returns a plain pointer, so
// an index expression must be parenthesised: r + (a - b), never r + a - b.
template <typename T>
struct RowPtr {
  T* p = nullptr;
  int j = 0, k = 0, n2 = 0, nic = 0, tiled = 0;  // (j,k) of a tiled row
  __device__ __forceinline__ T* operator+(long e) const { return tiled ? p + woff((int)e, j, k, n2, nic) : p + e; }
  __device__ __forceinline__ T& operator[](long e) const { return *(*this + e); }
  __device__ __forceinline__ explicit operator bool() const { return p != nullptr; }
};

// ---------------------------------------------------------------------------
// pass X: rows of the halo'd field are contiguous along x; rows of the work
// buffer are contiguous in chunks of ZT.
// rows r = j + n2*k (0-based); the complex line of team q carries rows 2q, 2q+1.
// ---------------------------------------------------------------------------
template <typename T>
struct XArgs {
  Layout L;
  T* fld[NBATCH];  // halo'd fields (origin), one per system of a batch (blockIdx.y)
  T* W[NBATCH];    // their z-tiled work buffers (woff)
  int N, logN, odd, Tn, pairs, kind;
  int nd;          // data length along x: N, or N/2 for K_ND / K_DN / K_DDF
  const cplx<double>* tw;
  const cplx<double>* mk;
  int div;         // forward only: build the rhs from the velocity instead of reading fld
  DivSrc<T> ds;
  int r4 = 0;      // radix-4 power-of-two stages (fft_line)
};

// L(corr[0] + ... ) at the cells c, c+1 of row (j,k): the 7-point operator of the solver
// (the eigenvalues and tridiagonal coefficients of initsolver.f90:85-187) on the sum of a
// refined solve's earlier corrections, whose ghost cells boundp has set
template <typename T>
__device__ __forceinline__ d2 corr_lap_pair(const DivSrc<T>& ds, const Layout& L, long c, int k) {
  auto sum2 = [&](long o) -> d2 {
    d2 v = ld2(ds.corr[0] + o);
    if (ds.ncorr > 1) {
      const d2 w = ld2(ds.corr[1] + o);
      v = d2(v.x + w.x, v.y + w.y);
    }
    return v;
  };
  auto sum1 = [&](long o) -> double {
    double v = (double)ds.corr[0][o];
    if (ds.ncorr > 1) v += (double)ds.corr[1][o];
    return v;
  };
  const d2 x0 = sum2(c), ym = sum2(c - L.sy), yp = sum2(c + L.sy), zm = sum2(c - L.sz), zp = sum2(c + L.sz);
  const double xm = sum1(c - 1), xp = sum1(c + 2);
  const double dzc = ds.dzci[k], dzm = ds.dzci[k - 1], dzf = ds.dzfi[k];
  const double l0 = (x0.y - 2. * x0.x + xm) * ds.dxi2 + (yp.x - 2. * x0.x + ym.x) * ds.dyi2 +
                    ((zp.x - x0.x) * dzc - (x0.x - zm.x) * dzm) * dzf;
  const double l1 = (xp - 2. * x0.y + x0.x) * ds.dxi2 + (yp.y - 2. * x0.y + ym.y) * ds.dyi2 +
                    ((zp.y - x0.y) * dzc - (x0.y - zm.y) * dzm) * dzf;
  return d2(l0, l1);
}

// two consecutive rhs values (elements e, e+1 of row r) of fillps.f90:38-50; RES: minus
// L(corrections), the residual of a refinement sweep
template <typename T, bool RES = false>
__device__ __forceinline__ d2 div_pair(const XArgs<T>& a, long r, int e) {
  const Layout& L = a.L;
  const int j = (int)(r % L.n2) + 1, k = (int)(r / L.n2) + 1;
  const long c = L.at(1, j, k) + e;
  if (a.ds.divu) {  // dti * (div(u) + div(e)); div(u) was formed by the momentum kernel
    const d2 dv = ld2(a.ds.divu + r * L.n1 + e);
    const d2 eu2 = ld2(a.ds.eu + c), ew2 = ld2(a.ds.ew + c), ewm = ld2(a.ds.ew + c - L.sz);
    const d2 ev2 = ld2(a.ds.ev + c), evm = ld2(a.ds.ev + c - L.sy);
    const double eum = (double)a.ds.eu[c - 1];
    const double dzfi = a.ds.dzfi[k], dyi = a.ds.dyi, dxi = a.ds.dxi, dti = a.ds.dti;
    double q0 = dti * (dv.x + ((ew2.x - ewm.x) * dzfi + (ev2.x - evm.x) * dyi + (eu2.x - eum) * dxi));
    double q1 = dti * (dv.y + ((ew2.y - ewm.y) * dzfi + (ev2.y - evm.y) * dyi + (eu2.y - eu2.x) * dxi));
    if constexpr (RES) {
      const d2 lp = corr_lap_pair(a.ds, L, c, k);
      q0 = q0 - lp.x;
      q1 = q1 - lp.y;
    }
    return d2(q0, q1);
  }
  d2 u2 = ld2(a.ds.u + c), w2 = ld2(a.ds.w + c), wm = ld2(a.ds.w + c - L.sz);
  d2 v2 = ld2(a.ds.v + c), vm = ld2(a.ds.v + c - L.sy);
  double um = a.ds.u[c - 1];
  if (a.ds.f) {  // the forced velocity u + f (rk.f90:307-311), rounded as the stored one would be
    const double fu = a.ds.f[0], fv = a.ds.f[1], fw = a.ds.f[2];
    u2 = d2(u2.x + fu, u2.y + fu);
    um = um + fu;
    v2 = d2(v2.x + fv, v2.y + fv);
    vm = d2(vm.x + fv, vm.y + fv);
    w2 = d2(w2.x + fw, w2.y + fw);
    wm = d2(wm.x + fw, wm.y + fw);
  }
  if (a.ds.eu) {  // velocity held as u + increment
    const d2 eu2 = ld2(a.ds.eu + c), ew2 = ld2(a.ds.ew + c), ewm = ld2(a.ds.ew + c - L.sz);
    const d2 ev2 = ld2(a.ds.ev + c), evm = ld2(a.ds.ev + c - L.sy);
    u2 = d2(u2.x + eu2.x, u2.y + eu2.y);
    w2 = d2(w2.x + ew2.x, w2.y + ew2.y);
    wm = d2(wm.x + ewm.x, wm.y + ewm.y);
    v2 = d2(v2.x + ev2.x, v2.y + ev2.y);
    vm = d2(vm.x + evm.x, vm.y + evm.y);
    um = um + (double)a.ds.eu[c - 1];
  }
  const double dti = a.ds.dti, dzfi = a.ds.dzfi[k], dy = a.ds.dtidyi, dx = a.ds.dtidxi;
  double p0 = ((w2.x - wm.x) * dti * dzfi + (v2.x - vm.x) * dy + (u2.x - um) * dx);
  double p1 = ((w2.y - wm.y) * dti * dzfi + (v2.y - vm.y) * dy + (u2.y - u2.x) * dx);
  if constexpr (RES) {
    const d2 lp = corr_lap_pair(a.ds, L, c, k);
    p0 = p0 - lp.x;
    p1 = p1 - lp.y;
  }
  return d2(p0, p1);
}

// DIR +1 forward (fld -> W), -1 backward (W -> fld); DV (forward only): the rhs is
// formed from the velocity (a.ds) instead of read from fld (DV = 2: minus L(corrections),
// the residual of a refinement sweep).  DV is a template flag
// so that the plain forward pass carries none of the divergence code's registers.
// bx: work-group index among this system's; sys: system of a batch.  Lanes past the
// last whole team (a.pairs teams of a.Tn) only take part in the barriers.
// FLIP (face-centred DN, K_DNF): the K_DD pair with swapped directions, so the forward pass
// runs the RODFT01 code and the backward pass the RODFT10 code.
template <int DIR, typename T, int DV = 0, bool FLIP = false, int R4 = 0>
__device__ __forceinline__ void xpass_body(const XArgs<T>& a, int bx, int sys) {
  T* const fld = a.fld[sys];
  T* const W = a.W[sys];
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* smem = (cplx<T>*)smem_raw;  // LDS line in the storage type
  const int N = a.N, Tn = a.Tn, kind = FLIP ? K_DD : a.kind;
  constexpr bool FWD = (DIR > 0) != FLIP;  // forward-type code (R2HC, REDFT10, RODFT10)
  const int team = threadIdx.x / Tn, lt = threadIdx.x % Tn;
  const bool act = team < a.pairs;
  const long nrows = (long)a.L.n2 * a.L.n3;
  const long q = (long)bx * a.pairs + team;
  const long ra = 2 * q, rb = 2 * q + 1;
  const bool va = act && ra < nrows, vb = act && rb < nrows;
  const IxPad ix{team * padded_len(N)};
  const Layout& L = a.L;
  auto frow = [&](bool v, long r) {
    RowPtr<T> q;
    if (v) q.p = fld + L.at(1, (int)(r % L.n2) + 1, (int)(r / L.n2) + 1);
    return q;
  };
  auto wrow = [&](bool v, long r) {
    RowPtr<T> q;
    if (v) {
      q.p = W;
      q.j = (int)(r % L.n2);
      q.k = (int)(r / L.n2);
      q.n2 = L.n2;
      q.nic = wnic(L.n1);
      q.tiled = 1;
    }
    return q;
  };
  const RowPtr<T> xa = DIR > 0 ? frow(va, ra) : wrow(va, ra);
  const RowPtr<T> xb = DIR > 0 ? frow(vb, rb) : wrow(vb, rb);
  const RowPtr<T> ya = DIR > 0 ? wrow(va, ra) : frow(va, ra);
  const RowPtr<T> yb = DIR > 0 ? wrow(vb, rb) : frow(vb, rb);
  if (FWD || k_double(kind)) {
    if (va) {
      if (N == 1) {
        if (lt == 0) smem[ix(0)] = narrow<T>((double)xa[0], xb ? (double)xb[0] : 0.);
      } else if (kind == K_DDF) {
        // odd extension: data element e (< nd-1) at point e+1 and its negative at N-1-e; points 0
        // and nd are zero; the boundary face nd-1 is not transformed and passes through
        const int nd = a.nd;
        if (lt == 0) {
          smem[ix(0)] = narrow<T>(0., 0.);
          smem[ix(nd)] = narrow<T>(0., 0.);
        }
        for (int e = 2 * lt; e < nd; e += 2 * Tn) {
          const d2 a2 = ld2(xa + e), b2 = xb ? ld2(xb + e) : d2(0., 0.);
          smem[ix(e + 1)] = narrow<T>(a2.x, b2.x);
          smem[ix(N - 1 - e)] = narrow<T>(-a2.x, -b2.x);
          if (e + 1 < nd - 1) {
            smem[ix(e + 2)] = narrow<T>(a2.y, b2.y);
            smem[ix(N - 2 - e)] = narrow<T>(-a2.y, -b2.y);
          } else {
            ya[e + 1] = (T)a2.y;
            if (yb) yb[e + 1] = (T)b2.y;
          }
        }
      } else if (kind == K_ND || kind == K_DN) {
        const int nd = a.nd;
        const bool dn = kind == K_DN;  // RODFT11 reads the reversed input
        for (int p = (nd + 1) / 2 + lt; p < N - nd / 2; p += Tn) smem[ix(p)] = narrow<T>(0., 0.);
        for (int e = 2 * lt; e < nd; e += 2 * Tn) {
          const d2 a2 = ld2(xa + e), b2 = xb ? ld2(xb + e) : d2(0., 0.);
          smem[ix(mk_pos(dn ? nd - 1 - e : e, N))] = narrow<T>(a2.x, b2.x);
          smem[ix(mk_pos(dn ? nd - 2 - e : e + 1, N))] = narrow<T>(a2.y, b2.y);
        }
      } else {
        // compile-time trip count (N = 8 Tn, or N < 8 with Tn = 1): the unrolled loop issues
        // all of the lane's row loads before its first LDS store
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int e = 2 * lt + 2 * r * Tn;
          if (e >= N) continue;
          d2 a2, b2;
          if constexpr (DV != 0) {
            a2 = div_pair<T, DV == 2>(a, ra, e);
            b2 = xb ? div_pair<T, DV == 2>(a, rb, e) : d2(0., 0.);
          } else {
            a2 = ld2(xa + e);
            b2 = xb ? ld2(xb + e) : d2(0., 0.);
          }
          if (kind == K_DD) {  // (-1)^j on the odd element
            a2.y = -a2.y;
            b2.y = -b2.y;
          }
          const int p0 = kind == K_PP ? e : mk_pos(e, N), p1 = kind == K_PP ? e + 1 : mk_pos(e + 1, N);
          smem[ix(p0)] = narrow<T>(a2.x, b2.x);
          smem[ix(p1)] = narrow<T>(a2.y, b2.y);
        }
      }
    }
    __syncthreads();
    fft_line<T, -1>(smem, ix, N, a.logN, lt, Tn, line_tw<T>(a.tw, N), a.odd, act, R4 != 0);
    if (!va) return;
    if (kind == K_PP) {
      const int H = N >> 1;
      if (N == 1) {
        if (lt == 0) {
          ya[0] = smem[ix(0)].re;
          if (yb) yb[0] = smem[ix(0)].im;
        }
        return;
      }
      for (int k = lt; k < H; k += Tn) {
        double are, aim, bre, bim;
        if (k == 0) {
          const cplx<double> z0 = widen(smem[ix(0)]), zh = widen(smem[ix(H)]);
          are = z0.re;
          aim = zh.re;
          bre = z0.im;
          bim = zh.im;
        } else {
          const cplx<double> zk = widen(smem[ix(k)]), zc = widen(smem[ix(N - k)]);
          split_pp(zk, zc, are, aim, bre, bim);
        }
        st2(ya + 2 * k, are, aim);
        if (yb) st2(yb + 2 * k, bre, bim);
      }
    } else if (kind == K_DDF) {
      // RODFT00 of the two lines: Y_A(k-1) = -Im Z_k, Y_B(k-1) = Re Z_k, k = 1 .. nd-1
      const int nd = a.nd;
      for (int m = 2 * lt; m < nd - 1; m += 2 * Tn) {
        const cplx<double> z0 = widen(smem[ix(m + 1)]);
        if (m + 1 < nd - 1) {
          const cplx<double> z1 = widen(smem[ix(m + 2)]);
          st2(ya + m, -z0.im, -z1.im);
          if (yb) st2(yb + m, z0.re, z1.re);
        } else {
          ya[m] = (T)(-z0.im);
          if (yb) yb[m] = (T)z0.re;
        }
      }
    } else if (kind == K_ND || kind == K_DN) {
      // outputs m, m+1 (m even) are spectral indices 2m+1, 2m+3 of the N-point REDFT10
      for (int m = 2 * lt; m < a.nd; m += 2 * Tn) {
        double y0a, y0b, y1a, y1b;
        split_dct(widen(smem[ix(2 * m + 1)]), widen(smem[ix(N - 2 * m - 1)]), a.mk[2 * m + 1], y0a, y0b);
        split_dct(widen(smem[ix(2 * m + 3)]), widen(smem[ix(N - 2 * m - 3)]), a.mk[2 * m + 3], y1a, y1b);
        if (kind == K_DN) {  // (-1)^k on the odd output
          y1a = -y1a;
          y1b = -y1b;
        }
        st2(ya + m, y0a, y1a);
        if (yb) st2(yb + m, y0b, y1b);
      }
    } else {
      for (int k = 2 * lt; k < N; k += 2 * Tn) {
        double y0a, y0b, y1a, y1b;
        split_dct(widen(smem[ix(k)]), widen(smem[ix(k ? N - k : 0)]), a.mk[k], y0a, y0b);
        split_dct(widen(smem[ix(k + 1)]), widen(smem[ix(N - k - 1)]), a.mk[k + 1], y1a, y1b);
        if (kind == K_DD) {  // output reversed: element N-1-k
          st2(ya + (N - 2 - k), y1a, y0a);
          if (yb) st2(yb + (N - 2 - k), y1b, y0b);
        } else {
          st2(ya + k, y0a, y1a);
          if (yb) st2(yb + k, y0b, y1b);
        }
      }
    }
  } else {
    // spectral indices k and N-k read the same elements: one thread takes both
    if (va) {
      const int H = N >> 1;
      if (N == 1) {
        if (lt == 0) smem[ix(0)] = narrow<T>(merge_pp((double)xa[0], 0., xb ? (double)xb[0] : 0., 0.));
      } else if (kind == K_PP) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int m = lt + r * Tn;
          if (m >= H) continue;
          const d2 a2 = ld2(xa + 2 * m);  // m = 0: (r0, r_{N/2})
          const d2 b2 = xb ? ld2(xb + 2 * m) : d2(0., 0.);
          if (m == 0) {
            smem[ix(0)] = narrow<T>(merge_pp(a2.x, 0., b2.x, 0.));
            smem[ix(H)] = narrow<T>(merge_pp(a2.y, 0., b2.y, 0.));
          } else {
            smem[ix(m)] = narrow<T>(merge_pp(a2.x, a2.y, b2.x, b2.y));
            smem[ix(N - m)] = narrow<T>(merge_pp(a2.x, -a2.y, b2.x, -b2.y));
          }
        }
      } else {
        const bool dd = kind == K_DD;  // RODFT01 reads the reversed input
#pragma unroll
        for (int r = 0; r < 5; r++) {
          const int m = lt + r * Tn;
          if (m > H) continue;
          if (m == 0) {
            const int i0 = dd ? N - 1 : 0;
            smem[ix(0)] = narrow<T>(merge_dct(a.mk[0], (double)xa[i0], 0., xb ? (double)xb[i0] : 0., 0.));
            continue;
          }
          const int ia = dd ? N - 1 - m : m, ib = dd ? m - 1 : N - m;
          const double xaa = (double)xa[ia], xab = (double)xa[ib];
          const double xba = xb ? (double)xb[ia] : 0., xbb = xb ? (double)xb[ib] : 0.;
          smem[ix(m)] = narrow<T>(merge_dct(a.mk[m], xaa, xab, xba, xbb));
          if (m != H) smem[ix(N - m)] = narrow<T>(merge_dct(a.mk[N - m], xab, xaa, xbb, xba));
        }
      }
    }
    __syncthreads();
    fft_line<T, +1>(smem, ix, N, a.logN, lt, Tn, line_tw<T>(a.tw, N), a.odd, act, R4 != 0);
    if (!va) return;
    if (N == 1) {
      if (lt == 0) {
        ya[0] = smem[ix(0)].re;
        if (yb) yb[0] = smem[ix(0)].im;
      }
      return;
    }
    for (int e = 2 * lt; e < N; e += 2 * Tn) {
      cplx<double> z0 = widen(smem[ix(kind == K_PP ? e : mk_pos(e, N))]);
      cplx<double> z1 = widen(smem[ix(kind == K_PP ? e + 1 : mk_pos(e + 1, N))]);
      if (kind == K_DD) {  // (-1)^k on the odd output
        z1.re = -z1.re;
        z1.im = -z1.im;
      }
      st2(ya + e, z0.re, z1.re);
      if (yb) st2(yb + e, z0.im, z1.im);
    }
  }
}

// one kernel per storage type (the FP32 and FP64 forms differ in LDS and register use); R4:
// the radix-4 stage split, a kernel of its own, so that neither form carries the other's registers
template <int DIR, int DV, bool FLIP = false, int R4 = 0>
__global__ void __launch_bounds__(512) k_xpass_f(XArgs<float> a) {
  xpass_body<DIR, float, DV, FLIP, R4>(a, blockIdx.x, blockIdx.y);
}
template <int DIR, int DV, bool FLIP = false, int R4 = 0>
__global__ void __launch_bounds__(512) k_xpass_d(XArgs<double> a) {
  xpass_body<DIR, double, DV, FLIP, R4>(a, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// pass Y: in place on W; a work-group owns LINES consecutive column pairs (2*LINES
// x-columns; after the last pair of a z-plane come the first of the next).
// thread t -> line c = t % LINES, team thread lt = t / LINES; LDS element (e, c) at [e*LINES + c].
// The two columns of a line are adjacent, so every global access is one 16-byte pair.
// ---------------------------------------------------------------------------
template <typename T>
struct YArgs {
  int n1, n2, n3;
  T* W[NBATCH];    // z-tiled work buffers (woff), one per system of a batch (blockIdx.y)
  int N, logN, odd, Tn, lines, kind;
  int nd;          // data length along y: N, or N/2 for K_ND / K_DN / K_DDF
  const cplx<double>* tw;
  const cplx<double>* mk;
  int r4 = 0;      // radix-4 power-of-two stages (fft_line)
};

// One tile of the y pass in three steps: ypass_load reads the elements this lane's
// pre-processing needs into registers (raw), all loads of a batch issued before any LDS
// store; ypass_pre forms the complex line in LDS from them; ypass_fft_post transforms it
// and stores the result.  The forward-type forms take their 8 elements in one batch; the
// inverse forms (two elements per spectral index m = lt + r Tn) may split r into batches
// [R0, R1) to bound the registers raw holds (one batch measured fastest: issuing every
// load before the first LDS store keeps the most bytes in flight).  Compile-time trip
// counts (N = 8 Tn, or N < 8 with Tn = 1) keep raw in registers.
constexpr int YRAW = 10;  // raw elements per lane: 8 (forward-type forms), up to 10 (inverse DCT/DST)

template <typename T>
struct YTile {
  int c, lt, k, ia;
  bool act, va, vb;
};

template <typename T>
__device__ __forceinline__ YTile<T> ytile(const YArgs<T>& a, long bx) {
  YTile<T> t;
  const int LN = a.lines;
  t.c = threadIdx.x % LN;
  t.lt = threadIdx.x / LN;
  t.act = t.lt < a.Tn;
  // column pairs are numbered plane by plane; a work-group's lines may span planes
  const int ncp = (a.n1 + 1) / 2;
  const long g = bx * LN + t.c;
  t.k = (int)(g / ncp);
  t.ia = 2 * (int)(g % ncp);
  t.va = t.act && t.k < a.n3;
  t.vb = t.va && t.ia + 1 < a.n1;
  return t;
}

template <typename T>
__device__ __forceinline__ T* ycol(const YArgs<T>& a, const YTile<T>& t, int sys, int e) {
  return a.W[sys] + woff(t.ia, e, t.k, a.n2, wnic(a.n1));
}

template <int DIR, typename T, bool FLIP, int R0, int R1>
__device__ __forceinline__ void ypass_load(const YArgs<T>& a, const YTile<T>& t, int sys,
                                           HIP_vector_type<T, 2>* raw) {
  typedef HIP_vector_type<T, 2> v2;
  const int N = a.N, Tn = a.Tn, lt = t.lt, kind = FLIP ? K_DD : a.kind;
  constexpr bool FWD = (DIR > 0) != FLIP;
  if (!t.va) return;
  auto ld = [&](int e) -> v2 {
    const T* p = ycol(a, t, sys, e);
    return t.vb ? *(const v2*)p : v2(*p, (T)0);
  };
  if (FWD || k_double(kind)) {
    if (R0 != 0) return;
    const int lim = kind == K_DDF ? a.nd - 1 : (kind == K_ND || kind == K_DN) ? a.nd : N;
#pragma unroll
    for (int r = 0; r < 8; r++) {
      const int e = lt + r * Tn;
      if (e < lim) raw[r] = ld(e);
    }
  } else if (N == 1) {
    if (R0 == 0 && lt == 0) raw[0] = ld(0);
  } else if (kind == K_PP) {
    const int H = N >> 1;
#pragma unroll
    for (int r = R0; r < R1 && r < 4; r++) {
      const int m = lt + r * Tn;
      if (m < H) {
        raw[2 * (r - R0)] = ld(2 * m);
        raw[2 * (r - R0) + 1] = ld(2 * m + 1);
      }
    }
  } else {
    const int H = N >> 1;
    const bool dd = kind == K_DD;  // RODFT01 reads the reversed input
#pragma unroll
    for (int r = R0; r < R1; r++) {
      const int m = lt + r * Tn;
      if (m > H) continue;
      if (m == 0) {
        raw[2 * (r - R0)] = ld(dd ? N - 1 : 0);
        continue;
      }
      raw[2 * (r - R0)] = ld(dd ? N - 1 - m : m);
      raw[2 * (r - R0) + 1] = ld(dd ? m - 1 : N - m);
    }
  }
}

template <int DIR, typename T, bool FLIP, int R0, int R1>
__device__ __forceinline__ void ypass_pre(const YArgs<T>& a, const YTile<T>& t, const HIP_vector_type<T, 2>* raw,
                                          cplx<T>* smem) {
  const int N = a.N, Tn = a.Tn, lt = t.lt, kind = FLIP ? K_DD : a.kind;
  constexpr bool FWD = (DIR > 0) != FLIP;
  const IxStride ix{t.c, a.lines};
  if (!t.va) return;
  auto w = [&](int r) -> d2 { return d2((double)raw[r].x, (double)raw[r].y); };
  if (FWD || k_double(kind)) {
    if (R0 != 0) return;
    if (kind == K_DDF) {
      // odd extension (as in the x pass); the boundary face nd-1 stays in place
      const int nd = a.nd;
      if (lt == 0) {
        smem[ix(0)] = narrow<T>(0., 0.);
        smem[ix(nd)] = narrow<T>(0., 0.);
      }
#pragma unroll
      for (int r = 0; r < 8; r++) {
        const int e = lt + r * Tn;
        if (e < nd - 1) {
          const d2 v = w(r);
          smem[ix(e + 1)] = narrow<T>(v.x, v.y);
          smem[ix(N - 1 - e)] = narrow<T>(-v.x, -v.y);
        }
      }
    } else if (kind == K_ND || kind == K_DN) {
      const int nd = a.nd;
      for (int p = (nd + 1) / 2 + lt; p < N - nd / 2; p += Tn) smem[ix(p)] = narrow<T>(0., 0.);
#pragma unroll
      for (int r = 0; r < 8; r++) {
        const int e = lt + r * Tn;
        if (e < nd) {
          const d2 v = w(r);
          smem[ix(mk_pos(kind == K_DN ? nd - 1 - e : e, N))] = narrow<T>(v.x, v.y);
        }
      }
    } else {
#pragma unroll
      for (int r = 0; r < 8; r++) {
        const int e = lt + r * Tn;
        if (e < N) {
          d2 v = w(r);
          if (kind == K_DD && (e & 1)) {
            v.x = -v.x;
            v.y = -v.y;
          }
          smem[ix(kind == K_PP ? e : mk_pos(e, N))] = narrow<T>(v.x, v.y);
        }
      }
    }
    return;
  }
  // spectral indices m and N-m read the same two rows: one thread takes both
  const int H = N >> 1;
  if (N == 1) {
    if (R0 == 0 && lt == 0) {
      const d2 v = w(0);
      smem[ix(0)] = narrow<T>(merge_pp(v.x, 0., v.y, 0.));
    }
  } else if (kind == K_PP) {
#pragma unroll
    for (int r = R0; r < R1 && r < 4; r++) {
      const int m = lt + r * Tn;
      if (m >= H) continue;
      const d2 re = w(2 * (r - R0)), im = w(2 * (r - R0) + 1);  // m = 0: r0 and r_{N/2}
      if (m == 0) {
        smem[ix(0)] = narrow<T>(merge_pp(re.x, 0., re.y, 0.));
        smem[ix(H)] = narrow<T>(merge_pp(im.x, 0., im.y, 0.));
      } else {
        smem[ix(m)] = narrow<T>(merge_pp(re.x, im.x, re.y, im.y));
        smem[ix(N - m)] = narrow<T>(merge_pp(re.x, -im.x, re.y, -im.y));
      }
    }
  } else {
#pragma unroll
    for (int r = R0; r < R1; r++) {
      const int m = lt + r * Tn;
      if (m > H) continue;
      if (m == 0) {
        const d2 x0 = w(2 * (r - R0));
        smem[ix(0)] = narrow<T>(merge_dct(a.mk[0], x0.x, 0., x0.y, 0.));
        continue;
      }
      const d2 xa = w(2 * (r - R0)), xb = w(2 * (r - R0) + 1);
      smem[ix(m)] = narrow<T>(merge_dct(a.mk[m], xa.x, xb.x, xa.y, xb.y));
      if (m != H) smem[ix(N - m)] = narrow<T>(merge_dct(a.mk[N - m], xb.x, xa.x, xb.y, xa.y));
    }
  }
}

// the line in LDS (after a barrier) -> transform -> post-processing -> stores
template <int DIR, typename T, bool FLIP, int R4 = 0>
__device__ __forceinline__ void ypass_fft_post(const YArgs<T>& a, const YTile<T>& t, int sys, cplx<T>* smem) {
  const int N = a.N, Tn = a.Tn, lt = t.lt, kind = FLIP ? K_DD : a.kind;
  constexpr bool FWD = (DIR > 0) != FLIP;
  const IxStride ix{t.c, a.lines};
  const bool vb = t.vb;
  auto st = [&](int e, double x, double y) {
    T* p = ycol(a, t, sys, e);
    if (vb) st2(p, x, y);
    else *p = (T)x;
  };
  if (FWD || k_double(kind)) {
    fft_line<T, -1>(smem, ix, N, a.logN, lt, Tn, line_tw<T>(a.tw, N), a.odd, t.act, R4 != 0);
    if (!t.va) return;
    if (kind == K_PP) {
      const int H = N >> 1;
      if (N == 1) {
        if (lt == 0) st(0, (double)smem[ix(0)].re, (double)smem[ix(0)].im);
        return;
      }
      for (int kk = lt; kk < H; kk += Tn) {
        double are, aim, bre, bim;
        if (kk == 0) {
          const cplx<double> z0 = widen(smem[ix(0)]), zh = widen(smem[ix(H)]);
          are = z0.re;
          aim = zh.re;
          bre = z0.im;
          bim = zh.im;
        } else {
          split_pp(widen(smem[ix(kk)]), widen(smem[ix(N - kk)]), are, aim, bre, bim);
        }
        st(2 * kk, are, bre);
        st(2 * kk + 1, aim, bim);
      }
    } else if (kind == K_DDF) {
      for (int m = lt; m < a.nd - 1; m += Tn) {
        const cplx<double> z = widen(smem[ix(m + 1)]);
        st(m, -z.im, z.re);
      }
    } else if (kind == K_ND || kind == K_DN) {
      for (int m = lt; m < a.nd; m += Tn) {
        double ya, yb;
        split_dct(widen(smem[ix(2 * m + 1)]), widen(smem[ix(N - 2 * m - 1)]), a.mk[2 * m + 1], ya, yb);
        if (kind == K_DN && (m & 1)) {
          ya = -ya;
          yb = -yb;
        }
        st(m, ya, yb);
      }
    } else {
      for (int kk = lt; kk < N; kk += Tn) {
        double ya, yb;
        split_dct(widen(smem[ix(kk)]), widen(smem[ix(kk ? N - kk : 0)]), a.mk[kk], ya, yb);
        st(kind == K_DD ? N - 1 - kk : kk, ya, yb);
      }
    }
  } else {
    fft_line<T, +1>(smem, ix, N, a.logN, lt, Tn, line_tw<T>(a.tw, N), a.odd, t.act, R4 != 0);
    if (!t.va) return;
    for (int e = lt; e < N; e += Tn) {
      cplx<double> z = widen(smem[ix(kind == K_PP ? e : mk_pos(e, N))]);
      if (kind == K_DD && (e & 1)) {
        z.re = -z.re;
        z.im = -z.im;
      }
      st(e, z.re, z.im);
    }
  }
}

// bx: work-group index among this system's; sys: system of a batch.  Lanes past
// a.lines * a.Tn only take part in the barriers.  FLIP: as in xpass_body.
template <int DIR, typename T, bool FLIP = false, int R4 = 0>
__device__ __forceinline__ void ypass_body(const YArgs<T>& a, int bx, int sys) {
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* smem = (cplx<T>*)smem_raw;  // LDS line in the storage type
  HIP_vector_type<T, 2> raw[YRAW];
  const YTile<T> t = ytile(a, bx);
  ypass_load<DIR, T, FLIP, 0, 5>(a, t, sys, raw);
  ypass_pre<DIR, T, FLIP, 0, 5>(a, t, raw, smem);
  __syncthreads();
  ypass_fft_post<DIR, T, FLIP, R4>(a, t, sys, smem);
}

// FP32 lines: 16 lines x N/8 threads = 1024 threads in 64 KB of LDS; at most 64 VGPRs
// lets two such work-groups share a CU.  FP64 lines: LDS already limits a CU to two
// 512-thread work-groups, which needs no register cap.
template <int DIR, bool FLIP = false, int R4 = 0>
__global__ void __launch_bounds__(1024, 8) k_ypass_f(YArgs<float> a) {
  ypass_body<DIR, float, FLIP, R4>(a, blockIdx.x, blockIdx.y);
}
template <int DIR, bool FLIP = false, int R4 = 0>
__global__ void __launch_bounds__(1024) k_ypass_d(YArgs<double> a) {
  ypass_body<DIR, double, FLIP, R4>(a, blockIdx.x, blockIdx.y);
}
template <int DIR, typename T>
void launch_ypass(dim3 g, int t, size_t shm, hipStream_t st, const YArgs<T>& a) {
  const bool flip = a.kind == K_DNF;
  if constexpr (sizeof(T) == sizeof(float)) {
    if (flip) k_ypass_f<DIR, true><<<g, t, shm, st>>>(a);
    else if (a.r4) k_ypass_f<DIR, false, 1><<<g, t, shm, st>>>(a);
    else k_ypass_f<DIR, false><<<g, t, shm, st>>>(a);
  } else {
    if (flip) k_ypass_d<DIR, true><<<g, t, shm, st>>>(a);
    else if (a.r4) k_ypass_d<DIR, false, 1><<<g, t, shm, st>>>(a);
    else k_ypass_d<DIR, false><<<g, t, shm, st>>>(a);
  }
}

// ---------------------------------------------------------------------------
// x and y passes of a batch whose systems use different transforms (the velocity
// components' Helmholtz systems next to walls: face-centred DST-I along each component's
// normal, cell-centred DST-II along the others).  blockIdx.y selects the system; each
// system's work-groups run the single-system pass body with that system's line length and
// kind, as many teams (x) or lines (y) as fit the common work-group size, the lanes left
// over joining only the barriers; work-groups past a system's own count return at once.
// ---------------------------------------------------------------------------
template <typename T>
struct HetArgs {
  Layout L;
  T* fld[NBATCH];
  T* W[NBATCH];
  int N[NBATCH], nd[NBATCH], logN[NBATCH], odd[NBATCH], Tn[NBATCH], kind[NBATCH];
  int per[NBATCH];     // x: complex lines (teams) per work-group; y: lines per work-group
  int blocks[NBATCH];  // work-groups of the system
  const cplx<double>* tw[NBATCH];
  const cplx<double>* mk[NBATCH];
  int r4[NBATCH];      // the system's plan runs radix-4 stages (its tuner's choice)
};

template <int DIR, typename T>
__global__ void __launch_bounds__(512) k_het_xpass(HetArgs<T> h) {
  const int s = blockIdx.y;
  if ((int)blockIdx.x >= h.blocks[s]) return;
  XArgs<T> a;
  a.L = h.L;
  a.fld[0] = h.fld[s];
  a.W[0] = h.W[s];
  a.N = h.N[s];
  a.nd = h.nd[s];
  a.logN = h.logN[s];
  a.odd = h.odd[s];
  a.Tn = h.Tn[s];
  a.pairs = h.per[s];
  a.kind = h.kind[s];
  a.tw = h.tw[s];
  a.mk = h.mk[s];
  a.div = 0;
  if (a.kind == K_DNF) xpass_body<DIR, T, 0, true>(a, blockIdx.x, 0);
  else if (h.r4[s]) xpass_body<DIR, T, 0, false, 1>(a, blockIdx.x, 0);
  else xpass_body<DIR, T, 0, false>(a, blockIdx.x, 0);
}

template <int DIR, typename T>
__global__ void __launch_bounds__(1024) k_het_ypass(HetArgs<T> h) {
  const int s = blockIdx.y;
  if ((int)blockIdx.x >= h.blocks[s]) return;
  YArgs<T> a;
  a.n1 = h.L.n1;
  a.n2 = h.L.n2;
  a.n3 = h.L.n3;
  a.W[0] = h.W[s];
  a.N = h.N[s];
  a.nd = h.nd[s];
  a.logN = h.logN[s];
  a.odd = h.odd[s];
  a.Tn = h.Tn[s];
  a.lines = h.per[s];
  a.kind = h.kind[s];
  a.tw = h.tw[s];
  a.mk = h.mk[s];
  if (a.kind == K_DNF) ypass_body<DIR, T, true>(a, blockIdx.x, 0);
  else if (h.r4[s]) ypass_body<DIR, T, false, 1>(a, blockIdx.x, 0);
  else ypass_body<DIR, T, false>(a, blockIdx.x, 0);
}

// ---------------------------------------------------------------------------
// pass Z: Thomas elimination per (kx,ky) line, gaussel (solver.f90:114-307)
// with identical arithmetic.  The backward sweep needs d(k) and the forward-
// modified rhs in reverse order; instead of storing them for every k
// (3 extra field passes), the forward sweep checkpoints them every CK rows and
// the backward sweep re-reads the rhs of one block, recomputes the block's
// d/rhs from its checkpoint (bit-identical: same operations) and substitutes.
// ---------------------------------------------------------------------------
// one system of a (batched) solve, as the caller gives it
template <typename T>
struct SolveSys {
  T* fld;
  double normfft;
  const double *lam, *a, *b, *c;
  const char* bc;      // cbc(0:1,1:3) of the variable
  const char* c_or_f;  // (3)
};

// one system of a (batched) z pass: its work buffer, rows, normalisation, coefficients
template <typename T>
struct ZSys {
  T* W;                     // z-tiled work buffer (woff)
  int n;                    // rows of the system (n3 - q)
  double norm;
  const double *a, *b, *c, *lam;
  double *ckd, *ckp, *ckq;  // checkpoints [nblk][nl]
};

template <typename T>
struct ZArgs {
  ZSys<T> sys[NBATCH];      // blockIdx.y selects the system
  long nl;                  // number of lines (n1*n2)
  int n1, n2, nic;          // tile geometry of the work buffer
  long PS;                  // elements per row block
  int periodic;
  double pin_eps;           // relative pivot tolerance of the zero-mode pin
  double* X2;               // periodic (single system): auxiliary solution p2 [nn][nl]
};

// one forward elimination row (solver.f90:147-170, with the p2 recurrence of 235-245)
struct ZRow {
  double d, pp, q2;
};

// The tridiagonal coefficients are read-only during the solve and indexed by the row
// only (the same address in every lane): reading them through the constant address
// space lets the compiler use scalar loads instead of one vector load per row.
typedef const __attribute__((address_space(4))) double* kcoef;

// 1/x.  FAST (used when the solve's storage is FP32): the hardware reciprocal
// refined by two Newton-Raphson steps, r <- r + r(1 - x r), below 1 ulp of the
// correctly rounded quotient; otherwise the IEEE division, as the reference.
template <bool FAST>
__device__ __forceinline__ double rcp64(double x) {
  if constexpr (FAST) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.);
    r = fma(r, e, r);
    e = fma(-x, r, 1.);
    return fma(r, e, r);
  } else {
    return 1. / x;
  }
}

// interior forward row 1 < k < nn of a non-periodic system: no pin test, same
// operations and order as z_row
template <bool FAST>
__device__ __forceinline__ void z_row_mid(double l, double r, double norm, double ak, double bk, double ck, double& d,
                                          double& pp) {
  const double bl = bk + l;
  const double ad = ak * d;
  const double zz = rcp64<FAST>(bl - ad);
  d = ck * zz;
  pp = (r * norm - ak * pp) * zz;
}

template <bool FAST = false>
__device__ __forceinline__ void z_row(int k, int nn, int periodic, double l, double r, double norm, kcoef a, kcoef b,
                                      kcoef c, double eps, ZRow& s) {
  if (k == 1) {
    const double zz = rcp64<FAST>(b[0] + l);
    s.d = c[0] * zz;
    s.pp = r * norm * zz;
    if (periodic) s.q2 = ((nn == 1) ? (-a[0] - c[nn - 1]) : -a[0]) * zz;
    return;
  }
  const double bl = b[k - 1] + l;
  const double ad = a[k - 1] * s.d;
  const double den = bl - ad;
  if (periodic) {
    const double zq = 1. / den;
    const double rq = (k == nn) ? -c[nn - 1] : 0.;
    s.q2 = (rq - a[k - 1] * s.q2) * zq;
  }
  const double tol = eps * fmax(fabs(bl), fabs(ad));
  if (k == nn && fabs(den) <= tol) {
    s.d = 0.;
    s.pp = 0.;
  } else {
    const double zz = rcp64<FAST>(den);
    s.d = c[k - 1] * zz;
    s.pp = (r * norm - a[k - 1] * s.pp) * zz;
  }
}

// PER: periodic z (solver.f90:203-307); a compile-time flag so that the regular
// kernel carries no Sherman-Morrison state in registers
template <typename T, bool PER>
__global__ void __launch_bounds__(256) k_gaussel(ZArgs<T> z) {
  const ZSys<T>& zs = z.sys[blockIdx.y];
  const long line = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= z.nl) return;
  const double l = zs.lam[line];
  const double eps = z.pin_eps;
  T* P = zs.W + woff((int)(line % z.n1), (int)(line / z.n1), 0, z.n2, z.nic);  // row 0
  const long PS = z.PS;
  constexpr int periodic = PER ? 1 : 0;
  const int nn = periodic ? zs.n - 1 : zs.n;
  const int nblk = (nn + CK - 1) / CK;
  const kcoef a = (kcoef)zs.a;
  const kcoef b = (kcoef)zs.b;
  const kcoef c = (kcoef)zs.c;
  const double norm = zs.norm;
  // ---- forward sweep (checkpoints only), next chunk's loads issued before this chunk's recurrence ----
  ZRow s = {0., 0., 0.};
  double rr[CK], rn[CK];
#pragma unroll
  for (int t = 0; t < CK; t++) rr[t] = (t + 1 <= nn) ? (double)P[zrow(t, PS)] : 0.;
  for (int bb = 0; bb < nblk; bb++) {
    const int k0 = bb * CK + 1;
#pragma unroll
    for (int t = 0; t < CK; t++) {
      const int k = k0 + CK + t;
      rn[t] = (k <= nn) ? (double)P[zrow(k - 1, PS)] : 0.;
    }
#pragma unroll
    for (int t = 0; t < CK; t++)
      if (k0 + t <= nn) z_row(k0 + t, nn, periodic, l, rr[t], norm, a, b, c, eps, s);
    const long o = (long)bb * z.nl + line;
    zs.ckd[o] = s.d;
    zs.ckp[o] = s.pp;
    if (periodic) zs.ckq[o] = s.q2;
#pragma unroll
    for (int t = 0; t < CK; t++) rr[t] = rn[t];
  }
  // ---- backward sweep: previous block's rhs prefetched while this block is recomputed and substituted ----
  double xn = 0., xq = 0.;
  double x1 = 0., xq1 = 0., xnn = 0., xqnn = 0.;
  {
    const int k0 = (nblk - 1) * CK + 1;
#pragma unroll
    for (int t = 0; t < CK; t++) rr[t] = (k0 + t <= nn) ? (double)P[zrow(k0 + t - 1, PS)] : 0.;
  }
  for (int bb = nblk - 1; bb >= 0; bb--) {
    const int k0 = bb * CK + 1;
    const int k1 = min(k0 + CK - 1, nn);
    if (bb > 0) {
#pragma unroll
      for (int t = 0; t < CK; t++) rn[t] = (double)P[zrow(k0 - CK + t - 1, PS)];
    }
    ZRow r0 = {0., 0., 0.};
    if (bb > 0) {
      const long o = (long)(bb - 1) * z.nl + line;
      r0.d = zs.ckd[o];
      r0.pp = zs.ckp[o];
      r0.q2 = periodic ? zs.ckq[o] : 0.;
    }
    double dd[CK], pv[CK], qv[CK];
#pragma unroll
    for (int t = 0; t < CK; t++) {
      if (k0 + t <= k1) {
        z_row(k0 + t, nn, periodic, l, rr[t], norm, a, b, c, eps, r0);
        dd[t] = r0.d;
        pv[t] = r0.pp;
        qv[t] = r0.q2;
      }
    }
#pragma unroll
    for (int t = CK - 1; t >= 0; t--) {
      const int k = k0 + t;
      if (k <= k1) {
        double x, x2 = 0.;
        if (k == nn) {
          x = pv[t];
          x2 = qv[t];
          xnn = x;
          xqnn = x2;
        } else {
          x = pv[t] - dd[t] * xn;
          if (periodic) x2 = qv[t] - dd[t] * xq;
        }
        P[zrow(k - 1, PS)] = (T)x;
        if (periodic) z.X2[(long)(k - 1) * z.nl + line] = x2;
        xn = x;
        xq = x2;
        if (k == 1) {
          x1 = x;
          xq1 = x2;
        }
      }
    }
#pragma unroll
    for (int t = 0; t < CK; t++) rr[t] = rn[t];
  }
  if (!periodic) return;
  // ---- periodic closure (solver.f90:268-296) and Sherman-Morrison correction ----
  const int m = nn + 1;  // closure row, 1-based
  const double den = b[m - 1] + l + c[m - 1] * xq1 + a[m - 1] * xqnn;
  const double tol = eps * fmax(fabs(b[m - 1] + l), fabs(c[m - 1] * xq1 + a[m - 1] * xqnn));
  double xm;
  if (fabs(den) <= tol) xm = 0.;
  else xm = ((double)P[zrow(m - 1, PS)] * norm - c[m - 1] * x1 - a[m - 1] * xnn) / den;
  P[zrow(m - 1, PS)] = (T)xm;
  for (int k = 1; k <= nn; k++) {
    const long o = zrow(k - 1, PS);
    P[o] = (T)((double)P[o] + z.X2[(long)(k - 1) * z.nl + line] * xm);
  }
}

// Non-periodic z (solver.f90:147-200): the same elimination and checkpoints as
// k_gaussel<T, false>, arranged so that every block but the one holding the last
// row runs without branches (no row-1, range or pin tests), which lets the
// compiler batch the coefficient loads and interleave independent work.  The
// right-hand-side prefetch registers hold the storage type.  Blocks are read
// whole: rows past the last one (up to CK - 2 of them) lie inside the work buffer
// (wbuf_elems) and their values are never used.  FP64 storage gives
// the same operations as k_gaussel; FAST replaces the divisions by rcp64<true>.
template <typename T, bool FAST>
__device__ __forceinline__ void thomas_body(const ZArgs<T>& z, const ZSys<T>& zs, long line) {
  const double l = zs.lam[line];
  const double eps = z.pin_eps, norm = zs.norm;
  T* P = zs.W + woff((int)(line % z.n1), (int)(line / z.n1), 0, z.n2, z.nic);  // row 0
  const long PS = z.PS;
  // first row of checkpoint block bb; its rows follow at stride ZT
  auto blk = [&](int bb) { return P + zrow(bb * CK, PS); };
  const int nn = zs.n;
  const int nblk = (nn + CK - 1) / CK;
  const kcoef a = (kcoef)zs.a;
  const kcoef b = (kcoef)zs.b;
  const kcoef c = (kcoef)zs.c;
  // ---- forward sweep: checkpoints (d, p') at the end of every block but the last ----
  double d = 0., pp = 0.;
  T rr[CK], rn[CK];
#pragma unroll
  for (int t = 0; t < CK; t++) rr[t] = P[t * ZT];
  for (int bb = 0; bb < nblk; bb++) {
    const int k0 = bb * CK + 1;
    if (bb + 1 < nblk) {
      const T* B = blk(bb + 1);
#pragma unroll
      for (int t = 0; t < CK; t++) rn[t] = B[t * ZT];
    }
    if (bb == nblk - 1) {
      ZRow s = {d, pp, 0.};
#pragma unroll
      for (int t = 0; t < CK; t++)
        if (k0 + t <= nn) z_row<FAST>(k0 + t, nn, 0, l, (double)rr[t], norm, a, b, c, eps, s);
      break;
    }
    if (bb == 0) {
      const double zz = rcp64<FAST>(b[0] + l);
      d = c[0] * zz;
      pp = (double)rr[0] * norm * zz;
#pragma unroll
      for (int t = 1; t < CK; t++) z_row_mid<FAST>(l, (double)rr[t], norm, a[t], b[t], c[t], d, pp);
    } else {
#pragma unroll
      for (int t = 0; t < CK; t++)
        z_row_mid<FAST>(l, (double)rr[t], norm, a[k0 + t - 1], b[k0 + t - 1], c[k0 + t - 1], d, pp);
    }
    const long o = (long)bb * z.nl + line;
    zs.ckd[o] = d;
    zs.ckp[o] = pp;
#pragma unroll
    for (int t = 0; t < CK; t++) rr[t] = rn[t];
  }
  // ---- backward sweep: recompute each block from its checkpoint, then substitute ----
  double xn = 0.;
  {
    const T* B = blk(nblk - 1);
#pragma unroll
    for (int t = 0; t < CK; t++) rr[t] = B[t * ZT];
  }
  for (int bb = nblk - 1; bb >= 0; bb--) {
    const int k0 = bb * CK + 1;
    T* const B = blk(bb);
    double d0 = 0., p0 = 0.;
    if (bb > 0) {
      const T* Bp = blk(bb - 1);
#pragma unroll
      for (int t = 0; t < CK; t++) rn[t] = Bp[t * ZT];
      const long o = (long)(bb - 1) * z.nl + line;
      d0 = zs.ckd[o];
      p0 = zs.ckp[o];
    }
    double dd[CK], pv[CK];
    if (bb == nblk - 1) {
      ZRow s = {d0, p0, 0.};
#pragma unroll
      for (int t = 0; t < CK; t++) {
        if (k0 + t <= nn) {
          z_row<FAST>(k0 + t, nn, 0, l, (double)rr[t], norm, a, b, c, eps, s);
          dd[t] = s.d;
          pv[t] = s.pp;
        }
      }
#pragma unroll
      for (int t = CK - 1; t >= 0; t--) {
        const int k = k0 + t;
        if (k <= nn) {
          const double x = (k == nn) ? pv[t] : pv[t] - dd[t] * xn;
          B[t * ZT] = (T)x;
          xn = x;
        }
      }
    } else {
      if (bb == 0) {
        const double zz = rcp64<FAST>(b[0] + l);
        d0 = c[0] * zz;
        p0 = (double)rr[0] * norm * zz;
        dd[0] = d0;
        pv[0] = p0;
#pragma unroll
        for (int t = 1; t < CK; t++) {
          z_row_mid<FAST>(l, (double)rr[t], norm, a[t], b[t], c[t], d0, p0);
          dd[t] = d0;
          pv[t] = p0;
        }
      } else {
#pragma unroll
        for (int t = 0; t < CK; t++) {
          z_row_mid<FAST>(l, (double)rr[t], norm, a[k0 + t - 1], b[k0 + t - 1], c[k0 + t - 1], d0, p0);
          dd[t] = d0;
          pv[t] = p0;
        }
      }
#pragma unroll
      for (int t = CK - 1; t >= 0; t--) {
        const double x = pv[t] - dd[t] * xn;
        B[t * ZT] = (T)x;
        xn = x;
      }
    }
#pragma unroll
    for (int t = 0; t < CK; t++) rr[t] = rn[t];
  }
}

template <typename T, bool FAST>
__global__ void __launch_bounds__(256) k_thomas(ZArgs<T> z) {
  const long line = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= z.nl) return;
  thomas_body<T, FAST>(z, z.sys[blockIdx.y], line);
}

// ---------------------------------------------------------------------------
// Distributed tridiagonal solve (gaussel_dtdma, solver.f90:309-517): the z rows are
// split into nslab consecutive slabs, as the ranks of a z decomposition hold them
// (slab s: n/nslab rows, the first n%nslab slabs one more).  On several GPUs each
// slab's elimination and back-substitution run on its own GPU and only the 2*nslab
// boundary rows per line travel; on one GPU the three phases show those kernels at
// their real shape, and the slabs give nslab times more independent sweeps.
//
// The arithmetic is the reference's: in each slab the lower diagonals are eliminated
// downwards (every row then refers to the slab's first unknown), the upper diagonals
// upwards (every row refers to the slab's last unknown), the first row is closed
// against the second, and the interior follows from the two boundary unknowns,
// x_k = p_k - aa_k x_first - cc_k x_last.  Instead of storing the eliminated
// (p, aa, cc) of every row (three FP64 arrays written and read twice), the kernels keep
// the elimination states only at the ends of DCK-row blocks and recompute a block's rows
// from its checkpoint and the right-hand side, as the plain sweep does: the right-hand
// side is read three times and the solution written once.
//
// Periodic z: a(1) couples row 1 to row n and c(n) row n to row 1, so the reduced
// system of the boundary rows is cyclic; it is closed as solver.f90:469-498 does.  The
// reference pins no zero mode here; a line whose final reduced pivot (or cyclic
// closure) vanishes to rounding (the constant mode of an all-Neumann/periodic
// operator) gets its last unknown pinned to zero, the gauge the plain sweep chooses.
// ---------------------------------------------------------------------------
constexpr int DCK = 16;  // rows per checkpoint block of the slab interiors

template <typename T>
struct DtArgs {
  T* W;
  long nl;
  int n1, n2, nic;
  long PS;
  int n, nslab, periodic, nbk;  // nbk: checkpoint blocks of the largest slab's interior
  double norm, pin_eps;
  const double *a, *b, *c, *lam;
  double *ckl, *cku;     // [nbk][3][nslab][nl]: lower / upper elimination state per block
  double *ra, *rc, *rp;  // [2*nslab][nl]: the reduced system
  double *rc0, *r2;      // periodic: [2*nslab][nl], the unmodified rc and the auxiliary solution
};

__device__ __forceinline__ void slab_range(int n, int nslab, int s, int& lo, int& cnt) {
  const int m = n / nslab, r = n % nslab;
  cnt = m + (s < r ? 1 : 0);
  lo = s * m + (s < r ? s : r);
}

struct DtRow {
  double p, a, c;
};

// one slab of one line: coefficients, the line's rhs rows and the checkpoint addressing
template <typename T>
struct DtSlab {
  long line, nl, PS;
  int s, nslab, lo, ns, nb, haslam;
  double l, norm;
  kcoef a, b, c;
  T* P;
  const double* ckl;
  __device__ DtSlab(const DtArgs<T>& z, long line_, int s_) : line(line_), nl(z.nl), PS(z.PS), s(s_) {
    nslab = z.nslab;
    slab_range(z.n, z.nslab, s, lo, ns);
    nb = (ns - 2 + DCK - 1) / DCK;
    haslam = z.lam != nullptr;
    l = haslam ? z.lam[line] : 0.;
    norm = z.norm;
    a = (kcoef)(z.a + lo);
    b = (kcoef)(z.b + lo);
    c = (kcoef)(z.c + lo);
    P = z.W + woff((int)(line % z.n1), (int)(line / z.n1), 0, z.n2, z.nic);
    ckl = z.ckl;
  }
  // lambda present: bb = b + lambda; absent: b itself (solver.f90:344-347, 395-398)
  __device__ __forceinline__ double bb(int k) const { return haslam ? b[k] + l : b[k]; }
  __device__ __forceinline__ T& row(int k) const { return P[zrow(lo + k, PS)]; }
  __device__ __forceinline__ long ck(int blk, int v) const { return (((long)blk * 3 + v) * nslab + s) * nl + line; }
  // rows 0 and 1 start the elimination (solver.f90:350-353)
  __device__ __forceinline__ DtRow first(int k, double r) const {
    const double zz = 1. / bb(k);
    return {r * norm * zz, a[k] * zz, c[k] * zz};
  }
  // elimination of the lower diagonal of row k >= 2 (solver.f90:357-362)
  __device__ __forceinline__ void lower(int k, double r, DtRow& L) const {
    const double zk = 1. / (bb(k) - a[k] * L.c);
    L.p = (r * norm - a[k] * L.p) * zk;
    L.a = -a[k] * L.a * zk;
    L.c = c[k] * zk;
  }
  // the interior block blk: rows r0 .. r0 + cnt - 1 (rows 1 .. ns-2 of the slab)
  __device__ __forceinline__ void block(int blk, int& r0, int& cnt) const {
    r0 = 1 + blk * DCK;
    cnt = min(DCK, ns - 1 - r0);
  }
  // the rhs rows of a block; rows past the slab's interior lie inside the work buffer
  // (at most DCK - 2 rows past row n-1) and are never used
  __device__ __forceinline__ void load(int blk, T* rr) const {
    const int r0 = 1 + blk * DCK;
#pragma unroll
    for (int t = 0; t < DCK; t++) rr[t] = row(r0 + t);
  }
  // the lower-elimination states of a block's rows, from the checkpoint before it
  __device__ __forceinline__ void recompute(int blk, const T* rr, DtRow* Lv) const {
    int r0, cnt;
    block(blk, r0, cnt);
    DtRow L;
    if (blk == 0) {
      L = first(1, (double)rr[0]);
    } else {
      L = {ckl[ck(blk - 1, 0)], ckl[ck(blk - 1, 1)], ckl[ck(blk - 1, 2)]};
      lower(r0, (double)rr[0], L);
    }
    Lv[0] = L;
#pragma unroll
    for (int t = 1; t < DCK; t++) {
      if (t < cnt) lower(r0 + t, (double)rr[t], L);
      Lv[t] = L;
    }
  }
  // elimination of the upper diagonal of row k <= ns-3 (solver.f90:366-370); row ns-2 starts it
  __device__ __forceinline__ void upper(int k, const DtRow& L, DtRow& U) const {
    if (k == ns - 2) {
      U = L;
    } else {
      U.p = L.p - L.c * U.p;
      U.a = L.a - L.c * U.a;
      U.c = -L.c * U.c;
    }
  }
};

// phase 1, one thread per (line, slab): the slab's first and last rows of the reduced system
template <typename T>
__global__ void __launch_bounds__(256) k_dtdma_local(DtArgs<T> z) {
  const long line = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= z.nl) return;
  const DtSlab<T> S(z, line, blockIdx.y);
  const int nb = S.nb, ns = S.ns;
  const DtRow R0 = S.first(0, (double)S.row(0));
  // downwards: lower elimination, checkpoints at the end of every interior block but the last
  T rr[DCK], rn[DCK];
  DtRow L = {0., 0., 0.};
  S.load(0, rr);
  for (int blk = 0; blk < nb; blk++) {
    if (blk + 1 < nb) S.load(blk + 1, rn);
    int r0, cnt;
    S.block(blk, r0, cnt);
#pragma unroll
    for (int t = 0; t < DCK; t++) {
      if (t < cnt) {
        if (r0 + t == 1)
          L = S.first(1, (double)rr[t]);
        else
          S.lower(r0 + t, (double)rr[t], L);
      }
    }
    if (blk + 1 < nb) {
      z.ckl[S.ck(blk, 0)] = L.p;
      z.ckl[S.ck(blk, 1)] = L.a;
      z.ckl[S.ck(blk, 2)] = L.c;
    }
#pragma unroll
    for (int t = 0; t < DCK; t++) rr[t] = rn[t];
  }
  S.lower(ns - 1, (double)S.row(ns - 1), L);  // the last row keeps its lower-eliminated form
  const long r1 = (long)(2 * S.s + 1) * z.nl + line;
  z.ra[r1] = L.a;
  z.rc[r1] = L.c;
  z.rp[r1] = L.p;
  // upwards: each block recomputed from its checkpoint, the upper elimination carried
  // through it; the state entering block blk (at the first row of block blk+1) is kept
  DtRow U = {0., 0., 0.};
  S.load(nb - 1, rr);
  for (int blk = nb - 1; blk >= 0; blk--) {
    if (blk > 0) S.load(blk - 1, rn);
    if (blk + 1 < nb) {
      z.cku[S.ck(blk, 0)] = U.p;
      z.cku[S.ck(blk, 1)] = U.a;
      z.cku[S.ck(blk, 2)] = U.c;
    }
    DtRow Lv[DCK];
    S.recompute(blk, rr, Lv);
    int r0, cnt;
    S.block(blk, r0, cnt);
#pragma unroll
    for (int t = DCK - 1; t >= 0; t--)
      if (t < cnt) S.upper(r0 + t, Lv[t], U);
#pragma unroll
    for (int t = 0; t < DCK; t++) rr[t] = rn[t];
  }
  // the first row against the second (solver.f90:372-375)
  const double zf = 1. / (1. - U.a * R0.c);
  const long q0 = (long)(2 * S.s) * z.nl + line;
  z.rp[q0] = (R0.p - R0.c * U.p) * zf;
  z.ra[q0] = R0.a * zf;
  z.rc[q0] = -R0.c * U.c * zf;
}

// phase 2, one thread per line: the reduced system of the 2*nslab boundary rows (unit
// diagonal, solver.f90:446-468), cyclic for a periodic z (solver.f90:469-498)
template <typename T>
__global__ void __launch_bounds__(256) k_dtdma_reduced(DtArgs<T> z) {
  const long line = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= z.nl) return;
  const long nl = z.nl;
  const int nr = 2 * z.nslab;
  const int nn = z.periodic ? nr - 1 : nr;
  double* ra = z.ra + line;
  double* rc = z.rc + line;
  double* rp = z.rp + line;
  double* rc0 = z.periodic ? z.rc0 + line : nullptr;
  if (z.periodic)
    for (int k = 0; k < nr; k++) rc0[(long)k * nl] = rc[(long)k * nl];
  double pprev = rp[0], cprev = rc[0];
  for (int k = 1; k < nn; k++) {
    const double ak = ra[(long)k * nl];
    const double den = 1. - ak * cprev;
    if (!z.periodic && k == nn - 1 && fabs(den) <= z.pin_eps * fmax(1., fabs(ak * cprev))) {
      pprev = 0.;  // the constant mode: pin the last unknown
      cprev = 0.;
    } else {
      const double zk = 1. / den;
      pprev = (rp[(long)k * nl] - ak * pprev) * zk;
      cprev = rc[(long)k * nl] * zk;
    }
    rp[(long)k * nl] = pprev;
    rc[(long)k * nl] = cprev;
  }
  for (int k = nn - 2; k >= 0; k--) {
    pprev = rp[(long)k * nl] - rc[(long)k * nl] * pprev;
    rp[(long)k * nl] = pprev;
  }
  if (!z.periodic) return;
  // the auxiliary solution for a unit last unknown, with the unmodified coefficients
  double* r2 = z.r2 + line;
  double qprev = -ra[0];
  cprev = rc0[0];
  if (nn == 1) qprev -= rc0[0];
  r2[0] = qprev;
  for (int k = 1; k < nn; k++) {
    const double ak = ra[(long)k * nl];
    const double zk = 1. / (1. - ak * cprev);
    const double rhs = k == nn - 1 ? -rc0[(long)k * nl] : 0.;
    qprev = (rhs - ak * qprev) * zk;
    cprev = rc0[(long)k * nl] * zk;
    r2[(long)k * nl] = qprev;
    rc0[(long)k * nl] = cprev;
  }
  for (int k = nn - 2; k >= 0; k--) {
    qprev = r2[(long)k * nl] - rc0[(long)k * nl] * qprev;
    r2[(long)k * nl] = qprev;
  }
  // the last unknown, then the correction of the others
  const double cl = rc[(long)nn * nl], al = ra[(long)nn * nl];  // row nn: never modified
  const double q0 = r2[0], qn = r2[(long)(nn - 1) * nl];
  const double den = 1. + cl * q0 + al * qn;
  const double xl = fabs(den) <= z.pin_eps * fmax(1., fabs(cl * q0 + al * qn))
                        ? 0.
                        : (rp[(long)nn * nl] - cl * rp[0] - al * rp[(long)(nn - 1) * nl]) / den;
  rp[(long)nn * nl] = xl;
  for (int k = 0; k < nn; k++) rp[(long)k * nl] += r2[(long)k * nl] * xl;
}

// phase 3, one thread per (line, slab): the boundary unknowns in, the interior rows
// x_k = p_k - aa_k x_first - cc_k x_last from each block's recomputed elimination (solver.f90:503-514)
template <typename T>
__global__ void __launch_bounds__(256) k_dtdma_back(DtArgs<T> z) {
  const long line = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= z.nl) return;
  const DtSlab<T> S(z, line, blockIdx.y);
  const int nb = S.nb, ns = S.ns;
  const double x0 = z.rp[(long)(2 * S.s) * z.nl + line], xl = z.rp[(long)(2 * S.s + 1) * z.nl + line];
  T rr[DCK], rn[DCK];
  S.load(0, rr);
  for (int blk = 0; blk < nb; blk++) {
    if (blk + 1 < nb) S.load(blk + 1, rn);
    DtRow Lv[DCK];
    S.recompute(blk, rr, Lv);
    DtRow U = {0., 0., 0.};
    if (blk + 1 < nb) U = {z.cku[S.ck(blk, 0)], z.cku[S.ck(blk, 1)], z.cku[S.ck(blk, 2)]};
    int r0, cnt;
    S.block(blk, r0, cnt);
#pragma unroll
    for (int t = DCK - 1; t >= 0; t--) {
      if (t < cnt) {
        S.upper(r0 + t, Lv[t], U);
        S.row(r0 + t) = (T)(U.p - U.a * x0 - U.c * xl);
      }
    }
#pragma unroll
    for (int t = 0; t < DCK; t++) rr[t] = rn[t];
  }
  S.row(0) = (T)x0;
  S.row(ns - 1) = (T)xl;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// W_N^t = exp(-2 pi i t/N), t < N, and the DCT/DST twiddles exp(-i pi k/(2N)), k <= N
static void make_tables(int N, std::vector<cplx<double>>& tw, std::vector<cplx<double>>& mk) {
  tw.assign(N > 0 ? N : 1, cplx<double>{1., 0.});
  mk.assign(N + 1, cplx<double>{1., 0.});
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
}

// bytes of a twiddle table as the device holds it: N double values, then N float ones
static size_t tw_bytes(size_t n) { return n * (sizeof(cplx<double>) + sizeof(cplx<float>)); }
static void tw_pack(const std::vector<cplx<double>>& tw, unsigned char* out) {
  std::memcpy(out, tw.data(), tw.size() * sizeof(cplx<double>));
  cplx<float>* f = (cplx<float>*)(out + tw.size() * sizeof(cplx<double>));
  for (size_t t = 0; t < tw.size(); t++) f[t] = {(float)tw[t].re, (float)tw[t].im};
}

// ---------------------------------------------------------------------------
// Transforms of any length: a direction whose length the FFT passes do not take
// (any factor other than 2, 3 and 5, more than three odd factors, an odd factor on
// a line shorter than 8, more than 4096 points, or an odd n1, whose rows the passes
// cannot move as element pairs) applies its transform as a dense n x n matrix.  The
// matrix holds FFTW's unnormalised definition of the direction's kind (fft.f90:260-313)
// with the output in the order the FFT passes produce (the packed real-FFT order of the
// reference's GPU build for PP, fft.f90:402-451 and initsolver.f90:98-117; natural
// order for the DCT/DST kinds), so lambdaxy, the z solve and the backward pass are
// those of any other length.  Each pass is one FP64 product of the matrix with a block
// of lines: 64 x 64 output tiles over 16-deep LDS slices of the matrix and the lines,
// a 4 x 4 register block per lane, computed in FP64 whatever the storage type.
// ---------------------------------------------------------------------------
template <typename T>
struct DenseArgs {
  const double* M;  // n x n, row-major: M[m * n + e] multiplies input e into output m
  int n;
  long nlines;
  Layout L;
  int n2, nic;
  const T* in;   // the field (x forward), or a work buffer
  T* out;        // a work buffer, or the field (x backward)
};

constexpr int DN_TM = 64, DN_TL = 64, DN_TK = 16;
constexpr int DENSE_MAX = 8192;  // longest dense direction (two 512 MB matrices)

// MODE 0: x forward (field rows -> work buffer); 1: x backward (work buffer -> field rows);
// 2: y (work buffer columns -> work buffer columns, out of place)
template <typename T, int MODE>
__device__ __forceinline__ long dense_off(const DenseArgs<T>& a, bool field, long l, int e) {
  if (MODE == 2) return woff((int)(l % a.L.n1), e, (int)(l / a.L.n1), a.n2, a.nic);
  const int j = (int)(l % a.L.n2), k = (int)(l / a.L.n2);
  return field ? a.L.at(1 + e, 1 + j, 1 + k) : woff(e, j, k, a.n2, a.nic);
}

template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_dense(DenseArgs<T> a) {
  __shared__ double As[DN_TK][DN_TM + 1], Bs[DN_TK][DN_TL + 1];
  const int t = threadIdx.x, tx = t % 16, ty = t / 16;
  const long l0 = (long)blockIdx.x * DN_TL;
  const int m0 = blockIdx.y * DN_TM;
  const int n = a.n;
  double acc[4][4] = {};
  for (int e0 = 0; e0 < n; e0 += DN_TK) {
    for (int q = t; q < DN_TM * DN_TK; q += 256) {  // matrix slice, e contiguous
      const int r = q / DN_TK, cc = q % DN_TK;
      const int m = m0 + r, e = e0 + cc;
      As[cc][r] = (m < n && e < n) ? a.M[(long)m * n + e] : 0.;
    }
    for (int q = t; q < DN_TK * DN_TL; q += 256) {  // lines: along e for x, along the lines for y
      const int r = MODE == 2 ? q / DN_TL : q % DN_TK, cc = MODE == 2 ? q % DN_TL : q / DN_TK;
      const int e = e0 + r;
      const long l = l0 + cc;
      Bs[r][cc] = (e < n && l < a.nlines) ? (double)a.in[dense_off<T, MODE>(a, MODE == 0, l, e)] : 0.;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < DN_TK; kk++) {
      double x[4], y[4];
#pragma unroll
      for (int i = 0; i < 4; i++) x[i] = As[kk][ty * 4 + i];
#pragma unroll
      for (int i = 0; i < 4; i++) y[i] = Bs[kk][tx + 16 * i];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int jj = 0; jj < 4; jj++) acc[i][jj] = fma(x[i], y[jj], acc[i][jj]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int m = m0 + ty * 4 + i;
    if (m >= n) continue;
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
      const long l = l0 + tx + 16 * jj;
      if (l < a.nlines) a.out[dense_off<T, MODE>(a, MODE == 1, l, m)] = (T)acc[i][jj];
    }
  }
}

template <typename T>
__global__ void k_copy(const T* __restrict__ s, T* __restrict__ d, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) d[i] = s[i];
}

// cos(pi p / q) and sin(pi p / q) from the integer angle p/q reduced to [0, 2)
static long double cospi_q(long p, long q) {
  p %= 2 * q;
  if (p < 0) p += 2 * q;
  return cosl(3.141592653589793238462643383279502884L * (long double)p / (long double)q);
}
static long double sinpi_q(long p, long q) {
  p %= 2 * q;
  if (p < 0) p += 2 * q;
  return sinl(3.141592653589793238462643383279502884L * (long double)p / (long double)q);
}

// the n x n matrix of a kind's forward (dir > 0) or backward transform, with FFTW's
// unnormalised definitions (fft.f90:260-313) and the FFT passes' output order
static void dense_matrix(int kind, int n, int dir, std::vector<double>& M) {
  M.assign((size_t)n * n, 0.);
  auto at = [&](int m, int e) -> double& { return M[(size_t)m * n + e]; };
  const long N = n;
  switch (kind) {
    case K_PP: {
      // R2HC r_k = sum x_j cos(2 pi jk/n), i_k = -sum x_j sin(2 pi jk/n), stored
      // (r0, r_{n/2} | i_{(n-1)/2}, r1, i1, ...); HC2R its unnormalised inverse
      const int h = n / 2;
      auto slot_r = [&](int k) { return (n % 2 == 0 && k == h) ? 1 : 2 * k; };  // r_k (k = 0 .. h)
      auto slot_i = [&](int k) { return (n % 2 == 1 && k == h) ? 1 : 2 * k + 1; };  // i_k (0 < k < n/2)
      for (int k = 0; k <= h; k++) {
        const bool has_i = k > 0 && 2 * k != n;
        for (int j = 0; j < n; j++) {
          const long pj = 2L * j * k;
          if (dir > 0) {
            at(slot_r(k), j) = (double)cospi_q(pj, N);
            if (has_i) at(slot_i(k), j) = -(double)sinpi_q(pj, N);
          } else {
            const double w = (k == 0 || 2 * k == n) ? 1. : 2.;
            at(j, slot_r(k)) = w * (double)cospi_q(pj, N);
            if (has_i) at(j, slot_i(k)) = -2. * (double)sinpi_q(pj, N);
          }
        }
      }
      break;
    }
    case K_NN:  // REDFT10 forward, REDFT01 backward
      for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++) {
          if (dir > 0) at(k, j) = 2. * (double)cospi_q((2L * j + 1) * k, 2 * N);
          else at(k, j) = j == 0 ? 1. : 2. * (double)cospi_q((long)j * (2L * k + 1), 2 * N);
        }
      break;
    case K_DD:    // RODFT10 forward, RODFT01 backward
    case K_DNF: { // RODFT01 forward, RODFT10 backward
      const bool ro10 = (kind == K_DD) == (dir > 0);
      for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++) {
          if (ro10) at(k, j) = 2. * (double)sinpi_q((2L * j + 1) * (k + 1), 2 * N);
          else at(k, j) = j == n - 1 ? ((k % 2) ? -1. : 1.) : 2. * (double)sinpi_q((j + 1L) * (2L * k + 1), 2 * N);
        }
      break;
    }
    case K_ND:  // REDFT11, its own inverse up to 2n
      for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++) at(k, j) = 2. * (double)cospi_q((2L * j + 1) * (2L * k + 1), 4 * N);
      break;
    case K_DN:  // RODFT11
      for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++) at(k, j) = 2. * (double)sinpi_q((2L * j + 1) * (2L * k + 1), 4 * N);
      break;
    default: {  // K_DDF: RODFT00 on the n-1 interior faces (fft.f90:80-82); face n passes through
      const long m1 = N;  // (n - 1) + 1
      for (int k = 0; k < n - 1; k++)
        for (int j = 0; j < n - 1; j++) at(k, j) = 2. * (double)sinpi_q((j + 1L) * (k + 1), m1);
      at(n - 1, n - 1) = 1.;
      break;
    }
  }
}

static int upload_dense(cans_plan* p, int d) {
  const int n = p->len[d];
  for (int dir = 0; dir < 2; dir++) {
    std::vector<double> M;
    dense_matrix(p->kind[d], n, dir == 0 ? +1 : -1, M);
    CANS_HIP_TRY(hipMalloc((void**)&p->dm[d][dir], M.size() * sizeof(double)));
    CANS_HIP_TRY(hipMemcpy(p->dm[d][dir], M.data(), M.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  return CANS_OK;
}

// one dense pass of direction d for one system: x (field <-> work buffer) or y (in the work buffer)
template <typename T>
static int dense_pass(cans_ctx* c, const cans_plan* p, T* fld, T* W, int d, int dir) {
  const Layout& L = c->L;
  DenseArgs<T> a;
  a.M = p->dm[d][dir > 0 ? 0 : 1];
  a.n = p->len[d];
  a.L = L;
  a.n2 = L.n2;
  a.nic = wnic(L.n1);
  a.nlines = d == 0 ? (long)L.n2 * L.n3 : (long)L.n1 * L.n3;
  const dim3 grid((unsigned)cdiv(a.nlines, DN_TL), (unsigned)cdiv(a.n, DN_TM));
  if (d == 0) {
    a.in = dir > 0 ? fld : W;
    a.out = dir > 0 ? W : fld;
    if (dir > 0) k_dense<T, 0><<<grid, 256, 0, c->stream>>>(a);
    else k_dense<T, 1><<<grid, 256, 0, c->stream>>>(a);
  } else {
    const long wel = wbuf_elems(L.n1, L.n2, L.n3);
    void* cp = nullptr;
    int rc = ensure_work(c, 11, (size_t)wel * sizeof(T), &cp);
    if (rc) return rc;
    k_copy<T><<<(unsigned)std::min<long>(cdiv(wel, 256), 65536), 256, 0, c->stream>>>(W, (T*)cp, wel);
    CANS_LAUNCH_CHECK("k_copy");
    a.in = (const T*)cp;
    a.out = W;
    k_dense<T, 2><<<grid, 256, 0, c->stream>>>(a);
  }
  CANS_LAUNCH_CHECK("k_dense");
  return CANS_OK;
}

static int upload_tables(cans_plan* p, int d) {
  std::vector<cplx<double>> tw, mk;
  make_tables(p->nfft[d], tw, mk);
  std::vector<unsigned char> twp(tw_bytes(tw.size()));
  tw_pack(tw, twp.data());
  CANS_HIP_TRY(hipMalloc((void**)&p->tw[d], twp.size()));
  CANS_HIP_TRY(hipMalloc((void**)&p->mk[d], mk.size() * sizeof(cplx<double>)));
  CANS_HIP_TRY(hipMemcpy(p->tw[d], twp.data(), twp.size(), hipMemcpyHostToDevice));
  CANS_HIP_TRY(hipMemcpy(p->mk[d], mk.data(), mk.size() * sizeof(cplx<double>), hipMemcpyHostToDevice));
  return CANS_OK;
}

template <typename T>
static int xpass(cans_ctx* c, const cans_plan* p, T* const* fld, T* const* W, int nsys, int dir,
                 const DivSrc<T>* div = nullptr) {
  XArgs<T> a;
  a.div = div != nullptr && dir > 0;
  if (a.div) a.ds = *div;
  a.L = c->L;
  for (int q = 0; q < nsys; q++) {
    a.fld[q] = fld[q];
    a.W[q] = W[q];
  }
  a.N = p->nfft[0];
  a.nd = p->len[0];
  a.logN = p->logn[0];
  a.odd = p->odd[0];
  a.Tn = a.N >= 8 ? a.N / 8 : 1;
  const int xt = p->xthreads[sizeof(T) == sizeof(double)] ? p->xthreads[sizeof(T) == sizeof(double)] : 256;
  a.pairs = a.Tn >= xt ? 1 : xt / a.Tn;
  a.kind = p->kind[0];
  a.tw = p->tw[0];
  a.mk = p->mk[0];
  a.r4 = p->radix4[0][sizeof(T) == sizeof(double)];
  const long nrows = (long)c->L.n2 * c->L.n3;
  const long npairs = (nrows + 1) / 2;
  dim3 grid(cdiv(npairs, a.pairs), nsys);
  const size_t shm = (size_t)a.pairs * padded_len(a.N) * sizeof(cplx<T>);
  const bool res = a.div && a.ds.ncorr > 0;
  const int nt = a.pairs * a.Tn;
  if (a.kind == K_DNF) {  // face-centred DN: the K_DD code with the directions swapped
    if (a.div) return fail(CANS_EINVAL, "k_xpass: the divergence input needs a periodic x transform");
    if constexpr (sizeof(T) == sizeof(float)) {
      if (dir < 0) k_xpass_f<-1, 0, true><<<grid, nt, shm, c->stream>>>(a);
      else k_xpass_f<1, 0, true><<<grid, nt, shm, c->stream>>>(a);
    } else {
      if (dir < 0) k_xpass_d<-1, 0, true><<<grid, nt, shm, c->stream>>>(a);
      else k_xpass_d<1, 0, true><<<grid, nt, shm, c->stream>>>(a);
    }
    CANS_LAUNCH_CHECK("k_xpass");
    return CANS_OK;
  }
  // radix-4 stages where the tuner chose them: the plain and divergence-fused passes (the
  // residual pass of the refinement keeps radix 8)
  const bool r4 = a.r4 && !res;
  if constexpr (sizeof(T) == sizeof(float)) {
    if (dir < 0) {
      if (r4) k_xpass_f<-1, 0, false, 1><<<grid, nt, shm, c->stream>>>(a);
      else k_xpass_f<-1, 0><<<grid, nt, shm, c->stream>>>(a);
    } else if (res) {
      k_xpass_f<1, 2><<<grid, nt, shm, c->stream>>>(a);
    } else if (a.div) {
      if (r4) k_xpass_f<1, 1, false, 1><<<grid, nt, shm, c->stream>>>(a);
      else k_xpass_f<1, 1><<<grid, nt, shm, c->stream>>>(a);
    } else {
      if (r4) k_xpass_f<1, 0, false, 1><<<grid, nt, shm, c->stream>>>(a);
      else k_xpass_f<1, 0><<<grid, nt, shm, c->stream>>>(a);
    }
  } else {
    if (res) return fail(CANS_EUNSUPPORTED, "k_xpass: refinement residuals are formed for FP32 storage only");
    if (dir < 0) {
      if (r4) k_xpass_d<-1, 0, false, 1><<<grid, nt, shm, c->stream>>>(a);
      else k_xpass_d<-1, 0><<<grid, nt, shm, c->stream>>>(a);
    } else if (a.div) {
      if (r4) k_xpass_d<1, 1, false, 1><<<grid, nt, shm, c->stream>>>(a);
      else k_xpass_d<1, 1><<<grid, nt, shm, c->stream>>>(a);
    } else {
      if (r4) k_xpass_d<1, 0, false, 1><<<grid, nt, shm, c->stream>>>(a);
      else k_xpass_d<1, 0><<<grid, nt, shm, c->stream>>>(a);
    }
  }
  CANS_LAUNCH_CHECK("k_xpass");
  return CANS_OK;
}

template <typename T>
static int ypass(cans_ctx* c, const cans_plan* p, T* const* W, int nsys, int dir) {
  YArgs<T> a;
  a.n1 = c->L.n1;
  a.n2 = c->L.n2;
  a.n3 = c->L.n3;
  for (int q = 0; q < nsys; q++) a.W[q] = W[q];
  a.N = p->nfft[1];
  a.nd = p->len[1];
  a.logN = p->logn[1];
  a.odd = p->odd[1];
  a.Tn = a.N >= 8 ? a.N / 8 : 1;
  int lines = 256 / a.Tn;
  if (lines < 8) lines = 8;
  if (sizeof(T) < sizeof(double) && lines < 16) lines = 16;  // 128-byte rows of FP32 columns
  while (lines > 1 && (size_t)lines * a.N * sizeof(cplx<T>) > 65536) lines /= 2;
  while ((long)lines * a.Tn > 1024) lines /= 2;
  if (p->ylines[sizeof(T) == sizeof(double)]) lines = p->ylines[sizeof(T) == sizeof(double)];
  a.lines = lines;
  a.kind = p->kind[1];
  a.tw = p->tw[1];
  a.mk = p->mk[1];
  a.r4 = p->radix4[1][sizeof(T) == sizeof(double)];
  dim3 grid(cdiv((long)a.n3 * ((a.n1 + 1) / 2), lines), nsys);
  const size_t shm = (size_t)lines * a.N * sizeof(cplx<T>);
  if (dir > 0) launch_ypass<1, T>(grid, lines * a.Tn, shm, c->stream, a);
  else launch_ypass<-1, T>(grid, lines * a.Tn, shm, c->stream, a);
  CANS_LAUNCH_CHECK("k_ypass");
  return CANS_OK;
}

// one x or y pass (d = 0, 1) of nsys systems with their own plans (HetArgs)
template <typename T>
static int het_pass(cans_ctx* c, const cans_plan* const* ps, T* const* fld, T* const* W, int nsys, int d, int dir) {
  HetArgs<T> h;
  h.L = c->L;
  const long nrows = (long)c->L.n2 * c->L.n3;
  int threads = 0, bmax = 0;
  size_t shm = 0;
  for (int q = 0; q < nsys; q++) {
    const cans_plan* p = ps[q];
    h.fld[q] = fld ? fld[q] : nullptr;
    h.W[q] = W[q];
    h.N[q] = p->nfft[d];
    h.nd[q] = p->len[d];
    h.logN[q] = p->logn[d];
    h.odd[q] = p->odd[d];
    h.Tn[q] = h.N[q] >= 8 ? h.N[q] / 8 : 1;
    h.kind[q] = p->kind[d];
    h.tw[q] = p->tw[d];
    h.mk[q] = p->mk[d];
    h.r4[q] = p->radix4[d][sizeof(T) == sizeof(double)];
  }
  int tmax = 1;
  for (int q = 0; q < nsys; q++) tmax = std::max(tmax, h.Tn[q]);
  if (d == 0) {
    const int xt = std::max(256, tmax);  // lanes per work-group: whole teams of every system
    for (int q = 0; q < nsys; q++) {
      h.per[q] = std::max(1, xt / h.Tn[q]);
      h.blocks[q] = (int)cdiv((nrows + 1) / 2, h.per[q]);
      threads = std::max(threads, h.per[q] * h.Tn[q]);
      shm = std::max(shm, (size_t)h.per[q] * padded_len(h.N[q]) * sizeof(cplx<T>));
      bmax = std::max(bmax, h.blocks[q]);
    }
    const dim3 g(bmax, nsys);
    if (dir > 0) k_het_xpass<1, T><<<g, threads, shm, c->stream>>>(h);
    else k_het_xpass<-1, T><<<g, threads, shm, c->stream>>>(h);
    CANS_LAUNCH_CHECK("k_het_xpass");
    return CANS_OK;
  }
  // y: every system's work-group holds yt lanes, yt / Tn lines of its length (the same LDS for
  // all lengths of 8 or more), and at least 16 lines in FP32 (128-byte rows)
  int yt = std::max(256, tmax);
  for (int q = 0; q < nsys; q++)
    while ((long)(yt / h.Tn[q]) * 2 * (long)sizeof(T) < 128 && yt < 1024) yt *= 2;
  const int ncp = (c->L.n1 + 1) / 2;
  for (int q = 0; q < nsys; q++) {
    h.per[q] = std::max(1, yt / h.Tn[q]);
    h.blocks[q] = (int)cdiv((long)c->L.n3 * ncp, h.per[q]);
    threads = std::max(threads, h.per[q] * h.Tn[q]);
    shm = std::max(shm, (size_t)h.per[q] * h.N[q] * sizeof(cplx<T>));
    bmax = std::max(bmax, h.blocks[q]);
  }
  if (threads > 1024) return fail(CANS_EUNSUPPORTED, "k_het_ypass: more than 1024 lanes per work-group");
  const dim3 g(bmax, nsys);
  if (dir > 0) k_het_ypass<1, T><<<g, threads, shm, c->stream>>>(h);
  else k_het_ypass<-1, T><<<g, threads, shm, c->stream>>>(h);
  CANS_LAUNCH_CHECK("k_het_ypass");
  return CANS_OK;
}

// do the systems' plans share direction d's transform (kind and length)?
bool plan_fft_x(const cans_plan* p) { return p && !p->dense[0]; }

static bool same_transform(const cans_plan* const* ps, int nsys, int d) {
  for (int q = 1; q < nsys; q++)
    if (ps[q]->kind[d] != ps[0]->kind[d] || ps[q]->len[d] != ps[0]->len[d]) return false;
  return true;
}

// Solve of nsys systems that share the x and y transforms (the plan p): each pass
// is one launch for all of them (blockIdx.y = system).  The z systems may differ
// (coefficients, c/f grid, normalisation); their z boundary pairs must agree on
// periodicity.  nsys = 1 is the plain solver.
// work buffer (rounded up to 256 bytes) and checkpoint bytes of nsys systems
static void solver_work_split(const Layout& L, int nsys, size_t es, bool periodic, size_t* wbytes,
                              size_t* ckbytes) {
  const long nl = (long)L.n1 * L.n2;
  const long nblk = (L.n3 + CK - 1) / CK;  // enough checkpoint blocks for any system of the batch
  *wbytes = ((size_t)nsys * wbuf_elems(L.n1, L.n2, L.n3) * es + 255) & ~(size_t)255;
  *ckbytes = (size_t)nsys * 3 * nblk * nl * sizeof(double) + (periodic ? (size_t)L.cells() * sizeof(double) : 0);
}

size_t solver_work_bytes(const cans_ctx* c, size_t es, bool periodic_z) {
  size_t w, k;
  solver_work_split(c->L, 1, es, periodic_z, &w, &k);
  return w + k;
}

template <typename T>
int run_solver_batch(cans_ctx* c, const cans_plan* p, int nsys, const SolveSys<T>* sy, const DivSrc<T>* div,
                     const SolveWork* ws = nullptr, const cans_plan* const* ps = nullptr) {
  const Layout& L = c->L;
  const long nl = (long)L.n1 * L.n2;
  if (nsys < 1 || nsys > NBATCH) return fail(CANS_EINVAL, "run_solver: batch size out of range");
  const int periodic = sy[0].bc[4] == 'P' && sy[0].bc[5] == 'P';
  if (nsys > 1 && (periodic || div || c->dtdma > 1))
    return fail(CANS_EUNSUPPORTED, "run_solver: batches need a non-periodic z and no fused divergence or slabs");
  if (ws && c->dtdma > 1) return fail(CANS_EUNSUPPORTED, "run_solver: a caller workspace does not cover DTDMA slabs");
  if ((periodic ? L.n3 - 1 : L.n3) < 1) return fail(CANS_EUNSUPPORTED, "cans_solver: fewer than 2 z rows");
  const long wel = wbuf_elems(L.n1, L.n2, L.n3);
  const long nblk = (L.n3 + CK - 1) / CK;
  const size_t ckper = (size_t)3 * nblk * nl;
  size_t wbytes, ckbytes;
  solver_work_split(L, nsys, sizeof(T), periodic, &wbytes, &ckbytes);
  void *W = nullptr, *ck = nullptr;
  int rc;
  if (ws) {
    if (ws->bytes < wbytes + ckbytes) return fail(CANS_EINVAL, "run_solver: workspace too small");
    W = ws->base;
    ck = (char*)ws->base + wbytes;
  } else {
    if ((rc = ensure_work(c, 1, wbytes, &W))) return rc;
    if ((rc = ensure_work(c, 2, ckbytes, &ck))) return rc;
  }
  T* fld[NBATCH];
  T* Wq[NBATCH];
  auto plan_of = [&](int q) { return ps ? ps[q] : p; };
  bool dense[2] = {false, false};  // a system with a dense direction: that direction system by system
  for (int q = 0; q < nsys; q++)
    for (int d = 0; d < 2; d++) dense[d] = dense[d] || plan_of(q)->dense[d];
  for (int q = 0; q < nsys; q++) {
    fld[q] = sy[q].fld;
    Wq[q] = (T*)W + q * wel;
    // the FFT passes move element pairs: interior rows must start on pair boundaries
    if (!plan_of(q)->dense[0] && (L.n1 % 2 || L.sy % 2 || ((uintptr_t)(fld[q] + 1) & (2 * sizeof(T) - 1))))
      return fail(CANS_EINVAL, "cans_solver: rows of p must start aligned to an element pair (n1 and ld even, "
                               "field from cans_field_alloc or with element (1,0,0) on a pair boundary)");
  }
  if (div && (p->kind[0] != K_PP || p->len[0] < 2 || p->dense[0]))
    return fail(CANS_EINVAL, "run_solver: the velocity-divergence input needs a periodic x transform of an FFT length");
  if (ws && dense[1]) return fail(CANS_EUNSUPPORTED, "run_solver: a caller workspace does not cover dense y transforms");
  // systems with their own plans (ps): a direction whose transforms differ runs the heterogeneous pass
  const bool hx = ps && !same_transform(ps, nsys, 0), hy = ps && !same_transform(ps, nsys, 1);
  if (hx && div) return fail(CANS_EINVAL, "run_solver: the divergence input needs one x transform");
  // one direction's pass for every system: batched, or system by system where a plan is dense
  auto pass = [&](int d, int dir) -> int {
    if (dense[d]) {
      for (int q = 0; q < nsys; q++) {
        const cans_plan* pq = plan_of(q);
        int r = pq->dense[d] ? dense_pass<T>(c, pq, fld[q], Wq[q], d, dir)
                : d == 0     ? xpass<T>(c, pq, &fld[q], &Wq[q], 1, dir)
                             : ypass<T>(c, pq, &Wq[q], 1, dir);
        if (r) return r;
      }
      return CANS_OK;
    }
    if (d == 0) return hx ? het_pass<T>(c, ps, fld, Wq, nsys, 0, dir) : xpass<T>(c, p, fld, Wq, nsys, dir, div);
    return hy ? het_pass<T>(c, ps, nullptr, Wq, nsys, 1, dir) : ypass<T>(c, p, Wq, nsys, dir);
  };
  if ((rc = pass(0, +1))) return rc;
  if ((rc = pass(1, +1))) return rc;
  ZArgs<T> z;
  z.nl = nl;
  z.n1 = L.n1;
  z.n2 = L.n2;
  z.nic = wnic(L.n1);
  z.PS = wplane(L.n1, L.n2);
  z.periodic = periodic;
  // The reference pins the last unknown of the zero-wavenumber line when its final
  // pivot is within one ulp of zero (solver.f90:151-160).  FP64 storage keeps that
  // test exactly.  With FP32 storage the rhs of that singular line is compatible only
  // to FP32 rounding, so an unpinned solve would add a constant ~1e8 times the
  // solution; the tolerance is widened to 4096 ulp (rounding accumulated over the
  // elimination), still ~1e8 below the smallest regular pivot of any other line.
  z.pin_eps = 2.220446049250313e-16 * (sizeof(T) == sizeof(double) ? 1. : 4096.);
  for (int q = 0; q < nsys; q++) {
    ZSys<T>& zs = z.sys[q];
    const int qq = (sy[q].c_or_f[2] == 'f' && sy[q].bc[5] == 'D') ? 1 : 0;
    zs.W = Wq[q];
    zs.n = L.n3 - qq;
    zs.norm = sy[q].normfft;
    zs.a = sy[q].a;
    zs.b = sy[q].b;
    zs.c = sy[q].c;
    zs.lam = sy[q].lam;
    zs.ckd = (double*)ck + q * ckper;
    zs.ckp = zs.ckd + nblk * nl;
    zs.ckq = zs.ckp + nblk * nl;
  }
  z.X2 = periodic ? z.sys[0].ckq + nblk * nl : nullptr;
  if (c->dtdma > 1) {
    // all n rows take part (periodic too: the wrap couplings enter the cyclic reduced system)
    const int nn = z.sys[0].n;
    if (nn / c->dtdma < 3) return fail(CANS_EINVAL, "cans_solver: distributed z solve needs 3 rows per slab");
    DtArgs<T> d;
    d.W = Wq[0];
    d.nl = nl;
    d.n1 = L.n1;
    d.n2 = L.n2;
    d.nic = z.nic;
    d.PS = z.PS;
    d.n = nn;
    d.nslab = c->dtdma;
    d.periodic = periodic;
    const int nsmax = nn / d.nslab + (nn % d.nslab ? 1 : 0);
    d.nbk = (nsmax - 2 + DCK - 1) / DCK;
    d.norm = sy[0].normfft;
    d.pin_eps = 4096. * 2.220446049250313e-16;
    d.a = sy[0].a;
    d.b = sy[0].b;
    d.c = sy[0].c;
    d.lam = sy[0].lam;
    void* dt = nullptr;
    const size_t ckn = (size_t)d.nbk * 3 * d.nslab * nl, rn = (size_t)2 * d.nslab * nl;
    rc = ensure_work(c, 6, (2 * ckn + (periodic ? 5 : 3) * rn) * sizeof(double), &dt);
    if (rc) return rc;
    d.ckl = (double*)dt;
    d.cku = d.ckl + ckn;
    d.ra = d.cku + ckn;
    d.rc = d.ra + rn;
    d.rp = d.rc + rn;
    d.rc0 = periodic ? d.rp + rn : nullptr;
    d.r2 = periodic ? d.rc0 + rn : nullptr;
    const dim3 gs(cdiv(nl, 256), d.nslab);
    k_dtdma_local<T><<<gs, 256, 0, c->stream>>>(d);
    CANS_LAUNCH_CHECK("k_dtdma_local");
    k_dtdma_reduced<T><<<cdiv(nl, 256), 256, 0, c->stream>>>(d);
    CANS_LAUNCH_CHECK("k_dtdma_reduced");
    k_dtdma_back<T><<<gs, 256, 0, c->stream>>>(d);
    CANS_LAUNCH_CHECK("k_dtdma_back");
  } else if (periodic) {
    k_gaussel<T, true><<<cdiv(nl, 256), 256, 0, c->stream>>>(z);
    CANS_LAUNCH_CHECK("k_gaussel");
  } else {
    k_thomas<T, sizeof(T) == sizeof(float)><<<dim3(cdiv(nl, 256), nsys), 256, 0, c->stream>>>(z);
    CANS_LAUNCH_CHECK("k_thomas");
  }
  if ((rc = pass(1, -1))) return rc;
  if ((rc = pass(0, -1))) return rc;
  return CANS_OK;
}

template <typename T>
int run_solver(cans_ctx* c, const cans_plan* p, double normfft, const double* lam, const double* a,
               const double* b, const double* cc, const char* bc, const char* c_or_f, T* fld,
               const DivSrc<T>* div, const SolveWork* ws) {
  SolveSys<T> s1 = {fld, normfft, lam, a, b, cc, bc, c_or_f};
  return run_solver_batch<T>(c, p, 1, &s1, div, ws);
}

// ---------------------------------------------------------------------------
// Helmholtz solves of the implicit diffusion (solve_helmholtz.f90:28-75) on an
// FP64 field: bb = b + 1/alpha, then the 3-D solver with normfft/alpha, or the
// z-only elimination of solver_gaussel_z (solver.f90:547-616).  The z-only system
// has no eigenvalue, so gaussel's lambda-absent branch computes z(k) = 1/(bb(k) -
// a(k)*d(k-1)) and d(k) = c(k)*z(k) once for all lines (solver.f90:170-185): one
// thread builds those tables, then one thread per (i,j) line sweeps the field in
// place (rhs -> modified rhs on the way up, solution on the way down).
// ---------------------------------------------------------------------------
__global__ void k_shift_diag(const double* b, double alphai, double* bb, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) bb[k] = b[k] + alphai;
}

// One work-group per row of 256 lines: the coefficients of bb = b + 1/alpha, a and c
// are staged in LDS, one thread runs the (line-independent) recurrence for z(k) and
// d(k) from them, then every thread sweeps its line in batches of ZB rows, the
// batch's loads issued before its dependent arithmetic.
constexpr int ZB = 8;

struct Z1Args {
  Layout L;
  double* p[NBATCH];
  const double *a[NBATCH], *b[NBATCH], *c[NBATCH];
  double alphai[NBATCH];
  int nn[NBATCH];
};

__global__ void __launch_bounds__(256) k_z1_solve(Z1Args z) {
  const int sys = blockIdx.z;
  const int nn = z.nn[sys];
  const double alphai = z.alphai[sys], norm = alphai;
  const double *a = z.a[sys], *b = z.b[sys], *c = z.c[sys];
  extern __shared__ double zsh[];
  double* la = zsh;           // a(k)
  double* lb = la + nn;       // bb(k), then z(k)
  double* lc = lb + nn;       // c(k), then d(k)
  for (int k = threadIdx.x; k < nn; k += blockDim.x) {
    la[k] = a[k];
    lb[k] = b[k] + alphai;
    lc[k] = c[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {  // solver.f90:170-185 (lambda absent): z and d once for all lines
    double zz = 1. / lb[0];
    lb[0] = zz;
    double d = lc[0] * zz;
    lc[0] = d;
    for (int k = 1; k < nn; k++) {
      zz = 1. / (lb[k] - la[k] * d);
      lb[k] = zz;
      d = lc[k] * zz;
      lc[k] = d;
    }
  }
  __syncthreads();
  const Layout& L = z.L;
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1;
  if (i > L.n1) return;
  double* P = z.p[sys] + L.at(i, j, 1);  // row k (1-based) at P + (k-1)*sz
  const long sz = L.sz;
  double prev = 0.;
  for (int k0 = 0; k0 < nn; k0 += ZB) {
    double r[ZB];
#pragma unroll
    for (int u = 0; u < ZB; u++) r[u] = (k0 + u < nn) ? P[(long)(k0 + u) * sz] : 0.;
#pragma unroll
    for (int u = 0; u < ZB; u++) {
      const int k = k0 + u;
      if (k < nn) {
        prev = (k == 0) ? r[u] * norm * lb[0] : (r[u] * norm - la[k] * prev) * lb[k];
        P[(long)k * sz] = prev;
      }
    }
  }
  // backward substitution from row nn-1 (1-based) down: p(k) = p(k) - d(k)*p(k+1)
  for (int k1 = nn - 2; k1 >= 0; k1 -= ZB) {
    double r[ZB];
#pragma unroll
    for (int u = 0; u < ZB; u++) r[u] = (k1 - u >= 0) ? P[(long)(k1 - u) * sz] : 0.;
#pragma unroll
    for (int u = 0; u < ZB; u++) {
      const int k = k1 - u;
      if (k >= 0) {
        prev = r[u] - lc[k] * prev;
        P[(long)k * sz] = prev;
      }
    }
  }
}

// batched Helmholtz solves of FP64 fields (the rhs in, the solution out): the z-only
// form (is_impdiff_1d) as one k_z1_solve launch for all systems; the 3-D form as one
// batched solver when the systems share their x/y transform kinds, else one by one
int run_helmholtz_batch(cans_ctx* c, int nsys, const HelmSys* hs) {
  const Layout& L = c->L;
  if (nsys < 1) return CANS_OK;
  if (nsys > NBATCH) {
    int rc = run_helmholtz_batch(c, NBATCH, hs);
    return rc ? rc : run_helmholtz_batch(c, nsys - NBATCH, hs + NBATCH);
  }
  if (c->imp == 2) {
    Z1Args z;
    z.L = L;
    int nnmax = 0;
    for (int q = 0; q < nsys; q++) {
      const char* bcz = hs[q].cbc + 4;
      if (bcz[0] == 'P' && bcz[1] == 'P')
        return fail(CANS_EUNSUPPORTED, "implicit z diffusion with periodic z is not built on the device");
      const int qq = (hs[q].c_or_f[2] == 'f' && bcz[1] == 'D') ? 1 : 0;
      z.p[q] = hs[q].fld;
      z.a[q] = hs[q].a;
      z.b[q] = hs[q].b;
      z.c[q] = hs[q].c;
      z.alphai[q] = 1. / hs[q].alpha;
      z.nn[q] = L.n3 - qq;
      nnmax = std::max(nnmax, z.nn[q]);
    }
    const size_t shm = (size_t)3 * nnmax * sizeof(double);
    if (shm > 160 * 1024) return fail(CANS_EUNSUPPORTED, "implicit z diffusion: too many z rows for LDS tables");
    k_z1_solve<<<dim3(cdiv(L.n1, 256), L.n2, nsys), 256, shm, c->stream>>>(z);
    CANS_LAUNCH_CHECK("k_z1_solve");
    return CANS_OK;
  }
  // one batch (one launch per pass) unless z is periodic (the periodic sweep solves one system)
  // or the z solve is distributed; systems whose x or y transforms differ take the
  // heterogeneous x/y passes (HetArgs)
  bool batch = hs[0].plan && hs[0].cbc[4] != 'P' && c->dtdma <= 1;
  for (int q = 1; q < nsys && batch; q++) batch = hs[q].plan && hs[q].cbc[4] != 'P';
  if (!batch) {
    for (int q = 0; q < nsys; q++) {
      int rc = run_helmholtz_3d<double>(c, hs[q].plan, hs[q].normfft, hs[q].alpha, hs[q].lam, hs[q].a, hs[q].b,
                                        hs[q].c, hs[q].cbc, hs[q].c_or_f, hs[q].fld);
      if (rc) return rc;
    }
    return CANS_OK;
  }
  void* tab = nullptr;
  int rc = ensure_work(c, 5, (size_t)NBATCH * L.n3 * sizeof(double), &tab);
  if (rc) return rc;
  SolveSys<double> sy[NBATCH];
  for (int q = 0; q < nsys; q++) {
    double* bb = (double*)tab + (long)q * L.n3;
    const double alphai = 1. / hs[q].alpha;
    k_shift_diag<<<cdiv(L.n3, 256), 256, 0, c->stream>>>(hs[q].b, alphai, bb, L.n3);
    CANS_LAUNCH_CHECK("k_shift_diag");
    sy[q] = SolveSys<double>{hs[q].fld, hs[q].normfft * alphai, hs[q].lam, hs[q].a, bb, hs[q].c, hs[q].cbc,
                             hs[q].c_or_f};
  }
  const cans_plan* ps[NBATCH];
  for (int q = 0; q < nsys; q++) ps[q] = hs[q].plan;
  return run_solver_batch<double>(c, hs[0].plan, nsys, sy, nullptr, nullptr, ps);
}

int run_helmholtz_z(cans_ctx* c, const double* a, const double* b, const double* cc, double alpha, const char* bcz,
                    char c_or_f_z, double* fld) {
  char cbc[6] = {'P', 'P', 'P', 'P', bcz[0], bcz[1]};
  char cf[3] = {'c', 'c', c_or_f_z};
  HelmSys h = {nullptr, 1., nullptr, a, b, cc, cbc, cf, fld, alpha};
  if (c->imp != 2) return fail(CANS_EINVAL, "run_helmholtz_z: z-only form needs is_impdiff_1d");
  return run_helmholtz_batch(c, 1, &h);
}

template <typename T>
int run_helmholtz_3d(cans_ctx* c, const cans_plan* p, double normfft, double alpha, const double* lam,
                     const double* a, const double* b, const double* cc, const char* cbc, const char* c_or_f,
                     T* fld) {
  const Layout& L = c->L;
  void* tab = nullptr;
  int rc = ensure_work(c, 5, (size_t)3 * L.n3 * sizeof(double), &tab);
  if (rc) return rc;
  double* bb = (double*)tab;
  const double alphai = 1. / alpha;
  k_shift_diag<<<cdiv(L.n3, 256), 256, 0, c->stream>>>(b, alphai, bb, L.n3);
  CANS_LAUNCH_CHECK("k_shift_diag");
  return run_solver<T>(c, p, normfft * alphai, lam, a, bb, cc, cbc, c_or_f, fld);
}
template int run_helmholtz_3d<double>(cans_ctx*, const cans_plan*, double, double, const double*, const double*,
                                      const double*, const double*, const char*, const char*, double*);
template int run_helmholtz_3d<float>(cans_ctx*, const cans_plan*, double, double, const double*, const double*,
                                     const double*, const double*, const char*, const char*, float*);

template int run_solver<double>(cans_ctx*, const cans_plan*, double, const double*, const double*, const double*,
                                const double*, const char*, const char*, double*, const DivSrc<double>*,
                                const SolveWork*);
template int run_solver<float>(cans_ctx*, const cans_plan*, double, const double*, const double*, const double*,
                               const double*, const char*, const char*, float*, const DivSrc<float>*,
                               const SolveWork*);

}  // namespace cans

using namespace cans;

namespace cans {

// ---------------------------------------------------------------------------
// Independent Poisson boxes of different sizes: one solver.f90:17-112 call per box
// (each with its own arrplan, normfft, lambdaxy, a, b, c), all boxes solved together.
// Each pass is one launch over every box: the work-groups of box b are
// off[b] .. off[b+1]-1 of the grid, found by a binary search over off, and run the
// single solve's pass body with that box's shapes.  Work-groups have one size for all
// boxes (128, 256 or 512 lanes, the fastest forward x and y passes at creation)
// lanes; a box uses as many whole transform teams as fit and the lanes left over
// only take part in the barriers.  Boxes are cell-centred with a non-periodic z.
// ---------------------------------------------------------------------------
constexpr int BOX_THREADS = 512;  // the largest work-group size

template <typename T>
struct BoxDesc {
  Layout L;
  T* fld;                        // halo'd field (origin)
  T* W;                          // z-tiled work buffer (woff)
  int N[2], logN[2], odd[2], Tn[2], kind[2];
  int nd[2];                     // data lengths (N[q], or N[q]/2 for K_ND / K_DN)
  int pairs, lines;              // x-pass teams, y-pass lines per work-group
  const cplx<double>* tw[2];
  const cplx<double>* mk[2];
  double norm;                   // normfft
  const double *lam, *a, *b, *c;
  double *ckd, *ckp;             // Thomas checkpoints [nblk][n1*n2]
};

template <typename T>
struct BoxArgs {
  const BoxDesc<T>* __restrict__ box;
  const int* __restrict__ off;   // [nbox + 1]: first work-group of each box in this pass
  int nbox;
  double pin_eps;
};

__device__ __forceinline__ int find_box(const int* __restrict__ off, int nbox, int g) {
  int lo = 0, hi = nbox;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

template <int DIR, typename T>
__global__ void __launch_bounds__(BOX_THREADS) k_box_xpass(BoxArgs<T> ba) {
  const int b = find_box(ba.off, ba.nbox, blockIdx.x);
  const BoxDesc<T>& d = ba.box[b];
  XArgs<T> a;
  a.L = d.L;
  a.fld[0] = d.fld;
  a.W[0] = d.W;
  a.N = d.N[0];
  a.nd = d.nd[0];
  a.logN = d.logN[0];
  a.odd = d.odd[0];
  a.Tn = d.Tn[0];
  a.pairs = d.pairs;
  a.kind = d.kind[0];
  a.tw = d.tw[0];
  a.mk = d.mk[0];
  a.div = 0;
  xpass_body<DIR, T, 0>(a, blockIdx.x - ba.off[b], 0);
}

template <int DIR, typename T>
__global__ void __launch_bounds__(BOX_THREADS) k_box_ypass(BoxArgs<T> ba) {
  const int b = find_box(ba.off, ba.nbox, blockIdx.x);
  const BoxDesc<T>& d = ba.box[b];
  YArgs<T> a;
  a.n1 = d.L.n1;
  a.n2 = d.L.n2;
  a.n3 = d.L.n3;
  a.W[0] = d.W;
  a.N = d.N[1];
  a.nd = d.nd[1];
  a.logN = d.logN[1];
  a.odd = d.odd[1];
  a.Tn = d.Tn[1];
  a.lines = d.lines;
  a.kind = d.kind[1];
  a.tw = d.tw[1];
  a.mk = d.mk[1];
  ypass_body<DIR, T>(a, blockIdx.x - ba.off[b], 0);
}

template <typename T, bool FAST>
__global__ void __launch_bounds__(256) k_box_thomas(BoxArgs<T> ba) {
  const int b = find_box(ba.off, ba.nbox, blockIdx.x);
  const BoxDesc<T>& d = ba.box[b];
  const long nl = (long)d.L.n1 * d.L.n2;
  const long line = (long)(blockIdx.x - ba.off[b]) * blockDim.x + threadIdx.x;
  if (line >= nl) return;
  ZArgs<T> z;
  z.nl = nl;
  z.n1 = d.L.n1;
  z.n2 = d.L.n2;
  z.nic = wnic(d.L.n1);
  z.PS = wplane(d.L.n1, d.L.n2);
  z.periodic = 0;
  z.pin_eps = ba.pin_eps;
  z.X2 = nullptr;
  ZSys<T> zs;
  zs.W = d.W;
  zs.n = d.L.n3;
  zs.norm = d.norm;
  zs.a = d.a;
  zs.b = d.b;
  zs.c = d.c;
  zs.lam = d.lam;
  zs.ckd = d.ckd;
  zs.ckp = d.ckp;
  zs.ckq = nullptr;
  thomas_body<T, FAST>(z, zs, line);
}

}  // namespace cans

struct cans_poisson_batch {
  cans_ctx* ctx;
  int nbox = 0;
  int es = 8;                    // element size of the boxes' fields
  void* mem = nullptr;           // one device allocation holding everything below
  const void* desc = nullptr;    // BoxDesc<T>[nbox]
  const int* off[3] = {};        // first work-group of each box: x passes, y passes, z pass
  int blocks[3] = {};            // work-groups per launch of each pass
  size_t shm[2] = {};            // LDS of the x and y launches
  int threads = 256;             // lanes per x/y work-group (the tuner's choice)
};

namespace cans {

// host staging of one device allocation: chunks at 256-byte aligned offsets
struct Arena {
  std::vector<unsigned char> host;
  size_t take(size_t bytes) {
    const size_t o = (host.size() + 255) & ~(size_t)255;
    host.resize(o + bytes);
    return o;
  }
};

template <typename T>
static BoxArgs<T> box_args(const cans_poisson_batch* B, int pass) {
  BoxArgs<T> a;
  a.box = (const BoxDesc<T>*)B->desc;
  a.off = B->off[pass];
  a.nbox = B->nbox;
  // the zero-wavenumber pin of run_solver_batch
  a.pin_eps = 2.220446049250313e-16 * (sizeof(T) == sizeof(double) ? 1. : 4096.);
  return a;
}

// the forward x and y passes of a batch solve
template <typename T>
static int batch_passes_fwd(cans_poisson_batch* B) {
  cans_ctx* c = B->ctx;
  k_box_xpass<1, T><<<B->blocks[0], B->threads, B->shm[0], c->stream>>>(box_args<T>(B, 0));
  CANS_LAUNCH_CHECK("k_box_xpass");
  k_box_ypass<1, T><<<B->blocks[1], B->threads, B->shm[1], c->stream>>>(box_args<T>(B, 1));
  CANS_LAUNCH_CHECK("k_box_ypass");
  return CANS_OK;
}

template <typename T>
static int batch_create(cans_ctx* c, int nbox, const cans_poisson_box* bx, double* normfft, cans_poisson_batch* B) {
  Arena ar;
  struct Fix {  // device addresses filled in once the allocation exists
    size_t desc, lam, a, b, c, tw[2], mk[2], W, ck;
    long nl, nblk;
  };
  std::vector<Fix> fx(nbox);
  std::vector<BoxDesc<T>> d(nbox);
  std::vector<int> off[3];
  for (int q = 0; q < 3; q++) off[q].assign(nbox + 1, 0);
  const size_t desc_at = ar.take(sizeof(BoxDesc<T>) * nbox);
  size_t off_at[3];
  for (int q = 0; q < 3; q++) off_at[q] = ar.take(sizeof(int) * (nbox + 1));
  long wtot = 0, cktot = 0;
  for (int i = 0; i < nbox; i++) {
    const cans_poisson_box& x = bx[i];
    const std::string who = "cans_poisson_batch_create: box " + std::to_string(i) + ": ";
    const int n1 = x.n[0], n2 = x.n[1], n3 = x.n[2];
    if (n1 < 2 || n2 < 1 || n3 < 1) return fail(CANS_EINVAL, who + "bad sizes");
    const long sy = x.sy ? x.sy : n1 + 2, sz = x.sz ? x.sz : sy * (n2 + 2);
    if (!x.p_d || !x.lambdaxy_h || !x.a_h || !x.b_h || !x.c_h) return fail(CANS_EINVAL, who + "null array");
    if (sy < n1 + 2 || sz < sy * (n2 + 2)) return fail(CANS_EINVAL, who + "strides smaller than the halo'd rows/planes");
    // the x pass moves element pairs: interior rows must start on pair boundaries
    if (n1 % 2 || sy % 2 || ((uintptr_t)((T*)x.p_d + 1) & (2 * sizeof(T) - 1)))
      return fail(CANS_EINVAL, who + "rows of p must start aligned to an element pair (n1 and sy even, element "
                                     "(1,0,0) on a pair boundary)");
    if (x.cbc[4] == 'P' || x.cbc[5] == 'P' || !std::strchr("ND", x.cbc[4]) || !std::strchr("ND", x.cbc[5]) ||
        !x.cbc[4] || !x.cbc[5])
      return fail(CANS_EUNSUPPORTED, who + "z must be non-periodic (N or D at each end)");
    BoxDesc<T>& b = d[i];
    b.L.n1 = n1;
    b.L.n2 = n2;
    b.L.n3 = n3;
    b.L.sy = sy;
    b.L.sz = sz;
    b.fld = (T*)x.p_d;
    double nf = 1.;
    for (int q = 0; q < 2; q++) {
      const char b0 = x.cbc[2 * q], b1 = x.cbc[2 * q + 1];
      if (b0 == 'P' && b1 == 'P') b.kind[q] = K_PP;
      else if (b0 == 'N' && b1 == 'N') b.kind[q] = K_NN;
      else if (b0 == 'D' && b1 == 'D') b.kind[q] = K_DD;
      else if (b0 == 'N' && b1 == 'D') b.kind[q] = K_ND;
      else if (b0 == 'D' && b1 == 'N') b.kind[q] = K_DN;
      else return fail(CANS_EINVAL, who + "x/y boundary pairs must be PP, NN, DD, ND or DN");
      b.nd[q] = x.n[q];
      b.N[q] = k_double(b.kind[q]) ? 2 * x.n[q] : x.n[q];
      if (!fft_factor(b.N[q], &b.logN[q], &b.odd[q]) || b.N[q] > 2048)
        return fail(CANS_EUNSUPPORTED, who + "x/y lengths must be 2^m, or 2^m (m >= 3) times at most three "
                                             "factors 3 and 5, at most 2048 (ND and DN pairs: the doubled length)");
      b.Tn[q] = b.N[q] >= 8 ? b.N[q] / 8 : 1;
      nf = nf * (b.kind[q] == K_PP ? 1. : 2.) * (x.n[q] + 0.);  // fft.f90:75,140
      std::vector<cplx<double>> tw, mk;
      make_tables(b.N[q], tw, mk);
      fx[i].tw[q] = ar.take(tw_bytes(tw.size()));
      tw_pack(tw, &ar.host[fx[i].tw[q]]);
      fx[i].mk[q] = ar.take(mk.size() * sizeof(cplx<double>));
      std::memcpy(&ar.host[fx[i].mk[q]], mk.data(), mk.size() * sizeof(cplx<double>));
    }
    b.norm = 1. / nf;
    if (normfft) normfft[i] = b.norm;
    const long nl = (long)n1 * n2;
    fx[i].nl = nl;
    fx[i].lam = ar.take(nl * sizeof(double));
    std::memcpy(&ar.host[fx[i].lam], x.lambdaxy_h, nl * sizeof(double));
    const double* abc[3] = {x.a_h, x.b_h, x.c_h};
    size_t* at[3] = {&fx[i].a, &fx[i].b, &fx[i].c};
    for (int q = 0; q < 3; q++) {
      *at[q] = ar.take(n3 * sizeof(double));
      std::memcpy(&ar.host[*at[q]], abc[q], n3 * sizeof(double));
    }
    fx[i].W = wtot;
    wtot += (wbuf_elems(n1, n2, n3) + 63) & ~63L;
    fx[i].nblk = (n3 + CK - 1) / CK;
    fx[i].ck = cktot;
    cktot += 2 * fx[i].nblk * nl;
    // work-groups of the z pass (those of the x and y passes depend on the work-group size)
    const long tot = (long)off[2][i] + (nl + 255) / 256;
    if (tot > INT32_MAX) return fail(CANS_EINVAL, "cans_poisson_batch_create: too many work-groups");
    off[2][i + 1] = (int)tot;
  }
  // x and y passes with work-groups of bt lanes: teams, lines, work-groups and LDS per launch;
  // false when a box's team does not fit or a launch would need more than 64 KB of LDS
  auto shape = [&](int bt) {
    B->shm[0] = B->shm[1] = 0;
    for (int i = 0; i < nbox; i++) {
      BoxDesc<T>& b = d[i];
      if (b.Tn[0] > bt || b.Tn[1] > bt) return false;
      b.pairs = bt / b.Tn[0];
      b.lines = bt / b.Tn[1];
      const long npairs = ((long)b.L.n2 * b.L.n3 + 1) / 2;
      const long g[2] = {(npairs + b.pairs - 1) / b.pairs,
                         ((long)b.L.n3 * ((b.L.n1 + 1) / 2) + b.lines - 1) / b.lines};
      for (int q = 0; q < 2; q++) {
        const long tot = (long)off[q][i] + g[q];
        if (tot > INT32_MAX) return false;
        off[q][i + 1] = (int)tot;
      }
      B->shm[0] = std::max(B->shm[0], (size_t)b.pairs * padded_len(b.N[0]) * sizeof(cplx<T>));
      B->shm[1] = std::max(B->shm[1], (size_t)b.lines * b.N[1] * sizeof(cplx<T>));
    }
    return B->shm[0] <= 65536 && B->shm[1] <= 65536;
  };
  if (!shape(256)) return fail(CANS_EINVAL, "cans_poisson_batch_create: transform lengths too large");
  // the work buffer and checkpoints follow the tables in the device allocation only
  const size_t W_at = (ar.host.size() + 255) & ~(size_t)255;
  const size_t ck_at = (W_at + (size_t)wtot * sizeof(T) + 255) & ~(size_t)255;
  const size_t total = ck_at + (size_t)cktot * sizeof(double);
  void* mem = nullptr;
  hipError_t e = hipMalloc(&mem, total);
  if (e != hipSuccess) return fail(CANS_ENOMEM, std::string("cans_poisson_batch_create: ") + hipGetErrorString(e));
  unsigned char* base = (unsigned char*)mem;
  for (int i = 0; i < nbox; i++) {
    BoxDesc<T>& b = d[i];
    for (int q = 0; q < 2; q++) {
      b.tw[q] = (const cplx<double>*)(base + fx[i].tw[q]);
      b.mk[q] = (const cplx<double>*)(base + fx[i].mk[q]);
    }
    b.lam = (const double*)(base + fx[i].lam);
    b.a = (const double*)(base + fx[i].a);
    b.b = (const double*)(base + fx[i].b);
    b.c = (const double*)(base + fx[i].c);
    b.W = (T*)(base + W_at) + fx[i].W;
    b.ckd = (double*)(base + ck_at) + fx[i].ck;
    b.ckp = b.ckd + fx[i].nblk * fx[i].nl;
  }
  // the work buffer and checkpoints need no initial values
  auto upload = [&]() -> int {
    std::memcpy(&ar.host[desc_at], d.data(), sizeof(BoxDesc<T>) * nbox);
    for (int q = 0; q < 3; q++) std::memcpy(&ar.host[off_at[q]], off[q].data(), sizeof(int) * (nbox + 1));
    CANS_HIP_TRY(hipMemcpy(mem, ar.host.data(), ar.host.size(), hipMemcpyHostToDevice));
    for (int q = 0; q < 3; q++) B->blocks[q] = off[q][nbox];
    return CANS_OK;
  };
  B->mem = mem;
  B->desc = base + desc_at;
  for (int q = 0; q < 3; q++) B->off[q] = (const int*)(base + off_at[q]);
  // launch-shape tuner: time the forward x and y passes (they read the fields and write only the
  // batch's own work buffers) for each work-group size and keep the fastest
  auto tune = [&]() -> int {
    hipEvent_t ev[2];
    CANS_HIP_TRY(hipEventCreate(&ev[0]));
    if (hipEventCreate(&ev[1]) != hipSuccess) {
      hipEventDestroy(ev[0]);
      return fail(CANS_EHIP, "cans_poisson_batch_create: hipEventCreate");
    }
    float best_ms = 0.f;
    int best = 0, r = CANS_OK;
    for (int bt : {128, 256, 512}) {
      if (!shape(bt)) continue;
      if ((r = upload())) break;
      B->threads = bt;
      float ms = 0.f;
      for (int rep = 0; rep < 2 && !r; rep++) {  // the first run warms up
        if (hipEventRecord(ev[0], c->stream) != hipSuccess) r = fail(CANS_EHIP, "cans_poisson_batch_create: timing");
        if (!r) r = batch_passes_fwd<T>(B);
        if (!r && (hipEventRecord(ev[1], c->stream) != hipSuccess || hipEventSynchronize(ev[1]) != hipSuccess ||
                   hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess))
          r = fail(CANS_EHIP, "cans_poisson_batch_create: timing");
      }
      if (r) break;
      if (!best || ms < best_ms) {
        best = bt;
        best_ms = ms;
      }
    }
    hipEventDestroy(ev[0]);
    hipEventDestroy(ev[1]);
    if (r) return r;
    if (!best || !shape(best)) return fail(CANS_EINVAL, "cans_poisson_batch_create: no launch shape fits");
    B->threads = best;
    return upload();
  };
  if (int rc = tune()) {
    hipFree(mem);
    B->mem = nullptr;
    return rc;
  }
  return CANS_OK;
}

template <typename T>
static int batch_solve(cans_poisson_batch* B) {
  cans_ctx* c = B->ctx;
  int rc = batch_passes_fwd<T>(B);
  if (rc) return rc;
  k_box_thomas<T, sizeof(T) == sizeof(float)><<<B->blocks[2], 256, 0, c->stream>>>(box_args<T>(B, 2));
  CANS_LAUNCH_CHECK("k_box_thomas");
  k_box_ypass<-1, T><<<B->blocks[1], B->threads, B->shm[1], c->stream>>>(box_args<T>(B, 1));
  CANS_LAUNCH_CHECK("k_box_ypass");
  k_box_xpass<-1, T><<<B->blocks[0], B->threads, B->shm[0], c->stream>>>(box_args<T>(B, 0));
  CANS_LAUNCH_CHECK("k_box_xpass");
  return CANS_OK;
}

}  // namespace cans

extern "C" int cans_poisson_batch_create(cans_ctx* c, int nbox, const cans_poisson_box* boxes, double* normfft,
                                         cans_poisson_batch** out) {
  using namespace cans;
  if (!c || !boxes || !out || nbox < 1) return fail(CANS_EINVAL, "cans_poisson_batch_create: bad argument");
  cans_poisson_batch* B = new cans_poisson_batch();
  B->ctx = c;
  B->nbox = nbox;
  B->es = c->precision == CANS_MIXED ? 4 : 8;
  int rc = B->es == 4 ? batch_create<float>(c, nbox, boxes, normfft, B) : batch_create<double>(c, nbox, boxes, normfft, B);
  if (rc) {
    delete B;
    return rc;
  }
  *out = B;
  return CANS_OK;
}

extern "C" int cans_poisson_batch_solve(cans_poisson_batch* B) {
  using namespace cans;
  if (!B) return fail(CANS_EINVAL, "cans_poisson_batch_solve: null batch");
  return B->es == 4 ? batch_solve<float>(B) : batch_solve<double>(B);
}

extern "C" int cans_poisson_batch_tuning(const cans_poisson_batch* B, int* threads) {
  if (!B || !threads) return fail(CANS_EINVAL, "cans_poisson_batch_tuning: null argument");
  *threads = B->threads;
  return CANS_OK;
}

extern "C" int cans_poisson_batch_destroy(cans_poisson_batch* B) {
  if (!B) return CANS_OK;
  hipFree(B->mem);
  delete B;
  return CANS_OK;
}

// ---------------------------------------------------------------------------
// Launch-shape tuner.  The x pass's work-group size (lines per group) and the y
// pass's lines per group trade LDS, registers and access width; the best choice
// depends on the transform lengths, the storage type and the part.  cans_fftini
// times the candidates on a zero-filled problem of the context's x-y shape with at
// most 128 z-planes (the passes treat planes independently) and keeps the fastest
// forward+backward pair per storage type.
// ---------------------------------------------------------------------------
// the tuner's radix choices: (direction, line length, kind, storage slot, device) -> radix 4
static std::map<std::tuple<int, int, int, int, int>, int> g_radix;
static std::mutex g_radix_mu;

template <typename T>
static int tune_type(cans_ctx* c, cans_plan* p) {
  const int slot = sizeof(T) == sizeof(double);
  cans_ctx cs = *c;  // the context's layout with fewer planes; its scratch is not used
  cs.L.n3 = std::min(c->L.n3, 128);
  const Layout& L = cs.L;
  const size_t fbytes = (size_t)L.sz * (L.n3 + 2) * sizeof(T) + 256;
  const size_t wbytes = (size_t)wbuf_elems(L.n1, L.n2, L.n3) * sizeof(T);
  void *fb = nullptr, *wb = nullptr;
  CANS_HIP_TRY(hipMalloc(&fb, fbytes));
  if (hipMalloc(&wb, wbytes) != hipSuccess) {
    hipFree(fb);
    return fail(CANS_ENOMEM, "cans_fftini: tuner scratch");
  }
  hipEvent_t e0, e1;
  int rc = CANS_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = fail(CANS_EHIP, "cans_fftini: events");
  T* fld = (T*)((char*)fb + 256 - sizeof(T));  // element (1,0,0) on a 256-byte boundary
  T* W = (T*)wb;
  if (!rc && (hipMemsetAsync(fb, 0, fbytes, cs.stream) != hipSuccess ||
              hipMemsetAsync(wb, 0, wbytes, cs.stream) != hipSuccess))
    rc = fail(CANS_EHIP, "cans_fftini: tuner memset");
  auto timed = [&](auto&& run) -> float {
    float best = 1e30f;
    for (int rep = 0; rep < 3 && !rc; rep++) {
      hipEventRecord(e0, cs.stream);
      if ((rc = run())) break;
      hipEventRecord(e1, cs.stream);
      hipEventSynchronize(e1);
      float ms = 0.f;
      hipEventElapsedTime(&ms, e0, e1);
      if (rep > 0) best = std::min(best, ms);  // the first run warms up
    }
    return best;
  };
  // x pass: shapes whose LDS exceeds the device's per-work-group limit are skipped
  int lds_max = 65536;
  if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) != hipSuccess)
    lds_max = 65536;
  const int tnx = p->nfft[0] >= 8 ? p->nfft[0] / 8 : 1;
  float bx = 1e30f;
  int xbest = 256;
  for (int xt : {128, 256, 512}) {
    if (p->dense[0]) break;  // a dense direction has no launch shape to choose
    if (xt < tnx && xt != 512) continue;
    const size_t xshm = (size_t)(tnx >= xt ? 1 : xt / tnx) * padded_len(p->nfft[0]) * sizeof(cplx<T>);
    if (xshm > (size_t)lds_max) continue;
    p->xthreads[slot] = xt;
    const float ms = timed([&]() -> int {
      int r = xpass<T>(&cs, p, &fld, &W, 1, +1);
      return r ? r : xpass<T>(&cs, p, &fld, &W, 1, -1);
    });
    if (ms < bx) bx = ms, xbest = xt;
  }
  p->xthreads[slot] = xbest;
  // y pass
  const int tny = p->nfft[1] >= 8 ? p->nfft[1] / 8 : 1;
  float by = 1e30f;
  int ybest = 0;
  for (int yl : {4, 8, 16, 32}) {
    if (p->dense[1]) break;
    if ((long)yl * tny > 1024 || (size_t)yl * p->nfft[1] * sizeof(cplx<T>) > (size_t)lds_max) continue;
    p->ylines[slot] = yl;
    const float ms = timed([&]() -> int {
      int r = ypass<T>(&cs, p, &W, 1, +1);
      return r ? r : ypass<T>(&cs, p, &W, 1, -1);
    });
    if (ms < by) by = ms, ybest = yl;
  }
  p->ylines[slot] = ybest;
  // radix-4 instead of radix-8 power-of-two stages, at the chosen shapes (lines of 2^m, m >= 3).
  // The split into stages changes the rounding, so the choice is made once per direction,
  // length, kind and storage type in a process and reused by every later plan: solves through
  // two plans of the same lengths agree bit for bit.
  for (int d = 0; d < 2 && !rc; d++) {
    const float base = d == 0 ? bx : by;
    if (p->dense[d] || p->logn[d] < 3 || !(base < 1e29f)) continue;
    const auto key = std::make_tuple(d, p->nfft[d], p->kind[d], slot, c->device);
    {
      std::lock_guard<std::mutex> lk(g_radix_mu);
      const auto it = g_radix.find(key);
      if (it != g_radix.end()) {
        p->radix4[d][slot] = it->second;
        continue;
      }
    }
    p->radix4[d][slot] = 1;
    const float ms = timed([&]() -> int {
      int r = d == 0 ? xpass<T>(&cs, p, &fld, &W, 1, +1) : ypass<T>(&cs, p, &W, 1, +1);
      if (r) return r;
      return d == 0 ? xpass<T>(&cs, p, &fld, &W, 1, -1) : ypass<T>(&cs, p, &W, 1, -1);
    });
    if (!(ms < base)) p->radix4[d][slot] = 0;
    if (!rc) {
      std::lock_guard<std::mutex> lk(g_radix_mu);
      g_radix.emplace(key, p->radix4[d][slot]);
    }
  }
  hipStreamSynchronize(cs.stream);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(fb);
  hipFree(wb);
  if (rc) {  // a failed trial keeps the heuristic shapes; the plan itself stays valid
    p->xthreads[slot] = 0;
    p->ylines[slot] = 0;
    p->radix4[0][slot] = p->radix4[1][slot] = 0;
    (void)hipGetLastError();
  }
  return CANS_OK;
}

// the kind of a direction's transforms (fft.f90:260-313) and the normalisation factor
// norm(1)*(n+norm(2)-ix) of fft.f90:75,140
static int bc_kind(char b0, char b1, bool face, int n, int* kind, double* nfac, const char* who) {
  if (b0 == 'P' && b1 == 'P') *kind = K_PP, *nfac = n;
  else if (b0 == 'N' && b1 == 'N' && !face) *kind = K_NN, *nfac = 2. * n;
  else if (b0 == 'D' && b1 == 'D') *kind = face ? K_DDF : K_DD, *nfac = 2. * n;
  else if (b0 == 'N' && b1 == 'D') *kind = face ? K_NN : K_ND, *nfac = 2. * n;
  else if (b0 == 'D' && b1 == 'N') *kind = face ? K_DNF : K_DN, *nfac = 2. * n;
  else if (b0 == 'N' && b1 == 'N')  // REDFT00: the reference forbids face-centred NN velocity walls
    // with implicit diffusion (sanity.f90:151-162), the only user of face-centred transforms
    return fail(CANS_EUNSUPPORTED, std::string(who) + ": face-centred NN (REDFT00) is not built on the device");
  else
    return fail(CANS_EINVAL, std::string(who) + ": boundary pairs must be PP, NN, DD, ND or DN");
  return CANS_OK;
}

extern "C" int cans_transform_matrix(int n, const char* bc, char c_or_f, int dir, double* m) {
  if (n < 1 || n > DENSE_MAX || !bc || !m || (c_or_f != 'c' && c_or_f != 'f') || (dir != 1 && dir != -1))
    return fail(CANS_EINVAL, "cans_transform_matrix: bad argument");
  int kind;
  double nfac;
  const int rc = bc_kind(bc[0], bc[1], c_or_f == 'f', n, &kind, &nfac, "cans_transform_matrix");
  if (rc) return rc;
  if (kind == K_DDF && n < 2) return fail(CANS_EINVAL, "cans_transform_matrix: face-centred DD needs 2 points");
  std::vector<double> M;
  dense_matrix(kind, n, dir, M);
  std::memcpy(m, M.data(), M.size() * sizeof(double));
  return CANS_OK;
}

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
    const bool face = c_or_f[d] == 'f';
    if (c_or_f[d] != 'c' && !face) {
      delete p;
      return fail(CANS_EINVAL, "cans_fftini: c_or_f must be 'c' or 'f'");
    }
    double nfac;
    const int krc = bc_kind(b0, b1, face, ng[d], &p->kind[d], &nfac, "cans_fftini");
    if (krc) {
      delete p;
      return krc;
    }
    if (p->kind[d] == K_DDF && ng[d] < 2) {
      delete p;
      return fail(CANS_EINVAL, "cans_fftini: face-centred DD needs at least 2 points");
    }
    p->len[d] = ng[d];
    p->nfft[d] = k_double(p->kind[d]) ? 2 * ng[d] : ng[d];
    // the FFT passes move x rows as element pairs and pair the columns of the y pass
    p->dense[d] = !fft_factor(p->nfft[d], &p->logn[d], &p->odd[d]) || ng[0] % 2 != 0;
    if (p->dense[d] && ng[d] > DENSE_MAX) {
      delete p;
      return fail(CANS_EUNSUPPORTED, "cans_fftini: a transform length outside 2^m 3^a 5^b (a + b <= 3, at most "
                                     "4096) must not exceed 8192 (its dense matrix)");
    }
    if (p->dense[d]) p->nfft[d] = ng[d];
    nf = nf * nfac;
    int rc = p->dense[d] ? upload_dense(p, d) : upload_tables(p, d);
    if (rc) {
      cans_fftend(p);
      return rc;
    }
  }
  p->normfft = 1. / nf;
  // tune the pass shapes for the storage types this context uses (FP32 for pressure-like
  // fields in the mixed mode; FP64 for everything else)
  int rc = tune_type<double>(c, p);
  if (!rc && c->precision == CANS_MIXED) rc = tune_type<float>(c, p);
  if (rc) {
    cans_fftend(p);
    return rc;
  }
  if (normfft) *normfft = p->normfft;
  *out = p;
  return CANS_OK;
}

extern "C" int cans_fft_tuning(const cans_plan* p, int* shapes) {
  if (!p || !shapes) return fail(CANS_EINVAL, "cans_fft_tuning: null argument");
  shapes[0] = p->xthreads[0];
  shapes[1] = p->xthreads[1];
  shapes[2] = p->ylines[0];
  shapes[3] = p->ylines[1];
  return CANS_OK;
}

extern "C" int cans_fft_radix(const cans_plan* p, int* radix) {
  if (!p || !radix) return fail(CANS_EINVAL, "cans_fft_radix: null argument");
  for (int d = 0; d < 2; d++)
    for (int slot = 0; slot < 2; slot++) radix[2 * d + slot] = p->dense[d] ? 0 : p->radix4[d][slot] ? 4 : 8;
  return CANS_OK;
}

extern "C" int cans_fft_set_radix(cans_plan* p, const int* radix) {
  if (!p || !radix) return fail(CANS_EINVAL, "cans_fft_set_radix: null argument");
  for (int q = 0; q < 4; q++)
    if (radix[q] != 4 && radix[q] != 8) return fail(CANS_EINVAL, "cans_fft_set_radix: radix 4 or 8");
  for (int d = 0; d < 2; d++)
    for (int slot = 0; slot < 2; slot++) p->radix4[d][slot] = radix[2 * d + slot] == 4;
  return CANS_OK;
}

extern "C" int cans_fftend(cans_plan* p) {
  if (!p) return CANS_OK;
  for (int d = 0; d < 2; d++) {
    hipFree(p->tw[d]);
    hipFree(p->mk[d]);
    hipFree(p->dm[d][0]);
    hipFree(p->dm[d][1]);
  }
  delete p;
  return CANS_OK;
}
