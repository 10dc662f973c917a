// Adjoint kernels of the RK3 substep on boxes whose directions are each periodic or
// bounded by walls (a periodic z on a uniform grid, walls in z on any grid).
//
// The momentum right-hand side of mom_xyz_ad (src/mom.f90:771-1019) is, for
// component c, with K^c_d(x) = alpha^c_d(k) q_d(x) + beta^c_d(k) q_d(x + e_c) (1/4 and
// 1/4, except rdzm(k)/4 and rdzp(k)/4 for w advected along x and y) and H^c_d = 1/dx,
// 1/dy along x, y and dzfi(k) (u, v) or dzci(k) (w) along z,
//   F_c(x) = -sum_d H^c_d(k) [K^c_d(x) q_c(x + e_d) - K^c_d(x - e_d) q_c(x - e_d)] + nu L_c q_c,
// quadratic in the velocity q.  With G^c_d(x) = H^c_d(k_x) g_c(x) and
// M^c_d(x) = G^c_d(x) q_c(x + e_d) - G^c_d(x + e_d) q_c(x), the transpose of its
// Jacobian applied to g is
//   (J^T g)_c(y) = -sum_d [K^c_d(y - e_d) G^c_d(y - e_d) - K^c_d(y) G^c_d(y + e_d)]
//                  - sum_c' [alpha^c'_c(k_y) M^c'_c(y) + beta^c'_c(k_y - [c' = z]) M^c'_c(y - e_c')]
//                  + nu L_c^T g_c(y).
// Every value it reads lies within one cell of y (diagonal neighbours included).  On
// a uniform periodic grid the coefficients are constants and halo'd fields with
// periodic ghost cells serve as they are.  Along a wall direction the sum also runs
// over the ghost layers (g is zero wherever F is not a degree of freedom: those
// layers and a normal velocity's wall face), and k_jt_fold folds each ghost layer
// into the adjacent interior one by the derivative of bound.f90's ghost value
// (u(0) = 2 ub - u(1): -1; a Neumann ghost: +1; the normal velocity's wall faces
// are constants).  bound.f90 sets the ghosts x, then y, then z, each over the
// others' ghost layers, so the folds run z, then y, then x.
#include "common.hpp"
#include "kernels.hpp"

namespace cans {

struct JtArgs {
  Layout L;
  const double* q[3];     // velocity at which the Jacobian is taken (halo'd, ghosts set)
  const double* g[3];     // vector the transpose applies to (halo'd, ghosts set); with COMB: wb
  const double* fb[3];    // COMB: g = f1 wb + fb (fb halo'd, ghosts set)
  const double* base[3];  // added to the result (own cell only: may be out)
  double* out[3];
  double* fbo[3];         // COMB: receives f2 wb
  double h[3];            // 1/dx, 1/dy, 1/dz (periodic z)
  double visc, f1, f2;
  const double* dzci;     // walls in z: dzci(0:n3+1), dzfi(0:n3+1)
  const double* dzfi;
};

// A work-group owns a JT_X x JT_Y tile of (i, j) and marches over JT_K planes of k.  The
// planes k-1, k, k+1 of q and g (with one ghost cell around the tile) sit in LDS in three
// rotating slots; each new plane is staged once, so every value is read from global
// memory about (JT_X+2)(JT_Y+2)/(JT_X JT_Y) times instead of once per use.  COMB: the
// adjoint substep's g = f1 wb + fb is formed while it is staged, and f2 wb is stored
// beside the result.  32 x 8 cells: 49 KB of LDS, so three work-groups share a CU (64 x 4
// needs 57 KB and allows two) and the ring re-reads are 1.33x instead of 1.55x.
constexpr int JT_X = 32, JT_Y = 8, JT_K = 16;
constexpr int JT_PX = JT_X + 2, JT_PLANE = (JT_X + 2) * (JT_Y + 2);

// WM: the directions bounded by walls (bit 0 x, bit 1 y, bit 2 z; 0: tri-periodic uniform).
// Along a wall direction the outputs include the ghost layers (i = 0 .. n1 + 1; a work-group's
// first output is blockIdx.x * JT_X) and the staged ring reaches -1 and n + 2, which read as
// zero.  g is read as zero off the degrees of freedom: the ghost layers of a wall direction and
// a normal velocity's wall face (u at i = n1, v at j = n2, w at k = n3).
template <bool COMB, int WM>
__global__ void __launch_bounds__(JT_X * JT_Y) k_mom_jt(JtArgs a) {
  constexpr bool WALL = WM != 0;
  constexpr int OX = WM & 1, OY = (WM >> 1) & 1, OZ = (WM >> 2) & 1;
  __shared__ double sq[3][3][JT_PLANE];  // [component][slot][cell of the padded plane]
  __shared__ double sg[3][3][JT_PLANE];
  const Layout& L = a.L;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * JT_X + tx;
  const int i0 = blockIdx.x * JT_X - OX, j0 = blockIdx.y * JT_Y - OY;  // padded cell (0, 0) is (i0, j0)
  const int kb = (int)blockIdx.z * JT_K - OZ;
  const int ke = min(kb + JT_K, L.n3 + OZ);
  // (i, j, kk) lies in the halo'd fields; g_c is a degree of freedom there
  auto qin = [&](int i, int j, int kk) {
    if (!WALL) return i <= L.n1 + 1 && j <= L.n2 + 1;
    return i >= 0 && i <= L.n1 + 1 && j >= 0 && j <= L.n2 + 1 && kk >= 0 && kk <= L.n3 + 1;
  };
  auto gin = [&](int c, int i, int j, int kk) {
    return (!OX || (i >= 1 && i <= L.n1 - (c == 0))) && (!OY || (j >= 1 && j <= L.n2 - (c == 1))) &&
           (!OZ || (kk >= 1 && kk <= L.n3 - (c == 2)));
  };
  auto stage = [&](int kk, int slot) {
    for (int e = tid; e < JT_PLANE; e += JT_X * JT_Y) {
      const int i = i0 + e % JT_PX, j = j0 + e / JT_PX;
      const bool in = qin(i, j, kk);
      const long o = in ? L.at(i, j, kk) : 0;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const bool ing = in && gin(c, i, j, kk);
        sq[c][slot][e] = in ? a.q[c][o] : 0.;
        if constexpr (COMB) sg[c][slot][e] = ing ? a.f1 * a.g[c][o] + a.fb[c][o] : 0.;
        else sg[c][slot][e] = ing ? a.g[c][o] : 0.;
      }
    }
  };
  // the plane after next is loaded into registers before the current plane's arithmetic and
  // stored to LDS after it, so its global-memory latency overlaps that work
  constexpr int NE = (JT_PLANE + JT_X * JT_Y - 1) / (JT_X * JT_Y);  // plane cells per lane
  double rq[NE][3], rg[NE][3];
  auto load = [&](int kk) {
#pragma unroll
    for (int r = 0; r < NE; r++) {
      const int e = tid + r * JT_X * JT_Y;
      const int i = i0 + e % JT_PX, j = j0 + e / JT_PX;
      const bool in = e < JT_PLANE && qin(i, j, kk);
      const long o = in ? L.at(i, j, kk) : 0;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const bool ing = in && gin(c, i, j, kk);
        rq[r][c] = in ? a.q[c][o] : 0.;
        if constexpr (COMB) rg[r][c] = ing ? a.f1 * a.g[c][o] + a.fb[c][o] : 0.;
        else rg[r][c] = ing ? a.g[c][o] : 0.;
      }
    }
  };
  auto store = [&](int slot) {
#pragma unroll
    for (int r = 0; r < NE; r++) {
      const int e = tid + r * JT_X * JT_Y;
      if (e < JT_PLANE) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
          sq[c][slot][e] = rq[r][c];
          sg[c][slot][e] = rg[r][c];
        }
      }
    }
  };
  stage(kb, 0);
  stage(kb + 1, 1);
  stage(kb + 2, 2);  // periodic z: kb < n3, so kb + 2 <= n3 + 1; walls: planes past n3 + 1 read as zero
  __syncthreads();
  const int i = i0 + tx + 1, j = j0 + ty + 1;
  const bool own = i <= L.n1 + OX && j <= L.n2 + OY;
  const bool inxy = i >= 1 && i <= L.n1 && j >= 1 && j <= L.n2;
  const int ec[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int k = kb + 1; k <= ke; k++) {
    const bool next = k + 2 <= ke + 1;
    if (next) load(k + 2);
    if (own) {
      const int sl[3] = {(k - 1 - kb) % 3, (k - kb) % 3, (k + 1 - kb) % 3};  // slots of k-1, k, k+1
      auto at = [&](int di, int dj) { return (ty + 1 + dj) * JT_PX + tx + 1 + di; };
      auto Q = [&](int c, const int* d) { return sq[c][sl[1 + d[2]]][at(d[0], d[1])]; };
      auto G = [&](int c, const int* d) { return sg[c][sl[1 + d[2]]][at(d[0], d[1])]; };
      const long o = L.at(i, j, k);
      if constexpr (WM == 0) {
#pragma unroll
        for (int m = 0; m < 3; m++) {
          const int* em = ec[m];
          const int z0[3] = {0, 0, 0};
          const double g0 = G(m, z0);
          double lap = 0.;
#pragma unroll
          for (int d = 0; d < 3; d++) {
            const int p[3] = {ec[d][0], ec[d][1], ec[d][2]}, n[3] = {-ec[d][0], -ec[d][1], -ec[d][2]};
            lap += a.h[d] * a.h[d] * (G(m, p) - 2. * g0 + G(m, n));
          }
          double acc = a.visc * lap;
          // g_m as the advected quantity: the coefficients K^m_d at y - e_d and y
#pragma unroll
          for (int d = 0; d < 3; d++) {
            const int p[3] = {ec[d][0], ec[d][1], ec[d][2]}, n[3] = {-ec[d][0], -ec[d][1], -ec[d][2]};
            const int nm[3] = {n[0] + em[0], n[1] + em[1], n[2] + em[2]};
            const double kmm = 0.25 * (Q(d, n) + Q(d, nm));
            const double km0 = 0.25 * (Q(d, z0) + Q(d, em));
            acc -= a.h[d] * (kmm * G(m, n) - km0 * G(m, p));
          }
          // q_m as the advecting velocity of every component c (direction m)
          double s = 0.;
#pragma unroll
          for (int c = 0; c < 3; c++) {
            const int n[3] = {-ec[c][0], -ec[c][1], -ec[c][2]};
            const int nm[3] = {n[0] + em[0], n[1] + em[1], n[2] + em[2]};
            s += G(c, z0) * Q(c, em) - G(c, em) * Q(c, z0);
            s += G(c, n) * Q(c, nm) - G(c, nm) * Q(c, n);
          }
          acc -= 0.25 * a.h[m] * s;
          const double b = a.base[m][o];
          a.out[m][o] = b + acc;
          if constexpr (COMB) a.fbo[m][o] = a.f2 * b;  // base is wb
        }
      } else {
        const int n3 = L.n3;
        auto zc = [&](int kk) { return kk >= 0 && kk <= n3 + 1 ? a.dzci[kk] : 0.; };
        auto zf = [&](int kk) { return kk >= 0 && kk <= n3 + 1 ? a.dzfi[kk] : 0.; };
        const double zcm = zc(k - 1), zc0 = zc(k), zcp = zc(k + 1);
        const double zfm = zf(k - 1), zf0 = zf(k), zfp = zf(k + 1);
        const bool in0 = k >= 1 && k <= n3, inm = !OZ || (k - 1 >= 1 && k - 1 <= n3);
        // rdzm(k) = dzci(k)/dzfi(k), rdzp(k) = dzci(k)/dzfi(k+1) (mom.f90's weights of u, v at w's
        // height; a periodic z reads them off the ghost entries); zero off the interior of a walled
        // z, where every term that would use them has g = 0
        const double rdzm0 = in0 ? zc0 / zf0 : 0., rdzp0 = in0 ? zc0 / zfp : 0., rdzpm = inm ? zcm / zf0 : 0.;
        // H^c_z on plane k + dk
        auto hz = [&](int c, int dk) { return c == 2 ? (dk < 0 ? zcm : dk > 0 ? zcp : zc0)
                                                     : (dk < 0 ? zfm : dk > 0 ? zfp : zf0); };
        auto H = [&](int c, int d, int dk) { return d < 2 ? a.h[d] : hz(c, dk); };
#pragma unroll
        for (int m = 0; m < 3; m++) {
          const int* em = ec[m];
          const int z0[3] = {0, 0, 0}, zp[3] = {0, 0, 1}, zm[3] = {0, 0, -1};
          const double g0 = G(m, z0);
          double lap = 0.;
#pragma unroll
          for (int d = 0; d < 2; d++) {
            const int p[3] = {ec[d][0], ec[d][1], ec[d][2]}, n[3] = {-ec[d][0], -ec[d][1], -ec[d][2]};
            lap += a.h[d] * a.h[d] * (G(m, p) - 2. * g0 + G(m, n));
          }
          if (m < 2) {  // L = dzfi(k) [dzci(k) d+ - dzci(k-1) d-]: L^T g = dzci(k) d+(dzfi g) - dzci(k-1) d-(dzfi g)
            const double gp = zfp * G(m, zp), gc = zf0 * g0, gm = zfm * G(m, zm);
            lap += zc0 * (gp - gc) - zcm * (gc - gm);
          } else {      // L = dzci(k) [dzfi(k+1) d+ - dzfi(k) d-]
            const double gp = zcp * G(m, zp), gc = zc0 * g0, gm = zcm * G(m, zm);
            lap += zfp * (gp - gc) - zf0 * (gc - gm);
          }
          double acc = a.visc * lap;
          // g_m as the advected quantity: K^m_d at y - e_d and y
#pragma unroll
          for (int d = 0; d < 3; d++) {
            const int p[3] = {ec[d][0], ec[d][1], ec[d][2]}, n[3] = {-ec[d][0], -ec[d][1], -ec[d][2]};
            const int nm[3] = {n[0] + em[0], n[1] + em[1], n[2] + em[2]};
            const bool wt = m == 2 && d < 2;  // same k at y - e_d and y
            const double al = wt ? 0.25 * rdzm0 : 0.25, be = wt ? 0.25 * rdzp0 : 0.25;
            const double kmm = al * Q(d, n) + be * Q(d, nm);
            const double km0 = al * Q(d, z0) + be * Q(d, em);
            acc -= kmm * H(m, d, -ec[d][2]) * G(m, n) - km0 * H(m, d, ec[d][2]) * G(m, p);
          }
          // q_m as the advecting velocity of every component c (direction m)
#pragma unroll
          for (int c = 0; c < 3; c++) {
            const int n[3] = {-ec[c][0], -ec[c][1], -ec[c][2]};
            const int nm[3] = {n[0] + em[0], n[1] + em[1], n[2] + em[2]};
            const int dkc = -ec[c][2], dkm = ec[m][2];
            const double m0 = H(c, m, 0) * G(c, z0) * Q(c, em) - H(c, m, dkm) * G(c, em) * Q(c, z0);
            const double m1 = H(c, m, dkc) * G(c, n) * Q(c, nm) - H(c, m, dkc + dkm) * G(c, nm) * Q(c, n);
            const bool wt = c == 2 && m < 2;
            acc -= (wt ? 0.25 * rdzm0 : 0.25) * m0 + (wt ? 0.25 * rdzpm : 0.25) * m1;
          }
          const double b = in0 && inxy ? a.base[m][o] : 0.;  // the ghost layers carry J^T g only
          a.out[m][o] = b + acc;
          if constexpr (COMB) a.fbo[m][o] = a.f2 * b;
        }
      }
    }
    __syncthreads();  // every lane is done with plane k-1
    if (next) {
      store((k - 1 - kb) % 3);
      __syncthreads();
    }
  }
}

static dim3 jt_grid(const Layout& L, int wm) {
  return dim3(cdiv(L.n1 + 2 * (wm & 1), JT_X), cdiv(L.n2 + 2 * (wm >> 1 & 1), JT_Y), cdiv(L.n3 + 2 * (wm >> 2 & 1), JT_K));
}

// Folds the ghost layers of a wall direction d that k_mom_jt<.., WM> wrote into the interior:
// for each tangential component m, q_m(1) += s0 q_m(0) and q_m(n) += s1 q_m(n + 1) along d (s = -1 for
// a Dirichlet wall, +1 for Neumann), the ghost layers then cleared; the normal component's wall
// face and ghost layers are cleared (constants).  One thread per point of the plane of the other
// two directions, over lo..hi of each (their ghost layers included while they are still to fold).
struct JtFold {
  int d;
  int lo[3], hi[3];
  double s[3][2];
};

__global__ void k_jt_fold(Layout L, double* u, double* v, double* w, JtFold f) {
  const int a = f.d == 0 ? 1 : 0, b = f.d == 2 ? 1 : 2;  // the plane's two directions
  int idx[3];
  idx[a] = f.lo[a] + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  idx[b] = f.lo[b] + (int)blockIdx.y;
  if (idx[a] > f.hi[a]) return;
  const int nd = f.d == 0 ? L.n1 : f.d == 1 ? L.n2 : L.n3;
  auto at = [&](int t) {
    int x[3] = {idx[0], idx[1], idx[2]};
    x[f.d] = t;
    return L.at(x[0], x[1], x[2]);
  };
  double* q[3] = {u, v, w};
  for (int m = 0; m < 3; m++) {
    if (m == f.d) {
      q[m][at(0)] = 0.;
      q[m][at(nd)] = 0.;
      q[m][at(nd + 1)] = 0.;
    } else {
      q[m][at(1)] += f.s[m][0] * q[m][at(0)];
      q[m][at(nd)] += f.s[m][1] * q[m][at(nd + 1)];
      q[m][at(0)] = 0.;
      q[m][at(nd + 1)] = 0.;
    }
  }
}

// The volume weights of the projection's transpose P^T = M P M^-1 (M: dzf(k) for u, v; dzc(k)
// for w) and the transposed bulk forcing, on the interior:
//   inv:  u *= dzfi(k), v *= dzfi(k), w *= dzci(k);
//   else: u = (u - mu) / dzfi(k), v = (v - mv) / dzfi(k), w = w / dzci(k)
// (dzfi, dzci null: no weights, only the means subtracted).  The normal velocities' wall faces
// of the directions in wm (u at i = n1, v at j = n2, w at k = n3) are cleared.
__global__ void k_adj_weigh(Layout L, double* u, double* v, double* w, const double* dzfi, const double* dzci,
                            int inv, double mu, double mv, int wm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1, k = blockIdx.z + 1;
  if (i > L.n1) return;
  const long o = L.at(i, j, k);
  if (inv) {
    if (dzfi) {
      u[o] *= dzfi[k];
      v[o] *= dzfi[k];
      w[o] *= dzci[k];
    }
  } else if (dzfi) {
    u[o] = (u[o] - mu) / dzfi[k];
    v[o] = (v[o] - mv) / dzfi[k];
    w[o] = w[o] / dzci[k];
  } else {
    u[o] -= mu;
    v[o] -= mv;
  }
  if ((wm & 1) && i == L.n1) u[o] = 0.;
  if ((wm & 2) && j == L.n2) v[o] = 0.;
  if ((wm & 4) && k == L.n3) w[o] = 0.;
}

// out += c L a on the interior, L the 7-point Laplacian of a tri-periodic uniform grid (h2: 1/dx^2,
// 1/dy^2, 1/dz^2), a halo'd with its periodic ghost cells set: the implicit-diffusion adjoint's
// (f1 + f2)/2 visc L hb
__global__ void k_adj_lap(Layout L, const double* a0, const double* a1, const double* a2, double* o0, double* o1,
                          double* o2, double c, double hx2, double hy2, double hz2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1, k = blockIdx.z + 1;
  if (i > L.n1) return;
  const long o = L.at(i, j, k), sy = L.sy, sz = L.sz;
  const double* a[3] = {a0, a1, a2};
  double* out[3] = {o0, o1, o2};
#pragma unroll
  for (int m = 0; m < 3; m++) {
    const double* x = a[m];
    const double x0 = x[o];
    const double lap = hx2 * (x[o + 1] - 2. * x0 + x[o - 1]) + hy2 * (x[o + sy] - 2. * x0 + x[o - sy]) +
                       hz2 * (x[o + sz] - 2. * x0 + x[o - sz]);
    out[m][o] += c * lap;
  }
}

int launch_adj_lap(cans_ctx* c, const double* const* a, double* const* out, double coef, const double* h) {
  const Layout& L = c->L;
  k_adj_lap<<<dim3(cdiv(L.n1, 64), L.n2, L.n3), 64, 0, c->stream>>>(L, a[0], a[1], a[2], out[0], out[1], out[2],
                                                                    coef, h[0] * h[0], h[1] * h[1], h[2] * h[2]);
  CANS_LAUNCH_CHECK("k_adj_lap");
  return CANS_OK;
}

// ---------------------------------------------------------------------------
// Scalars (explicit diffusion; each direction periodic or walled).  The scalar
// right-hand side of scal.f90:41-98 is
//   S(x) = -sum_d H_d(k) [F_d(x) - F_d(x - e_d)] + alpha L s,  F_d(x) = (s(x + e_d) + s(x))/2 q_d(x),
// H = dxi, dyi, dzfi(k); with R_d(x) = H_d(k_x) sigma(x) - H_d(k_x+e_d) sigma(x + e_d) its
// transposed Jacobians are
//   (J_s^T sigma)(y)   = -1/2 sum_d [q_d(y - e_d) R_d(y - e_d) + q_d(y) R_d(y)] + alpha L^T sigma(y),
//   (J_q^T sigma)_d(y) = -1/2 (s(y + e_d) + s(y)) R_d(y).
// k_scal_jt forms sigma = f1 sh + fsb as it reads it (zero off the scalar's degrees of
// freedom: the ghost layers of the wall directions), writes sbo = sh + J_s^T sigma (on those
// ghost layers too, which k_scal_fold folds, z first, then y, then x: boundp sets the ghosts x,
// y, z), fsbo = f2 sh, and adds J_q^T sigma to the velocity adjoint at the velocity's degrees of
// freedom (not on a normal velocity's wall face).  One thread per cell.
struct ScalJtArgs {
  Layout L;
  const double* q[3];  // the velocity the forward substep started from (halo'd, ghosts set)
  const double* s;     // the scalar it started from (halo'd, ghosts set)
  const double* sh;    // the updated scalar's adjoint (halo'd, periodic ghosts set)
  const double* fsb;   // the stored scalar right-hand side's adjoint (halo'd, periodic ghosts set)
  double *sbo, *fsbo;
  double* ubo[3];
  const double *dzci, *dzfi;  // (0:n3+1)
  double dxi, dyi, alpha, f1, f2;
  int wm;  // the wall directions, bit 0 x, 1 y, 2 z
};

__global__ void k_scal_jt(ScalJtArgs a) {
  const Layout& L = a.L;
  const int ox = a.wm & 1, oy = a.wm >> 1 & 1, oz = a.wm >> 2 & 1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1 - ox, j = (int)blockIdx.y + 1 - oy;
  const int k = (int)blockIdx.z + 1 - oz;
  if (i > L.n1 + ox) return;
  const int n[3] = {L.n1, L.n2, L.n3};
  // (ii, jj, kk) lies in the halo'd fields; sigma is a degree of freedom there
  auto inh = [&](int ii, int jj, int kk) {
    return ii >= 0 && ii <= n[0] + 1 && jj >= 0 && jj <= n[1] + 1 && kk >= 0 && kk <= n[2] + 1;
  };
  auto dof = [&](int ii, int jj, int kk) {
    return (!ox || (ii >= 1 && ii <= n[0])) && (!oy || (jj >= 1 && jj <= n[1])) && (!oz || (kk >= 1 && kk <= n[2]));
  };
  auto sig = [&](int di, int dj, int dk) -> double {
    const int ii = i + di, jj = j + dj, kk = k + dk;
    if (!inh(ii, jj, kk) || !dof(ii, jj, kk)) return 0.;
    const long o = L.at(ii, jj, kk);
    return a.f1 * a.sh[o] + a.fsb[o];
  };
  auto Q = [&](int d, int di, int dj, int dk) -> double {
    const int ii = i + di, jj = j + dj, kk = k + dk;
    return inh(ii, jj, kk) ? a.q[d][L.at(ii, jj, kk)] : 0.;
  };
  auto zf = [&](int kk) { return kk >= 0 && kk <= n[2] + 1 ? a.dzfi[kk] : 0.; };
  auto zc = [&](int kk) { return kk >= 0 && kk <= n[2] + 1 ? a.dzci[kk] : 0.; };
  auto H = [&](int d, int dk) { return d == 0 ? a.dxi : d == 1 ? a.dyi : zf(k + dk); };
  // R_d at the cell (i + di, j + dj, k + dk)
  auto R = [&](int d, int di, int dj, int dk) {
    const int ex = d == 0, ey = d == 1, ez = d == 2;
    return H(d, dk) * sig(di, dj, dk) - H(d, dk + ez) * sig(di + ex, dj + ey, dk + ez);
  };
  double js = 0.;
  double rd[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int ex = d == 0, ey = d == 1, ez = d == 2;
    rd[d] = R(d, 0, 0, 0);
    js -= 0.5 * (Q(d, -ex, -ey, -ez) * R(d, -ex, -ey, -ez) + Q(d, 0, 0, 0) * rd[d]);
  }
  const double s0 = sig(0, 0, 0);
  const double lxy = a.dxi * a.dxi * (sig(1, 0, 0) - 2. * s0 + sig(-1, 0, 0)) +
                     a.dyi * a.dyi * (sig(0, 1, 0) - 2. * s0 + sig(0, -1, 0));
  // L = dzfi(k) [dzci(k) d+ - dzci(k-1) d-]: L^T sigma = dzci(k) d+(dzfi sigma) - dzci(k-1) d-(dzfi sigma)
  const double gp = zf(k + 1) * sig(0, 0, 1), gc = zf(k) * s0, gm = zf(k - 1) * sig(0, 0, -1);
  js += a.alpha * (lxy + zc(k) * (gp - gc) - zc(k - 1) * (gc - gm));
  const long o = L.at(i, j, k);
  const bool in = i >= 1 && i <= n[0] && j >= 1 && j <= n[1] && k >= 1 && k <= n[2];
  const double base = in ? a.sh[o] : 0.;
  a.sbo[o] = base + js;
  if (!in) return;
  a.fsbo[o] = a.f2 * base;
  const double sc = a.s[o];
  const double sn[3] = {a.s[o + 1], a.s[o + L.sy], a.s[o + L.sz]};
  const int idx[3] = {i, j, k};
#pragma unroll
  for (int d = 0; d < 3; d++)
    if (!((a.wm >> d & 1) && idx[d] == n[d])) a.ubo[d][o] += -0.5 * (sn[d] + sc) * rd[d];  // not a wall face
}

// The scalar's ghost layers of wall direction d folded into the interior: s(1) += s0 s(0), s(n) +=
// s1 s(n + 1) (-1 Dirichlet, +1 Neumann), then cleared; one thread per point of the plane of the
// other two directions, over lo..hi of each (their ghost layers included while still to fold).
struct ScalFold {
  int d;
  int lo[3], hi[3];
  double s0, s1;
};

__global__ void k_scal_fold(Layout L, double* sb, ScalFold f) {
  const int a = f.d == 0 ? 1 : 0, b = f.d == 2 ? 1 : 2;
  int idx[3];
  idx[a] = f.lo[a] + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  idx[b] = f.lo[b] + (int)blockIdx.y;
  if (idx[a] > f.hi[a]) return;
  const int nd = f.d == 0 ? L.n1 : f.d == 1 ? L.n2 : L.n3;
  auto at = [&](int t) {
    int x[3] = {idx[0], idx[1], idx[2]};
    x[f.d] = t;
    return L.at(x[0], x[1], x[2]);
  };
  sb[at(1)] += f.s0 * sb[at(0)];
  sb[at(nd)] += f.s1 * sb[at(nd + 1)];
  sb[at(0)] = 0.;
  sb[at(nd + 1)] = 0.;
}

// sh += the Boussinesq term's transpose (rk.f90:163-177: u_c -= coef_c (s(x + e_c) + s(x))):
//   sh(y) -= sum_c coef_c (wb_c(y - e_c) + wb_c(y)),
// wb read as zero off the velocity's degrees of freedom (a normal velocity's wall faces; its
// terms at the scalar's ghost cells all vanish, so the interior is all it changes).
__global__ void k_scal_buoy_t(Layout L, double* sh, const double* wu, const double* wv, const double* ww,
                              double c0, double c1, double c2, int wm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1, k = blockIdx.z + 1;
  if (i > L.n1) return;
  const long o = L.at(i, j, k);
  const double um = ((wm & 1) && i == 1) ? 0. : wu[o - 1], u0 = ((wm & 1) && i == L.n1) ? 0. : wu[o];
  const double vm = ((wm & 2) && j == 1) ? 0. : wv[o - L.sy], v0 = ((wm & 2) && j == L.n2) ? 0. : wv[o];
  const double wmv = ((wm & 4) && k == 1) ? 0. : ww[o - L.sz], w0 = ((wm & 4) && k == L.n3) ? 0. : ww[o];
  sh[o] -= c0 * (um + u0) + c1 * (vm + v0) + c2 * (wmv + w0);
}

int launch_scal_adjoint(cans_ctx* c, const ScalAdj& sa) {
  const Layout& L = c->L;
  if (sa.buoy) {
    k_scal_buoy_t<<<dim3(cdiv(L.n1, 64), L.n2, L.n3), 64, 0, c->stream>>>(L, sa.sh, sa.wb[0], sa.wb[1], sa.wb[2],
                                                                          sa.coef[0], sa.coef[1], sa.coef[2],
                                                                          sa.wm);
    CANS_LAUNCH_CHECK("k_scal_buoy_t");
  }
  return CANS_OK;
}

// the bulk forcing s <- s + (scalf - sum_k gvr_f(k) s) has the Jacobian I - 1 gvr_f^T: its transpose
// subtracts gvr_f(k) times the plain sum of the adjoint
__global__ void k_scal_force_t(Layout L, double* sh, const double* gvr_f, double sum) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1, k = blockIdx.z + 1;
  if (i > L.n1) return;
  sh[L.at(i, j, k)] -= gvr_f[k] * sum;
}

int launch_scal_force_t(cans_ctx* c, double* sh, const double* gvr_f, double sum) {
  const Layout& L = c->L;
  k_scal_force_t<<<dim3(cdiv(L.n1, 64), L.n2, L.n3), 64, 0, c->stream>>>(L, sh, gvr_f, sum);
  CANS_LAUNCH_CHECK("k_scal_force_t");
  return CANS_OK;
}

int launch_scal_jt(cans_ctx* c, const ScalAdj& sa) {
  const Layout& L = c->L;
  ScalJtArgs a;
  a.L = L;
  for (int d = 0; d < 3; d++) {
    a.q[d] = sa.q[d];
    a.ubo[d] = sa.ubo[d];
  }
  a.s = sa.s;
  a.sh = sa.sh;
  a.fsb = sa.fsb;
  a.sbo = sa.sbo;
  a.fsbo = sa.fsbo;
  a.dzci = sa.dzci;
  a.dzfi = sa.dzfi;
  a.dxi = sa.dxi;
  a.dyi = sa.dyi;
  a.alpha = sa.alpha;
  a.f1 = sa.f1;
  a.f2 = sa.f2;
  a.wm = sa.wm;
  const int ox = sa.wm & 1, oy = sa.wm >> 1 & 1, oz = sa.wm >> 2 & 1;
  k_scal_jt<<<dim3(cdiv(L.n1 + 2 * ox, 64), L.n2 + 2 * oy, L.n3 + 2 * oz), 64, 0, c->stream>>>(a);
  CANS_LAUNCH_CHECK("k_scal_jt");
  const int n[3] = {L.n1, L.n2, L.n3};
  for (int d = 2; d >= 0; d--) {  // B^T in the reverse of boundp's x, y, z order
    if (!(sa.wm >> d & 1)) continue;
    ScalFold f;
    f.d = d;
    for (int e = 0; e < 3; e++) {
      const bool ghosts = e < d && (sa.wm >> e & 1);
      f.lo[e] = ghosts ? 0 : 1;
      f.hi[e] = ghosts ? n[e] + 1 : n[e];
    }
    f.s0 = sa.sign[d][0];
    f.s1 = sa.sign[d][1];
    const int pa = d == 0 ? 1 : 0, pb = d == 2 ? 1 : 2;
    k_scal_fold<<<dim3(cdiv(f.hi[pa] - f.lo[pa] + 1, 64), f.hi[pb] - f.lo[pb] + 1), 64, 0, c->stream>>>(L, sa.sbo,
                                                                                                          f);
    CANS_LAUNCH_CHECK("k_scal_fold");
  }
  return CANS_OK;
}

static JtArgs jt_args(cans_ctx* c, const double* const* q, const double* const* g, const double* const* fb,
                      const double* const* base, double* const* out, double* const* fbo, const double* h,
                      double visc, double f1, double f2, const double* dzci, const double* dzfi) {
  JtArgs a;
  a.L = c->L;
  for (int m = 0; m < 3; m++) {
    a.q[m] = q[m];
    a.g[m] = g[m];
    a.fb[m] = fb ? fb[m] : nullptr;
    a.base[m] = base[m];
    a.out[m] = out[m];
    a.fbo[m] = fbo ? fbo[m] : nullptr;
    a.h[m] = h[m];
  }
  a.visc = visc;
  a.f1 = f1;
  a.f2 = f2;
  a.dzci = dzci;
  a.dzfi = dzfi;
  return a;
}

int launch_mom_jt(cans_ctx* c, const double* const* q, const double* const* g, const double* const* base,
                  double* const* out, const double* h, double visc) {
  const JtArgs a = jt_args(c, q, g, nullptr, base, out, nullptr, h, visc, 0., 0., nullptr, nullptr);
  k_mom_jt<false, 0><<<jt_grid(c->L, 0), dim3(JT_X, JT_Y), 0, c->stream>>>(a);
  CANS_LAUNCH_CHECK("k_mom_jt");
  return CANS_OK;
}

template <int WM>
static void jt_adj_launch(cans_ctx* c, const JtArgs& a) {
  k_mom_jt<true, WM><<<jt_grid(c->L, WM), dim3(JT_X, JT_Y), 0, c->stream>>>(a);
}

int launch_mom_jt_adj(cans_ctx* c, const double* const* q, const double* const* wb, const double* const* fb,
                      double* const* out, double* const* fbo, const double* h, double visc, double f1, double f2,
                      const AdjWalls* wall) {
  const int wm = wall ? wall->mask & 7 : 0;
  const JtArgs a = jt_args(c, q, wb, fb, wb, out, fbo, h, visc, f1, f2, wm ? wall->dzci : nullptr,
                           wm ? wall->dzfi : nullptr);
  switch (wm) {
    case 0: jt_adj_launch<0>(c, a); break;
    case 1: jt_adj_launch<1>(c, a); break;
    case 2: jt_adj_launch<2>(c, a); break;
    case 3: jt_adj_launch<3>(c, a); break;
    case 4: jt_adj_launch<4>(c, a); break;
    case 5: jt_adj_launch<5>(c, a); break;
    case 6: jt_adj_launch<6>(c, a); break;
    default: jt_adj_launch<7>(c, a); break;
  }
  CANS_LAUNCH_CHECK("k_mom_jt");
  // B^T in the reverse of bound's x, y, z order: z first; a direction still to fold keeps its
  // ghost layers in the plane of the later folds
  const Layout& L = c->L;
  const int n[3] = {L.n1, L.n2, L.n3};
  for (int d = 2; d >= 0; d--) {
    if (!(wm >> d & 1)) continue;
    JtFold f;
    f.d = d;
    for (int e = 0; e < 3; e++) {
      const bool ghosts = e < d && (wm >> e & 1);
      f.lo[e] = ghosts ? 0 : 1;
      f.hi[e] = ghosts ? n[e] + 1 : n[e];
      f.s[e][0] = wall->sign[e][d][0];
      f.s[e][1] = wall->sign[e][d][1];
    }
    const int pa = d == 0 ? 1 : 0, pb = d == 2 ? 1 : 2;
    const int na = f.hi[pa] - f.lo[pa] + 1, nb = f.hi[pb] - f.lo[pb] + 1;
    k_jt_fold<<<dim3(cdiv(na, 64), nb), 64, 0, c->stream>>>(L, out[0], out[1], out[2], f);
    CANS_LAUNCH_CHECK("k_jt_fold");
  }
  return CANS_OK;
}

int launch_adj_weigh(cans_ctx* c, double* const* f, const double* dzfi, const double* dzci, bool inv, double mu,
                     double mv, int wm) {
  const Layout& L = c->L;
  k_adj_weigh<<<dim3(cdiv(L.n1, 64), L.n2, L.n3), 64, 0, c->stream>>>(L, f[0], f[1], f[2], dzfi, dzci, inv, mu, mv,
                                                                      wm);
  CANS_LAUNCH_CHECK("k_adj_weigh");
  return CANS_OK;
}

}  // namespace cans
