// Adjoint kernels of the RK3 substep on tri-periodic uniform grids.
//
// The momentum right-hand side of mom_xyz_ad (src/mom.f90:771-1019) is, for
// component c and with K^c_d(x) = (q_d(x) + q_d(x + e_c))/4,
//   F_c(x) = -sum_d h_d [K^c_d(x) q_c(x + e_d) - K^c_d(x - e_d) q_c(x - e_d)] + nu L q_c,
// quadratic in the velocity q.  With M^c_d(x) = g_c(x) q_c(x + e_d) - g_c(x + e_d) q_c(x),
// the transpose of its Jacobian applied to g is
//   (J^T g)_c(y) = -sum_d h_d [K^c_d(y - e_d) g_c(y - e_d) - K^c_d(y) g_c(y + e_d)]
//                  - h_c/4 sum_c' [M^c'_c(y) + M^c'_c(y - e_c')] + nu L g_c(y).
// Every value it reads lies within one cell of y (diagonal neighbours included), so
// halo'd fields with periodic ghost cells serve as they are.
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
  double h[3];            // 1/dx, 1/dy, 1/dz
  double visc, f1, f2;
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

template <bool COMB>
__global__ void __launch_bounds__(JT_X * JT_Y) k_mom_jt(JtArgs a) {
  __shared__ double sq[3][3][JT_PLANE];  // [component][slot][cell of the padded plane]
  __shared__ double sg[3][3][JT_PLANE];
  const Layout& L = a.L;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * JT_X + tx;
  const int i0 = blockIdx.x * JT_X, j0 = blockIdx.y * JT_Y;  // padded cell (0, 0) is (i0, j0)
  const int kb = blockIdx.z * JT_K, ke = min(kb + JT_K, L.n3);
  auto stage = [&](int kk, int slot) {
    for (int e = tid; e < JT_PLANE; e += JT_X * JT_Y) {
      const int i = i0 + e % JT_PX, j = j0 + e / JT_PX;
      const bool in = i <= L.n1 + 1 && j <= L.n2 + 1;
      const long o = L.at(i, j, kk);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        sq[c][slot][e] = in ? a.q[c][o] : 0.;
        if constexpr (COMB) sg[c][slot][e] = in ? a.f1 * a.g[c][o] + a.fb[c][o] : 0.;
        else sg[c][slot][e] = in ? a.g[c][o] : 0.;
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
      const bool in = e < JT_PLANE && i <= L.n1 + 1 && j <= L.n2 + 1;
      const long o = in ? L.at(i, j, kk) : 0;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        rq[r][c] = in ? a.q[c][o] : 0.;
        if constexpr (COMB) rg[r][c] = in ? a.f1 * a.g[c][o] + a.fb[c][o] : 0.;
        else rg[r][c] = in ? a.g[c][o] : 0.;
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
  stage(kb + 2, 2);  // kb < n3, so kb + 2 <= n3 + 1
  __syncthreads();
  const int i = i0 + tx + 1, j = j0 + ty + 1;
  const bool own = i <= L.n1 && j <= L.n2;
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
    }
    __syncthreads();  // every lane is done with plane k-1
    if (next) {
      store((k - 1 - kb) % 3);
      __syncthreads();
    }
  }
}

static dim3 jt_grid(const Layout& L) { return dim3(cdiv(L.n1, JT_X), cdiv(L.n2, JT_Y), cdiv(L.n3, JT_K)); }

int launch_mom_jt(cans_ctx* c, const double* const* q, const double* const* g, const double* const* base,
                  double* const* out, const double* h, double visc) {
  JtArgs a;
  a.L = c->L;
  for (int m = 0; m < 3; m++) {
    a.q[m] = q[m];
    a.g[m] = g[m];
    a.fb[m] = nullptr;
    a.base[m] = base[m];
    a.out[m] = out[m];
    a.fbo[m] = nullptr;
    a.h[m] = h[m];
  }
  a.visc = visc;
  a.f1 = a.f2 = 0.;
  k_mom_jt<false><<<jt_grid(c->L), dim3(JT_X, JT_Y), 0, c->stream>>>(a);
  CANS_LAUNCH_CHECK("k_mom_jt");
  return CANS_OK;
}

int launch_mom_jt_adj(cans_ctx* c, const double* const* q, const double* const* wb, const double* const* fb,
                      double* const* out, double* const* fbo, const double* h, double visc, double f1, double f2) {
  JtArgs a;
  a.L = c->L;
  for (int m = 0; m < 3; m++) {
    a.q[m] = q[m];
    a.g[m] = wb[m];
    a.fb[m] = fb[m];
    a.base[m] = wb[m];
    a.out[m] = out[m];
    a.fbo[m] = fbo[m];
    a.h[m] = h[m];
  }
  a.visc = visc;
  a.f1 = f1;
  a.f2 = f2;
  k_mom_jt<true><<<jt_grid(c->L), dim3(JT_X, JT_Y), 0, c->stream>>>(a);
  CANS_LAUNCH_CHECK("k_mom_jt");
  return CANS_OK;
}

}  // namespace cans
