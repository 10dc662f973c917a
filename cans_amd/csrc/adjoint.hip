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

// one thread per cell, all three components.  COMB: the adjoint substep's g = f1 wb + fb
// is formed where it is read, and f2 wb is stored beside the result (no separate pass)
template <bool COMB>
__global__ void __launch_bounds__(256) k_mom_jt(JtArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > a.L.n1 || j > a.L.n2) return;
  const long o = a.L.at(i, j, k);
  const long st[3] = {1, a.L.sy, a.L.sz};
  auto G = [&](int c, long off) {
    if constexpr (COMB) return a.f1 * a.g[c][o + off] + a.fb[c][o + off];
    else return a.g[c][o + off];
  };
#pragma unroll
  for (int m = 0; m < 3; m++) {
    const long em = st[m];
    const double g0 = G(m, 0);
    double lap = 0.;
#pragma unroll
    for (int d = 0; d < 3; d++) lap += a.h[d] * a.h[d] * (G(m, st[d]) - 2. * g0 + G(m, -st[d]));
    double acc = a.visc * lap;
    // g_m as the advected quantity: the coefficients K^m_d at y - e_d and y
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const long ed = st[d];
      const double* qd = a.q[d];
      const double kmm = 0.25 * (qd[o - ed] + qd[o - ed + em]);
      const double km0 = 0.25 * (qd[o] + qd[o + em]);
      acc -= a.h[d] * (kmm * G(m, -ed) - km0 * G(m, ed));
    }
    // q_m as the advecting velocity of every component c (direction m)
    double s = 0.;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const long ec = st[c];
      const double* qc = a.q[c];
      s += G(c, 0) * qc[o + em] - G(c, em) * qc[o];
      s += G(c, -ec) * qc[o - ec + em] - G(c, -ec + em) * qc[o - ec];
    }
    acc -= 0.25 * a.h[m] * s;
    a.out[m][o] = a.base[m][o] + acc;
    if constexpr (COMB) a.fbo[m][o] = a.f2 * a.g[m][o];
  }
}

static dim3 jt_grid(const Layout& L) { return dim3(cdiv(L.n1, 64), cdiv(L.n2, 4), L.n3); }

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
  k_mom_jt<false><<<jt_grid(c->L), dim3(64, 4), 0, c->stream>>>(a);
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
  k_mom_jt<true><<<jt_grid(c->L), dim3(64, 4), 0, c->stream>>>(a);
  CANS_LAUNCH_CHECK("k_mom_jt");
  return CANS_OK;
}

}  // namespace cans
