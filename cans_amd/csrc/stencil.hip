// Explicit-step kernels: momentum right-hand side + RK3 update, bulk forcing,
// divergence, projection, pressure update, boundary planes and reductions.
//
// Compiled with -ffp-contract=off: every expression keeps the reference's
// association and rounding (Fortran evaluates left to right, no fused
// multiply-add in the reference's -O3 CPU build), so these kernels reproduce the
// CPU arithmetic bit for bit on identical inputs.
//
// Velocity fields are always double.  Pressure-like arrays (p, pp, RHS_old and
// the RK increment of the mixed mode) have the storage type S: double, or float
// in the mixed-precision mode.  Values are widened to double on load and all
// arithmetic is double; only a store to S rounds.  With S = double the code is
// the plain FP64 path.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace cans {

// ---------------------------------------------------------------------------
// momentum RHS (mom.f90:771-1019, mom_xyz_ad) fused with the RK update
// (rk.f90:156-186) and the RHS_old copy (rk.f90:266-276).  Reads u,v,w at
// the old time level and writes the updated velocity to (uo,vo,wo), which must
// not alias (u,v,w).  Per-block partial sums of the updated velocity weighted
// by the grid-volume ratio feed the bulk forcing (utils.f90:15-42).
// ---------------------------------------------------------------------------
template <typename S>
struct MomArgs {
  Layout L;
  const double *u, *v, *w;
  const S* p;
  double *uo, *vo, *wo;          // updated velocity (delta == 0)
  S *eu, *ev, *ew;               // or its increment over (u,v,w) (delta == 1)
  S* divu;                       // nullptr, or div(u) at the old level, (1:n1,1:n2,1:n3)
  S *ru, *rv, *rw;               // dudtrko (1:n1,1:n2,1:n3), in: old RHS, out: new RHS
  const double *dzci, *dzfi;     // (0:n3+1)
  const double *rdzm, *rdzp;     // (0:n3+1): dzci(k)/dzfi(k), dzci(k)/dzfi(k+1)
  const double *gvr_c, *gvr_f;   // grid volume ratios (0:n3+1)
  double dxi, dyi, visc;
  double factor1, factor2, factor12;
  double bforce[3];
  int sum_mask;                  // bit c: accumulate the bulk mean of component c
  int read_ro, write_ro;
  int delta;
  double* partials;              // [3][nblocks]
  const double* sb;              // nullptr, or the buoyancy scalar (halo'd)
  double bq[3];                  // factor12*gacc(d)*beta*0.5
  int imp;                       // 0 explicit, 1 implicit diffusion, 2 implicit in z only
  double hf12;                   // 0.5*factor12
  // FOLD kernels: the previous substep's correc + bounduvw + updatep (MomFold)
  const S* fpp;
  const S *feu, *fev, *few;
  const double* ff;
  double fdt, fdtdxi, fdtdyi;
  S* pout;
  const double *gu, *gv, *gw;    // the projected velocity at the ghost cells (k_fold_ghosts)
};

// Tiling of the momentum kernel: a work-group owns a TX x TY column of cells and
// marches up KC planes in z.  Each plane of u,v,w (with a one-cell ring) is staged
// in LDS once per tile, two planes ahead of the arithmetic; three rotating plane
// buffers leave one barrier per plane; the z-neighbours at k-1 travel in registers.
constexpr int MOM_TX = 64, MOM_TY = 8, MOM_KC = 32;
constexpr int MOM_PX = MOM_TX + 2, MOM_PY = MOM_TY + 2;

// rdzm(k) = dzci(k)/dzfi(k), rdzp(k) = dzci(k)/dzfi(k+1) for k = 1..n3 (mom.f90:801-809)
__global__ void k_rdz(int n3, const double* dzci, const double* dzfi, double* tab) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > n3 + 1) return;
  const bool in = k >= 1 && k <= n3;
  tab[k] = in ? dzci[k] / dzfi[k] : 0.;
  tab[n3 + 2 + k] = in ? dzci[k] / dzfi[k + 1] : 0.;
}

// DELTA: the output is the increment over (u,v,w); DIVU: div(u) is written too; BUOY: the
// Boussinesq term of the scalar sb.  Compile-time
// choices, so that each form keeps only the pointers it uses in scalar registers (the
// arguments otherwise spill to vector-register lanes inside the plane loop).
// FOLD: the velocity this kernel advances is the previous substep's intermediate velocity
// projected (correc.f90:13-61) with that substep's pp, with the ghost values bounduvw
// (keep_norm, upper normal halo only, no tangential exchange; bound.f90:17-121) gives it
// after correc; the pressure is p + pp (updatep), which is also stored to pout.  The values
// are formed with the same operations in the same order as those routines, so the result is
// the one of the unfused sequence, bit for bit.  Each ghost's value is correc's at a source
// cell, then the wall rules of the directions that have one, in x, y, z order:
//  * periodic: the ghosts that bounduvw copies (along x every tangential ghost and the normal
//    component's upper one; along y and z the normal component's upper one) take their
//    periodic image; every other ghost keeps correc's own value there;
//  * at a wall a tangential component's ghost is factor + sgn * (the adjacent value);
//  * the normal component keeps correc's value at the ghost (keep_norm): the formula at
//    indices 0..n, the intermediate velocity itself at n+1.
// With DELTA the increments written are over the kernel's u, v, w (the step's initial
// velocity), not over the projected one.
template <typename S, bool DELTA, bool DIVU, bool BUOY, bool FOLD = false>
__global__ void __launch_bounds__(MOM_TX* MOM_TY, 4) k_mom_rk(MomArgs<S> a) {
  __shared__ double su[3][MOM_PY][MOM_PX], sv[3][MOM_PY][MOM_PX], sw[3][MOM_PY][MOM_PX];
  // FOLD: p + pp of the staged planes, so that the plane loop reads no pressure from memory
  __shared__ double sp[FOLD ? 3 : 1][FOLD ? MOM_PY : 1][FOLD ? MOM_PX : 1];
  const Layout& L = a.L;
  const int ntx = (L.n1 + MOM_TX - 1) / MOM_TX, nty = (L.n2 + MOM_TY - 1) / MOM_TY;
  const int nkc = (L.n3 + MOM_KC - 1) / MOM_KC;
  const int nb = ntx * nty * nkc;
  const int tile = xcd_tile(blockIdx.x, nb);
  const int tx = tile % ntx, ty = (tile / ntx) % nty, tk = tile / (ntx * nty);
  const int i0 = tx * MOM_TX, j0 = ty * MOM_TY;  // cell (i0+1, j0+1) is the first of the tile
  const int kb = tk * MOM_KC + 1, ke = min(kb + MOM_KC - 1, L.n3);
  const int li = threadIdx.x, lj = threadIdx.y;
  const int t = li + MOM_TX * lj;
  const int i = i0 + li + 1, j = j0 + lj + 1;
  const bool in = i <= L.n1 && j <= L.n2;
  const long sy = L.sy, sz = L.sz;
  // FOLD: the projected velocity.  At an interior cell it is correc's value there (k_correc's
  // expression); ghost cells were formed by k_fold_ghosts (the bounduvw rules) into gu, gv, gw.
  auto fval = [&](int cm, int gi, int gj, int gk, long o) -> double {
    if (gi < 1 || gi > L.n1 || gj < 1 || gj > L.n2 || gk < 1 || gk > L.n3)
      return cm == 0 ? a.gu[o] : cm == 1 ? a.gv[o] : a.gw[o];
    const double* b = cm == 0 ? a.u : cm == 1 ? a.v : a.w;
    const S* e = cm == 0 ? a.feu : cm == 1 ? a.fev : a.few;
    const double fc = a.ff[cm];
    const double tu = e ? b[o] + (double)e[o] + fc : b[o] + fc;
    const double pc = (double)a.fpp[o];
    if (cm == 0) return tu - a.fdtdxi * ((double)a.fpp[o + 1] - pc);
    if (cm == 1) return tu - a.fdtdyi * ((double)a.fpp[o + sy] - pc);
    return tu - a.fdt * a.dzci[gk] * ((double)a.fpp[o + sz] - pc);
  };
  // FOLD: the three components at one cell (the pressure correction loaded once for all)
  auto fval3 = [&](int gi, int gj, int gk, long o, double& xu, double& xv, double& xw) {
    if (gi < 1 || gi > L.n1 || gj < 1 || gj > L.n2 || gk < 1 || gk > L.n3) {
      xu = a.gu[o];
      xv = a.gv[o];
      xw = a.gw[o];
      return;
    }
    const double bu = a.u[o], bv = a.v[o], bw = a.w[o];
    const double pc = (double)a.fpp[o], px = (double)a.fpp[o + 1], py = (double)a.fpp[o + sy],
                 pz = (double)a.fpp[o + sz];
    const double dz = a.dzci[gk];
    double tu, tv, tw;
    if (a.feu) {
      tu = bu + (double)a.feu[o] + a.ff[0];
      tv = bv + (double)a.fev[o] + a.ff[1];
      tw = bw + (double)a.few[o] + a.ff[2];
    } else {
      tu = bu + a.ff[0];
      tv = bv + a.ff[1];
      tw = bw + a.ff[2];
    }
    xu = tu - a.fdtdxi * (px - pc);
    xv = tv - a.fdtdyi * (py - pc);
    xw = tw - a.fdt * dz * (pz - pc);
  };
  // the pressure the kernel reads: p, or (FOLD) p + pp rounded to the storage type
  auto pval = [&](long o) -> double {
    if constexpr (FOLD) return (double)(S)((double)a.p[o] + (double)a.fpp[o]);
    else return (double)a.p[o];
  };
  auto uval = [&](int cm, int gi, int gj, int gk, long o) -> double {
    if constexpr (FOLD) return fval(cm, gi, gj, gk, o);
    else return cm == 0 ? a.u[o] : cm == 1 ? a.v[o] : a.w[o];
  };
  // stage plane kk of the tile (+ ring) into buffer slot
  auto stage = [&](int kk, int slot) {
    for (int q = t; q < MOM_PX * MOM_PY; q += MOM_TX * MOM_TY) {
      const int pi = q % MOM_PX, pj = q / MOM_PX;
      const int gi = i0 + pi, gj = j0 + pj;  // halo'd indices (i0+pi is cell i0+pi)
      if (gi <= L.n1 + 1 && gj <= L.n2 + 1) {
        const long o = L.at(gi, gj, kk);
        if constexpr (FOLD) {
          double xu, xv, xw;
          fval3(gi, gj, kk, o, xu, xv, xw);
          su[slot][pj][pi] = xu;
          sv[slot][pj][pi] = xv;
          sw[slot][pj][pi] = xw;
          sp[slot][pj][pi] = pval(o);
        } else {
          su[slot][pj][pi] = a.u[o];
          sv[slot][pj][pi] = a.v[o];
          sw[slot][pj][pi] = a.w[o];
        }
      }
    }
  };
  const double dxi = a.dxi, dyi = a.dyi, visc = a.visc;
  // z-neighbours below the first plane
  double u_ccm = 0., v_ccm = 0., w_ccm = 0., w_pcm = 0., w_cpm = 0., p_c = 0.;
  if (in) {
    const long c0 = L.at(i, j, kb);
    u_ccm = uval(0, i, j, kb - 1, c0 - sz);
    v_ccm = uval(1, i, j, kb - 1, c0 - sz);
    w_ccm = uval(2, i, j, kb - 1, c0 - sz);
    w_pcm = uval(2, i + 1, j, kb - 1, c0 - sz + 1);
    w_cpm = uval(2, i, j + 1, kb - 1, c0 - sz + sy);
    if constexpr (!FOLD) p_c = pval(c0);
  }
  // plane k+2 is loaded into registers before plane k's arithmetic and stored to its slot
  // after it (the slot plane k-1 used, free since the previous barrier), so the loads'
  // latency overlaps that work; one barrier per plane
  constexpr int NQ = (MOM_PX * MOM_PY + MOM_TX * MOM_TY - 1) / (MOM_TX * MOM_TY);  // ring cells per lane
  double nu[NQ], nv[NQ], nw[NQ], np[FOLD ? NQ : 1];
  auto load = [&](int kk) {
#pragma unroll
    for (int r = 0; r < NQ; r++) {
      const int q = t + r * MOM_TX * MOM_TY;
      const int gi = i0 + q % MOM_PX, gj = j0 + q / MOM_PX;
      if (q < MOM_PX * MOM_PY && gi <= L.n1 + 1 && gj <= L.n2 + 1) {
        const long o = L.at(gi, gj, kk);
        if constexpr (FOLD) {
          fval3(gi, gj, kk, o, nu[r], nv[r], nw[r]);
          np[r] = pval(o);
        } else {
          nu[r] = a.u[o];
          nv[r] = a.v[o];
          nw[r] = a.w[o];
        }
      }
    }
  };
  auto store = [&](int slot) {
#pragma unroll
    for (int r = 0; r < NQ; r++) {
      const int q = t + r * MOM_TX * MOM_TY;
      const int pi = q % MOM_PX, pj = q / MOM_PX;
      if (q < MOM_PX * MOM_PY && i0 + pi <= L.n1 + 1 && j0 + pj <= L.n2 + 1) {
        su[slot][pj][pi] = nu[r];
        sv[slot][pj][pi] = nv[r];
        sw[slot][pj][pi] = nw[r];
        if constexpr (FOLD) sp[slot][pj][pi] = np[r];
      }
    }
  };
  // the z-grid values of the tile's planes, read from LDS into vector registers in the plane
  // loop (as uniform scalar loads they would take ~16 scalar registers, which this kernel's
  // arguments already fill)
  __shared__ double zg[MOM_KC][8];
  for (int q = t; q < (ke - kb + 1) * 8; q += MOM_TX * MOM_TY) {
    const int kk = kb + q / 8;
    double v;
    switch (q % 8) {
      case 0: v = a.dzci[kk]; break;
      case 1: v = a.dzci[kk - 1]; break;
      case 2: v = a.dzfi[kk]; break;
      case 3: v = a.dzfi[kk + 1]; break;
      case 4: v = a.rdzm[kk]; break;
      case 5: v = a.rdzp[kk]; break;
      case 6: v = a.gvr_f[kk]; break;
      default: v = a.gvr_c[kk]; break;
    }
    zg[q / 8][q % 8] = v;
  }
  stage(kb, kb % 3);
  stage(kb + 1, (kb + 1) % 3);  // kb <= n3, so kb + 1 <= n3 + 1
  __syncthreads();
  double s0 = 0., s1 = 0., s2 = 0.;
  for (int k = kb; k <= ke; k++) {
    const bool next = k + 2 <= ke + 1;
    if (next) load(k + 2);
    const int b0 = k % 3, b1 = (k + 1) % 3;
    const int pi = li + 1, pj = lj + 1;
    if (in) {
      const long c = L.at(i, j, k);
      // plane k
      const double u_ccc = su[b0][pj][pi], u_pcc = su[b0][pj][pi + 1], u_mcc = su[b0][pj][pi - 1],
                   u_cpc = su[b0][pj + 1][pi], u_cmc = su[b0][pj - 1][pi], u_mpc = su[b0][pj + 1][pi - 1];
      const double v_ccc = sv[b0][pj][pi], v_pcc = sv[b0][pj][pi + 1], v_mcc = sv[b0][pj][pi - 1],
                   v_cpc = sv[b0][pj + 1][pi], v_cmc = sv[b0][pj - 1][pi], v_pmc = sv[b0][pj - 1][pi + 1];
      const double w_ccc = sw[b0][pj][pi], w_pcc = sw[b0][pj][pi + 1], w_mcc = sw[b0][pj][pi - 1],
                   w_cpc = sw[b0][pj + 1][pi], w_cmc = sw[b0][pj - 1][pi];
      // plane k+1
      const double u_ccp = su[b1][pj][pi], u_mcp = su[b1][pj][pi - 1];
      const double v_ccp = sv[b1][pj][pi], v_cmp = sv[b1][pj - 1][pi];
      const double w_ccp = sw[b1][pj][pi];
      const double* zk = zg[k - kb];
      const double dzci_k = zk[0], dzci_km = zk[1], dzfi_k = zk[2], dzfi_kp = zk[3];
      const double rdzm = zk[4], rdzp = zk[5];
      // x
      const double dudxp = (u_pcc - u_ccc) * dxi, dudxm = (u_ccc - u_mcc) * dxi;
      const double dudyp = (u_cpc - u_ccc) * dyi, dudym = (u_ccc - u_cmc) * dyi;
      const double dudzp = (u_ccp - u_ccc) * dzci_k, dudzm = (u_ccc - u_ccm) * dzci_km;
      const double uuip = 0.25 * (u_ccc + u_pcc) * u_pcc, uuim = 0.25 * (u_ccc + u_mcc) * u_mcc;
      const double vujp = 0.25 * (v_ccc + v_pcc) * u_cpc, vujm = 0.25 * (v_cmc + v_pmc) * u_cmc;
      const double wukp = 0.25 * (w_ccc + w_pcc) * u_ccp, wukm = 0.25 * (w_ccm + w_pcm) * u_ccm;
      double dudt = -(uuip - uuim) * dxi - (vujp - vujm) * dyi - (wukp - wukm) * dzfi_k;
      // diffusion (mom.f90:990-1016): explicit, or split off for the implicit solve
      const double dudxy = visc * (dudxp - dudxm) * dxi + visc * (dudyp - dudym) * dyi;
      const double dudz = visc * (dudzp - dudzm) * dzfi_k;
      double dudtd = 0.;
      if (a.imp == 0) dudt = dudt + dudxy + dudz;
      else if (a.imp == 2) dudt = dudt + dudxy, dudtd = dudz;
      else dudtd = dudxy + dudz;
      // y
      const double dvdxp = (v_pcc - v_ccc) * dxi, dvdxm = (v_ccc - v_mcc) * dxi;
      const double dvdyp = (v_cpc - v_ccc) * dyi, dvdym = (v_ccc - v_cmc) * dyi;
      const double dvdzp = (v_ccp - v_ccc) * dzci_k, dvdzm = (v_ccc - v_ccm) * dzci_km;
      const double uvip = 0.25 * (u_ccc + u_cpc) * v_pcc, uvim = 0.25 * (u_mcc + u_mpc) * v_mcc;
      const double vvjp = 0.25 * (v_ccc + v_cpc) * v_cpc, vvjm = 0.25 * (v_ccc + v_cmc) * v_cmc;
      const double wvkp = 0.25 * (w_ccc + w_cpc) * v_ccp, wvkm = 0.25 * (w_ccm + w_cpm) * v_ccm;
      double dvdt = -(uvip - uvim) * dxi - (vvjp - vvjm) * dyi - (wvkp - wvkm) * dzfi_k;
      const double dvdxy = visc * (dvdxp - dvdxm) * dxi + visc * (dvdyp - dvdym) * dyi;
      const double dvdz = visc * (dvdzp - dvdzm) * dzfi_k;
      double dvdtd = 0.;
      if (a.imp == 0) dvdt = dvdt + dvdxy + dvdz;
      else if (a.imp == 2) dvdt = dvdt + dvdxy, dvdtd = dvdz;
      else dvdtd = dvdxy + dvdz;
      // z
      const double dwdxp = (w_pcc - w_ccc) * dxi, dwdxm = (w_ccc - w_mcc) * dxi;
      const double dwdyp = (w_cpc - w_ccc) * dyi, dwdym = (w_ccc - w_cmc) * dyi;
      const double dwdzp = (w_ccp - w_ccc) * dzfi_kp, dwdzm = (w_ccc - w_ccm) * dzfi_k;
      const double uwip = 0.25 * (rdzm * u_ccc + rdzp * u_ccp) * w_pcc;
      const double uwim = 0.25 * (rdzm * u_mcc + rdzp * u_mcp) * w_mcc;
      const double vwjp = 0.25 * (rdzm * v_ccc + rdzp * v_ccp) * w_cpc;
      const double vwjm = 0.25 * (rdzm * v_cmc + rdzp * v_cmp) * w_cmc;
      const double wwkp = 0.25 * (w_ccc + w_ccp) * w_ccp, wwkm = 0.25 * (w_ccc + w_ccm) * w_ccm;
      double dwdt = -(uwip - uwim) * dxi - (vwjp - vwjm) * dyi - (wwkp - wwkm) * dzci_k;
      const double dwdxy = visc * (dwdxp - dwdxm) * dxi + visc * (dwdyp - dwdym) * dyi;
      const double dwdz = visc * (dwdzp - dwdzm) * dzci_k;
      double dwdtd = 0.;
      if (a.imp == 0) dwdt = dwdt + dwdxy + dwdz;
      else if (a.imp == 2) dwdt = dwdt + dwdxy, dwdtd = dwdz;
      else dwdtd = dwdxy + dwdz;
      // RK update (rk.f90:156-186) and RHS_old <- RHS (rk.f90:266-276).  With rkpar(2) = 0
      // (first substep) the RHS_old term is factor2*RHS_old = 0 and is not read.
      const long o = (long)(i - 1) + (long)L.n1 * ((j - 1) + (long)L.n2 * (k - 1));
      double p_ip, p_jp, p_kp;
      if constexpr (FOLD) {
        p_c = sp[b0][pj][pi];
        p_ip = sp[b0][pj][pi + 1];
        p_jp = sp[b0][pj + 1][pi];
        p_kp = sp[b1][pj][pi];
      } else {
        p_ip = pval(c + 1);
        p_jp = pval(c + sy);
        p_kp = pval(c + sz);
      }
      const double ru = a.read_ro ? (double)a.ru[o] : 0., rv = a.read_ro ? (double)a.rv[o] : 0.,
                   rw = a.read_ro ? (double)a.rw[o] : 0.;
      double un = a.read_ro ? u_ccc + a.factor1 * dudt + a.factor2 * ru +
                                  a.factor12 * (a.bforce[0] - dxi * (p_ip - p_c))
                            : u_ccc + a.factor1 * dudt + a.factor12 * (a.bforce[0] - dxi * (p_ip - p_c));
      double vn = a.read_ro ? v_ccc + a.factor1 * dvdt + a.factor2 * rv +
                                  a.factor12 * (a.bforce[1] - dyi * (p_jp - p_c))
                            : v_ccc + a.factor1 * dvdt + a.factor12 * (a.bforce[1] - dyi * (p_jp - p_c));
      double wn = a.read_ro ? w_ccc + a.factor1 * dwdt + a.factor2 * rw +
                                  a.factor12 * (a.bforce[2] - dzci_k * (p_kp - p_c))
                            : w_ccc + a.factor1 * dwdt + a.factor12 * (a.bforce[2] - dzci_k * (p_kp - p_c));
      if constexpr (BUOY) {  // Boussinesq buoyancy of the first scalar (rk.f90:163-177)
        const double s_c = a.sb[c];
        un = un - a.bq[0] * (a.sb[c + 1] + s_c);
        vn = vn - a.bq[1] * (a.sb[c + sy] + s_c);
        wn = wn - a.bq[2] * (a.sb[c + sz] + s_c);
      }
      // implicit diffusion (rk.f90:180-184, 312-326): the update takes + factor12*dudtd; the
      // bulk mean is of that velocity; the stored value is the Helmholtz rhs u - 0.5*factor12*dudtd
      double uout = un, vout = vn, wout = wn;
      if (a.imp) {
        un = un + a.factor12 * dudtd;
        vn = vn + a.factor12 * dvdtd;
        wn = wn + a.factor12 * dwdtd;
        uout = un - a.hf12 * dudtd;
        vout = vn - a.hf12 * dvdtd;
        wout = wn - a.hf12 * dwdtd;
      }
      if constexpr (DELTA && FOLD) {  // increments over the step's initial velocity
        a.eu[c] = (S)(uout - a.u[c]);
        a.ev[c] = (S)(vout - a.v[c]);
        a.ew[c] = (S)(wout - a.w[c]);
      } else if constexpr (DELTA) {
        a.eu[c] = (S)(uout - u_ccc);
        a.ev[c] = (S)(vout - v_ccc);
        a.ew[c] = (S)(wout - w_ccc);
      } else {
        a.uo[c] = uout;
        a.vo[c] = vout;
        a.wo[c] = wout;
      }
      if (a.write_ro) {
        a.ru[o] = (S)dudt;
        a.rv[o] = (S)dvdt;
        a.rw[o] = (S)dwdt;
      }
      if constexpr (FOLD) a.pout[c] = (S)p_c;  // updatep: p + pp
      if constexpr (DIVU)  // fillps.f90:38-50 association, without the 1/dt factor
        a.divu[o] = (S)((w_ccc - w_ccm) * dzfi_k + (v_ccc - v_cmc) * dyi + (u_ccc - u_mcc) * dxi);
      if (a.sum_mask) {
        s0 += un * zk[6];
        s1 += vn * zk[6];
        s2 += wn * zk[7];
      }
      // march: plane k becomes plane k-1
      u_ccm = u_ccc;
      v_ccm = v_ccc;
      w_ccm = w_ccc;
      w_pcm = w_pcc;
      w_cpm = w_cpc;
      p_c = p_kp;
    }
    if (next) store((k + 2) % 3);
    __syncthreads();
  }
  if (a.sum_mask) {
    __syncthreads();
    block_sum3<MOM_TX * MOM_TY>(s0, s1, s2, a.partials + blockIdx.x, a.partials + nb + blockIdx.x,
                                a.partials + 2 * nb + blockIdx.x);
  }
}

// The projected velocity of a folded substep (MomFold) at the ghost cells, the six face planes
// of the halo'd box, with the rules k_mom_rk<FOLD> states: each ghost's value is correc's at a
// source cell (k_correc's expression, including its copy branch at index n+1), then the wall
// rules of the directions that have one, in x, y, z order (bounduvw applies them direction by
// direction).  Periodic: the ghosts bounduvw copies after correc (along x every tangential
// ghost and the normal component's upper one, is_bound on one GPU; along y and z the normal
// component's upper one, the only exchanged halo) take their periodic image, every other
// ghost keeps correc's own value; walls: a tangential ghost is factor + sgn * (the adjacent
// value), the normal component keeps correc's value at the ghost (keep_norm).
template <typename S>
struct FoldGhostArgs {
  Layout L;
  const double *u, *v, *w;
  const S *eu, *ev, *ew, *pp;
  const double *f, *dzci;
  double dt, dtdxi, dtdyi;
  int per;  // bit d: direction d periodic
  double factor[3][3][2], sgn[3][3][2];
  double *gu, *gv, *gw;
};

template <typename S>
__global__ void __launch_bounds__(256) k_fold_ghosts(FoldGhostArgs<S> a) {
  const Layout& L = a.L;
  const int face = blockIdx.y, dir = face >> 1, side = face & 1;
  const int e1 = dir == 0 ? L.n2 + 2 : L.n1 + 2;
  const int e2 = dir == 2 ? L.n2 + 2 : L.n3 + 2;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)e1 * e2) return;
  const int pa = (int)(idx % e1), pb = (int)(idx / e1);
  int gi, gj, gk;
  if (dir == 0) gi = side ? L.n1 + 1 : 0, gj = pa, gk = pb;
  else if (dir == 1) gi = pa, gj = side ? L.n2 + 1 : 0, gk = pb;
  else gi = pa, gj = pb, gk = side ? L.n3 + 1 : 0;
  const long go = L.at(gi, gj, gk);
  for (int cm = 0; cm < 3; cm++) {
    int si = gi, sj = gj, sk = gk;
    bool ox = false, oy = false, oz = false;
    int xs = 0, ys = 0, zs = 0;
    if (gi < 1 || gi > L.n1) {
      xs = gi > L.n1;
      if (a.per & 1) {
        if (cm != 0 || xs) si = xs ? 1 : L.n1;
      } else if (cm != 0) {
        si = xs ? L.n1 : 1, ox = true;
      }
    }
    if (gj < 1 || gj > L.n2) {
      ys = gj > L.n2;
      if (a.per & 2) {
        if (cm == 1 && ys) sj = 1;
      } else if (cm != 1) {
        sj = ys ? L.n2 : 1, oy = true;
      }
    }
    if (gk < 1 || gk > L.n3) {
      zs = gk > L.n3;
      if (a.per & 4) {
        if (cm == 2 && zs) sk = 1;
      } else if (cm != 2) {
        sk = zs ? L.n3 : 1, oz = true;
      }
    }
    const long o = L.at(si, sj, sk);
    const double* b = cm == 0 ? a.u : cm == 1 ? a.v : a.w;
    const S* e = cm == 0 ? a.eu : cm == 1 ? a.ev : a.ew;
    const double fc = a.f[cm];
    const double tu = e ? b[o] + (double)e[o] + fc : b[o] + fc;
    const double pc = (double)a.pp[o];
    double x;
    if (cm == 0) x = si <= L.n1 ? tu - a.dtdxi * ((double)a.pp[o + 1] - pc) : tu;
    else if (cm == 1) x = sj <= L.n2 ? tu - a.dtdyi * ((double)a.pp[o + L.sy] - pc) : tu;
    else x = sk <= L.n3 ? tu - a.dt * a.dzci[sk] * ((double)a.pp[o + L.sz] - pc) : tu;
    if (ox) x = a.factor[cm][0][xs] + a.sgn[cm][0][xs] * x;
    if (oy) x = a.factor[cm][1][ys] + a.sgn[cm][1][ys] * x;
    if (oz) x = a.factor[cm][2][zs] + a.sgn[cm][2][zs] * x;
    (cm == 0 ? a.gu : cm == 1 ? a.gv : a.gw)[go] = x;
  }
}

__global__ void k_fill(double* p, int n, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Deterministic two-stage reduction of per-block partials (fixed summation order).
// stage A: block g reduces partials[c*nb + [g*chunk, (g+1)*chunk)) -> stage[c*G + g]
template <int OP>  // 0 sum, 1 max
__global__ void k_reduce_a(const double* partials, int nb, int chunk, double* stage) {
  __shared__ double red[256];
  const int c = blockIdx.y, g = blockIdx.x, G = gridDim.x;
  const int b0 = g * chunk, b1 = min(nb, b0 + chunk);
  double s = 0.;
  for (int b = b0 + (int)threadIdx.x; b < b1; b += blockDim.x) {
    const double x = partials[(long)c * nb + b];
    s = OP ? fmax(s, x) : s + x;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] = OP ? fmax(red[threadIdx.x], red[threadIdx.x + st])
                                                     : red[threadIdx.x] + red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) stage[c * G + g] = red[0];
}

// stage B: one block per component reduces the G stage values.
// mode 0: out[c] = result; mode 1 (bulk forcing, rk.f90:493-518): out[c] = bit c of mask ? velf[c] - mean : 0
template <int OP>
__global__ void k_reduce_b(const double* stage, int G, double* out, int mode, int mask, double v0, double v1,
                           double v2) {
  __shared__ double red[256];
  const int c = blockIdx.x;
  double s = 0.;
  for (int g = threadIdx.x; g < G; g += blockDim.x) s = OP ? fmax(s, stage[c * G + g]) : s + stage[c * G + g];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] = OP ? fmax(red[threadIdx.x], red[threadIdx.x + st])
                                                     : red[threadIdx.x] + red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (mode == 0) out[c] = red[0];
    else {
      const double velf = c == 0 ? v0 : c == 1 ? v1 : v2;
      out[c] = (mask >> c & 1) ? velf - red[0] : 0.;
    }
  }
}

// partials (ncomp x nb, component-major) -> out[0..ncomp); stage scratch follows the partials
// `base` holds ncomp x nb partials inside c->partials; `stage` is scratch for ncomp*1024 values
static int reduce_partials(cans_ctx* c, const double* base, int nb, int ncomp, int op, double* out, double* stage,
                           int mode = 0, int mask = 0, const double* velf = nullptr) {
  const int G = (int)std::min<long>(1024, std::max<long>(1, (nb + 1023) / 1024));
  const int chunk = (nb + G - 1) / G;
  dim3 ga(G, ncomp);
  if (op) k_reduce_a<1><<<ga, 256, 0, c->stream>>>(base, nb, chunk, stage);
  else k_reduce_a<0><<<ga, 256, 0, c->stream>>>(base, nb, chunk, stage);
  CANS_LAUNCH_CHECK("k_reduce_a");
  const double v0 = velf ? velf[0] : 0., v1 = velf ? velf[1] : 0., v2 = velf ? velf[2] : 0.;
  if (op) k_reduce_b<1><<<ncomp, 256, 0, c->stream>>>(stage, G, out, mode, mask, v0, v1, v2);
  else k_reduce_b<0><<<ncomp, 256, 0, c->stream>>>(stage, G, out, mode, mask, v0, v1, v2);
  CANS_LAUNCH_CHECK("k_reduce_b");
  return CANS_OK;
}

static int ensure_partials(cans_ctx* c, size_t n) {
  n += 4096;  // room for the reduction stages (3 x 1024)
  if (n <= c->partials_n) return CANS_OK;
  CANS_HIP_TRY(hipStreamSynchronize(c->stream));
  hipFree(c->partials);
  c->partials_n = n;
  CANS_HIP_TRY(hipMalloc((void**)&c->partials, n * sizeof(double)));
  return CANS_OK;
}

// ---------------------------------------------------------------------------
// Scalar transport, explicit diffusion (scal.f90:41-98, the default loop), fused
// with the RK update of rk_scal (rk.f90:427-441) and RHS_old <- RHS (rk.f90:481-489):
//   so = s + factor1*dsdt + factor2*dsdtrko + factor12*ssource,  dsdtrko <- dsdt.
// With a forcing, block partials of so*grid_vol_ratio_f feed the bulk mean
// (rk.f90:471-474).  One thread per cell of a row; neighbours come through the caches.
// ---------------------------------------------------------------------------
struct ScalArgs {
  Layout L;
  const double *u, *v, *w, *s;
  double* so;                 // updated scalar (interior written)
  double* ro;                 // dsdtrko (1:n1,1:n2,1:n3)
  const double *dzci, *dzfi, *gvr_f;
  double dxi, dyi, alpha, factor1, factor2, factor12, source;
  int read_ro;
  double* partials;           // nullptr: no bulk mean
  int imp;                    // 0 explicit, 1 implicit diffusion, 2 implicit in z only
  double hf12;                // 0.5*factor12
};

__global__ void __launch_bounds__(256) k_scal_rk(ScalArgs a) {
  const Layout& L = a.L;
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1, k = blockIdx.z + 1;
  double acc = 0.;
  if (i <= L.n1) {
    const long c = L.at(i, j, k), sy = L.sy, sz = L.sz;
    const double* s = a.s;
    const double sc = s[c], sxm = s[c - 1], sxp = s[c + 1], sym = s[c - sy], syp = s[c + sy];
    const double szm = s[c - sz], szp = s[c + sz];
    const double dxi = a.dxi, dyi = a.dyi, alpha = a.alpha, dzfi_k = a.dzfi[k];
    const double usim = 0.5 * (sxm + sc) * a.u[c - 1];
    const double usip = 0.5 * (sxp + sc) * a.u[c];
    const double vsjm = 0.5 * (sym + sc) * a.v[c - sy];
    const double vsjp = 0.5 * (syp + sc) * a.v[c];
    const double wskm = 0.5 * (szm + sc) * a.w[c - sz];
    const double wskp = 0.5 * (szp + sc) * a.w[c];
    const double dsdxp = (sxp - sc) * dxi, dsdxm = (sc - sxm) * dxi;
    const double dsdyp = (syp - sc) * dyi, dsdym = (sc - sym) * dyi;
    const double dsdzp = (szp - sc) * a.dzci[k], dsdzm = (sc - szm) * a.dzci[k - 1];
    double dsdt = dxi * (-usip + usim) + dyi * (-vsjp + vsjm) + dzfi_k * (-wskp + wskm);
    const double dxy = (dsdxp - dsdxm) * alpha * dxi + (dsdyp - dsdym) * alpha * dyi;
    const double dz = (dsdzp - dsdzm) * alpha * dzfi_k;
    double dsdtd = 0.;  // scal.f90:86-97: the implicit part, if any
    if (a.imp == 0) dsdt = dsdt + dxy + dz;
    else if (a.imp == 2) dsdt = dsdt + dxy, dsdtd = dz;
    else dsdtd = dxy + dz;
    const long o = (long)(i - 1) + (long)L.n1 * ((j - 1) + (long)L.n2 * (k - 1));
    double sn = a.read_ro ? sc + a.factor1 * dsdt + a.factor2 * a.ro[o] + a.factor12 * a.source
                          : sc + a.factor1 * dsdt + a.factor12 * a.source;
    if (a.imp) sn = sn + a.factor12 * dsdtd;  // rk.f90:437-440
    // the bulk mean is of sn; with implicit diffusion the stored value is the Helmholtz rhs (rk.f90:476-486)
    a.so[c] = a.imp ? sn - a.hf12 * dsdtd : sn;
    a.ro[o] = dsdt;
    acc = sn * a.gvr_f[k];
  }
  if (a.partials) {
    const long b = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
    double d0 = 0., d1 = 0.;
    block_sum3<256>(acc, d0, d1, a.partials + b, nullptr, nullptr, true);
  }
}

// u(1:n) += f(c) (mom.f90:725-769); src may differ from dst (copy + forcing)
template <typename T>
__global__ void k_add_forcing(Layout L, const T* src, T* dst, const double* f, int comp, int apply) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1, k = blockIdx.z + 1;
  if (i > L.n1) return;
  const long c = L.at(i, j, k);
  dst[c] = apply ? (T)((double)src[c] + f[comp]) : src[c];
}

// velocity value u[o] + fc, or u[o] + e[o] + fc when the field is held as u plus an increment e
// (fc: the bulk forcing of rk.f90:307-311 not yet added to the stored field, or 0)
template <typename S>
__device__ __forceinline__ double vel(const double* u, const S* e, double fc, long o) {
  return e ? u[o] + (double)e[o] + fc : u[o] + fc;
}

// fillps.f90:13-51 (velocity u,v,w, optionally plus increments eu,ev,ew and forcing f)
template <typename S>
__global__ void k_fillps(Layout L, const double* u, const double* v, const double* w, const S* eu, const S* ev,
                         const S* ew, const double* f, S* p, const double* dzfi, double dti, double dtidxi,
                         double dtidyi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1, k = blockIdx.z + 1;
  if (i > L.n1) return;
  const long c = L.at(i, j, k);
  const double fu = f ? f[0] : 0., fv = f ? f[1] : 0., fw = f ? f[2] : 0.;
  const double wc = vel(w, ew, fw, c), wm = vel(w, ew, fw, c - L.sz), vc = vel(v, ev, fv, c);
  const double vm = vel(v, ev, fv, c - L.sy), uc = vel(u, eu, fu, c), um = vel(u, eu, fu, c - 1);
  p[c] = (S)((wc - wm) * dti * dzfi[k] + (vc - vm) * dtidyi + (uc - um) * dtidxi);
}

// correc.f90:13-61 over the whole halo'd box; (su,sv,sw) -> (du,dv,dw), which may alias.
// With pu != nullptr the pressure update p += pp of updatep.f90:78-86 rides along
// (it needs only pp and p, both final at this point of the substep).
// The source velocity is (su,sv,sw), plus increments (eu,ev,ew) when those are given;
// the destination may alias the source (each element is read and written by one thread).
// A thread owns one (i,j) column of the halo'd box and marches CORREC_KC planes up in
// z, so p(k+1), loaded for the z-gradient, is reused as p(k) of the next plane; the
// x and y neighbours of p come from adjacent lanes and rows of the same block.  The
// loads of plane k+1 are issued before the stores of plane k (two planes in flight).
// Lane tx of x-block bx takes i = 1 + 64 bx + tx, so a wave's row segment starts at
// interior element 1 (256-byte aligned in fields of cans_field_alloc); the one index
// past the box, n1 + 2, wraps to the halo cell i = 0.
constexpr int CORREC_TX = 64, CORREC_TY = 4, CORREC_KC = 32;

template <typename S>
struct CorrecIn {
  double tu, tv, tw, px, py, pz, pu;
};

// NP pressure fields p, p1, p2 (NP > 1: the corrections of a refined solve), read as their FP64 sum
template <typename S, int NP>
__global__ void __launch_bounds__(CORREC_TX* CORREC_TY)
    k_correc(Layout L, const S* p0, const double* su, const double* sv, const double* sw, const S* eu, const S* ev,
             const S* ew, const double* f, double* du, double* dv, double* dw, const double* dzci, double dt,
             double factori, double factorj, S* pu, const S* p1, const S* p2) {
  struct PSum {
    const S *a, *b, *c;
    __device__ __forceinline__ double operator[](long o) const {
      double v = (double)a[o];
      if constexpr (NP > 1) v += (double)b[o];
      if constexpr (NP > 2) v += (double)c[o];
      return v;
    }
  };
  const PSum p = {p0, p1, p2};
  int i = 1 + blockIdx.x * CORREC_TX + threadIdx.x;
  if (i == L.n1 + 2) i = 0;
  const int j = blockIdx.y * CORREC_TY + threadIdx.y;
  if (i > L.n1 + 1 || j > L.n2 + 1) return;
  const int kb = blockIdx.z * CORREC_KC, ke = min(kb + CORREC_KC - 1, L.n3 + 1);
  const double fu = f ? f[0] : 0., fv = f ? f[1] : 0., fw = f ? f[2] : 0.;
  const bool in_ij = i >= 1 && i <= L.n1 && j >= 1 && j <= L.n2;
  const bool xin = i <= L.n1, yin = j <= L.n2;
  auto fetch = [&](long c, int k) {
    CorrecIn<S> q;
    q.tu = vel(su, eu, fu, c);
    q.tv = vel(sv, ev, fv, c);
    q.tw = vel(sw, ew, fw, c);
    q.px = xin ? p[c + 1] : 0.;
    q.py = yin ? p[c + L.sy] : 0.;
    q.pz = (k <= L.n3) ? p[c + L.sz] : 0.;
    q.pu = (pu && in_ij && k >= 1 && k <= L.n3) ? (double)pu[c] : 0.;
    return q;
  };
  long c = L.at(i, j, kb);
  double pc = p[c];
  CorrecIn<S> cur = fetch(c, kb);
  for (int k = kb; k <= ke; k++, c += L.sz) {
    CorrecIn<S> nxt = cur;
    if (k < ke) nxt = fetch(c + L.sz, k + 1);
    du[c] = xin ? cur.tu - factori * (cur.px - pc) : cur.tu;
    dv[c] = yin ? cur.tv - factorj * (cur.py - pc) : cur.tv;
    dw[c] = (k <= L.n3) ? cur.tw - dt * dzci[k] * (cur.pz - pc) : cur.tw;
    if (pu && in_ij && k >= 1 && k <= L.n3) pu[c] = (S)(cur.pu + pc);
    pc = cur.pz;
    cur = nxt;
  }
}

// updatep.f90:14-88
template <typename T>
__global__ void k_updatep(Layout L, const T* pp, T* p, const double* dzci, const double* dzfi, double alpha,
                          double dxi2, double dyi2, int impdiff) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1, k = blockIdx.z + 1;
  if (i > L.n1) return;
  const long c = L.at(i, j, k);
  const double q = pp[c];
  if (!impdiff) {
    p[c] = (T)((double)p[c] + q);
    return;
  }
  const double qxp = pp[c + 1], qxm = pp[c - 1], qyp = pp[c + L.sy], qym = pp[c - L.sy];
  const double qzp = pp[c + L.sz], qzm = pp[c - L.sz];
  double lap = 0.;
  if (impdiff != 2)  // is_impdiff_1d: the z part only
    lap = lap + (qxp - 2. * q + qxm) * dxi2 + (qyp - 2. * q + qym) * dyi2;
  lap = lap + ((qzp - q) * dzci[k] - (q - qzm) * dzci[k - 1]) * dzfi[k];
  p[c] = (T)((double)p[c] + q + alpha * lap);
}

// one boundary plane per op (bound.f90:182-463 and the single-rank halo copy of 600-713)
template <typename T>
__global__ void k_bound(Layout L, PlaneOps ops) {
  const PlaneOp op = ops.op[blockIdx.y];
  const int e1 = op.dir == 0 ? L.n2 + 2 : L.n1 + 2;
  const int e2 = op.dir == 2 ? L.n2 + 2 : L.n3 + 2;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)e1 * e2) return;
  const int a = (int)(idx % e1), b = (int)(idx / e1);
  long base, st;  // base: element at normal index 0; st: stride along the normal
  if (op.dir == 0) { base = L.at(0, a, b); st = 1; }
  else if (op.dir == 1) { base = L.at(a, 0, b); st = L.sy; }
  else { base = L.at(a, b, 0); st = L.sz; }
  T* f = (T*)op.f;
  const int nd = op.nd;
  if (op.base) {  // f holds increments over op.base (+ fadd): apply the rule to the sum, store the increment
    const double* g = op.base;
    const double fc = op.fadd ? *op.fadd : 0.;
    auto t = [&](long o) { return g[o] + (double)f[o] + fc; };
    auto put = [&](long o, double x) { f[o] = (T)(x - g[o] - fc); };
    switch (op.type) {
      case BT_PERIODIC:
        if (op.side == 0) put(base, t(base + st * nd));
        else put(base + st * (nd + 1), t(base + st));
        break;
      case BT_CENTERED:
        if (op.side == 0) put(base, op.factor + op.sgn * t(base + st));
        else put(base + st * (nd + 1), op.factor + op.sgn * t(base + st * nd));
        break;
      case BT_STAG_D:
        if (op.side == 0) put(base, op.factor);
        else {
          put(base + st * (nd + 1), t(base + st * (nd - 1)));
          put(base + st * nd, op.factor);
        }
        break;
      case BT_STAG_N:
        if (op.side == 0) put(base, 1. * op.factor + t(base + st));
        else {
          put(base + st * (nd + 1), t(base + st * nd));
          put(base + st * nd, 1. * op.factor + t(base + st * (nd - 1)));
        }
        break;
    }
    return;
  }
  switch (op.type) {
    case BT_PERIODIC:
      if (op.side == 0) f[base] = f[base + st * nd];
      else f[base + st * (nd + 1)] = f[base + st];
      break;
    case BT_CENTERED:
      if (op.side == 0) f[base] = (T)(op.factor + op.sgn * (double)f[base + st]);
      else f[base + st * (nd + 1)] = (T)(op.factor + op.sgn * (double)f[base + st * nd]);
      break;
    case BT_STAG_D:
      if (op.side == 0) f[base] = (T)op.factor;
      else {
        f[base + st * (nd + 1)] = f[base + st * (nd - 1)];
        f[base + st * nd] = (T)op.factor;
      }
      break;
    case BT_STAG_N:
      if (op.side == 0) f[base] = (T)(1. * op.factor + (double)f[base + st]);
      else {
        f[base + st * (nd + 1)] = f[base + st * nd];
        f[base + st * nd] = (T)(1. * op.factor + (double)f[base + st * (nd - 1)]);
      }
      break;
  }
}

// updt_rhs_b (bound.f90:514-598): p(plane) += rhsb(:,:,side)*norm
template <typename T>
__global__ void k_rhs_b(Layout L, T* p, const double* r, int dir, int plane, double norm) {
  const int e1 = dir == 0 ? L.n2 : L.n1;
  const int e2 = dir == 2 ? L.n2 : L.n3;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)e1 * e2) return;
  const int a = (int)(idx % e1) + 1, b = (int)(idx / e1) + 1;
  long c;
  if (dir == 0) c = L.at(plane, a, b);
  else if (dir == 1) c = L.at(a, plane, b);
  else c = L.at(a, b, plane);
  p[c] = (T)((double)p[c] + r[idx] * norm);
}

// divergence norms (chkdiv.f90:15-54): partial sum of |div|*vol and max |div|
template <typename T>
__global__ void k_chkdiv(Layout L, const T* u, const T* v, const T* w, const double* dzfi, double dxi, double dyi,
                         double* partials, int nb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1, k = blockIdx.z + 1;
  double tot = 0., mx = 0.;
  if (i <= L.n1) {
    const long c = L.at(i, j, k);
    const double wc = w[c], wm = w[c - L.sz], vc = v[c], vm = v[c - L.sy], uc = u[c], um = u[c - 1];
    const double div = (wc - wm) * dzfi[k] + (vc - vm) * dyi + (uc - um) * dxi;
    mx = fabs(div);
    tot = fabs(div) / (dxi * dyi * dzfi[k]);
  }
  const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  block_sum_max(tot, mx, partials + bid, partials + nb + bid);
}

// CFL estimate (chkdt.f90:16-78): partial max of dti
template <typename T>
__global__ void k_chkdt(Layout L, const T* u, const T* v, const T* w, const double* dzci, const double* dzfi,
                        double dxi, double dyi, double* partials) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1, k = blockIdx.z + 1;
  double dti = 0.;
  if (i <= L.n1) {
    const long c = L.at(i, j, k);
    const long sy = L.sy, sz = L.sz;
    auto U = [&](long o) { return (double)u[o]; };
    auto V = [&](long o) { return (double)v[o]; };
    auto W = [&](long o) { return (double)w[o]; };
    const double ux = fabs(U(c));
    const double vx = 0.25 * fabs(V(c) + V(c - sy) + V(c + 1) + V(c + 1 - sy));
    const double wx = 0.25 * fabs(W(c) + W(c - sz) + W(c + 1) + W(c + 1 - sz));
    const double dtix = ux * dxi + vx * dyi + wx * dzfi[k];
    const double uy = 0.25 * fabs(U(c) + U(c + sy) + U(c - 1 + sy) + U(c - 1));
    const double vy = fabs(V(c));
    const double wy = 0.25 * fabs(W(c) + W(c + sy) + W(c + sy - sz) + W(c - sz));
    const double dtiy = uy * dxi + vy * dyi + wy * dzfi[k];
    const double uz = 0.25 * fabs(U(c) + U(c - 1) + U(c - 1 + sz) + U(c + sz));
    const double vz = 0.25 * fabs(V(c) + V(c - sy) + V(c - sy + sz) + V(c + sz));
    const double wz = fabs(W(c));
    const double dtiz = uz * dxi + vz * dyi + wz * dzci[k];
    dti = fmax(dti, fmax(dtix, fmax(dtiy, dtiz)));
  }
  const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  block_max(dti, partials + bid);
}

// bulk mean of one field (utils.f90:15-42)
template <typename T>
__global__ void k_bulk_mean(Layout L, const T* q, const double* gvr, double* partials) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1, k = blockIdx.z + 1;
  double s = 0.;
  if (i <= L.n1) s = gvr ? (double)q[L.at(i, j, k)] * gvr[k] : (double)q[L.at(i, j, k)];  // gvr null: a plain sum
  const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  double d1, d2;
  block_sum3(s, 0., 0., partials + bid, &d1, &d2, true);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static dim3 row_grid(const Layout& L, int bx) { return dim3(cdiv(L.n1, bx), L.n2, L.n3); }

template <typename S>
int launch_mom_rk(cans_ctx* c, const MomLaunch<S>& m) {
  const Layout& L = c->L;
  MomArgs<S> a;
  a.L = L;
  a.u = m.u; a.v = m.v; a.w = m.w; a.p = m.p;
  a.uo = m.uo; a.vo = m.vo; a.wo = m.wo;
  a.eu = m.eu; a.ev = m.ev; a.ew = m.ew;
  a.delta = m.eu != nullptr;
  a.divu = m.divu;
  a.ru = m.ru; a.rv = m.rv; a.rw = m.rw;
  a.dzci = m.dzci; a.dzfi = m.dzfi;
  a.gvr_c = m.gvr_c; a.gvr_f = m.gvr_f;
  a.dxi = m.dxi; a.dyi = m.dyi; a.visc = m.visc;
  a.factor1 = m.rkpar[0] * m.dt;
  a.factor2 = m.rkpar[1] * m.dt;
  a.factor12 = a.factor1 + a.factor2;
  for (int d = 0; d < 3; d++) a.bforce[d] = m.bforce[d];
  a.sum_mask = (m.is_forced[0] ? 1 : 0) | (m.is_forced[1] ? 2 : 0) | (m.is_forced[2] ? 4 : 0);
  // rdzm/rdzp tables (mom.f90:801-809), kept in context scratch slot 3
  void* tab = nullptr;
  int rc = ensure_work(c, 3, 2 * (L.n3 + 2) * sizeof(double), &tab);
  if (rc) return rc;
  a.rdzm = (const double*)tab;
  a.rdzp = (const double*)tab + (L.n3 + 2);
  k_rdz<<<cdiv(L.n3 + 2, 256), 256, 0, c->stream>>>(L.n3, m.dzci, m.dzfi, (double*)tab);
  CANS_LAUNCH_CHECK("k_rdz");
  const long nb = (long)cdiv(L.n1, MOM_TX) * cdiv(L.n2, MOM_TY) * cdiv(L.n3, MOM_KC);
  dim3 grid((unsigned)nb);
  if (a.sum_mask && (rc = ensure_partials(c, 3 * nb))) return rc;
  a.partials = c->partials;
  a.read_ro = m.read_ro && m.rkpar[1] != 0.;
  a.write_ro = m.write_ro;
  // the buoyancy coefficient in the reference's left-to-right order: factor12*gacc*beta*0.5
  a.sb = m.sb;
  for (int d = 0; d < 3; d++) a.bq[d] = a.factor12 * m.gacc[d] * m.beta * 0.5;
  a.imp = m.imp;
  a.hf12 = 0.5 * a.factor12;
  const dim3 blk(MOM_TX, MOM_TY);
  if (m.fold) {
    const MomFold<S>& f = *m.fold;
    if (a.divu || a.sb || a.imp || !f.pp || !f.f || !f.pout)
      return fail(CANS_EINVAL, "k_mom_rk: the folded projection excludes div(u), buoyancy and implicit diffusion");
    a.fpp = f.pp;
    a.feu = f.eu; a.fev = f.ev; a.few = f.ew;
    a.ff = f.f;
    a.fdt = f.dt;
    a.fdtdxi = f.dt * m.dxi;  // correc's factori = dt*dli(1)
    a.fdtdyi = f.dt * m.dyi;
    a.pout = f.pout;
    // the ghost cells' projected values, three halo'd FP64 fields of which only the faces are used
    void* gb = nullptr;
    const size_t fe = (size_t)L.sz * (L.n3 + 2);
    if ((rc = ensure_work(c, 10, 3 * fe * sizeof(double), &gb))) return rc;
    FoldGhostArgs<S> g;
    g.L = L;
    g.u = m.u; g.v = m.v; g.w = m.w;
    g.eu = f.eu; g.ev = f.ev; g.ew = f.ew;
    g.pp = f.pp;
    g.f = f.f;
    g.dzci = m.dzci;
    g.dt = f.dt;
    g.dtdxi = a.fdtdxi;
    g.dtdyi = a.fdtdyi;
    g.per = (f.periodic[0] ? 1 : 0) | (f.periodic[1] ? 2 : 0) | (f.periodic[2] ? 4 : 0);
    for (int q = 0; q < 3; q++)
      for (int d = 0; d < 3; d++)
        for (int sd = 0; sd < 2; sd++) {
          g.factor[q][d][sd] = f.factor[q][d][sd];
          g.sgn[q][d][sd] = f.sgn[q][d][sd];
        }
    g.gu = (double*)gb;
    g.gv = g.gu + fe;
    g.gw = g.gv + fe;
    a.gu = g.gu; a.gv = g.gv; a.gw = g.gw;
    const long emax = (long)std::max(L.n1, L.n2) + 2;
    const long plane = emax * (std::max(L.n2, L.n3) + 2);
    k_fold_ghosts<S><<<dim3(cdiv(plane, 256), 6), 256, 0, c->stream>>>(g);
    CANS_LAUNCH_CHECK("k_fold_ghosts");
    if (a.delta) k_mom_rk<S, true, false, false, true><<<grid, blk, 0, c->stream>>>(a);
    else k_mom_rk<S, false, false, false, true><<<grid, blk, 0, c->stream>>>(a);
    CANS_LAUNCH_CHECK("k_mom_rk");
    if (a.sum_mask)
      return reduce_partials(c, c->partials, (int)nb, 3, 0, c->dscal, c->partials + 3 * nb, 1, a.sum_mask, m.velf);
    k_fill<<<1, 64, 0, c->stream>>>(c->dscal, 3, 0.);
    CANS_LAUNCH_CHECK("k_fill");
    return CANS_OK;
  }
  const int form = (a.delta ? 4 : 0) | (a.divu ? 2 : 0) | (a.sb ? 1 : 0);
  switch (form) {
    case 0: k_mom_rk<S, false, false, false><<<grid, blk, 0, c->stream>>>(a); break;
    case 1: k_mom_rk<S, false, false, true><<<grid, blk, 0, c->stream>>>(a); break;
    case 2: k_mom_rk<S, false, true, false><<<grid, blk, 0, c->stream>>>(a); break;
    case 3: k_mom_rk<S, false, true, true><<<grid, blk, 0, c->stream>>>(a); break;
    case 4: k_mom_rk<S, true, false, false><<<grid, blk, 0, c->stream>>>(a); break;
    case 5: k_mom_rk<S, true, false, true><<<grid, blk, 0, c->stream>>>(a); break;
    case 6: k_mom_rk<S, true, true, false><<<grid, blk, 0, c->stream>>>(a); break;
    default: k_mom_rk<S, true, true, true><<<grid, blk, 0, c->stream>>>(a); break;
  }
  CANS_LAUNCH_CHECK("k_mom_rk");
  if (a.sum_mask)
    return reduce_partials(c, c->partials, (int)nb, 3, 0, c->dscal, c->partials + 3 * nb, 1, a.sum_mask, m.velf);
  k_fill<<<1, 64, 0, c->stream>>>(c->dscal, 3, 0.);  // no memset node if the caller captures a graph
  CANS_LAUNCH_CHECK("k_fill");
  return CANS_OK;
}

template <typename T>
int launch_add_forcing(cans_ctx* c, const T* src, T* dst, int comp, int apply) {
  const Layout& L = c->L;
  k_add_forcing<T><<<row_grid(L, 256), 256, 0, c->stream>>>(L, src, dst, c->dscal, comp, apply);
  CANS_LAUNCH_CHECK("k_add_forcing");
  return CANS_OK;
}

int launch_fill(cans_ctx* c, double* p, int n, double v) {
  k_fill<<<cdiv(n, 64), 64, 0, c->stream>>>(p, n, v);
  CANS_LAUNCH_CHECK("k_fill");
  return CANS_OK;
}

int launch_scal_rk(cans_ctx* c, const ScalLaunch& m) {
  const Layout& L = c->L;
  ScalArgs a;
  a.L = L;
  a.u = m.u; a.v = m.v; a.w = m.w; a.s = m.s;
  a.so = m.so; a.ro = m.ro;
  a.dzci = m.dzci; a.dzfi = m.dzfi; a.gvr_f = m.gvr_f;
  a.dxi = m.dxi; a.dyi = m.dyi; a.alpha = m.alpha;
  a.factor1 = m.rkpar[0] * m.dt;
  a.factor2 = m.rkpar[1] * m.dt;
  a.factor12 = a.factor1 + a.factor2;
  a.source = m.source;
  a.read_ro = m.rkpar[1] != 0.;
  a.imp = m.imp;
  a.hf12 = 0.5 * a.factor12;
  const dim3 grid = row_grid(L, 256);
  const long nb = (long)grid.x * grid.y * grid.z;
  int rc;
  if (m.is_forced && (rc = ensure_partials(c, nb))) return rc;
  a.partials = m.is_forced ? c->partials : nullptr;
  k_scal_rk<<<grid, 256, 0, c->stream>>>(a);
  CANS_LAUNCH_CHECK("k_scal_rk");
  // f = scalf - mean (rk.f90:471-474), or 0, into dscal[m.fslot]
  if (m.is_forced) {
    const double sf[3] = {m.scalf, 0., 0.};
    return reduce_partials(c, c->partials, (int)nb, 1, 0, c->dscal + m.fslot, c->partials + nb, 1, 1, sf);
  }
  k_fill<<<1, 64, 0, c->stream>>>(c->dscal + m.fslot, 1, 0.);
  CANS_LAUNCH_CHECK("k_fill");
  return CANS_OK;
}

template <typename S>
int launch_fillps(cans_ctx* c, const double* u, const double* v, const double* w, const S* eu, const S* ev,
                  const S* ew, const double* f, S* p, const double* dzfi, double dti, const double* dli) {
  const Layout& L = c->L;
  k_fillps<S><<<row_grid(L, 256), 256, 0, c->stream>>>(L, u, v, w, eu, ev, ew, f, p, dzfi, dti, dti * dli[0],
                                                         dti * dli[1]);
  CANS_LAUNCH_CHECK("k_fillps");
  return CANS_OK;
}

template <typename S>
int launch_correc(cans_ctx* c, const S* p, const double* su, const double* sv, const double* sw, const S* eu,
                  const S* ev, const S* ew, const double* f, double* du, double* dv, double* dw, const double* dzci,
                  double dt, const double* dli, S* p_update, const S* p1, const S* p2) {
  const Layout& L = c->L;
  dim3 grid(cdiv(L.n1 + 2, CORREC_TX), cdiv(L.n2 + 2, CORREC_TY), cdiv(L.n3 + 2, CORREC_KC));
  const dim3 blk(CORREC_TX, CORREC_TY);
  if (p2 && !p1) return fail(CANS_EINVAL, "launch_correc: p2 without p1");
  if (p2)
    k_correc<S, 3><<<grid, blk, 0, c->stream>>>(L, p, su, sv, sw, eu, ev, ew, f, du, dv, dw, dzci, dt, dt * dli[0],
                                                dt * dli[1], p_update, p1, p2);
  else if (p1)
    k_correc<S, 2><<<grid, blk, 0, c->stream>>>(L, p, su, sv, sw, eu, ev, ew, f, du, dv, dw, dzci, dt, dt * dli[0],
                                                dt * dli[1], p_update, p1, p2);
  else
    k_correc<S, 1><<<grid, blk, 0, c->stream>>>(L, p, su, sv, sw, eu, ev, ew, f, du, dv, dw, dzci, dt, dt * dli[0],
                                                dt * dli[1], p_update, p1, p2);
  CANS_LAUNCH_CHECK("k_correc");
  return CANS_OK;
}

// fillps.f90:13-51 minus the solver's 7-point operator applied to c0 + c1 (c1 may be nullptr):
// the residual of a refinement sweep when the divergence is not formed inside the first x pass
template <typename S>
__global__ void k_fillps_residual(Layout L, const double* u, const double* v, const double* w, const S* eu,
                                  const S* ev, const S* ew, const double* f, S* p, const double* dzfi, double dti,
                                  double dtidxi, double dtidyi, const S* c0, const S* c1, const double* dzci,
                                  double dxi2, double dyi2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1, k = blockIdx.z + 1;
  if (i > L.n1) return;
  const long c = L.at(i, j, k);
  const double fu = f ? f[0] : 0., fv = f ? f[1] : 0., fw = f ? f[2] : 0.;
  const double wc = vel(w, ew, fw, c), wm = vel(w, ew, fw, c - L.sz), vc = vel(v, ev, fv, c);
  const double vm = vel(v, ev, fv, c - L.sy), uc = vel(u, eu, fu, c), um = vel(u, eu, fu, c - 1);
  const double b = (wc - wm) * dti * dzfi[k] + (vc - vm) * dtidyi + (uc - um) * dtidxi;
  auto x = [&](long o) { return c1 ? (double)c0[o] + (double)c1[o] : (double)c0[o]; };
  const double xc = x(c);
  const double lap = (x(c + 1) - 2. * xc + x(c - 1)) * dxi2 + (x(c + L.sy) - 2. * xc + x(c - L.sy)) * dyi2 +
                     ((x(c + L.sz) - xc) * dzci[k] - (xc - x(c - L.sz)) * dzci[k - 1]) * dzfi[k];
  p[c] = (S)(b - lap);
}

template <typename S>
int launch_fillps_residual(cans_ctx* c, const double* u, const double* v, const double* w, const S* eu,
                           const S* ev, const S* ew, const double* f, S* p, const double* dzfi, double dti,
                           const double* dli, const S* c0, const S* c1, const double* dzci) {
  const Layout& L = c->L;
  k_fillps_residual<S><<<row_grid(L, 256), 256, 0, c->stream>>>(L, u, v, w, eu, ev, ew, f, p, dzfi, dti,
                                                                  dti * dli[0], dti * dli[1], c0, c1, dzci,
                                                                  dli[0] * dli[0], dli[1] * dli[1]);
  CANS_LAUNCH_CHECK("k_fillps_residual");
  return CANS_OK;
}

template <typename T>
int launch_updatep(cans_ctx* c, const T* pp, T* p, const double* dzci, const double* dzfi, double alpha,
                   const double* dli, int impdiff) {
  const Layout& L = c->L;
  k_updatep<T><<<row_grid(L, 256), 256, 0, c->stream>>>(L, pp, p, dzci, dzfi, alpha, dli[0] * dli[0],
                                                          dli[1] * dli[1], impdiff);
  CANS_LAUNCH_CHECK("k_updatep");
  return CANS_OK;
}

template <typename T>
int launch_bound_dir(cans_ctx* c, const PlaneOps& ops) {
  if (ops.nops == 0) return CANS_OK;
  const Layout& L = c->L;
  const int dir = ops.op[0].dir;
  const long e1 = dir == 0 ? L.n2 + 2 : L.n1 + 2;
  const long e2 = dir == 2 ? L.n2 + 2 : L.n3 + 2;
  dim3 grid(cdiv(e1 * e2, 256), ops.nops);
  k_bound<T><<<grid, 256, 0, c->stream>>>(L, ops);
  CANS_LAUNCH_CHECK("k_bound");
  return CANS_OK;
}

template <typename T>
int launch_rhs_b(cans_ctx* c, T* p, const double* r, int dir, int plane, double norm) {
  const Layout& L = c->L;
  const long e1 = dir == 0 ? L.n2 : L.n1;
  const long e2 = dir == 2 ? L.n2 : L.n3;
  k_rhs_b<T><<<cdiv(e1 * e2, 256), 256, 0, c->stream>>>(L, p, r, dir, plane, norm);
  CANS_LAUNCH_CHECK("k_rhs_b");
  return CANS_OK;
}

template <typename T>
int run_chkdiv(cans_ctx* c, const T* u, const T* v, const T* w, const double* dzfi, const double* dli, double* tot,
               double* mx) {
  const Layout& L = c->L;
  dim3 grid = row_grid(L, 256);
  const long nb = (long)grid.x * grid.y * grid.z;
  int rc = ensure_partials(c, 2 * nb);
  if (rc) return rc;
  k_chkdiv<T><<<grid, 256, 0, c->stream>>>(L, u, v, w, dzfi, dli[0], dli[1], c->partials, (int)nb);
  CANS_LAUNCH_CHECK("k_chkdiv");
  if ((rc = reduce_partials(c, c->partials, (int)nb, 1, 0, c->dscal + 8, c->partials + 2 * nb))) return rc;
  if ((rc = reduce_partials(c, c->partials + nb, (int)nb, 1, 1, c->dscal + 9, c->partials + 2 * nb + 1024)))
    return rc;
  double h[2];
  CANS_HIP_TRY(hipMemcpyAsync(h, c->dscal + 8, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CANS_HIP_TRY(hipStreamSynchronize(c->stream));
  *tot = h[0];
  *mx = h[1];
  return CANS_OK;
}

template <typename T>
int run_chkdt(cans_ctx* c, const T* u, const T* v, const T* w, const double* dzci, const double* dzfi,
              const double* dl, double* dti_out) {
  const Layout& L = c->L;
  dim3 grid = row_grid(L, 256);
  const long nb = (long)grid.x * grid.y * grid.z;
  int rc = ensure_partials(c, nb);
  if (rc) return rc;
  k_chkdt<T><<<grid, 256, 0, c->stream>>>(L, u, v, w, dzci, dzfi, 1. / dl[0], 1. / dl[1], c->partials);
  CANS_LAUNCH_CHECK("k_chkdt");
  if ((rc = reduce_partials(c, c->partials, (int)nb, 1, 1, c->dscal + 10, c->partials + nb))) return rc;
  CANS_HIP_TRY(hipMemcpyAsync(dti_out, c->dscal + 10, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CANS_HIP_TRY(hipStreamSynchronize(c->stream));
  return CANS_OK;
}

template <typename T>
int run_bulk_mean(cans_ctx* c, const T* q, const double* gvr, double* mean) {
  const Layout& L = c->L;
  dim3 grid = row_grid(L, 256);
  const long nb = (long)grid.x * grid.y * grid.z;
  int rc = ensure_partials(c, nb);
  if (rc) return rc;
  k_bulk_mean<T><<<grid, 256, 0, c->stream>>>(L, q, gvr, c->partials);
  CANS_LAUNCH_CHECK("k_bulk_mean");
  if ((rc = reduce_partials(c, c->partials, (int)nb, 1, 0, c->dscal + 11, c->partials + nb))) return rc;
  CANS_HIP_TRY(hipMemcpyAsync(mean, c->dscal + 11, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CANS_HIP_TRY(hipStreamSynchronize(c->stream));
  return CANS_OK;
}

// Residual of the pressure equation for the iterative refinement of a Poisson solve:
// r = b - L x at interior cells, L the 7-point operator whose eigenvalues and
// tridiagonal coefficients the solver uses (initsolver.f90:85-187: second differences
// with dxi^2, dyi^2 along x, y; dzfi(k) (dzci(k) (x(k+1) - x(k)) - dzci(k-1) (x(k) -
// x(k-1))) along z), the ghost values homogeneous by boundary type (bound.f90:193-206
// for cell-centred BCs): P the periodic image, N x(0) = x(1), D x(0) = -x(1).  x = nullptr
// gives r = b.  bt[2d+s]: 0 P, 1 N, 2 D for direction d, side s.
template <typename S>
__global__ void k_residual(Layout L, const double* b, const double* x, S* r, const double* dzci,
                           const double* dzfi, double dxi2, double dyi2, int4 bt0, int2 bt1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1, k = blockIdx.z + 1;
  if (i > L.n1) return;
  const long o = L.at(i, j, k);
  if (!x) {
    r[o] = (S)b[o];
    return;
  }
  const int bt[6] = {bt0.x, bt0.y, bt0.z, bt0.w, bt1.x, bt1.y};
  const double xc = x[o];
  // neighbour along direction d (index pos in 1..n, stride st) on side s: interior, periodic image or ghost
  auto nbd = [&](int d, int pos, int n, long st, int side) {
    const int t = bt[2 * d + side];
    if (side == 0 && pos > 1) return x[o - st];
    if (side == 1 && pos < n) return x[o + st];
    if (t == 0) return side == 0 ? x[o + (long)(n - 1) * st] : x[o - (long)(n - 1) * st];
    return t == 1 ? xc : -xc;
  };
  const double xm = nbd(0, i, L.n1, 1, 0), xp = nbd(0, i, L.n1, 1, 1);
  const double ym = nbd(1, j, L.n2, L.sy, 0), yp = nbd(1, j, L.n2, L.sy, 1);
  const double zm = nbd(2, k, L.n3, L.sz, 0), zp = nbd(2, k, L.n3, L.sz, 1);
  const double lap = (xp - 2. * xc + xm) * dxi2 + (yp - 2. * xc + ym) * dyi2 +
                     ((zp - xc) * dzci[k] - (xc - zm) * dzci[k - 1]) * dzfi[k];
  r[o] = (S)(b[o] - lap);
}

// x = (first ? 0 : x) + d at interior cells
template <typename S>
__global__ void k_accumulate(Layout L, double* x, const S* d, int first) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1, k = blockIdx.z + 1;
  if (i > L.n1) return;
  const long o = L.at(i, j, k);
  x[o] = first ? (double)d[o] : x[o] + (double)d[o];
}

template <typename S>
int launch_residual(cans_ctx* c, const double* b, const double* x, S* r, const double* dzci, const double* dzfi,
                    const double* dli, const int* bt) {
  const Layout& L = c->L;
  k_residual<S><<<row_grid(L, 256), 256, 0, c->stream>>>(L, b, x, r, dzci, dzfi, dli[0] * dli[0], dli[1] * dli[1],
                                                          make_int4(bt[0], bt[1], bt[2], bt[3]),
                                                          make_int2(bt[4], bt[5]));
  CANS_LAUNCH_CHECK("k_residual");
  return CANS_OK;
}

template <typename S>
int launch_accumulate(cans_ctx* c, double* x, const S* d, int first) {
  const Layout& L = c->L;
  k_accumulate<S><<<row_grid(L, 256), 256, 0, c->stream>>>(L, x, d, first);
  CANS_LAUNCH_CHECK("k_accumulate");
  return CANS_OK;
}

#define CANS_STENCIL_INSTANTIATE(S)                                                                              \
  template int launch_mom_rk<S>(cans_ctx*, const MomLaunch<S>&);                                                \
  template int launch_add_forcing<S>(cans_ctx*, const S*, S*, int, int);                                        \
  template int launch_fillps<S>(cans_ctx*, const double*, const double*, const double*, const S*, const S*,     \
                                const S*, const double*, S*, const double*, double, const double*);             \
  template int launch_correc<S>(cans_ctx*, const S*, const double*, const double*, const double*, const S*,     \
                                const S*, const S*, const double*, double*, double*, double*, const double*,    \
                                double, const double*, S*, const S*, const S*);                                  \
  template int launch_fillps_residual<S>(cans_ctx*, const double*, const double*, const double*, const S*,       \
                                         const S*, const S*, const double*, S*, const double*, double,           \
                                         const double*, const S*, const S*, const double*);                      \
  template int launch_updatep<S>(cans_ctx*, const S*, S*, const double*, const double*, double, const double*, int); \
  template int launch_bound_dir<S>(cans_ctx*, const PlaneOps&);                                                  \
  template int launch_rhs_b<S>(cans_ctx*, S*, const double*, int, int, double);                               \
  template int launch_residual<S>(cans_ctx*, const double*, const double*, S*, const double*, const double*,   \
                                  const double*, const int*);                                                  \
  template int launch_accumulate<S>(cans_ctx*, double*, const S*, int);
CANS_STENCIL_INSTANTIATE(double)
CANS_STENCIL_INSTANTIATE(float)
// reductions read the velocity, which is always double
template int run_chkdiv<double>(cans_ctx*, const double*, const double*, const double*, const double*,
                                const double*, double*, double*);
template int run_chkdt<double>(cans_ctx*, const double*, const double*, const double*, const double*,
                               const double*, const double*, double*);
template int run_bulk_mean<double>(cans_ctx*, const double*, const double*, double*);

}  // namespace cans
