// Internal launch interfaces shared by the translation units of the library.
#pragma once
#include "common.hpp"

namespace cans {

// boundary plane operations (bound.f90:182-463, single-rank halo copies 600-713)
enum { BT_PERIODIC = 0, BT_CENTERED = 1, BT_STAG_D = 2, BT_STAG_N = 3 };
struct PlaneOp {
  void* f;              // field (element type chosen at launch)
  const double* base;   // nullptr, or: f holds increments over this velocity field,
  const double* fadd;   // plus this device scalar (nullptr = 0): value = base + f + *fadd
  int dir;    // 0,1,2 = x,y,z
  int side;   // 0 lower, 1 upper
  int type;   // BT_*
  int nd;     // interior extent along dir
  double factor, sgn;
};
struct PlaneOps {
  PlaneOp op[8];
  int nops;
};

// The previous substep's projection, applied while the momentum kernel stages its
// velocity: u = correc(intermediate velocity, pp) with the ghost values bounduvw gives
// after correc (bound.f90, keep_norm), and p + pp (updatep) for the pressure term.
template <typename S>
struct MomFold {
  const S* pp;                    // its pressure correction (halo'd, ghosts set by boundp)
  const S *eu, *ev, *ew;          // mixed: its increments over the kernel's u, v, w; FP64: nullptr
                                  // (u, v, w then hold its intermediate velocity)
  const double* f;                // its bulk forcing (device, 3)
  double dt;                      // its dtrk
  S* pout;                        // p + pp, written for the interior cells
  int periodic[3];                // direction periodic
  // tangential components at non-periodic faces: ghost = factor + sgn * adjacent value
  double factor[3][3][2], sgn[3][3][2];  // [component][direction][side]
};

template <typename S>
struct MomLaunch {
  const double *u, *v, *w;
  const S* p;
  const MomFold<S>* fold = nullptr;  // fold the previous substep's projection (see MomFold)
  double *uo = nullptr, *vo = nullptr, *wo = nullptr;  // updated velocity, or
  S *eu = nullptr, *ev = nullptr, *ew = nullptr;       // its increment over (u,v,w)
  S* divu = nullptr;                                   // optional: div(u) per cell, (n1,n2,n3)
  S *ru, *rv, *rw;
  const double *dzci, *dzfi, *gvr_c, *gvr_f;
  double dxi, dyi, visc, dt;
  double rkpar[2];
  double bforce[3], velf[3];
  int is_forced[3];
  int read_ro = 1, write_ro = 1;  // RHS_old input/output needed (see launch_mom_rk)
  const double* sb = nullptr;     // Boussinesq buoyancy: the first scalar (halo'd), or none
  double gacc[3] = {0., 0., 0.}, beta = 0.;
  int imp = 0;                    // 0 explicit, 1 implicit diffusion, 2 implicit in z only
};

// scalar transport + RK update of one scalar (rk.f90:335-491, explicit diffusion);
// the bulk forcing f (scalf - mean, or 0) goes to c->dscal[fslot]
struct ScalLaunch {
  const double *u, *v, *w, *s;
  double* so;  // updated scalar
  double* ro;  // dsdtrko
  const double *dzci, *dzfi, *gvr_f;
  double dxi, dyi, alpha, dt, source, scalf;
  double rkpar[2];
  int is_forced;
  int fslot;
  int imp = 0;  // 0 explicit, 1 implicit diffusion, 2 implicit in z only
};
int launch_scal_rk(cans_ctx* c, const ScalLaunch& m);
// implicit-diffusion Helmholtz solves of FP64 fields (solver.hip): one system each, or a batch
struct HelmSys {
  const cans_plan* plan;  // nullptr in the z-only form
  double normfft;
  const double *lam, *a, *b, *c;
  const char *cbc, *c_or_f;
  double* fld;
  double alpha;
};
int run_helmholtz_batch(cans_ctx* c, int nsys, const HelmSys* hs);
int run_helmholtz_z(cans_ctx* c, const double* a, const double* b, const double* cc, double alpha, const char* bcz,
                    char c_or_f_z, double* fld);
template <typename T>
int run_helmholtz_3d(cans_ctx* c, const cans_plan* p, double normfft, double alpha, const double* lam,
                     const double* a, const double* b, const double* cc, const char* cbc, const char* c_or_f,
                     T* fld);
int launch_fill(cans_ctx* c, double* p, int n, double v);  // p[0..n) = v, on the context stream
int field_elem_size(const cans_ctx* c, const void* field);

// XCD-aware tile order: the hardware hands work-group b to XCD b % 8; remap so that
// each XCD walks a contiguous range of tiles (neighbouring tiles then share its L2)
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int x = b & 7, q = b >> 3;
  const int base = nb >> 3, rem = nb & 7;
  const int start = x * base + (x < rem ? x : rem);
  return start + q;
}

// deterministic block reductions for NT-thread blocks
template <int NT = 256>
__device__ __forceinline__ void block_sum3(double a, double b, double c, double* oa, double* ob, double* oc,
                                           bool only_a = false) {
  __shared__ double r[3][NT];
  const int t = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  r[0][t] = a;
  r[1][t] = b;
  r[2][t] = c;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (t < s) {
      r[0][t] += r[0][t + s];
      if (!only_a) {
        r[1][t] += r[1][t + s];
        r[2][t] += r[2][t + s];
      }
    }
    __syncthreads();
  }
  if (t == 0) {
    *oa = r[0][0];
    if (!only_a) {
      *ob = r[1][0];
      *oc = r[2][0];
    }
  }
}

__device__ __forceinline__ void block_sum_max(double a, double m, double* oa, double* om) {
  __shared__ double r[2][256];
  const int t = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  r[0][t] = a;
  r[1][t] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      r[0][t] += r[0][t + s];
      r[1][t] = fmax(r[1][t], r[1][t + s]);
    }
    __syncthreads();
  }
  if (t == 0) {
    *oa = r[0][0];
    *om = r[1][0];
  }
}

__device__ __forceinline__ void block_max(double m, double* om) {
  __shared__ double r[256];
  const int t = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  r[t] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) r[t] = fmax(r[t], r[t + s]);
    __syncthreads();
  }
  if (t == 0) *om = r[0];
}

int ensure_work(cans_ctx* c, int slot, size_t bytes, void** ptr);

// stencil.hip.  Velocity arrays are double; S is the storage type of the
// pressure-like arrays (double, or float in the mixed-precision mode).
template <typename S>
int launch_mom_rk(cans_ctx* c, const MomLaunch<S>& m);
template <typename T>
int launch_add_forcing(cans_ctx* c, const T* src, T* dst, int comp, int apply);
// Where increments e are given, the velocity is u + e + f[comp] with f a device
// vector of 3 (nullptr = 0): the bulk forcing, folded in instead of stored.
template <typename S>
int launch_fillps(cans_ctx* c, const double* u, const double* v, const double* w, const S* eu, const S* ev,
                  const S* ew, const double* f, S* p, const double* dzfi, double dti, const double* dli);
// p1, p2 (optional): further pressure corrections; the gradient and the update of p then use
// p + p1 + p2, summed in FP64 (the corrections of a refined solve)
template <typename S>
int launch_correc(cans_ctx* c, const S* p, const double* su, const double* sv, const double* sw, const S* eu,
                  const S* ev, const S* ew, const double* f, double* du, double* dv, double* dw, const double* dzci,
                  double dt, const double* dli, S* p_update, const S* p1 = nullptr, const S* p2 = nullptr);
// fillps of the velocity minus L(c0 + c1) (the residual of a refined solve; c1 may be nullptr)
template <typename S>
int launch_fillps_residual(cans_ctx* c, const double* u, const double* v, const double* w, const S* eu,
                           const S* ev, const S* ew, const double* f, S* p, const double* dzfi, double dti,
                           const double* dli, const S* c0, const S* c1, const double* dzci);
template <typename S>
int launch_updatep(cans_ctx* c, const S* pp, S* p, const double* dzci, const double* dzfi, double alpha,
                   const double* dli, int impdiff);
template <typename T>
int launch_bound_dir(cans_ctx* c, const PlaneOps& ops);
template <typename S>
int launch_rhs_b(cans_ctx* c, S* p, const double* r, int dir, int plane, double norm);
// iterative refinement of a Poisson solve: r = b - L x (x = nullptr: r = b) with homogeneous
// cell-centred ghost values by boundary type bt[6] (0 P, 1 N, 2 D); x = (first ? 0 : x) + d
template <typename S>
int launch_residual(cans_ctx* c, const double* b, const double* x, S* r, const double* dzci, const double* dzfi,
                    const double* dli, const int* bt);
template <typename S>
int launch_accumulate(cans_ctx* c, double* x, const S* d, int first);
template <typename T>
int run_chkdiv(cans_ctx* c, const T* u, const T* v, const T* w, const double* dzfi, const double* dli, double* tot,
               double* mx);
template <typename T>
int run_chkdt(cans_ctx* c, const T* u, const T* v, const T* w, const double* dzci, const double* dzfi,
              const double* dl, double* dti_out);
// sum over the interior of q times gvr(k) (utils.f90:15-42); gvr null: the plain sum
template <typename T>
int run_bulk_mean(cans_ctx* c, const T* q, const double* gvr, double* mean);

// adjoint.hip (FP64; each direction periodic or walled; z periodic on a uniform grid or walled on any grid)
// out = base + J(q)^T g, J the Jacobian of mom_xyz_ad's right-hand side at q; h = (1/dx, 1/dy, 1/dz)
// (tri-periodic uniform grids)
int launch_mom_jt(cans_ctx* c, const double* const* q, const double* const* g, const double* const* base,
                  double* const* out, const double* h, double visc);
// walls: mask = the directions bounded by walls (bit 0 x, 1 y, 2 z), the device arrays dzci,
// dzfi (0:n3+1), and sign[m][d][end] the derivative of the tangential velocity m's ghost value
// at end 0/1 of wall direction d with respect to the adjacent interior value (-1 Dirichlet, +1
// Neumann)
struct AdjWalls {
  int mask;
  const double *dzci, *dzfi;
  double sign[3][3][2];
};
// the adjoint substep's last part: out = wb + J(q)^T (f1 wb + fb), fbo = f2 wb (wb, fb halo'd with
// periodic ghosts along the periodic directions; out and fbo distinct from every input; with walls,
// out's ghost layers are folded into the interior and the normal velocities' wall faces cleared)
int launch_mom_jt_adj(cans_ctx* c, const double* const* q, const double* const* wb, const double* const* fb,
                      double* const* out, double* const* fbo, const double* h, double visc, double f1, double f2,
                      const AdjWalls* wall);
// out += coef L a (tri-periodic uniform grid; h = 1/dx, 1/dy, 1/dz; a with periodic ghosts)
int launch_adj_lap(cans_ctx* c, const double* const* a, double* const* out, double coef, const double* h);
// one scalar's part of the adjoint substep (explicit diffusion): sh is the updated
// scalar's adjoint (in place: + the Boussinesq transpose of wb when buoy), then sbo = sh + J_s^T
// (f1 sh + fsb) (folded at the walls), fsbo = f2 sh, ubo += J_q^T (f1 sh + fsb)
struct ScalAdj {
  const double* q[3];
  const double* s;
  double* sh;
  const double* fsb;
  double *sbo, *fsbo;
  double* ubo[3];
  const double *wb[3];
  const double *dzci, *dzfi;
  double dxi, dyi, alpha, f1, f2;
  double coef[3];     // f12 gacc_c beta / 2
  double sign[3][2];  // per wall direction and end: -1 Dirichlet, +1 Neumann
  int wm, buoy;       // wm: the wall directions, bit 0 x, 1 y, 2 z
};
int launch_scal_adjoint(cans_ctx* c, const ScalAdj& sa);  // the buoyancy transpose into sh
// the transposed scalar forcing (scal.f90:313-329, rk.f90:474-475): sh -= gvr_f(k) * sum
int launch_scal_force_t(cans_ctx* c, double* sh, const double* gvr_f, double sum);
int launch_scal_jt(cans_ctx* c, const ScalAdj& sa);
// the volume weights of P^T = M P M^-1 and the transposed bulk forcing (k_adj_weigh)
int launch_adj_weigh(cans_ctx* c, double* const* f, const double* dzfi, const double* dzci, bool inv, double mu,
                     double mv, int wm);

// solver.hip
// right-hand side of the pressure equation computed from the velocity (plus its
// increments eu,ev,ew when given) inside the first transform pass
// (fillps.f90:38-50); p then receives only the solution
// With divu given (increments mode), the divergence is divu + div(e): u is not read
// (a constant f has no divergence, and the increments' halos carry the BCs).
template <typename S>
struct DivSrc {
  const double *u, *v, *w;
  const S *eu, *ev, *ew;
  const S* divu;
  const double* dzfi;
  double dti, dtidxi, dtidyi, dxi, dyi;
  const double* f = nullptr;  // FP64 velocity (no increments): bulk forcing added where the velocity is read
  // Iterative refinement: the rhs becomes the residual b - L(corr[0] + ... + corr[ncorr-1]), L the
  // solver's 7-point operator (second differences with dxi2, dyi2; dzci/dzfi along z) applied to
  // the sum of the earlier corrections, halo'd fields of S with their ghost cells set by boundp.
  const S* corr[2] = {nullptr, nullptr};
  int ncorr = 0;
  const double* dzci = nullptr;
  double dxi2 = 0., dyi2 = 0.;
};
// A caller-provided workspace for one solve: the transform work buffer and the z pass's
// checkpoints are carved from [base, base + bytes) instead of the context's scratch slots,
// so the solve allocates nothing.
struct SolveWork {
  void* base;
  size_t bytes;
};
// bytes of workspace one solve of the context's grid needs with element size es
size_t solver_work_bytes(const cans_ctx* c, size_t es, bool periodic_z);
// device address of the fixed scratch (CANS_NOALLOC_SCRATCH_BYTES) on a device that has a
// context, or nullptr
void* noalloc_scratch(int device);
// whether the plan's x direction runs the FFT passes (which can form the divergence), not a dense matrix
bool plan_fft_x(const cans_plan* p);
template <typename S>
int run_solver(cans_ctx* c, const cans_plan* plan, double normfft, const double* lam, const double* a,
               const double* b, const double* cc, const char* bc, const char* c_or_f, S* p,
               const DivSrc<S>* div = nullptr, const SolveWork* ws = nullptr);

}  // namespace cans
