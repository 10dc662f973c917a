// Internal launch interfaces shared by the translation units of the library.
#pragma once
#include "common.hpp"

namespace cans {

// boundary plane operations (bound.f90:182-463, single-rank halo copies 600-713)
enum { BT_PERIODIC = 0, BT_CENTERED = 1, BT_STAG_D = 2, BT_STAG_N = 3 };
struct PlaneOp {
  double* f;
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

struct MomLaunch {
  const double *u, *v, *w, *p;
  double *uo, *vo, *wo;
  double *ru, *rv, *rw;
  const double *dzci, *dzfi, *gvr_c, *gvr_f;
  double dxi, dyi, visc, dt;
  double rkpar[2];
  double bforce[3], velf[3];
  int is_forced[3];
  int read_ro = 1, write_ro = 1;  // RHS_old input/output needed (see launch_mom_rk)
};

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

// stencil.hip
int launch_mom_rk(cans_ctx* c, const MomLaunch& m);
int launch_add_forcing(cans_ctx* c, const double* src, double* dst, int comp, int apply);
int launch_fillps(cans_ctx* c, const double* u, const double* v, const double* w, double* p, const double* dzfi,
                  double dti, const double* dli);
int launch_correc(cans_ctx* c, const double* p, const double* su, const double* sv, const double* sw, double* du,
                  double* dv, double* dw, const double* dzci, double dt, const double* dli,
                  double* p_update = nullptr);
int launch_updatep(cans_ctx* c, const double* pp, double* p, const double* dzci, const double* dzfi, double alpha,
                   const double* dli, int impdiff);
int launch_bound_dir(cans_ctx* c, const PlaneOps& ops);
int launch_rhs_b(cans_ctx* c, double* p, const double* r, int dir, int plane, double norm);
int run_chkdiv(cans_ctx* c, const double* u, const double* v, const double* w, const double* dzfi, const double* dli,
               double* tot, double* mx);
int run_chkdt(cans_ctx* c, const double* u, const double* v, const double* w, const double* dzci, const double* dzfi,
              const double* dl, double* dti_out);
int run_bulk_mean(cans_ctx* c, const double* q, const double* gvr, double* mean);

// solver.hip
// right-hand side of the pressure equation computed from the velocity inside the
// first transform pass (fillps.f90:38-50); p then receives only the solution
struct DivSrc {
  const double *u, *v, *w, *dzfi;
  double dti, dtidxi, dtidyi;
};
int run_solver(cans_ctx* c, const cans_plan* plan, double normfft, const double* lam, const double* a,
               const double* b, const double* cc, const char* bc, const char* c_or_f, double* p,
               const DivSrc* div = nullptr);

}  // namespace cans
