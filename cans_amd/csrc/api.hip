// C-ABI entry points: argument checking, boundary-plane planning (host) and
// launch sequencing.  Kernels live in stencil.hip and solver.hip.  Velocity
// fields are double; pressure-like arrays (p, pp, dudtrko) have the context's
// scalar storage type (double for CANS_FP64, float for CANS_MIXED), bound to T
// by CANS_DISPATCH.
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

using namespace cans;

namespace {

bool same_n(const cans_ctx* c, const int* n) { return n && n[0] == c->L.n1 && n[1] == c->L.n2 && n[2] == c->L.n3; }

#define CHECK_CTX(name) \
  if (!c) return fail(CANS_EINVAL, std::string(name) + ": null context");
#define CHECK_N(name, n) \
  if (!same_n(c, n)) return fail(CANS_EINVAL, std::string(name) + ": n differs from the context grid");

// set_bc (bound.f90:182-463) for one face of one field -> plane op
PlaneOp make_setbc(void* f, char ctype, int ib, int idir, bool centered, double rvalue, double dr, int nd) {
  PlaneOp op;
  op.f = f;
  op.base = nullptr;
  op.fadd = nullptr;
  op.dir = idir;
  op.side = ib;
  op.nd = nd;
  op.factor = rvalue;
  op.sgn = 0.;
  if (ctype == 'P') {
    op.type = BT_PERIODIC;
  } else if (ctype == 'D' && centered) {
    op.type = BT_CENTERED;
    op.factor = 2. * rvalue;
    op.sgn = -1.;
  } else if (ctype == 'N') {
    op.type = centered ? BT_CENTERED : BT_STAG_N;
    op.factor = ib == 0 ? -dr * rvalue : dr * rvalue;
    op.sgn = 1.;
  } else {
    op.type = BT_STAG_D;
  }
  return op;
}

PlaneOp make_copy(void* f, int idir, int side, int nd) {
  PlaneOp op;
  op.f = f;
  op.base = nullptr;
  op.fadd = nullptr;
  op.dir = idir;
  op.side = side;
  op.nd = nd;
  op.type = BT_PERIODIC;
  op.factor = op.sgn = 0.;
  return op;
}

void push(PlaneOps& ops, const PlaneOp& op) { ops.op[ops.nops++] = op; }

// halo exchange of one field along a decomposed direction (bound.f90:600-713):
// on one GPU it is a periodic copy when the neighbour exists (periodic), else nothing
void push_exchange(PlaneOps& ops, void* f, int idir, int nd, int nb0, int nb1, int halo_side) {
  if (idir == 0) return;  // x is the pencil axis: no exchange
  bool up_lower = true, up_upper = true;
  if (halo_side == CANS_HALO_LOWER) up_upper = false;
  else if (halo_side == CANS_HALO_NONE) up_lower = up_upper = false;
  else if (halo_side == CANS_HALO_UPPER) up_lower = false;
  if (up_upper && nb1 >= 0) push(ops, make_copy(f, idir, 1, nd));
  if (up_lower && nb0 >= 0) push(ops, make_copy(f, idir, 0, nd));
}

// bounduvw on (u,v,w) (double), or, with base != nullptr, on base + (u,v,w) + fadd[comp]
// where u,v,w are increments of the scalar storage type (fadd: device vector of 3 or nullptr)
int plan_bounduvw(cans_ctx* c, const char* cbc, const double* bc, const int* nb, const int* is_bound,
                  const double* dl, const double* dzc, const double* dzf, void* u, void* v, void* w,
                  int keep_norm_values, int halos_norm, int halos_tang, const double* const* base = nullptr,
                  const double* fadd = nullptr) {
  const Layout& L = c->L;
  const int n[3] = {L.n1, L.n2, L.n3};
#define CBC(ib, d, cc) cbc[(ib) + 2 * (d) + 6 * (cc)]
#define BCV(ib, d, cc) bc[(ib) + 2 * (d) + 6 * (cc)]
  const bool keep_norm = keep_norm_values > 0;
  bool updt_norm[2] = {true, true};
  if (halos_norm == CANS_HALO_LOWER) updt_norm[1] = false;
  else if (halos_norm == CANS_HALO_NONE) updt_norm[0] = updt_norm[1] = false;
  else if (halos_norm == CANS_HALO_UPPER) updt_norm[0] = false;
  void* f[3] = {u, v, w};
  // The reference exchanges y then z halos first and applies set_bc x, y, z next.
  // Periodic copies commute with the other directions' plane updates (each
  // direction's update reads only interior-or-final values along that direction),
  // so the same final state is produced direction by direction: x, y, z.
  for (int d = 0; d < 3; d++) {
    PlaneOps ops;
    ops.nops = 0;
    for (int cc = 0; cc < 3; cc++)
      push_exchange(ops, f[cc], d, n[d], nb[2 * d], nb[2 * d + 1], cc == d ? halos_norm : halos_tang);
    const bool pp = CBC(0, d, d) == 'P' && CBC(1, d, d) == 'P';
    for (int ib = 0; ib < 2; ib++) {
      if (!is_bound[ib + 2 * d]) continue;
      const bool apply_norm = !keep_norm || (pp && updt_norm[ib]);
      for (int cc = 0; cc < 3; cc++) {
        const bool normal = cc == d;
        if (normal && !apply_norm) continue;
        double dr;
        if (d < 2) dr = dl[d];
        else dr = normal ? (ib == 0 ? dzf[0] : dzf[L.n3]) : (ib == 0 ? dzc[0] : dzc[L.n3]);
        push(ops, make_setbc(f[cc], CBC(ib, d, cc), ib, d, !normal, BCV(ib, d, cc), dr, n[d]));
      }
    }
    if (base) {
      for (int q = 0; q < ops.nops; q++)
        for (int cc = 0; cc < 3; cc++)
          if (ops.op[q].f == f[cc]) {
            ops.op[q].base = base[cc];
            ops.op[q].fadd = fadd ? fadd + cc : nullptr;
          }
    }
    int rc = base ? CANS_DISPATCH(c, launch_bound_dir<T>(c, ops)) : launch_bound_dir<double>(c, ops);
    if (rc) return rc;
  }
#undef CBC
#undef BCV
  return CANS_OK;
}

int plan_boundp(cans_ctx* c, const char* cbc, const double* bc, const int* nb, const int* is_bound, const double* dl,
                const double* dzc, void* p, int halo_side, int es = 0) {
  const Layout& L = c->L;
  const int n[3] = {L.n1, L.n2, L.n3};
  bool update_halo[2] = {true, true};
  if (halo_side == CANS_HALO_LOWER) update_halo[1] = false;
  else if (halo_side == CANS_HALO_NONE) update_halo[0] = update_halo[1] = false;
  else if (halo_side == CANS_HALO_UPPER) update_halo[0] = false;
  for (int d = 0; d < 3; d++) {
    PlaneOps ops;
    ops.nops = 0;
    push_exchange(ops, p, d, n[d], nb[2 * d], nb[2 * d + 1], halo_side);
    for (int ib = 0; ib < 2; ib++) {
      if (!(is_bound[ib + 2 * d] && update_halo[ib])) continue;
      const double dr = d < 2 ? dl[d] : (ib == 0 ? dzc[0] : dzc[L.n3]);
      push(ops, make_setbc(p, cbc[ib + 2 * d], ib, d, true, bc[ib + 2 * d], dr, n[d]));
    }
    int rc = es == (int)sizeof(double) ? launch_bound_dir<double>(c, ops)
             : es == (int)sizeof(float) ? launch_bound_dir<float>(c, ops)
                                        : CANS_DISPATCH(c, launch_bound_dir<T>(c, ops));
    if (rc) return rc;
  }
  return CANS_OK;
}

int do_updt_rhs_b(cans_ctx* c, const char* c_or_f, const char* cbc, const int* is_bound, const double* rx,
                  const double* ry, const double* rz, void* p, double norm, int es = 0) {
  const Layout& L = c->L;
  const int n[3] = {L.n1, L.n2, L.n3};
  const double* r[3] = {rx, ry, rz};
  for (int d = 0; d < 3; d++) {
    if (!r[d]) continue;
    const int q = (c_or_f[d] == 'f' && cbc[1 + 2 * d] == 'D') ? 1 : 0;
    const long plane = (long)(d == 0 ? L.n2 : L.n1) * (d == 2 ? L.n2 : L.n3);
    int rc;
    // es: element size of p (0: the context's pressure-like type)
    auto one = [&](const double* rr, int pl) -> int {
      if (es == (int)sizeof(double)) return launch_rhs_b<double>(c, (double*)p, rr, d, pl, norm);
      if (es == (int)sizeof(float)) return launch_rhs_b<float>(c, (float*)p, rr, d, pl, norm);
      return CANS_DISPATCH(c, launch_rhs_b<T>(c, (T*)p, rr, d, pl, norm));
    };
    if (is_bound[2 * d] && (rc = one(r[d], 1))) return rc;
    if (is_bound[2 * d + 1] && (rc = one(r[d] + plane, n[d] - q))) return rc;
  }
  return CANS_OK;
}

// three halo'd scratch fields of element type T (room for double), aligned like cans_field_alloc's;
// set 1 is the second set cans_rk3_step alternates with
template <typename T>
int ensure_tmp_fields(cans_ctx* c, T** t, int set = 0) {
  void* base = nullptr;
  const size_t fe = (size_t)c->L.sz * (c->L.n3 + 2);
  int rc = ensure_work(c, set ? 8 : 0, 3 * fe * sizeof(double) + 256, &base);
  if (rc) return rc;
  for (int k = 0; k < 3; k++) t[k] = (T*)((char*)base + 256 - sizeof(T)) + k * fe;
  return CANS_OK;
}

const double RKCOEFF[3][2] = {{32. / 60., 0.}, {25. / 60., -17. / 60.}, {45. / 60., -25. / 60.}};

}  // namespace

// momentum + RK update; the result goes to (uo,vo,wo) or, when eo is given, as
// increments over (u,v,w) to eo[0..2]
template <typename S>
static int rk_core(cans_ctx* c, const double* rkpar, const double* dli, const double* dzci, const double* dzfi,
                   const double* gvr_c, const double* gvr_f, double dt, double visc, const S* p,
                   const int* is_forced, const double* velf, const double* bforce, S* ru, S* rv, S* rw,
                   const double* u, const double* v, const double* w, double* uo, double* vo, double* wo,
                   int write_ro = 1, S* const* eo = nullptr, S* divu = nullptr, const double* sb = nullptr,
                   const double* gacc = nullptr, double beta = 0., const MomFold<S>* fold = nullptr) {
  if (c->rk_first) {  // rk.f90:72-100: RHS_old starts at zero on the first call
    c->rk_first = false;
    const size_t bytes = (size_t)c->L.cells() * sizeof(S);
    CANS_HIP_TRY(hipMemsetAsync(ru, 0, bytes, c->stream));
    CANS_HIP_TRY(hipMemsetAsync(rv, 0, bytes, c->stream));
    CANS_HIP_TRY(hipMemsetAsync(rw, 0, bytes, c->stream));
  }
  MomLaunch<S> m;
  m.u = u; m.v = v; m.w = w; m.p = p;
  m.fold = fold;
  m.uo = uo; m.vo = vo; m.wo = wo;
  if (eo) {
    m.eu = eo[0]; m.ev = eo[1]; m.ew = eo[2];
  }
  m.divu = divu;
  m.ru = ru; m.rv = rv; m.rw = rw;
  m.dzci = dzci; m.dzfi = dzfi; m.gvr_c = gvr_c; m.gvr_f = gvr_f;
  m.dxi = dli[0]; m.dyi = dli[1]; m.visc = visc; m.dt = dt;
  m.rkpar[0] = rkpar[0]; m.rkpar[1] = rkpar[1];
  m.read_ro = 1;
  m.write_ro = write_ro;
  m.imp = c->imp;
  for (int d = 0; d < 3; d++) {
    m.bforce[d] = bforce[d];
    m.velf[d] = velf[d];
    m.is_forced[d] = is_forced[d];
  }
  if (sb) {
    m.sb = sb;
    for (int d = 0; d < 3; d++) m.gacc[d] = gacc[d];
    m.beta = beta;
  }
  return launch_mom_rk<S>(c, m);
}

// rk_scal (rk.f90:335-491) of one scalar: the update goes to the double scratch t,
// f = scalf - <s> (or 0) to dscal[DSCAL_SCAL + slot]; then t -> s, plus f when
// apply_forcing (bulk_forcing_s, scal.f90:313-329, fused)
static int scal_core(cans_ctx* c, const double* rkpar, const double* dli, const double* dzci, const double* dzfi,
                     const double* gvr_f, double alpha, double dt, const double* u, const double* v,
                     const double* w, int is_forced, double scalf, double ssource, double* dsdtrko, double* s,
                     int slot, int apply_forcing) {
  double* t[3];
  int rc = ensure_tmp_fields<double>(c, t);
  if (rc) return rc;
  ScalLaunch m;
  m.u = u; m.v = v; m.w = w; m.s = s;
  m.so = t[0];
  m.ro = dsdtrko;
  m.dzci = dzci; m.dzfi = dzfi; m.gvr_f = gvr_f;
  m.dxi = dli[0]; m.dyi = dli[1]; m.alpha = alpha; m.dt = dt;
  m.source = ssource; m.scalf = scalf;
  m.rkpar[0] = rkpar[0]; m.rkpar[1] = rkpar[1];
  m.is_forced = is_forced ? 1 : 0;
  m.fslot = DSCAL_SCAL + slot;
  m.imp = c->imp;
  if ((rc = launch_scal_rk(c, m))) return rc;
  return launch_add_forcing<double>(c, t[0], s, DSCAL_SCAL + slot, apply_forcing && is_forced ? 1 : 0);
}

extern "C" int cans_rk(cans_ctx* c, const double* rkpar, const int* n, const double* dli, const double* dzci,
                       const double* dzfi, const double* gvr_c, const double* gvr_f, double dt, double visc,
                       const void* p, const int* is_forced, const double* velf, const double* bforce,
                       const double* gacc, double beta, const void* s, void* ru, void* rv, void* rw, void* u, void* v,
                       void* w, double* f_h) {
  CHECK_CTX("cans_rk");
  CHECK_N("cans_rk", n);
  if (!rkpar || !dli || !dzci || !dzfi || !gvr_c || !gvr_f || !p || !is_forced || !velf || !bforce || !ru || !rv ||
      !rw || !u || !v || !w || (s && !gacc))
    return fail(CANS_EINVAL, "cans_rk: null argument");
  int rc = CANS_DISPATCH(c, [&]() -> int {
    double* t[3];
    int r = ensure_tmp_fields<double>(c, t);
    if (r) return r;
    r = rk_core<T>(c, rkpar, dli, dzci, dzfi, gvr_c, gvr_f, dt, visc, (const T*)p, is_forced, velf, bforce, (T*)ru,
                   (T*)rv, (T*)rw, (const double*)u, (const double*)v, (const double*)w, t[0], t[1], t[2], 1,
                   nullptr, nullptr, (const double*)s, gacc, beta);
    if (r) return r;
    double* dst[3] = {(double*)u, (double*)v, (double*)w};
    for (int k = 0; k < 3; k++)
      if ((r = launch_add_forcing<double>(c, t[k], dst[k], k, is_forced[k] ? 1 : 0))) return r;
    return CANS_OK;
  }());
  if (rc) return rc;
  if (f_h) {
    CANS_HIP_TRY(hipMemcpyAsync(f_h, c->dscal, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CANS_HIP_TRY(hipStreamSynchronize(c->stream));
  }
  return CANS_OK;
}

extern "C" int cans_rk_scal(cans_ctx* c, const double* rkpar, const int* n, const double* dli, const double* l,
                            const double* dzci, const double* dzfi, const double* gvr_f, double alpha, double dt,
                            const int* is_bound, const void* u, const void* v, const void* w, int is_forced,
                            double scalf, double ssource, double* fluxo, void* dsdtrko, void* s, double* f_h) {
  (void)l;
  (void)is_bound;
  (void)fluxo;
  CHECK_CTX("cans_rk_scal");
  CHECK_N("cans_rk_scal", n);
  if (!rkpar || !dli || !dzci || !dzfi || !gvr_f || !u || !v || !w || !dsdtrko || !s)
    return fail(CANS_EINVAL, "cans_rk_scal: null argument");
  int rc = scal_core(c, rkpar, dli, dzci, dzfi, gvr_f, alpha, dt, (const double*)u, (const double*)v,
                     (const double*)w, is_forced, scalf, ssource, (double*)dsdtrko, (double*)s, 0, 0);
  if (rc) return rc;
  if (f_h) {
    CANS_HIP_TRY(hipMemcpyAsync(f_h, c->dscal + DSCAL_SCAL, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CANS_HIP_TRY(hipStreamSynchronize(c->stream));
  }
  return CANS_OK;
}

extern "C" int cans_bulk_forcing_s(cans_ctx* c, const int* n, int is_forced, double ff, void* s) {
  CHECK_CTX("cans_bulk_forcing_s");
  CHECK_N("cans_bulk_forcing_s", n);
  if (!s) return fail(CANS_EINVAL, "cans_bulk_forcing_s: null argument");
  if (!is_forced) return CANS_OK;
  // ff through a device slot, so the kernel is the one of the fused forcing
  int rc = launch_fill(c, c->dscal + DSCAL_SCAL, 1, ff);
  if (rc) return rc;
  return launch_add_forcing<double>(c, (const double*)s, (double*)s, DSCAL_SCAL, 1);
}

extern "C" int cans_fillps(cans_ctx* c, const int* n, const double* dli, const double* dzfi, double dti,
                           const void* u, const void* v, const void* w, void* p) {
  CHECK_CTX("cans_fillps");
  CHECK_N("cans_fillps", n);
  if (!dli || !dzfi || !u || !v || !w || !p) return fail(CANS_EINVAL, "cans_fillps: null argument");
  return CANS_DISPATCH(c, launch_fillps<T>(c, (const double*)u, (const double*)v, (const double*)w, (const T*)nullptr,
                                            (const T*)nullptr, (const T*)nullptr, nullptr, (T*)p, dzfi, dti,
                                            dli));
}

extern "C" int cans_correc(cans_ctx* c, const int* n, const double* dli, const double* dzci, double dt,
                           const void* p, void* u, void* v, void* w) {
  CHECK_CTX("cans_correc");
  CHECK_N("cans_correc", n);
  if (!dli || !dzci || !p || !u || !v || !w) return fail(CANS_EINVAL, "cans_correc: null argument");
  return CANS_DISPATCH(c, launch_correc<T>(c, (const T*)p, (const double*)u, (const double*)v, (const double*)w,
                                            (const T*)nullptr, (const T*)nullptr, (const T*)nullptr, nullptr,
                                            (double*)u, (double*)v, (double*)w, dzci, dt, dli, (T*)nullptr));
}

extern "C" int cans_updatep(cans_ctx* c, const int* n, const double* dli, const double* dzci, const double* dzfi,
                            double alpha, const void* pp, void* p, int is_impdiff) {
  CHECK_CTX("cans_updatep");
  CHECK_N("cans_updatep", n);
  if (!dli || !dzci || !dzfi || !pp || !p) return fail(CANS_EINVAL, "cans_updatep: null argument");
  // updatep.f90:16-76: with is_impdiff_1d (context numerics) only the z part of the Laplacian
  const int mode = is_impdiff ? (c->imp == 2 ? 2 : 1) : 0;
  return CANS_DISPATCH(c, launch_updatep<T>(c, (const T*)pp, (T*)p, dzci, dzfi, alpha, dli, mode));
}

extern "C" int cans_bounduvw(cans_ctx* c, const char* cbc, const int* n, const double* bc, const int* nb,
                             const int* is_bound, const double* dl, const double* dzc, const double* dzf, void* u,
                             void* v, void* w, int keep_norm, int halos_norm, int halos_tang) {
  CHECK_CTX("cans_bounduvw");
  CHECK_N("cans_bounduvw", n);
  if (!cbc || !bc || !nb || !is_bound || !dl || !dzc || !dzf || !u || !v || !w)
    return fail(CANS_EINVAL, "cans_bounduvw: null argument");
  return plan_bounduvw(c, cbc, bc, nb, is_bound, dl, dzc, dzf, u, v, w, keep_norm, halos_norm, halos_tang);
}

extern "C" int cans_boundp(cans_ctx* c, const char* cbc, const int* n, const double* bc, const int* nb,
                           const int* is_bound, const double* dl, const double* dzc, void* p, int halo_side) {
  CHECK_CTX("cans_boundp");
  CHECK_N("cans_boundp", n);
  if (!cbc || !bc || !nb || !is_bound || !dl || !dzc || !p) return fail(CANS_EINVAL, "cans_boundp: null argument");
  // a field from cans_field_alloc_kind carries its element type (a scalar is FP64 in
  // every mode); other pointers are taken as pressure-like
  return plan_boundp(c, cbc, bc, nb, is_bound, dl, dzc, p, halo_side, field_elem_size(c, p));
}

extern "C" int cans_updt_rhs_b(cans_ctx* c, const char* c_or_f, const char* cbc, const int* n, const int* is_bound,
                               const double* rx, const double* ry, const double* rz, void* p, const double* alpha) {
  CHECK_CTX("cans_updt_rhs_b");
  CHECK_N("cans_updt_rhs_b", n);
  if (!c_or_f || !cbc || !is_bound || !p) return fail(CANS_EINVAL, "cans_updt_rhs_b: null argument");
  return do_updt_rhs_b(c, c_or_f, cbc, is_bound, rx, ry, rz, p, alpha ? *alpha : 1.);
}

extern "C" int cans_solver(cans_ctx* c, const int* n, const int* ng, const cans_plan* plan, double normfft,
                           const double* lam, const double* a, const double* b, const double* cc, const char* bc,
                           const char* c_or_f, void* p) {
  CHECK_CTX("cans_solver");
  CHECK_N("cans_solver", n);
  CHECK_N("cans_solver", ng);
  if (!plan || !lam || !a || !b || !cc || !bc || !c_or_f || !p) return fail(CANS_EINVAL, "cans_solver: null argument");
  return CANS_DISPATCH(c, run_solver<T>(c, plan, normfft, lam, a, b, cc, bc, c_or_f, (T*)p));
}

// cans_solver on the fixed device scratch: the same passes, their work buffer and
// checkpoints carved from this device's array; no allocation on any path of the call
extern "C" int cans_solver_noalloc(cans_ctx* c, const int* n, const int* ng, const cans_plan* plan, double normfft,
                                   const double* lam, const double* a, const double* b, const double* cc,
                                   const char* bc, const char* c_or_f, void* p) {
  CHECK_CTX("cans_solver_noalloc");
  CHECK_N("cans_solver_noalloc", n);
  CHECK_N("cans_solver_noalloc", ng);
  if (!plan || !lam || !a || !b || !cc || !bc || !c_or_f || !p)
    return fail(CANS_EINVAL, "cans_solver_noalloc: null argument");
  if (c->dtdma > 1) return fail(CANS_EUNSUPPORTED, "cans_solver_noalloc: the distributed z solve is not covered");
  const bool per = bc[4] == 'P' && bc[5] == 'P';
  const size_t es = c->precision == CANS_MIXED ? sizeof(float) : sizeof(double);
  const size_t need = solver_work_bytes(c, es, per);
  if (need > (size_t)CANS_NOALLOC_SCRATCH_BYTES)
    return fail(CANS_EINVAL, "cans_solver_noalloc: this grid's solve needs " + std::to_string(need) +
                                 " bytes of scratch, above the fixed " +
                                 std::to_string((long)CANS_NOALLOC_SCRATCH_BYTES));
  void* base = noalloc_scratch(c->device);
  if (!base) return fail(CANS_EHIP, "cans_solver_noalloc: no scratch address for the context's device");
  const SolveWork ws{base, (size_t)CANS_NOALLOC_SCRATCH_BYTES};
  return CANS_DISPATCH(c, run_solver<T>(c, plan, normfft, lam, a, b, cc, bc, c_or_f, (T*)p, nullptr, &ws));
}

extern "C" int cans_solve_helmholtz(cans_ctx* c, const int* n, const int* ng, const cans_plan* plan, double normfft,
                                    double alpha, const double* lam, const double* a, const double* b,
                                    const double* cc, const double* rx, const double* ry, const double* rz,
                                    const int* is_bound, const char* cbc, const char* c_or_f, void* p) {
  CHECK_CTX("cans_solve_helmholtz");
  CHECK_N("cans_solve_helmholtz", n);
  if (!a || !b || !cc || !is_bound || !cbc || !c_or_f || !p)
    return fail(CANS_EINVAL, "cans_solve_helmholtz: null argument");
  // the field's element type: a registered field's own, else the context's pressure-like type
  int es = field_elem_size(c, p);
  if (!es) es = c->precision == CANS_MIXED ? (int)sizeof(float) : (int)sizeof(double);
  // solve_helmholtz.f90:58-73: rhs boundary terms scaled by alpha, bb = b + 1/alpha, norm/alpha
  int rc = do_updt_rhs_b(c, c_or_f, cbc, is_bound, rx, ry, rz, p, alpha, es);
  if (rc) return rc;
  if (c->imp == 2) {  // is_impdiff_1d: solver_gaussel_z
    if (es != (int)sizeof(double)) return fail(CANS_EUNSUPPORTED, "cans_solve_helmholtz: z-only form needs an FP64 field");
    return run_helmholtz_z(c, a, b, cc, alpha, cbc + 4, c_or_f[2], (double*)p);
  }
  if (!plan || !lam) return fail(CANS_EINVAL, "cans_solve_helmholtz: null plan or eigenvalues");
  if (es == (int)sizeof(double))
    return run_helmholtz_3d<double>(c, plan, normfft, alpha, lam, a, b, cc, cbc, c_or_f, (double*)p);
  return run_helmholtz_3d<float>(c, plan, normfft, alpha, lam, a, b, cc, cbc, c_or_f, (float*)p);
}

static int helmholtz(cans_ctx* c, int n, const cans_helmholtz* const* h, const int* is_bound, const double* alpha,
                     double* const* f);

extern "C" int cans_solve_helmholtz_batch(cans_ctx* c, int nsys, const cans_helmholtz* h, const double* alpha,
                                          const int* is_bound, void* const* f) {
  CHECK_CTX("cans_solve_helmholtz_batch");
  if (nsys < 1 || nsys > CANS_MAX_SCALARS + 3 || !h || !alpha || !is_bound || !f)
    return fail(CANS_EINVAL, "cans_solve_helmholtz_batch: bad argument");
  const cans_helmholtz* hp[CANS_MAX_SCALARS + 3];
  double* fp[CANS_MAX_SCALARS + 3];
  for (int q = 0; q < nsys; q++) {
    if (!f[q] || !h[q].a_d || !h[q].b_d || !h[q].c_d)
      return fail(CANS_EINVAL, "cans_solve_helmholtz_batch: null field or coefficients");
    if (field_elem_size(c, f[q]) == (int)sizeof(float))
      return fail(CANS_EINVAL, "cans_solve_helmholtz_batch: the fields are FP64");
    hp[q] = &h[q];
    fp[q] = (double*)f[q];
  }
  return helmholtz(c, nsys, hp, is_bound, alpha, fp);
}

extern "C" int cans_chkdiv(cans_ctx* c, const int* lo, const int* hi, const double* l, const double* dli,
                           const double* dzfi, const void* u, const void* v, const void* w, double* divtot,
                           double* divmax) {
  CHECK_CTX("cans_chkdiv");
  if (!lo || !hi || !l || !dli || !dzfi || !u || !v || !w || !divtot || !divmax)
    return fail(CANS_EINVAL, "cans_chkdiv: null argument");
  const int n[3] = {hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1};
  CHECK_N("cans_chkdiv", n);
  double tot, mx;
  int rc = run_chkdiv<double>(c, (const double*)u, (const double*)v, (const double*)w, dzfi, dli, &tot, &mx);
  if (rc) return rc;
  *divtot = tot / (l[0] * l[1] * l[2]);
  *divmax = mx;
  return CANS_OK;
}

extern "C" int cans_chkdt(cans_ctx* c, const int* n, const double* dl, const double* dzci, const double* dzfi,
                          double visc, double alpha, const void* u, const void* v, const void* w, int is_impdiff,
                          int is_impdiff_1d, double* dtmax) {
  CHECK_CTX("cans_chkdt");
  CHECK_N("cans_chkdt", n);
  if (!dl || !dzci || !dzfi || !u || !v || !w || !dtmax) return fail(CANS_EINVAL, "cans_chkdt: null argument");
  double dti;
  int rc = run_chkdt<double>(c, (const double*)u, (const double*)v, (const double*)w, dzci, dzfi, dl, &dti);
  if (rc) return rc;
  // dlmin over dl(1:2) and 1/dzfi(0:n3+1) (chkdt.f90:37-44)
  std::vector<double> hz(c->L.n3 + 2);
  CANS_HIP_TRY(hipMemcpy(hz.data(), dzfi, hz.size() * sizeof(double), hipMemcpyDeviceToHost));
  double dlmin = fmin(dl[0], dl[1]);
  if (!is_impdiff_1d)
    for (double x : hz) dlmin = fmin(dlmin, 1. / x);
  if (dti < 2.220446049250313e-16) dti = 1.;
  const double sqrt3 = (double)sqrtf(3.f);    // sqrt(3.) is a default-real expression
  const double c165 = (double)(1.65f / 12.f);  // so is 1.65/12.
  if (is_impdiff && !is_impdiff_1d) *dtmax = sqrt3 / dti;
  else *dtmax = fmin(c165 / fmax(visc, alpha) * (dlmin * dlmin), sqrt3 / dti);
  return CANS_OK;
}

extern "C" int cans_bulk_mean(cans_ctx* c, const int* n, const double* gvr, const void* p, double* mean) {
  CHECK_CTX("cans_bulk_mean");
  CHECK_N("cans_bulk_mean", n);
  if (!gvr || !p || !mean) return fail(CANS_EINVAL, "cans_bulk_mean: null argument");
  return run_bulk_mean<double>(c, (const double*)p, gvr, mean);
}

// solve_helmholtz (solve_helmholtz.f90:28-75) of n FP64 fields f[q] with the systems h[q]:
// each system's boundary rhs, then one batched solve (the passes carry all systems)
static int helmholtz(cans_ctx* c, int n, const cans_helmholtz* const* h, const int* is_bound, const double* alpha,
                     double* const* f) {
  const bool z_only = c->imp == 2;
  HelmSys hs[CANS_MAX_SCALARS + 3];
  for (int q = 0; q < n; q++) {
    // the z-only form has no x/y boundary planes (main.f90:317-329 deallocates them)
    int rc = do_updt_rhs_b(c, h[q]->c_or_f, h[q]->cbc, is_bound, z_only ? nullptr : h[q]->rhsbx_d,
                           z_only ? nullptr : h[q]->rhsby_d, h[q]->rhsbz_d, f[q], alpha[q], (int)sizeof(double));
    if (rc) return rc;
    if (!z_only && (!h[q]->plan || !h[q]->lambdaxy_d))
      return fail(CANS_EINVAL, "cans_rk3_substep: Helmholtz system without a plan");
    hs[q] = HelmSys{h[q]->plan, h[q]->normfft, h[q]->lambdaxy_d, h[q]->a_d, h[q]->b_d, h[q]->c_d,
                    h[q]->cbc, h[q]->c_or_f, f[q], alpha[q]};
  }
  return run_helmholtz_batch(c, n, hs);
}

// One substep.  With S = double (FP64) the updated velocity of rk lives in double
// scratch fields t, as in the reference.  With S = float (mixed) the scratch holds
// its increment e over (u,v,w) in FP32 instead: e is ~dt times the right-hand side,
// so its rounding is far below that of the velocity, which stays FP64 throughout.
// The intermediate velocity is then u + e + f, with the bulk forcing f kept as a
// device scalar (no forcing pass), and the momentum kernel also leaves div(u), so
// the first transform pass forms div(u + e + f) = div(u) + div(e) without reading u.
// With implicit diffusion the mixed mode keeps the FP64 form of the velocity (the forced
// intermediate velocity is the rhs of its Helmholtz systems, solved in FP64 on the FP64
// velocity fields); the pressure side (p, pp, RHS_old, the Poisson solve) stays FP32.
// The boundary rules the folded projection needs (MomFold): which directions are periodic,
// and, at the faces of the others, set_bc's centred rule for the tangential components
// (bound.f90:193-206; make_setbc's factor and sign)
template <typename S>
static void fold_rules(const cans_step_params* s, const cans_ctx* c, MomFold<S>& f) {
  for (int d = 0; d < 3; d++) f.periodic[d] = s->cbcvel[2 * d + 6 * d] == 'P' && s->cbcvel[1 + 2 * d + 6 * d] == 'P';
  for (int cc = 0; cc < 3; cc++)
    for (int d = 0; d < 3; d++)
      for (int ib = 0; ib < 2; ib++) {
        const char t = s->cbcvel[ib + 2 * d + 6 * cc];
        const double bv = s->bcvel[ib + 2 * d + 6 * cc];
        const double dr = d < 2 ? s->dl[d] : (ib == 0 ? s->dzc_h[0] : s->dzc_h[c->L.n3]);
        if (t == 'D') {
          f.factor[cc][d][ib] = 2. * bv;
          f.sgn[cc][d][ib] = -1.;
        } else {
          f.factor[cc][d][ib] = ib == 0 ? -dr * bv : dr * bv;
          f.sgn[cc][d][ib] = 1.;
        }
      }
}

// Folding one substep's projection into the next (cans_rk3_step): a deferring substep
// stops after its pressure solve (and boundp of pp), leaving its intermediate velocity in
// scratch set `set`; the next substep's momentum kernel then forms the projected velocity
// and p + pp while it stages its tiles (MomFold), reading its intermediate velocity from
// the other set and writing p + pp to the other pressure buffer.
struct StepFold {
  bool defer = false;    // this substep leaves its projection to the next one
  bool fold_in = false;  // this substep folds the previous substep's projection
  int set = 0;           // scratch set of this substep's intermediate velocity
  double dtrk_in = 0.;   // the folded substep's dtrk
  void* p_in = nullptr;  // the pressure buffer holding p^(k-1)
  void* p_out = nullptr; // fold_in: the buffer that receives p^(k-1) + pp
};

template <typename S>
static int rk3_substep(cans_ctx* c, const cans_step_params* s, int irk, double dt, cans_state* st,
                       const StepFold* fo = nullptr) {
  const bool mixed = sizeof(S) != sizeof(double);
  const bool delta = mixed && !c->imp && c->mixed_incr;
  const double* rkpar = RKCOEFF[irk - 1];
  const double dtrk = (rkpar[0] + rkpar[1]) * dt;  // main.f90:438-439
  const double dtrki = 1. / dtrk;
  double* u = (double*)st->u;
  double* v = (double*)st->v;
  double* w = (double*)st->w;
  S* p = (S*)st->p;
  S* pp = (S*)st->pp;
  const bool defer = fo && fo->defer, fold_in = fo && fo->fold_in;
  const int set = fo ? fo->set : 0;
  double* t[3];
  S* e[3];
  int rc = ensure_tmp_fields<double>(c, t, set);
  if (rc) return rc;
  if ((rc = ensure_tmp_fields<S>(c, e, set))) return rc;  // same scratch, viewed as S
  // the pressure buffers: p^(k-1) is read from pin; with fold_in, p^(k-1) + pp goes to pnew
  // (the other buffer) and the substep continues with it
  S* pin = (S*)(fo ? fo->p_in : st->p);
  S* pnew = pin;
  MomFold<S> mf;
  const double* sbase[3] = {u, v, w};  // the velocity the momentum kernel reads
  if (fold_in) {
    double* tp[3];
    S* ep[3];
    if ((rc = ensure_tmp_fields<double>(c, tp, 1 - set))) return rc;
    if ((rc = ensure_tmp_fields<S>(c, ep, 1 - set))) return rc;
    pnew = (S*)fo->p_out;
    mf.pp = pp;
    mf.eu = delta ? ep[0] : nullptr;
    mf.ev = delta ? ep[1] : nullptr;
    mf.ew = delta ? ep[2] : nullptr;
    mf.f = c->dscal;
    mf.dt = fo->dtrk_in;
    mf.pout = pnew;
    fold_rules<S>(s, c, mf);
    if (!delta) {  // FP64: the stencil reads the folded substep's intermediate velocity
      sbase[0] = tp[0];
      sbase[1] = tp[1];
      sbase[2] = tp[2];
    }
  }
  p = pin;
  const double* base[3] = {u, v, w};
  // bulk forcing (rk writes it to dscal[0..2], 0 for unforced components): added where the
  // intermediate velocity is read (divergence, correc) rather than stored, except before the
  // implicit-diffusion solve, which needs the forced field.  Forcing acts only along periodic
  // directions (sanity.f90:chk_forcing), so the ghost cells that bounduvw fills from the unforced
  // field and the divergence reads (u_d along d) differ from the forced ones by exactly f; the
  // other ghosts correc reads are rewritten by the final bounduvw.
  const double* fvec = (delta || !c->imp) ? c->dscal : nullptr;
  // With homogeneous pressure BCs and a periodic x transform the divergence is formed
  // inside the first transform pass (fillps -> updt_rhs_b -> solver otherwise)
  const bool fuse_div = !(s->rhsbx_d || s->rhsby_d || s->rhsbz_d) && s->cbcpre[0] == 'P' && s->cbcpre[1] == 'P' &&
                        c->L.n1 >= 2 && c->L.n1 % 2 == 0 && plan_fft_x(s->plan);
  void* divu_v = nullptr;
  if (delta && fuse_div && (rc = ensure_work(c, 4, (size_t)c->L.cells() * sizeof(S), &divu_v))) return rc;
  S* divu = (S*)divu_v;
  // refinement sweeps of this substep's pressure solve (cans_ctx_set_refinement; mixed mode):
  // their corrections live in two halo'd fields of S, each field's element 1 256-byte aligned
  const int nsweep = (mixed && ((c->refine_mask >> (irk - 1)) & 1)) ? c->refine_sweeps : 0;
  S* corr[2] = {nullptr, nullptr};
  if (nsweep) {
    if (s->rhsbx_d || s->rhsby_d || s->rhsbz_d)
      return fail(CANS_EUNSUPPORTED, "cans_rk3_substep: refinement needs homogeneous pressure boundary values");
    const size_t fe = ((size_t)c->L.sz * (c->L.n3 + 2) + 63) & ~(size_t)63;
    void* cb = nullptr;
    if ((rc = ensure_work(c, 7, 2 * fe * sizeof(S) + 256, &cb))) return rc;
    corr[0] = (S*)((char*)cb + 256 - sizeof(S));
    corr[1] = corr[0] + fe;
  }
  // scalars first, from the old velocity (main.f90:440-450): rk_scal, bulk_forcing_s,
  // [solve_helmholtz], boundp.  A scalar's step reads only the velocity and itself, so
  // all of them run rk_scal first and then share one batched Helmholtz solve.
  if (s->nscal > CANS_MAX_SCALARS) return fail(CANS_EINVAL, "cans_rk3_substep: too many scalars");
  const cans_helmholtz* hsc[CANS_MAX_SCALARS];
  double asc[CANS_MAX_SCALARS];
  double* fsc[CANS_MAX_SCALARS];
  for (int i = 0; i < s->nscal; i++) {
    const cans_scalar& sc = s->scal[i];
    if ((rc = scal_core(c, rkpar, s->dli, s->dzci_d, s->dzfi_d, s->grid_vol_ratio_f_d, sc.alpha, dt, u, v, w,
                        sc.is_forced, sc.scalf, sc.source, (double*)sc.dsdtrko, (double*)sc.val, i, 1)))
      return rc;
    hsc[i] = &sc.h;
    asc[i] = -0.5 * sc.alpha * dtrk;
    fsc[i] = (double*)sc.val;
  }
  if (c->imp && s->nscal && (rc = helmholtz(c, s->nscal, hsc, s->is_bound, asc, fsc))) return rc;
  for (int i = 0; i < s->nscal; i++) {
    const cans_scalar& sc = s->scal[i];
    if ((rc = plan_boundp(c, sc.cbc, sc.bc, s->nb, s->is_bound, s->dl, s->dzc_h, sc.val, CANS_HALO_ABSENT,
                          (int)sizeof(double))))
      return rc;
  }
  const double* sb = (s->is_buoyancy && s->nscal > 0) ? (const double*)s->scal[0].val : nullptr;
  // rk: updated velocity (or its increment) into the scratch; u,v,w keep the old level for the stencil
  // (a folding substep keeps div(u) of the step's initial velocity, written by substep 1: its
  // increments are over that velocity)
  rc = rk_core<S>(c, rkpar, s->dli, s->dzci_d, s->dzfi_d, s->grid_vol_ratio_c_d, s->grid_vol_ratio_f_d, dt,
                  s->visc, p, s->is_forced, s->velf, s->bforce, (S*)st->dudtrko, (S*)st->dvdtrko, (S*)st->dwdtrko,
                  sbase[0], sbase[1], sbase[2], t[0], t[1], t[2], irk != 3, delta ? e : nullptr,
                  fold_in ? nullptr : divu, sb, s->gacc, s->beta, fold_in ? &mf : nullptr);
  if (rc) return rc;
  if (fold_in) {  // p + pp of the folded substep: its upper ghosts, as updatep + boundp leave them
    p = pnew;
    if ((rc = plan_boundp(c, s->cbcpre, s->bcpre, s->nb, s->is_bound, s->dl, s->dzc_h, p, CANS_HALO_UPPER)))
      return rc;
  }
  for (int k = 0; k < 3 && !fvec; k++)
    if (s->is_forced[k] && (rc = launch_add_forcing<double>(c, t[k], t[k], k, 1))) return rc;
  // implicit diffusion (main.f90:453-462): the three components' Helmholtz systems, alpha = -visc*dtrk/2,
  // solved as one batch
  if (c->imp) {
    const cans_helmholtz* hv[3] = {&s->hvel[0], &s->hvel[1], &s->hvel[2]};
    const double av = -.5 * s->visc * dtrk;
    const double aa3[3] = {av, av, av};
    if ((rc = helmholtz(c, 3, hv, s->is_bound, aa3, t))) return rc;
  }
  if (delta)
    rc = plan_bounduvw(c, s->cbcvel, s->bcvel, s->nb, s->is_bound, s->dl, s->dzc_h, s->dzf_h, e[0], e[1], e[2], 0,
                       CANS_HALO_LOWER, CANS_HALO_ABSENT, base, fvec);
  else
    rc = plan_bounduvw(c, s->cbcvel, s->bcvel, s->nb, s->is_bound, s->dl, s->dzc_h, s->dzf_h, t[0], t[1], t[2], 0,
                       CANS_HALO_LOWER, CANS_HALO_ABSENT);
  if (rc) return rc;
  // the intermediate velocity: t, or (u,v,w) + increments e + forcing f
  const double* su = delta ? u : t[0];
  const double* sv = delta ? v : t[1];
  const double* sw = delta ? w : t[2];
  const S* eu = delta ? e[0] : nullptr;
  const S* ev = delta ? e[1] : nullptr;
  const S* ew = delta ? e[2] : nullptr;
  if (fuse_div) {
    DivSrc<S> ds = {su, sv, sw, eu, ev, ew, divu, s->dzfi_d, dtrki, dtrki * s->dli[0], dtrki * s->dli[1],
                    s->dli[0], s->dli[1]};
    if (!delta) ds.f = fvec;
    if ((rc = run_solver<S>(c, s->plan, s->normfft, s->lambdaxy_d, s->a_d, s->b_d, s->c_d, s->cbcpre, "ccc", pp, &ds)))
      return rc;
  } else {
    if ((rc = launch_fillps<S>(c, su, sv, sw, eu, ev, ew, fvec, pp, s->dzfi_d, dtrki, s->dli))) return rc;
    if (s->rhsbx_d || s->rhsby_d || s->rhsbz_d)
      if ((rc = do_updt_rhs_b(c, "ccc", s->cbcpre, s->is_bound, s->rhsbx_d, s->rhsby_d, s->rhsbz_d, pp, 1.)))
        return rc;
    if ((rc = run_solver<S>(c, s->plan, s->normfft, s->lambdaxy_d, s->a_d, s->b_d, s->c_d, s->cbcpre, "ccc", pp)))
      return rc;
  }
  if ((rc = plan_boundp(c, s->cbcpre, s->bcpre, s->nb, s->is_bound, s->dl, s->dzc_h, pp, CANS_HALO_ABSENT))) return rc;
  // Refinement sweeps: solve L x_q = b - L(pp + x_1 + ... + x_{q-1}) with FP32 storage, the
  // residual formed in FP64 from the velocity (b is fillps's rhs) and the boundp'd corrections
  // so far.  correc then projects with pp + x_1 + x_2, so div(u) after this substep is dtrk times
  // the residual left by the last sweep; pp keeps the first solve's part only.
  // The residual's b is formed from the velocity itself (u + e + f, FP64), not from the stored
  // FP32 div(u) of the step's initial velocity, whose rounding the sweeps could not remove.
  for (int q = 1; q <= nsweep; q++) {
    S* xq = corr[q - 1];
    if (fuse_div) {
      DivSrc<S> ds = {su, sv, sw, eu, ev, ew, nullptr, s->dzfi_d, dtrki, dtrki * s->dli[0], dtrki * s->dli[1],
                      s->dli[0], s->dli[1]};
      ds.f = fvec;
      ds.corr[0] = pp;
      ds.corr[1] = q > 1 ? corr[0] : nullptr;
      ds.ncorr = q;
      ds.dzci = s->dzci_d;
      ds.dxi2 = s->dli[0] * s->dli[0];
      ds.dyi2 = s->dli[1] * s->dli[1];
      if ((rc = run_solver<S>(c, s->plan, s->normfft, s->lambdaxy_d, s->a_d, s->b_d, s->c_d, s->cbcpre, "ccc", xq,
                              &ds)))
        return rc;
    } else {
      if ((rc = launch_fillps_residual<S>(c, su, sv, sw, eu, ev, ew, fvec, xq, s->dzfi_d, dtrki, s->dli, pp,
                                          q > 1 ? corr[0] : nullptr, s->dzci_d)))
        return rc;
      if ((rc = run_solver<S>(c, s->plan, s->normfft, s->lambdaxy_d, s->a_d, s->b_d, s->c_d, s->cbcpre, "ccc", xq)))
        return rc;
    }
    if ((rc = plan_boundp(c, s->cbcpre, s->bcpre, s->nb, s->is_bound, s->dl, s->dzc_h, xq, CANS_HALO_ABSENT)))
      return rc;
  }
  if (defer) return CANS_OK;  // the next substep's momentum kernel projects (StepFold)
  // correc reads the intermediate velocity and writes the caller's fields (whole halo'd
  // box); updatep (p += pp) is independent of the velocity halos and rides along
  // with implicit diffusion updatep also adds alpha*lap(pp) (updatep.f90:16-33): its own pass
  if ((rc = launch_correc<S>(c, pp, su, sv, sw, eu, ev, ew, fvec, u, v, w, s->dzci_d, dtrk, s->dli,
                             c->imp ? nullptr : p, nsweep > 0 ? corr[0] : nullptr, nsweep > 1 ? corr[1] : nullptr)))
    return rc;
  if (c->imp && (rc = launch_updatep<S>(c, pp, p, s->dzci_d, s->dzfi_d, -.5 * s->visc * dtrk, s->dli, c->imp)))
    return rc;
  if ((rc = plan_bounduvw(c, s->cbcvel, s->bcvel, s->nb, s->is_bound, s->dl, s->dzc_h, s->dzf_h, u, v, w, 1,
                          CANS_HALO_UPPER, CANS_HALO_NONE)))
    return rc;
  return plan_boundp(c, s->cbcpre, s->bcpre, s->nb, s->is_bound, s->dl, s->dzc_h, p, CANS_HALO_UPPER);
}

extern "C" int cans_ctx_set_refinement(cans_ctx* c, int sweeps, int substep_mask) {
  CHECK_CTX("cans_ctx_set_refinement");
  if (sweeps < 0 || sweeps > 2) return fail(CANS_EINVAL, "cans_ctx_set_refinement: sweeps must be 0, 1 or 2");
  if (substep_mask < 0 || substep_mask > 7) return fail(CANS_EINVAL, "cans_ctx_set_refinement: substep_mask is 3 bits");
  if (sweeps && c->precision != CANS_MIXED)
    return fail(CANS_EINVAL, "cans_ctx_set_refinement: refinement applies to CANS_MIXED contexts");
  c->refine_sweeps = sweeps;
  c->refine_mask = sweeps ? substep_mask : 0;
  return CANS_OK;
}

extern "C" int cans_ctx_set_mixed_increments(cans_ctx* c, int on) {
  CHECK_CTX("cans_ctx_set_mixed_increments");
  if (on && c->precision != CANS_MIXED)
    return fail(CANS_EINVAL, "cans_ctx_set_mixed_increments: increments apply to CANS_MIXED contexts");
  c->mixed_incr = on != 0;
  return CANS_OK;
}

extern "C" int cans_rk3_substep(cans_ctx* c, const cans_step_params* s, int irk, double dt, cans_state* st) {
  CHECK_CTX("cans_rk3_substep");
  if (!s || !st || irk < 1 || irk > 3) return fail(CANS_EINVAL, "cans_rk3_substep: bad argument");
  if (!st->u || !st->v || !st->w || !st->p || !st->pp || !st->dudtrko || !st->dvdtrko || !st->dwdtrko)
    return fail(CANS_EINVAL, "cans_rk3_substep: null field");
  if (!s->plan || !s->lambdaxy_d || !s->a_d || !s->b_d || !s->c_d || !s->dzci_d || !s->dzfi_d || !s->dzc_h ||
      !s->dzf_h || !s->grid_vol_ratio_c_d || !s->grid_vol_ratio_f_d)
    return fail(CANS_EINVAL, "cans_rk3_substep: null parameter array");
  return CANS_DISPATCH(c, rk3_substep<T>(c, s, irk, dt, st));
}

// the three substeps of one step; substep k's projection (correc, bounduvw, updatep,
// boundp of p) is folded into substep k+1's momentum kernel for k = 1, 2 where that
// applies: explicit diffusion, no scalars, homogeneous pressure boundary values, and no
// refinement of substeps 1 and 2.  The state between the substeps lives in the context's
// scratch (two sets of intermediate fields and a second pressure field); after the call
// st holds what three cans_rk3_substep calls leave, bit for bit.
template <typename S>
static int rk3_step(cans_ctx* c, const cans_step_params* s, double dt, cans_state* st) {
  bool fold = !c->imp && s->nscal == 0 && !(s->rhsbx_d || s->rhsby_d || s->rhsbz_d) && !(c->refine_mask & 3);
  for (int q = 0; q < 6 && fold; q++) fold = s->bcpre[q] == 0. || s->cbcpre[q] == 'P';
  if (!fold) {
    for (int irk = 1; irk <= 3; irk++) {
      int rc = rk3_substep<S>(c, s, irk, dt, st);
      if (rc) return rc;
    }
    return CANS_OK;
  }
  // the second pressure buffer: a halo'd field of S, element 1 on a 256-byte boundary
  void* pb = nullptr;
  const size_t fe = (size_t)c->L.sz * (c->L.n3 + 2);
  int rc = ensure_work(c, 9, fe * sizeof(S) + 256, &pb);
  if (rc) return rc;
  void* p2 = (char*)pb + 256 - sizeof(S);
  const double* rkp = RKCOEFF[0];
  StepFold f1;  // substep 1: intermediate velocity in set 0, projection deferred
  f1.defer = true;
  f1.set = 0;
  f1.p_in = st->p;
  if ((rc = rk3_substep<S>(c, s, 1, dt, st, &f1))) return rc;
  StepFold f2;  // substep 2: folds substep 1 (set 0, p -> p2), intermediate in set 1, deferred
  f2.defer = true;
  f2.fold_in = true;
  f2.set = 1;
  f2.dtrk_in = (rkp[0] + rkp[1]) * dt;
  f2.p_in = st->p;
  f2.p_out = p2;
  if ((rc = rk3_substep<S>(c, s, 2, dt, st, &f2))) return rc;
  rkp = RKCOEFF[1];
  StepFold f3;  // substep 3: folds substep 2 (set 1, p2 -> p), intermediate in set 0, projects
  f3.fold_in = true;
  f3.set = 0;
  f3.dtrk_in = (rkp[0] + rkp[1]) * dt;
  f3.p_in = p2;
  f3.p_out = st->p;
  return rk3_substep<S>(c, s, 3, dt, st, &f3);
}

extern "C" int cans_rk3_step(cans_ctx* c, const cans_step_params* s, double dt, cans_state* st) {
  CHECK_CTX("cans_rk3_step");
  if (!s || !st) return fail(CANS_EINVAL, "cans_rk3_step: bad argument");
  if (!st->u || !st->v || !st->w || !st->p || !st->pp || !st->dudtrko || !st->dvdtrko || !st->dwdtrko)
    return fail(CANS_EINVAL, "cans_rk3_step: null field");
  if (!s->plan || !s->lambdaxy_d || !s->a_d || !s->b_d || !s->c_d || !s->dzci_d || !s->dzfi_d || !s->dzc_h ||
      !s->dzf_h || !s->grid_vol_ratio_c_d || !s->grid_vol_ratio_f_d)
    return fail(CANS_EINVAL, "cans_rk3_step: null parameter array");
  return CANS_DISPATCH(c, rk3_step<T>(c, s, dt, st));
}

// ---------------------------------------------------------------------------
// Adjoint of the RK3 substep (no counterpart in the reference; for gradients of a
// loss on the final velocity).  Scope: FP64 contexts, tri-periodic uniform grids,
// explicit diffusion, no forcing, scalars or buoyancy.  There the substep maps
// (u, Fo) to (P(u + f1 F(u) + f2 Fo), F(u)) (main.f90:437-473, rk.f90:24-333), and
// P = I - G L^-1 D is symmetric (G = -D^T), so its adjoint is
//   wb = P ub,   ub <- wb + J(u)^T (f1 wb + fb),   fb <- f2 wb.
// ---------------------------------------------------------------------------
namespace {

// the directions bounded by walls, bit 0 x, 1 y, 2 z
int adjoint_walls(const cans_step_params* s) {
  int wm = 0;
  for (int d = 0; d < 3; d++)
    if (s->cbcpre[2 * d] != 'P' || s->cbcpre[2 * d + 1] != 'P') wm |= 1 << d;
  return wm;
}

int adjoint_scope(cans_ctx* c, const cans_step_params* s, const char* who, bool scalars = false) {
  const std::string w(who);
  if (c->precision != CANS_FP64 && c->precision != CANS_MIXED)
    return fail(CANS_EUNSUPPORTED, w + ": the adjoint needs a CANS_FP64 or CANS_MIXED context");
  if (!s) {
    if (c->imp) return fail(CANS_EUNSUPPORTED, w + ": the transposed Jacobian covers explicit diffusion");
    return CANS_OK;
  }
  if (s->nscal && !scalars)
    return fail(CANS_EINVAL, w + ": a step with scalars runs backwards through cans_rk3_substep_adjoint_scal");
  if (s->nscal) {
    // scalars: explicit diffusion; periodic where the box is, Dirichlet or Neumann at its walls
    if (c->imp) return fail(CANS_EUNSUPPORTED, w + ": with scalars the adjoint covers explicit diffusion");
    if (!s->scal || s->nscal > CANS_MAX_SCALARS) return fail(CANS_EINVAL, w + ": bad scalar list");
    const int swm = adjoint_walls(s);
    for (int i = 0; i < s->nscal; i++) {
      const cans_scalar& q = s->scal[i];
      if (q.is_forced && !s->grid_vol_ratio_f_d) return fail(CANS_EINVAL, w + ": null grid_vol_ratio_f");
      for (int e = 0; e < 6; e++) {
        const bool wall = swm >> (e / 2) & 1;
        if (wall ? (q.cbc[e] != 'D' && q.cbc[e] != 'N') : q.cbc[e] != 'P')
          return fail(CANS_EUNSUPPORTED, w + ": a scalar must be periodic where the box is, D or N at its walls");
      }
    }
  }
  // implicit diffusion: the 3-D form on tri-periodic uniform grids, where the Helmholtz operator
  // commutes with the gradient, so that the pressure still drops out (P H^-1 G p = 0)
  if (c->imp == 2) return fail(CANS_EUNSUPPORTED, w + ": the adjoint covers 3-D implicit diffusion, not z-only");
  if (c->imp && adjoint_walls(s))
    return fail(CANS_EUNSUPPORTED, w + ": with implicit diffusion the adjoint covers tri-periodic grids");
  const int n[3] = {c->L.n1, c->L.n2, c->L.n3};
  const int wm = adjoint_walls(s);
  for (int d = 0; d < 3; d++) {
    const char* cp = s->cbcpre + 2 * d;
    if (!(wm >> d & 1)) {
      for (int m = 0; m < 3; m++)
        if (s->cbcvel[6 * m + 2 * d] != 'P' || s->cbcvel[6 * m + 2 * d + 1] != 'P')
          return fail(CANS_EUNSUPPORTED, w + ": velocity and pressure disagree on a periodic direction");
      continue;
    }
    // walls: Neumann pressure, Dirichlet normal velocity (its wall faces constants), Dirichlet or
    // Neumann tangential velocity
    if (cp[0] != 'N' || cp[1] != 'N' || s->cbcvel[6 * d + 2 * d] != 'D' || s->cbcvel[6 * d + 2 * d + 1] != 'D')
      return fail(CANS_EUNSUPPORTED, w + ": walls need Neumann pressure and a Dirichlet normal velocity");
    for (int m = 0; m < 3; m++)
      for (int e = 0; e < 2; e++) {
        const char b = s->cbcvel[6 * m + 2 * d + e];
        if (m != d && b != 'D' && b != 'N')
          return fail(CANS_EUNSUPPORTED, w + ": walls need a Dirichlet or Neumann tangential velocity");
      }
    if (n[d] < 2) return fail(CANS_EUNSUPPORTED, w + ": walls need at least two cells across");
  }
  if (s->is_forced[2]) return fail(CANS_EUNSUPPORTED, w + ": the adjoint covers forcing of u and v");
  for (int m = 0; m < 2; m++)
    if (s->is_forced[m] && (wm >> m & 1))
      return fail(CANS_EUNSUPPORTED, w + ": forcing needs its direction periodic");
  if (!(wm & 4)) {
    const double dz = s->dzc_h[1];
    for (int k = 0; k <= c->L.n3 + 1; k++)
      if (std::fabs(s->dzc_h[k] - dz) > 1e-12 * dz || std::fabs(s->dzf_h[k] - dz) > 1e-12 * dz)
        return fail(CANS_EUNSUPPORTED, w + ": with a periodic z the adjoint covers uniform grids");
  }
  if (!s->grid_vol_ratio_f_d && (s->is_forced[0] || s->is_forced[1]))
    return fail(CANS_EINVAL, w + ": null grid_vol_ratio_f");
  return CANS_OK;
}

}  // namespace

extern "C" int cans_mom_jt(cans_ctx* c, const int* n, const double* dli, double visc, const void* u, const void* v,
                           const void* w, const void* gu, const void* gv, const void* gw, const void* bu,
                           const void* bv, const void* bw, void* ou, void* ov, void* ow) {
  CHECK_CTX("cans_mom_jt");
  CHECK_N("cans_mom_jt", n);
  int rc = adjoint_scope(c, nullptr, "cans_mom_jt");
  if (rc) return rc;
  if (!dli || !u || !v || !w || !gu || !gv || !gw || !bu || !bv || !bw || !ou || !ov || !ow)
    return fail(CANS_EINVAL, "cans_mom_jt: null argument");
  const void* rd[6] = {u, v, w, gu, gv, gw};
  void* wr[3] = {ou, ov, ow};
  for (void* o : wr)
    for (const void* r : rd)
      if (o == r) return fail(CANS_EINVAL, "cans_mom_jt: the output may not be a velocity or g field");
  const double* q[3] = {(const double*)u, (const double*)v, (const double*)w};
  const double* g[3] = {(const double*)gu, (const double*)gv, (const double*)gw};
  const double* b[3] = {(const double*)bu, (const double*)bv, (const double*)bw};
  double* o[3] = {(double*)ou, (double*)ov, (double*)ow};
  return launch_mom_jt(c, q, g, b, o, dli, visc);
}

static int substep_adjoint_vel(cans_ctx* c, const cans_step_params* s, int irk, double dt, const void* u,
                               const void* v, const void* w, void* ub, void* vb, void* wb, void* fu, void* fv,
                               void* fw, void* ubo, void* vbo, void* wbo, void* fuo, void* fvo, void* fwo, void* pp,
                               bool scalars) {
  if (!s || irk < 1 || irk > 3) return fail(CANS_EINVAL, "cans_rk3_substep_adjoint: bad argument");
  const void* all[16] = {u, v, w, ub, vb, wb, fu, fv, fw, ubo, vbo, wbo, fuo, fvo, fwo, pp};
  for (int q = 0; q < 16; q++) {
    if (!all[q]) return fail(CANS_EINVAL, "cans_rk3_substep_adjoint: null field");
    for (int r = 0; r < q; r++)
      if (all[q] == all[r]) return fail(CANS_EINVAL, "cans_rk3_substep_adjoint: every field must be distinct");
  }
  if (!s->plan || !s->lambdaxy_d || !s->a_d || !s->b_d || !s->c_d || !s->dzci_d || !s->dzfi_d || !s->dzc_h ||
      !s->dzf_h)
    return fail(CANS_EINVAL, "cans_rk3_substep_adjoint: null parameter array");
  int rc = adjoint_scope(c, s, "cans_rk3_substep_adjoint", scalars);
  if (rc) return rc;
  const double* rkpar = RKCOEFF[irk - 1];
  const double f1 = rkpar[0] * dt, f2 = rkpar[1] * dt;
  double* a[3] = {(double*)ub, (double*)vb, (double*)wb};
  double* fb[3] = {(double*)fu, (double*)fv, (double*)fw};
  const int wm = adjoint_walls(s);
  const bool zwall = wm & 4;
  // the linear part of bound.f90's ghost map (boundary values zero): periodic images, and at walls
  // u(0) = -u(1) (Dirichlet) or u(0) = u(1) (Neumann) tangentially, 0 on the normal wall faces
  const double zero[18] = {};
  auto ghosts = [&](double* const* f) {
    return plan_bounduvw(c, s->cbcvel, zero, s->nb, s->is_bound, s->dl, s->dzc_h, s->dzf_h, f[0], f[1], f[2], -1,
                         CANS_HALO_ABSENT, CANS_HALO_ABSENT);
  };
  // wb = Fr'^T P^T ub in place, P^T = M P M^-1 (M the cells' dz: dzf for u and v, dzc for w; a
  // multiple of I with a uniform z) and Fr'^T b = b - gvr_f sum(b) the transposed bulk forcing,
  // applied as dzf (P a - mean_f(P a)).  P: fillps with dti = 1, solver, boundp, correc with dt = 1
  // (correc reads each cell's input before it writes that cell)
  if (wm && (rc = launch_adj_weigh(c, a, zwall ? s->dzfi_d : nullptr, zwall ? s->dzci_d : nullptr, true, 0., 0., wm)))
    return rc;
  if ((rc = ghosts(a))) return rc;
  // the projection in the context's pressure storage (CANS_MIXED: FP32 pp and transforms, FP64
  // elimination and velocity, as the forward substep projects)
  auto project = [&](auto tag) -> int {
    using S = decltype(tag);
    S* P = (S*)pp;
    int r;
    if ((r = launch_fillps<S>(c, a[0], a[1], a[2], nullptr, nullptr, nullptr, nullptr, P, s->dzfi_d, 1., s->dli)))
      return r;
    if ((r = run_solver<S>(c, s->plan, s->normfft, s->lambdaxy_d, s->a_d, s->b_d, s->c_d, s->cbcpre, "ccc", P)))
      return r;
    if ((r = plan_boundp(c, s->cbcpre, zero, s->nb, s->is_bound, s->dl, s->dzc_h, P, CANS_HALO_ABSENT))) return r;
    return launch_correc<S>(c, P, a[0], a[1], a[2], nullptr, nullptr, nullptr, nullptr, a[0], a[1], a[2], s->dzci_d,
                            1., s->dli, nullptr);
  };
  if ((rc = c->precision == CANS_MIXED ? project(float{}) : project(double{}))) return rc;
  // implicit diffusion (main.f90:453-462): hb = H^-1 P ub, H = I - visc dtrk/2 L symmetric here, one
  // batched solve of the three components' systems with homogeneous boundary values
  const double dtrk = f1 + f2;
  if (c->imp) {
    HelmSys hs[3];
    for (int q = 0; q < 3; q++) {
      const cans_helmholtz& hq = s->hvel[q];
      if (!hq.plan || !hq.lambdaxy_d || !hq.a_d || !hq.b_d || !hq.c_d)
        return fail(CANS_EINVAL, "cans_rk3_substep_adjoint: Helmholtz system without a plan");
      hs[q] = HelmSys{hq.plan, hq.normfft, hq.lambdaxy_d, hq.a_d, hq.b_d, hq.c_d, hq.cbc, hq.c_or_f, a[q],
                      -.5 * s->visc * dtrk};
    }
    if ((rc = run_helmholtz_batch(c, 3, hs))) return rc;
  }
  double mean[2] = {0., 0.};
  for (int m = 0; m < 2; m++)
    if (s->is_forced[m] && (rc = run_bulk_mean<double>(c, a[m], s->grid_vol_ratio_f_d, &mean[m]))) return rc;
  if (wm || s->is_forced[0] || s->is_forced[1]) {
    if ((rc = launch_adj_weigh(c, a, zwall ? s->dzfi_d : nullptr, zwall ? s->dzci_d : nullptr, false, mean[0],
                               mean[1], wm)))
      return rc;
  }
  // (ubo, fbo) = (wb + J(u)^T (f1 wb + fb), f2 wb): one pass over wb and fb with their ghosts
  if ((rc = ghosts(a)) || (rc = ghosts(fb))) return rc;
  const double h[3] = {s->dli[0], s->dli[1], 1. / s->dzc_h[1]};
  const double* q[3] = {(const double*)u, (const double*)v, (const double*)w};
  double* o[3] = {(double*)ubo, (double*)vbo, (double*)wbo};
  double* fo[3] = {(double*)fuo, (double*)fvo, (double*)fwo};
  if (!wm) {
    // implicit diffusion: J is the advection's alone, and the diffusion's explicit half adds
    // (f1 + f2)/2 visc L hb
    if ((rc = launch_mom_jt_adj(c, q, a, fb, o, fo, h, c->imp ? 0. : s->visc, f1, f2, nullptr))) return rc;
    if (c->imp) return launch_adj_lap(c, a, o, .5 * dtrk * s->visc, h);
    return CANS_OK;
  }
  AdjWalls aw;
  aw.mask = wm;
  aw.dzci = s->dzci_d;
  aw.dzfi = s->dzfi_d;
  for (int m = 0; m < 3; m++)
    for (int d = 0; d < 3; d++)
      for (int e = 0; e < 2; e++) aw.sign[m][d][e] = s->cbcvel[6 * m + 2 * d + e] == 'N' ? 1. : -1.;
  return launch_mom_jt_adj(c, q, a, fb, o, fo, h, s->visc, f1, f2, &aw);
}

extern "C" int cans_rk3_substep_adjoint(cans_ctx* c, const cans_step_params* s, int irk, double dt, const void* u,
                                        const void* v, const void* w, void* ub, void* vb, void* wb, void* fu,
                                        void* fv, void* fw, void* ubo, void* vbo, void* wbo, void* fuo, void* fvo,
                                        void* fwo, void* pp) {
  CHECK_CTX("cans_rk3_substep_adjoint");
  return substep_adjoint_vel(c, s, irk, dt, u, v, w, ub, vb, wb, fu, fv, fw, ubo, vbo, wbo, fuo, fvo, fwo, pp, false);
}

// The adjoint substep with scalars (explicit diffusion): the velocity part as
// cans_rk3_substep_adjoint, then per scalar i (main.f90:440-450: the scalars advance first, their
// bulk forcing is added, and the velocity's rk reads the first one's new value in the Boussinesq
// term, rk.f90:163-177) sh_i = Fs'^T (sb_i (+ that term's transpose of wb for i = 0)), sbo_i = sh_i +
// J_s^T (f1 sh_i + fsb_i), fsbo_i = f2 sh_i, ubo += J_q^T (f1 sh_i + fsb_i).
extern "C" int cans_rk3_substep_adjoint_scal(cans_ctx* c, const cans_step_params* s, int irk, double dt,
                                             const void* u, const void* v, const void* w, void* ub, void* vb,
                                             void* wb, void* fu, void* fv, void* fw, void* ubo, void* vbo, void* wbo,
                                             void* fuo, void* fvo, void* fwo, void* pp, const void* const* sc,
                                             void* const* sb, void* const* fsb, void* const* sbo,
                                             void* const* fsbo) {
  CHECK_CTX("cans_rk3_substep_adjoint_scal");
  if (!s || !s->nscal || !sc || !sb || !fsb || !sbo || !fsbo)
    return fail(CANS_EINVAL, "cans_rk3_substep_adjoint_scal: null argument or no scalars");
  if (s->nscal > CANS_MAX_SCALARS) return fail(CANS_EINVAL, "cans_rk3_substep_adjoint_scal: too many scalars");
  {
    std::vector<const void*> all = {u, v, w, ub, vb, wb, fu, fv, fw, ubo, vbo, wbo, fuo, fvo, fwo, pp};
    for (int i = 0; i < s->nscal; i++) {
      all.push_back(sc[i]);
      all.push_back(sb[i]);
      all.push_back(fsb[i]);
      all.push_back(sbo[i]);
      all.push_back(fsbo[i]);
    }
    for (size_t q = 16; q < all.size(); q++) {
      if (!all[q]) return fail(CANS_EINVAL, "cans_rk3_substep_adjoint_scal: null field");
      for (size_t r = 0; r < q; r++)
        if (all[q] == all[r]) return fail(CANS_EINVAL, "cans_rk3_substep_adjoint_scal: every field must be distinct");
    }
  }
  int rc = substep_adjoint_vel(c, s, irk, dt, u, v, w, ub, vb, wb, fu, fv, fw, ubo, vbo, wbo, fuo, fvo, fwo, pp, true);
  if (rc) return rc;
  const double* rkpar = RKCOEFF[irk - 1];
  const double f1 = rkpar[0] * dt, f2 = rkpar[1] * dt;
  const int swm = adjoint_walls(s);
  const double zero6[6] = {};
  ScalAdj sa;
  sa.q[0] = (const double*)u;
  sa.q[1] = (const double*)v;
  sa.q[2] = (const double*)w;
  sa.ubo[0] = (double*)ubo;
  sa.ubo[1] = (double*)vbo;
  sa.ubo[2] = (double*)wbo;
  sa.wb[0] = (const double*)ub;  // ub..wb hold wb = Fr'^T P^T ub with its ghosts (above)
  sa.wb[1] = (const double*)vb;
  sa.wb[2] = (const double*)wb;
  sa.dzci = s->dzci_d;
  sa.dzfi = s->dzfi_d;
  sa.dxi = s->dli[0];
  sa.dyi = s->dli[1];
  sa.f1 = f1;
  sa.f2 = f2;
  sa.wm = swm;
  for (int cc = 0; cc < 3; cc++) sa.coef[cc] = (f1 + f2) * s->gacc[cc] * s->beta * 0.5;
  for (int i = 0; i < s->nscal; i++) {
    const cans_scalar& qs = s->scal[i];
    sa.s = (const double*)sc[i];
    sa.sh = (double*)sb[i];
    sa.fsb = (const double*)fsb[i];
    sa.sbo = (double*)sbo[i];
    sa.fsbo = (double*)fsbo[i];
    sa.alpha = qs.alpha;
    for (int d = 0; d < 3; d++)
      for (int e = 0; e < 2; e++) sa.sign[d][e] = qs.cbc[2 * d + e] == 'N' ? 1. : -1.;
    sa.buoy = i == 0 && s->is_buoyancy;
    if ((rc = launch_scal_adjoint(c, sa))) return rc;
    if (qs.is_forced) {  // the forcing acts before the buoyancy reads the scalar: its transpose after
      double sum = 0.;
      if ((rc = run_bulk_mean<double>(c, sa.sh, nullptr, &sum))) return rc;
      if ((rc = launch_scal_force_t(c, sa.sh, s->grid_vol_ratio_f_d, sum))) return rc;
    }
    // periodic images of sh and fsb (the wall ghosts are read as zero)
    if ((rc = plan_boundp(c, qs.cbc, zero6, s->nb, s->is_bound, s->dl, s->dzc_h, sb[i], CANS_HALO_ABSENT,
                          (int)sizeof(double))))
      return rc;
    if ((rc = plan_boundp(c, qs.cbc, zero6, s->nb, s->is_bound, s->dl, s->dzc_h, fsb[i], CANS_HALO_ABSENT,
                          (int)sizeof(double))))
      return rc;
    if ((rc = launch_scal_jt(c, sa))) return rc;
  }
  return CANS_OK;
}

// ---------------------------------------------------------------------------
// Poisson solve with iterative refinement (no counterpart in the reference, whose
// solves run in one precision): p (FP64) holds b on entry and receives x with
// L x = b.  The inner solves run in the context's pressure storage type (float
// under CANS_MIXED: FP32 transforms and storage, FP64 elimination), the residuals
// b - L x and the accumulation of x in FP64:
//   x = S(b);  repeat sweeps times: x += S(b - L x)
// with S the solver (solver.f90:17-112) and L its 7-point operator.  Cell-centred
// (c_or_f "ccc").
// ---------------------------------------------------------------------------
extern "C" int cans_solver_refined(cans_ctx* c, const int* n, const int* ng, const cans_plan* plan, double normfft,
                                   const double* lambdaxy, const double* a, const double* b, const double* cc,
                                   const char* bc, const char* c_or_f, const double* dli, const double* dzci,
                                   const double* dzfi, int sweeps, void* p) {
  CHECK_CTX("cans_solver_refined");
  CHECK_N("cans_solver_refined", n);
  if (!ng || !plan || !lambdaxy || !a || !b || !cc || !bc || !c_or_f || !dli || !dzci || !dzfi || !p)
    return fail(CANS_EINVAL, "cans_solver_refined: null argument");
  if (std::strncmp(c_or_f, "ccc", 3)) return fail(CANS_EUNSUPPORTED, "cans_solver_refined: cell-centred systems only");
  if (sweeps < 0 || sweeps > 8) return fail(CANS_EINVAL, "cans_solver_refined: sweeps must be in 0..8");
  int bt[6];
  for (int q = 0; q < 6; q++) {
    if (bc[q] == 'P') bt[q] = 0;
    else if (bc[q] == 'N') bt[q] = 1;
    else if (bc[q] == 'D') bt[q] = 2;
    else return fail(CANS_EINVAL, "cans_solver_refined: boundary types are P, N or D");
  }
  // scratch: the FP64 rhs b in tmp field 0; the inner solve's field in tmp field 2, viewed in the
  // storage type with element 1 on a pair boundary (the float view lies inside tmp field 1's bytes)
  double* td[3];
  int rc = ensure_tmp_fields<double>(c, td);
  if (rc) return rc;
  const size_t fbytes = (size_t)c->L.sz * (c->L.n3 + 2) * sizeof(double);
  double* x = (double*)p;
  CANS_HIP_TRY(hipMemcpyAsync(td[0], x, fbytes, hipMemcpyDeviceToDevice, c->stream));
  return CANS_DISPATCH(c, [&]() -> int {
    T* tt[3];
    int r = ensure_tmp_fields<T>(c, tt);
    if (r) return r;
    T* f = tt[2];
    for (int s = 0; s <= sweeps; s++) {
      if ((r = launch_residual<T>(c, td[0], s ? x : nullptr, f, dzci, dzfi, dli, bt))) return r;
      if ((r = run_solver<T>(c, plan, normfft, lambdaxy, a, b, cc, bc, "ccc", f))) return r;
      if ((r = launch_accumulate<T>(c, x, f, s == 0))) return r;
    }
    return CANS_OK;
  }());
}
