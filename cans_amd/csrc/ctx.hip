// Context, device memory and host-side set-up helpers.
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "common.hpp"

namespace cans {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// Fixed device scratch of the allocation-free solve (cans_solver_noalloc): the transform
// work buffer and the z pass's checkpoints of one solve.  Every GPU the library runs on
// holds its own copy, at its own address, until the process ends.  Host code never
// passes the array's name to a kernel (on the host it only names a stand-in): the
// device address is looked up once per device, with that device current, when the
// first context on it is created, and kept in g_scratch_ptr.
static __device__ __align__(256) unsigned char scratch[CANS_NOALLOC_SCRATCH_BYTES];
constexpr int MAX_DEVICES = 64;
static void* g_scratch_ptr[MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

// device address of this device's scratch (the device must be current)
static int scratch_setup(int device) {
  if (device < 0 || device >= MAX_DEVICES) return fail(CANS_EUNSUPPORTED, "cans_ctx_create: device id above 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return CANS_OK;
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(CANS_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = ptr;
  return CANS_OK;
}

void* noalloc_scratch(int device) {
  if (device < 0 || device >= MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

// origin pointer -> (base pointer, element size) of every field allocated through a context
struct FieldRec {
  void* base;
  size_t es;
};
static std::mutex g_mu;
static std::map<void*, FieldRec> g_fields;

// element size of a field kind in a context: velocity is double in every mode;
// pressure-like arrays are float in the mixed mode
static size_t kind_size(const cans_ctx* c, int kind) {
  return (kind == CANS_FIELD_PRESSURE && c->precision == CANS_MIXED) ? sizeof(float) : sizeof(double);
}

}  // namespace cans

using namespace cans;

extern "C" const char* cans_last_error(void) { return g_err.c_str(); }

extern "C" int cans_ctx_create(const int* n, int ld, int precision, void* stream, cans_ctx** out) {
  if (!n || !out) return fail(CANS_EINVAL, "cans_ctx_create: null argument");
  if (n[0] < 1 || n[1] < 1 || n[2] < 1) return fail(CANS_EINVAL, "cans_ctx_create: sizes must be >= 1");
  if (precision != CANS_FP64 && precision != CANS_MIXED)
    return fail(CANS_EINVAL, "cans_ctx_create: unknown precision mode");
  if (ld == 0) ld = n[0] + 2;
  if (ld < n[0] + 2) return fail(CANS_EINVAL, "cans_ctx_create: ld must be >= n1+2");
  cans_ctx* c = new cans_ctx();
  c->L.n1 = n[0];
  c->L.n2 = n[1];
  c->L.n3 = n[2];
  c->L.sy = ld;
  c->L.sz = (long)ld * (n[1] + 2);
  c->precision = precision;
  c->stream = (hipStream_t)stream;
  hipError_t e = hipGetDevice(&c->device);
  if (e != hipSuccess) {
    delete c;
    return fail(CANS_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  }
  if (int rc = scratch_setup(c->device)) {  // c->device is the current device
    delete c;
    return rc;
  }
  e = hipMalloc((void**)&c->dscal, 64 * sizeof(double));
  if (e == hipSuccess) e = hipMemsetAsync(c->dscal, 0, 64 * sizeof(double), c->stream);
  c->partials_n = 1 << 16;
  if (e == hipSuccess) e = hipMalloc((void**)&c->partials, c->partials_n * sizeof(double));
  if (e != hipSuccess) {
    hipFree(c->dscal);
    delete c;
    return fail(CANS_ENOMEM, std::string("cans_ctx_create: ") + hipGetErrorString(e));
  }
  *out = c;
  return CANS_OK;
}

extern "C" int cans_ctx_destroy(cans_ctx* c) {
  if (!c) return CANS_OK;
  hipStreamSynchronize(c->stream);
  for (int s = 0; s < cans_ctx::NWORK; s++) hipFree(c->work[s]);
  hipFree(c->dscal);
  hipFree(c->partials);
  delete c;
  return CANS_OK;
}

extern "C" int cans_ctx_layout(const cans_ctx* c, long* sy, long* sz, long* elems) {
  if (!c) return fail(CANS_EINVAL, "cans_ctx_layout: null context");
  if (sy) *sy = c->L.sy;
  if (sz) *sz = c->L.sz;
  if (elems) *elems = c->L.sz * (c->L.n3 + 2);
  return CANS_OK;
}

extern "C" int cans_ctx_sync(cans_ctx* c) {
  if (!c) return fail(CANS_EINVAL, "cans_ctx_sync: null context");
  CANS_HIP_TRY(hipStreamSynchronize(c->stream));
  return CANS_OK;
}

extern "C" int cans_ctx_forcing(cans_ctx* c, const double** f_d) {
  if (!c || !f_d) return fail(CANS_EINVAL, "cans_ctx_forcing: null argument");
  *f_d = c->dscal;
  return CANS_OK;
}

namespace cans {
// grow-on-demand scratch slot owned by the context
int ensure_work(cans_ctx* c, int slot, size_t bytes, void** ptr) {
  if (c->work_bytes[slot] < bytes) {
    CANS_HIP_TRY(hipStreamSynchronize(c->stream));
    hipFree(c->work[slot]);
    c->work[slot] = nullptr;
    c->work_bytes[slot] = 0;
    hipError_t e = hipMalloc(&c->work[slot], bytes);
    if (e != hipSuccess) return fail(CANS_ENOMEM, std::string("scratch allocation: ") + hipGetErrorString(e));
    c->work_bytes[slot] = bytes;
    CANS_HIP_TRY(hipMemsetAsync(c->work[slot], 0, bytes, c->stream));  // defined contents from the start
  }
  *ptr = c->work[slot];
  return CANS_OK;
}
}  // namespace cans

extern "C" int cans_ctx_elem_size(const cans_ctx* c, int kind, int* bytes) {
  if (!c || !bytes) return fail(CANS_EINVAL, "cans_ctx_elem_size: null argument");
  if (kind != CANS_FIELD_VELOCITY && kind != CANS_FIELD_PRESSURE)
    return fail(CANS_EINVAL, "cans_ctx_elem_size: unknown field kind");
  *bytes = (int)kind_size(c, kind);
  return CANS_OK;
}

// Fields: the origin (element (0,0,0)) sits one element before a 256-byte
// boundary so that interior rows start aligned when ld is a multiple of 32.
extern "C" int cans_field_alloc(cans_ctx* c, void** field) { return cans_field_alloc_kind(c, CANS_FIELD_VELOCITY, field); }

extern "C" int cans_field_alloc_kind(cans_ctx* c, int kind, void** field) {
  if (!c || !field) return fail(CANS_EINVAL, "cans_field_alloc: null argument");
  if (kind != CANS_FIELD_VELOCITY && kind != CANS_FIELD_PRESSURE)
    return fail(CANS_EINVAL, "cans_field_alloc_kind: unknown field kind");
  const size_t es = kind_size(c, kind);
  size_t bytes = (size_t)c->L.sz * (c->L.n3 + 2) * es + 256;
  void* base = nullptr;
  hipError_t e = hipMalloc(&base, bytes);
  if (e != hipSuccess) return fail(CANS_ENOMEM, std::string("cans_field_alloc: ") + hipGetErrorString(e));
  e = hipMemsetAsync(base, 0, bytes, c->stream);
  if (e != hipSuccess) {
    hipFree(base);
    return fail(CANS_EHIP, std::string("cans_field_alloc: ") + hipGetErrorString(e));
  }
  void* origin = (char*)base + 256 - es;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_fields[origin] = FieldRec{base, es};
  }
  *field = origin;
  return CANS_OK;
}

extern "C" int cans_field_free(cans_ctx* c, void* field) {
  if (!field) return CANS_OK;
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_fields.find(field);
    if (it == g_fields.end()) return fail(CANS_EINVAL, "cans_field_free: not a context field");
    base = it->second.base;
    g_fields.erase(it);
  }
  if (!base) return CANS_OK;  // registered by the caller (cans_field_register): not ours to free
  if (c) hipStreamSynchronize(c->stream);
  CANS_HIP_TRY(hipFree(base));
  return CANS_OK;
}

extern "C" int cans_ctx_set_numerics(cans_ctx* c, int is_impdiff, int is_impdiff_1d) {
  if (!c) return fail(CANS_EINVAL, "cans_ctx_set_numerics: null context");
  // param.f90:331: is_impdiff_1d implies is_impdiff
  c->imp = is_impdiff_1d ? 2 : is_impdiff ? 1 : 0;
  return CANS_OK;
}

extern "C" int cans_ctx_set_dtdma(cans_ctx* c, int nslab) {
  if (!c) return fail(CANS_EINVAL, "cans_ctx_set_dtdma: null context");
  if (nslab < 0) return fail(CANS_EINVAL, "cans_ctx_set_dtdma: negative slab count");
  c->dtdma = nslab;
  return CANS_OK;
}

extern "C" int cans_field_register(cans_ctx* c, void* field, int kind) {
  if (!c || !field) return fail(CANS_EINVAL, "cans_field_register: null argument");
  if (kind != CANS_FIELD_VELOCITY && kind != CANS_FIELD_PRESSURE)
    return fail(CANS_EINVAL, "cans_field_register: unknown field kind");
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_fields.find(field);
  if (it != g_fields.end() && it->second.base)
    return fail(CANS_EINVAL, "cans_field_register: field allocated by cans_field_alloc_kind");
  g_fields[field] = FieldRec{nullptr, kind_size(c, kind)};
  return CANS_OK;
}

namespace cans {
// element size of a registered field (cans_field_alloc*), 0 if unknown
int field_elem_size(const cans_ctx* c, const void* field) {
  (void)c;
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_fields.find(const_cast<void*>(field));
  return it == g_fields.end() ? 0 : (int)it->second.es;
}
}  // namespace cans

// Element size of a field: from the registry for fields of cans_field_alloc*, else
// double (accepted only where every field is double, i.e. in an FP64 context).
static int field_es(const cans_ctx* c, const void* field, size_t* es, const char* who) {
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_fields.find(const_cast<void*>(field));
    if (it != g_fields.end()) {
      *es = it->second.es;
      return CANS_OK;
    }
  }
  if (c->precision != CANS_FP64)
    return fail(CANS_EINVAL, std::string(who) + ": field not from cans_field_alloc_kind (element type unknown)");
  *es = sizeof(double);
  return CANS_OK;
}

// Host arrays are always double; FP32 storage (pressure-like fields of a
// CANS_MIXED context) rounds on upload (to nearest) and widens exactly on download.
extern "C" int cans_field_upload(cans_ctx* c, void* field, const double* host) {
  if (!c || !field || !host) return fail(CANS_EINVAL, "cans_field_upload: null argument");
  const Layout& L = c->L;
  const size_t rows = (size_t)(L.n2 + 2) * (L.n3 + 2), nx = (size_t)(L.n1 + 2);
  size_t es;
  int rc = field_es(c, field, &es, "cans_field_upload");
  if (rc) return rc;
  if (es == sizeof(double)) {
    const size_t w = nx * sizeof(double);
    CANS_HIP_TRY(hipMemcpy2DAsync(field, L.sy * sizeof(double), host, w, w, rows, hipMemcpyHostToDevice, c->stream));
  } else {
    std::vector<float> tmp(nx * rows);
    for (size_t q = 0; q < tmp.size(); q++) tmp[q] = (float)host[q];
    const size_t w = nx * sizeof(float);
    CANS_HIP_TRY(hipMemcpy2DAsync(field, L.sy * sizeof(float), tmp.data(), w, w, rows, hipMemcpyHostToDevice,
                                  c->stream));
    CANS_HIP_TRY(hipStreamSynchronize(c->stream));  // tmp must outlive the copy
  }
  CANS_HIP_TRY(hipStreamSynchronize(c->stream));
  return CANS_OK;
}

extern "C" int cans_field_download(cans_ctx* c, const void* field, double* host) {
  if (!c || !field || !host) return fail(CANS_EINVAL, "cans_field_download: null argument");
  const Layout& L = c->L;
  const size_t rows = (size_t)(L.n2 + 2) * (L.n3 + 2), nx = (size_t)(L.n1 + 2);
  size_t es;
  int rc = field_es(c, field, &es, "cans_field_download");
  if (rc) return rc;
  if (es == sizeof(double)) {
    const size_t w = nx * sizeof(double);
    CANS_HIP_TRY(hipMemcpy2DAsync(host, w, field, L.sy * sizeof(double), w, rows, hipMemcpyDeviceToHost, c->stream));
    CANS_HIP_TRY(hipStreamSynchronize(c->stream));
  } else {
    std::vector<float> tmp(nx * rows);
    const size_t w = nx * sizeof(float);
    CANS_HIP_TRY(hipMemcpy2DAsync(tmp.data(), w, field, L.sy * sizeof(float), w, rows, hipMemcpyDeviceToHost,
                                  c->stream));
    CANS_HIP_TRY(hipStreamSynchronize(c->stream));
    for (size_t q = 0; q < tmp.size(); q++) host[q] = (double)tmp[q];
  }
  return CANS_OK;
}

extern "C" int cans_device_alloc(cans_ctx* c, long bytes, void** ptr) {
  if (!c || !ptr || bytes < 0) return fail(CANS_EINVAL, "cans_device_alloc: bad argument");
  *ptr = nullptr;
  hipError_t e = hipMalloc(ptr, bytes > 0 ? (size_t)bytes : 1);
  if (e != hipSuccess) return fail(CANS_ENOMEM, std::string("cans_device_alloc: ") + hipGetErrorString(e));
  return CANS_OK;
}

extern "C" int cans_device_free(cans_ctx* c, void* ptr) {
  if (!ptr) return CANS_OK;
  if (c) hipStreamSynchronize(c->stream);
  CANS_HIP_TRY(hipFree(ptr));
  return CANS_OK;
}

extern "C" int cans_memcpy_h2d(cans_ctx* c, void* dst, const void* src, long bytes) {
  if (!c || !dst || !src || bytes < 0) return fail(CANS_EINVAL, "cans_memcpy_h2d: bad argument");
  CANS_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  CANS_HIP_TRY(hipStreamSynchronize(c->stream));
  return CANS_OK;
}

extern "C" int cans_memcpy_d2h(cans_ctx* c, void* dst, const void* src, long bytes) {
  if (!c || !dst || !src || bytes < 0) return fail(CANS_EINVAL, "cans_memcpy_d2h: bad argument");
  CANS_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  CANS_HIP_TRY(hipStreamSynchronize(c->stream));
  return CANS_OK;
}

// ---------------------------------------------------------------------------
// host set-up helpers (initgrid.f90, initsolver.f90).  Fortran default-real
// expressions of the reference are evaluated in single precision, as there.
// ---------------------------------------------------------------------------
static double gp_stretch(int gtype, double alpha, double z0) {
  const double eps = 2.220446049250313e-16;
  if (!(alpha > eps)) return z0;
  switch (gtype) {
    case 2:  // initgrid.f90:119-131
      return 1.0 * (1. + std::tanh((z0 - 1.0) * alpha) / std::tanh(alpha / 1.));
    case 3:  // initgrid.f90:132-144
      return 1. - 1.0 * (1. + std::tanh((1. - z0 - 1.0) * alpha) / std::tanh(alpha / 1.));
    case 4:  // initgrid.f90:145-161
      if (z0 <= 0.5) return 0.5 * (1. - 1. + std::tanh(2. * alpha * (z0 - 0.)) / std::tanh(alpha));
      return 0.5 * (1. + 1. + std::tanh(2. * alpha * (z0 - 1.)) / std::tanh(alpha));
    default:  // initgrid.f90:105-118
      return 0.5 * (1. + std::tanh((z0 - 0.5) * alpha) / std::tanh(alpha / 2.));
  }
}

// gridpoint_natural (initgrid.f90:162-203): the wall-resolved channel spacing of Pirozzoli &
// Orlandi (JCP 439, 110408, 2021), clustered at both walls; kb, alpha, c_eta and dyp at the
// reference's defaults.  All FP64, as there.
static double gp_natural(int kg, int nzg) {
  const double pi = std::acos(-1.);
  const double kb = 32., alpha = pi / 1.5, c_eta = 0.8, dyp = 0.05;
  const double n = nzg / 2.;
  const double rn = n / kb;
  const double retau = 1. / (1. + rn * rn) * (dyp * n + std::pow(3. / 4. * alpha * c_eta * n, 4. / 3.) * (rn * rn));
  const double k = 1. * std::min(kg, nzg - kg);
  const double rk = k / kb;
  double z = 1. / (1. + rk * rk) * (dyp * k + std::pow(3. / 4. * alpha * c_eta * k, 4. / 3.) * (rk * rk)) /
             (2. * retau);
  if (kg > nzg - kg) z = 1. - z;
  return z;
}

static int initgrid(int gtype, int n, double gr, double lz, int is_periodic, bool natural, double* dzc,
                    double* dzf, double* zc, double* zf) {
  if (n < 1 || !dzc || !dzf || !zc || !zf) return fail(CANS_EINVAL, "cans_initgrid: bad argument");
  zf[0] = 0.;
  for (int k = 1; k <= n; k++) {
    if (natural) {
      zf[k] = gp_natural(k, n) * lz;
    } else {
      float z0 = ((float)k - 0.f) / (1.f * (float)n);
      zf[k] = gp_stretch(gtype, gr, (double)z0) * lz;
    }
  }
  if (std::fabs(gr) < 2.220446049250313e-16 && !natural) {
    for (int k = 0; k <= n + 1; k++) dzf[k] = dzc[k] = lz / (double)(1.f * (float)n);
  } else {
    for (int k = 1; k <= n; k++) dzf[k] = zf[k] - zf[k - 1];
    dzf[0] = is_periodic ? dzf[n] : dzf[1];
    dzf[n + 1] = is_periodic ? dzf[1] : dzf[n];
    for (int k = 0; k <= n; k++) dzc[k] = .5 * (dzf[k] + dzf[k + 1]);
    dzc[n + 1] = is_periodic ? dzc[1] : dzc[n];
  }
  zc[0] = -dzc[0] / 2.;
  zf[0] = 0.;
  for (int k = 1; k <= n + 1; k++) {
    zc[k] = zc[k - 1] + dzc[k - 1];
    zf[k] = zf[k - 1] + dzf[k];
  }
  return CANS_OK;
}

extern "C" int cans_initgrid(int gtype, int n, double gr, double lz, int is_periodic, double* dzc,
                             double* dzf, double* zc, double* zf) {
  return initgrid(gtype, n, gr, lz, is_periodic, false, dzc, dzf, zc, zf);
}

extern "C" int cans_initgrid_natural(int n, double lz, int is_periodic, double* dzc, double* dzf, double* zc,
                                     double* zf) {
  return initgrid(1, n, 0., lz, is_periodic, true, dzc, dzf, zc, zf);
}

extern "C" int cans_eigenvalues(int n, const char* bc, char c_or_f, double* lam) {
  if (n < 1 || !bc || !lam) return fail(CANS_EINVAL, "cans_eigenvalues: bad argument");
  const double pi = 3.14159265358979323846;
  const float fn = 1.f * (float)n;
  std::vector<double> l(n);
  char b0 = bc[0], b1 = bc[1];
  if (b0 == 'P' && b1 == 'P') {
    for (int i = 1; i <= n; i++) l[i - 1] = -2. * (1. - std::cos((2 * (i - 1)) * pi / (double)fn));
    // device (packed real-FFT) ordering: (r0, r_{n/2}, r1, i1, ...), initsolver.f90:98-117
    std::vector<int> isw(n + 1, 0);
    int nh = (n + 1) / 2;
    isw[1] = 1;
    if (n >= 2) isw[2] = nh + (1 - n % 2);
    for (int i = 2; i <= n - 1; i++) {
      if (i <= nh) isw[2 * i - 1] = i;
      else isw[n - 2 * (i - (nh + 1)) - n % 2] = i + 1;
    }
    for (int i = 1; i <= n; i++) lam[i - 1] = l[isw[i] - 1];
    return CANS_OK;
  }
  if (b0 == 'N' && b1 == 'N') {
    float fm = c_or_f == 'c' ? fn : 1.f * (float)(n - 1 + 1);
    for (int i = 1; i <= n; i++) lam[i - 1] = -2. * (1. - std::cos((i - 1) * pi / (double)fm));
  } else if (b0 == 'D' && b1 == 'D') {
    if (c_or_f == 'c') {
      for (int i = 1; i <= n; i++) lam[i - 1] = -2. * (1. - std::cos(i * pi / (double)fn));
    } else {
      float fm = 1.f * (float)(n + 1 - 1);
      for (int i = 1; i <= n - 1; i++) lam[i - 1] = -2. * (1. - std::cos(i * pi / (double)fm));
      lam[n - 1] = 0.;
    }
  } else if ((b0 == 'N' && b1 == 'D') || (b0 == 'D' && b1 == 'N')) {
    float f2n = 2.f * (float)n;
    for (int i = 1; i <= n; i++) lam[i - 1] = -2. * (1. - std::cos((2 * i - 1) * pi / (double)f2n));
  } else {
    return fail(CANS_EINVAL, "cans_eigenvalues: unsupported boundary pair");
  }
  return CANS_OK;
}

extern "C" int cans_bc_rhs(const char* cbc, const double* bc, const double* dlc, const double* dlf, char c_or_f,
                           double* rhs) {
  if (!cbc || !bc || !dlc || !dlf || !rhs || (c_or_f != 'c' && c_or_f != 'f'))
    return fail(CANS_EINVAL, "cans_bc_rhs: bad argument");
  // initsolver.f90:189-232: plane ib of the boundary rhs is factor(ib)/dlc(ib)/dlf(ib)
  for (int ib = 0; ib < 2; ib++) {
    double factor = 0.;
    if (cbc[ib] == 'D') factor = c_or_f == 'c' ? -2. * bc[ib] : -bc[ib];
    else if (cbc[ib] == 'N') factor = (ib == 0 ? 1. : -1.) * (c_or_f == 'c' ? dlc[ib] : dlf[ib]) * bc[ib];
    else if (cbc[ib] != 'P') return fail(CANS_EINVAL, "cans_bc_rhs: unknown boundary type");
    rhs[ib] = factor / dlc[ib] / dlf[ib];
  }
  return CANS_OK;
}

extern "C" int cans_tridmatrix(const char* bc, int n, const double* dzci, const double* dzfi, char c_or_f,
                               double* a, double* b, double* c) {
  if (n < 1 || !bc || !dzci || !dzfi || !a || !b || !c) return fail(CANS_EINVAL, "cans_tridmatrix: bad argument");
  for (int k = 1; k <= n; k++) {
    if (c_or_f == 'c') {
      a[k - 1] = dzfi[k] * dzci[k - 1];
      c[k - 1] = dzfi[k] * dzci[k];
    } else {
      a[k - 1] = dzfi[k] * dzci[k];
      c[k - 1] = dzfi[k + 1] * dzci[k];
    }
    b[k - 1] = -(a[k - 1] + c[k - 1]);
  }
  double f[2];
  for (int ib = 0; ib < 2; ib++) f[ib] = bc[ib] == 'P' ? 0. : bc[ib] == 'D' ? -1. : 1.;
  if (c_or_f == 'c') {
    b[0] = b[0] + f[0] * a[0];
    b[n - 1] = b[n - 1] + f[1] * c[n - 1];
  } else {
    if (bc[0] == 'N') b[0] = b[0] + f[0] * a[0];
    if (bc[1] == 'N') b[n - 1] = b[n - 1] + f[1] * c[n - 1];
  }
  return CANS_OK;
}
