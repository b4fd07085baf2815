// extern "C" boundary of the library (declared in include/qepw.h): argument
// checks, plan construction and error reporting.  No C++ exception crosses
// this boundary; failures return a negative code and set a per-thread
// message.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "qepw_internal.h"

struct qepw_plan {
  qepw::Plan p;
};

namespace {

thread_local std::string g_error;

int fail(int code, const char* msg) {
  g_error = msg;
  return code;
}

int fail_hip(hipError_t e, const char* where) {
  g_error = std::string(where) + ": " + hipGetErrorString(e);
  return QEPW_ERR_HIP;
}

int finish(int rc, const char* where) {
  if (rc == QEPW_OK) return rc;
  switch (rc) {
    case QEPW_ERR_ARG: g_error = std::string(where) + ": invalid argument"; break;
    case QEPW_ERR_UNSUPPORTED:
      g_error = std::string(where) + ": shape outside the kernels' supported range";
      break;
    case QEPW_ERR_WORKSPACE: g_error = std::string(where) + ": workspace too small"; break;
    default: {
      hipError_t e = hipGetLastError();
      g_error = std::string(where) + ": HIP failure (" + hipGetErrorString(e) + ")";
    }
  }
  return rc;
}

bool line_plan(int n, qepw::LinePlan* lp) {
  if (n < 1 || n > qepw::kMaxLine) return false;
  lp->n = n;
  lp->nstages = 0;
  int m = n;
  const int radices[4] = {4, 2, 3, 5};
  for (int r : radices) {
    while (m % r == 0) {
      if (lp->nstages >= qepw::kMaxStages) return false;
      lp->radix[lp->nstages++] = r;
      m /= r;
    }
  }
  for (int s = lp->nstages; s < qepw::kMaxStages; ++s) lp->radix[s] = 1;
  return m == 1;
}

template <typename T>
hipError_t upload(T** dst, const std::vector<T>& src) {
  *dst = nullptr;
  if (src.empty()) return hipSuccess;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(dst), src.size() * sizeof(T));
  if (e != hipSuccess) return e;
  return hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
}

void free_tables(qepw::StickTables& t) {
  hipFree(t.d_off);
  hipFree(t.d_gpos);
  hipFree(t.d_plane);
  hipFree(t.d_xcols);
  hipFree(t.d_stickq);
  hipFree(t.d_colmap);
  hipFree(t.d_colstart);
  hipFree(t.d_colpos);
  t.d_off = t.d_gpos = t.d_plane = t.d_xcols = t.d_stickq = t.d_colmap = nullptr;
  t.d_colstart = t.d_colpos = nullptr;
}

void free_plan(qepw::Plan& p) {
  free_tables(p.sphere);
  free_tables(p.full);
  for (int a = 0; a < 3; ++a) {
    hipFree(p.d_tw64[a]);
    hipFree(p.d_tw32[a]);
    p.d_tw64[a] = nullptr;
    p.d_tw32[a] = nullptr;
  }
  hipFree(p.d_perm);
  p.d_perm = nullptr;
}

struct TableHost {
  std::vector<int> off, gpos, plane, xcols, stickq, colmap, colstart, colpos;
};

hipError_t upload_tables(qepw::StickTables& t, const TableHost& h) {
  hipError_t e;
  if ((e = upload(&t.d_off, h.off)) != hipSuccess) return e;
  if ((e = upload(&t.d_gpos, h.gpos)) != hipSuccess) return e;
  if ((e = upload(&t.d_plane, h.plane)) != hipSuccess) return e;
  if ((e = upload(&t.d_stickq, h.stickq)) != hipSuccess) return e;
  if ((e = upload(&t.d_colmap, h.colmap)) != hipSuccess) return e;
  if ((e = upload(&t.d_colstart, h.colstart)) != hipSuccess) return e;
  if ((e = upload(&t.d_colpos, h.colpos)) != hipSuccess) return e;
  return upload(&t.d_xcols, h.xcols);
}

int pmod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

}  // namespace

extern "C" {

const char* qepw_last_error(void) { return g_error.c_str(); }

int qepw_abi_version(void) { return 1; }

static int plan_create_impl(const int32_t* miller, int64_t npw, int n1, int n2, int n3,
                            int device, qepw_plan** out) {
  if (!out) return fail(QEPW_ERR_ARG, "qepw_plan_create: out is NULL");
  *out = nullptr;
  if (!miller || npw < 1) return fail(QEPW_ERR_ARG, "qepw_plan_create: empty basis");
  if (npw > 0x7fffffff) return fail(QEPW_ERR_UNSUPPORTED, "qepw_plan_create: npw too large");
  qepw::LinePlan lp[3];
  const int dims[3] = {n1, n2, n3};
  for (int a = 0; a < 3; ++a)
    if (!line_plan(dims[a], &lp[a]))
      return fail(QEPW_ERR_UNSUPPORTED,
                  "qepw_plan_create: FFT lengths must be 2,3,5-smooth and <= 256");
  const int64_t ngrid = (int64_t)n1 * n2 * n3;
  if (ngrid > 0x7fffffff) return fail(QEPW_ERR_UNSUPPORTED, "qepw_plan_create: grid too large");

  // slots and stick-major order
  std::vector<int64_t> key(npw);
  std::vector<int> zs(npw), stick_key(npw);
  for (int64_t g = 0; g < npw; ++g) {
    const int x = pmod(miller[3 * g], n1), y = pmod(miller[3 * g + 1], n2),
              z = pmod(miller[3 * g + 2], n3);
    stick_key[g] = x * n2 + y;
    zs[g] = z;
    key[g] = (int64_t)stick_key[g] * n3 + z;
  }
  std::vector<int> perm(npw);
  for (int64_t g = 0; g < npw; ++g) perm[g] = (int)g;
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return key[a] < key[b]; });
  for (int64_t d = 1; d < npw; ++d)
    if (key[perm[d]] == key[perm[d - 1]])
      return fail(QEPW_ERR_ARG, "qepw_plan_create: two plane waves share one FFT slot "
                                "(grid too small for the basis)");

  TableHost sphere;
  std::vector<int> xs_seen(n1, 0);
  sphere.gpos.resize(npw);
  int nst = 0;
  for (int64_t d = 0; d < npw; ++d) {
    const int g = perm[d];
    if (d == 0 || stick_key[g] != stick_key[perm[d - 1]]) {
      sphere.off.push_back((int)d);
      sphere.plane.push_back(stick_key[g]);
      xs_seen[stick_key[g] / n2] = 1;
      ++nst;
    }
    sphere.gpos[d] = (nst - 1) * n3 + zs[g];
  }
  sphere.off.push_back((int)npw);
  for (int x = 0; x < n1; ++x)
    if (xs_seen[x]) sphere.xcols.push_back(x);
  sphere.colmap.assign(n1, -1);
  for (size_t c = 0; c < sphere.xcols.size(); ++c) sphere.colmap[sphere.xcols[c]] = (int)c;
  sphere.stickq.resize(nst);
  sphere.colstart.assign(sphere.xcols.size() + 1, nst);
  for (int s = nst - 1; s >= 0; --s) {
    const int x = sphere.plane[s] / n2, y = sphere.plane[s] % n2;
    sphere.stickq[s] = sphere.colmap[x] * n2 + y;
    sphere.colstart[sphere.colmap[x]] = s;  // sticks are sorted by x, then y
  }
  sphere.colpos.assign(sphere.xcols.size() * (size_t)n2, -1);
  for (int s = 0; s < nst; ++s) sphere.colpos[sphere.stickq[s]] = s;

  TableHost full;
  const int nfull = n1 * n2;
  full.off.resize(nfull + 1);
  full.plane.resize(nfull);
  full.gpos.resize(ngrid);
  for (int s = 0; s <= nfull; ++s) full.off[s] = s * n3;
  for (int s = 0; s < nfull; ++s) full.plane[s] = s;
  for (int64_t i = 0; i < ngrid; ++i) full.gpos[i] = (int)i;
  for (int x = 0; x < n1; ++x) full.xcols.push_back(x);
  full.stickq.resize(nfull);
  for (int s = 0; s < nfull; ++s) full.stickq[s] = s;  // column x = s / n2, y = s % n2
  full.colstart.resize(n1 + 1);
  for (int x = 0; x <= n1; ++x) full.colstart[x] = x * n2;
  full.colpos.resize(nfull);
  for (int s = 0; s < nfull; ++s) full.colpos[s] = s;

  qepw_plan* plan = new (std::nothrow) qepw_plan();
  if (!plan) return fail(QEPW_ERR_HIP, "qepw_plan_create: out of host memory");
  qepw::Plan& p = plan->p;
  p.n1 = n1;
  p.n2 = n2;
  p.n3 = n3;
  p.npw = npw;
  p.device = device;
  for (int a = 0; a < 3; ++a) p.line[a] = lp[a];
  p.sphere.nst = nst;
  p.sphere.nxcols = (int)sphere.xcols.size();
  if (n1 == n2) {
    std::vector<int> stick_y(nst);
    for (int s = 0; s < nst; ++s) stick_y[s] = sphere.plane[s] % n2;
    p.sphere.slot_mask = qepw::team_slot_mask(n1, sphere.xcols, stick_y);
  }
  p.sphere.npoints = npw;
  p.sphere.group = qepw::stick_group_for(n3, nst);
  p.sphere.ngroups = (nst + p.sphere.group - 1) / p.sphere.group;
  p.full.nst = nfull;
  p.full.nxcols = n1;
  p.full.npoints = ngrid;
  p.full.group = qepw::stick_group_for(n3, nfull);
  p.full.ngroups = (nfull + p.full.group - 1) / p.full.group;
  p.h_perm = perm;

  int prev = 0;
  hipGetDevice(&prev);
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = upload_tables(p.sphere, sphere);
  if (e == hipSuccess) e = upload_tables(p.full, full);
  if (e == hipSuccess) e = upload(&p.d_perm, perm);
  for (int a = 0; a < 3 && e == hipSuccess; ++a) {
    const int n = dims[a];
    std::vector<double2> t64(n);
    std::vector<float2> t32(n);
    for (int k = 0; k < n; ++k) {
      const long double ang = -2.0L * 3.141592653589793238462643383279502884L * k / n;
      t64[k].x = (double)cosl(ang);
      t64[k].y = (double)sinl(ang);
      t32[k].x = (float)cosl(ang);
      t32[k].y = (float)sinl(ang);
    }
    e = upload(&p.d_tw64[a], t64);
    if (e == hipSuccess) e = upload(&p.d_tw32[a], t32);
  }
  hipSetDevice(prev);
  if (e != hipSuccess) {
    free_plan(p);
    delete plan;
    return fail_hip(e, "qepw_plan_create");
  }
  *out = plan;
  return QEPW_OK;
}

// host allocations of the tables can throw: no exception leaves the C ABI
int qepw_plan_create(const int32_t* miller, int64_t npw, int n1, int n2, int n3, int device,
                     qepw_plan** out) {
  try {
    return plan_create_impl(miller, npw, n1, n2, n3, device, out);
  } catch (const std::exception& e) {
    if (out) *out = nullptr;
    return fail(QEPW_ERR_HIP, (std::string("qepw_plan_create: host allocation: ") + e.what()).c_str());
  }
}

int qepw_plan_destroy(qepw_plan* plan) {
  if (!plan) return QEPW_OK;
  int prev = 0;
  hipGetDevice(&prev);
  hipSetDevice(plan->p.device);
  free_plan(plan->p);
  hipSetDevice(prev);
  delete plan;
  return QEPW_OK;
}

int qepw_plan_query(const qepw_plan* plan, int64_t* info) {
  if (!plan || !info) return fail(QEPW_ERR_ARG, "qepw_plan_query: NULL argument");
  const qepw::Plan& p = plan->p;
  info[0] = p.npw;
  info[1] = p.n1;
  info[2] = p.n2;
  info[3] = p.n3;
  info[4] = p.sphere.nst;
  info[5] = p.sphere.nxcols;
  info[6] = p.sphere.group;
  info[7] = p.full.nst;
  return QEPW_OK;
}

int qepw_plan_permutation(const qepw_plan* plan, int32_t* perm) {
  if (!plan || !perm) return fail(QEPW_ERR_ARG, "qepw_plan_permutation: NULL argument");
  std::memcpy(perm, plan->p.h_perm.data(), plan->p.h_perm.size() * sizeof(int32_t));
  return QEPW_OK;
}

const int32_t* qepw_plan_device_permutation(const qepw_plan* plan) {
  return plan ? plan->p.d_perm : nullptr;
}

size_t qepw_apply_local_workspace(const qepw_plan* plan, int precision, int nb) {
  if (!plan || nb < 0 || (precision != QEPW_C128 && precision != QEPW_C64)) return 0;
  return qepw::apply_local_band_bytes(plan->p, precision) * (size_t)nb;
}

namespace {
// the message of a local-potential plan whose planes the kernels cannot hold
int unsupported_grid(const qepw::Plan& p, const char* where) {
  g_error = std::string(where) + ": the " + std::to_string(p.n1) + " x " + std::to_string(p.n2) +
            " x " + std::to_string(p.n3) +
            " grid is outside the FFT kernels' range (an xy plane's active columns and "
            "exchange lines, or a z stick group, must fit one workgroup's 160 KB of LDS)";
  return QEPW_ERR_UNSUPPORTED;
}
}  // namespace

int qepw_plan_supported(const qepw_plan* plan, int precision) {
  if (!plan) return fail(QEPW_ERR_ARG, "qepw_plan_supported: NULL plan");
  if (precision != QEPW_C128 && precision != QEPW_C64)
    return fail(QEPW_ERR_ARG, "qepw_plan_supported: unknown precision");
  // no bands: only the capability checks of qepw_apply_local run
  const int rc = qepw::apply_local(plan->p, precision, nullptr, 0, 0, nullptr, nullptr, nullptr,
                                   0, nullptr, 0, nullptr);
  if (rc == QEPW_ERR_UNSUPPORTED) return unsupported_grid(plan->p, "qepw_plan_supported");
  return finish(rc, "qepw_plan_supported");
}

int qepw_apply_local(const qepw_plan* plan, int precision, const void* vectors, int64_t ldv,
                     int nb, const void* potential, const void* kinetic, void* result,
                     int64_t ldr, void* workspace, size_t ws_bytes, void* stream) {
  if (!plan) return fail(QEPW_ERR_ARG, "qepw_apply_local: NULL plan");
  if (nb < 0) return fail(QEPW_ERR_ARG, "qepw_apply_local: negative band count");
  if (nb == 0) return QEPW_OK;
  if (!vectors || !potential || !result || !workspace)
    return fail(QEPW_ERR_ARG, "qepw_apply_local: NULL buffer");
  if (ldv < plan->p.npw || ldr < plan->p.npw)
    return fail(QEPW_ERR_ARG, "qepw_apply_local: leading dimension smaller than npw");
  if (vectors == result) return fail(QEPW_ERR_ARG, "qepw_apply_local: result aliases input");
  const int rc = qepw::apply_local(plan->p, precision, vectors, ldv, nb, potential, kinetic,
                                   result, ldr, workspace, ws_bytes,
                                   static_cast<hipStream_t>(stream));
  if (rc == QEPW_ERR_UNSUPPORTED) return unsupported_grid(plan->p, "qepw_apply_local");
  return finish(rc, "qepw_apply_local");
}

int qepw_apply_local_timed(const qepw_plan* plan, int precision, const void* vectors,
                           int64_t ldv, int nb, const void* potential, const void* kinetic,
                           void* result, int64_t ldr, void* workspace, size_t ws_bytes,
                           void* start_event, void* stop_event, void* stream) {
  if (!start_event || !stop_event) return fail(QEPW_ERR_ARG, "qepw_apply_local_timed: NULL event");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = hipEventRecord(static_cast<hipEvent_t>(start_event), s);
  if (e != hipSuccess) return fail_hip(e, "qepw_apply_local_timed: hipEventRecord");
  const int rc = qepw_apply_local(plan, precision, vectors, ldv, nb, potential, kinetic, result,
                                  ldr, workspace, ws_bytes, stream);
  if (rc != QEPW_OK) return rc;
  e = hipEventRecord(static_cast<hipEvent_t>(stop_event), s);
  if (e != hipSuccess) return fail_hip(e, "qepw_apply_local_timed: hipEventRecord");
  return QEPW_OK;
}

static int apply_local_group_impl(int count, const qepw_plan* const* plans, int precision,
                                  const void* const* vectors, const int64_t* ldv,
                                  const int* nb, const void* const* potentials,
                                  const void* const* kinetics, void* const* results,
                                  const int64_t* ldr, void* const* workspaces,
                                  const size_t* ws_bytes, void* stream) {
  if (count < 0 || (count > 0 && (!plans || !vectors || !ldv || !nb || !potentials ||
                                  !kinetics || !results || !ldr || !workspaces || !ws_bytes)))
    return fail(QEPW_ERR_ARG, "qepw_apply_local_group: NULL argument");
  if (precision != QEPW_C128 && precision != QEPW_C64)
    return fail(QEPW_ERR_ARG, "qepw_apply_local_group: unknown precision");
  std::vector<const qepw::Plan*> ps((size_t)count);
  for (int i = 0; i < count; ++i) {
    if (!plans[i]) return fail(QEPW_ERR_ARG, "qepw_apply_local_group: NULL plan");
    if (nb[i] < 0) return fail(QEPW_ERR_ARG, "qepw_apply_local_group: negative band count");
    if (nb[i] > 0 && (!vectors[i] || !potentials[i] || !results[i] || !workspaces[i] ||
                      ldv[i] < plans[i]->p.npw || ldr[i] < plans[i]->p.npw))
      return fail(QEPW_ERR_ARG, "qepw_apply_local_group: bad problem arguments");
    ps[(size_t)i] = &plans[i]->p;
  }
  return finish(qepw::apply_local_group(count, ps.data(), precision, vectors, ldv, nb,
                                        potentials, kinetics, results, ldr, workspaces,
                                        ws_bytes, static_cast<hipStream_t>(stream)),
                "qepw_apply_local_group");
}

int qepw_apply_local_group(int count, const qepw_plan* const* plans, int precision,
                           const void* const* vectors, const int64_t* ldv, const int* nb,
                           const void* const* potentials, const void* const* kinetics,
                           void* const* results, const int64_t* ldr, void* const* workspaces,
                           const size_t* ws_bytes, void* stream) {
  try {
    return apply_local_group_impl(count, plans, precision, vectors, ldv, nb, potentials,
                                  kinetics, results, ldr, workspaces, ws_bytes, stream);
  } catch (const std::exception& e) {
    return fail(QEPW_ERR_HIP,
                (std::string("qepw_apply_local_group: host allocation: ") + e.what()).c_str());
  }
}

size_t qepw_inverse_real_workspace(const qepw_plan* plan) {
  return plan ? qepw::inverse_real_bytes(plan->p) : 0;
}

int qepw_inverse_real(const qepw_plan* plan, const void* grid_g, double* out, void* workspace,
                      size_t ws_bytes, void* stream) {
  if (!plan || !grid_g || !out || !workspace)
    return fail(QEPW_ERR_ARG, "qepw_inverse_real: NULL argument");
  return finish(qepw::inverse_real(plan->p, grid_g, out, workspace, ws_bytes,
                                   static_cast<hipStream_t>(stream)),
                "qepw_inverse_real");
}

size_t qepw_gemm_workspace(int precision, int op_a, int64_t m, int64_t n, int64_t k) {
  return qepw::gemm_workspace_bytes(precision, op_a, m, n, k);
}

int qepw_gemm(int precision, int op_a, int64_t m, int64_t n, int64_t k, const double* alpha,
              const void* a, int64_t lda, const void* b, int64_t ldb, const double* beta,
              void* c, int64_t ldc, void* workspace, size_t ws_bytes, void* stream) {
  if (!alpha || !beta) return fail(QEPW_ERR_ARG, "qepw_gemm: NULL alpha/beta");
  if ((m > 0 && n > 0) && (!c || (k > 0 && (!a || !b))))
    return fail(QEPW_ERR_ARG, "qepw_gemm: NULL matrix");
  return finish(qepw::gemm(precision, op_a, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                           workspace, ws_bytes, static_cast<hipStream_t>(stream)),
                "qepw_gemm");
}

int qepw_gemm_variant_count(void) { return qepw::gemm_variant_count(); }

int qepw_gemm_plan(int precision, int op_a, int64_t m, int64_t n, int64_t k, int64_t* out) {
  if (!out || m < 0 || n < 0 || k < 0) return fail(QEPW_ERR_ARG, "qepw_gemm_plan: bad argument");
  return finish(qepw::gemm_plan(precision, op_a, m, n, k, out), "qepw_gemm_plan");
}

int qepw_gemm_set_rule(int precision, int op_a, const int64_t* ranges, int variant,
                       int64_t splits) {
  const int rc = qepw::gemm_set_rule(precision, op_a, ranges, variant, splits);
  if (rc == QEPW_ERR_UNSUPPORTED) return fail(rc, "qepw_gemm_set_rule: rule table full");
  return finish(rc, "qepw_gemm_set_rule");
}

int qepw_gemm_clear_rules(void) {
  qepw::gemm_clear_rules();
  return QEPW_OK;
}

}  // extern "C"

// Scratch of qepw_gemm_static: the split-K partial tiles live in this
// file-scope device array instead of a caller or library allocation.  Each
// GPU the library runs on holds its own copy (QEPW_GEMM_STATIC_SCRATCH
// bytes) until the process ends.  Host code never uses the array's name as an
// address: the device address of each GPU's copy comes from
// hipGetSymbolAddress, taken once per device with that device current.
namespace {
constexpr size_t kStaticScratchBytes = QEPW_GEMM_STATIC_SCRATCH;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the calling thread's current device's scratch (or an
// error code from the lookup).
int static_scratch(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail_hip(e, "qepw_gemm_static: hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices)
    return fail(QEPW_ERR_UNSUPPORTED, "qepw_gemm_static: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail_hip(e != hipSuccess ? e : hipErrorInvalidSymbol,
                      "qepw_gemm_static: hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return QEPW_OK;
}
}  // namespace

extern "C" {

size_t qepw_gemm_static_scratch_bytes(void) { return kStaticScratchBytes; }

int qepw_gemm_static(int precision, int op_a, int64_t m, int64_t n, int64_t k,
                     const double* alpha, const void* a, int64_t lda, const void* b,
                     int64_t ldb, const double* beta, void* c, int64_t ldc, void* stream) {
  if (!alpha || !beta) return fail(QEPW_ERR_ARG, "qepw_gemm_static: NULL alpha/beta");
  if ((m > 0 && n > 0) && (!c || (k > 0 && (!a || !b))))
    return fail(QEPW_ERR_ARG, "qepw_gemm_static: NULL matrix");
  if (m < 0 || n < 0 || k < 0) return fail(QEPW_ERR_ARG, "qepw_gemm_static: negative size");
  if (precision != QEPW_C128 && precision != QEPW_C64)
    return fail(QEPW_ERR_ARG, "qepw_gemm_static: unknown precision");
  const size_t need = qepw::gemm_workspace_bytes(precision, op_a, m, n, k);
  if (need > kStaticScratchBytes) {
    char msg[192];
    std::snprintf(msg, sizeof msg,
                  "qepw_gemm_static: this product needs %zu bytes of split-K scratch, more "
                  "than the static %zu (use qepw_gemm with a workspace)",
                  need, kStaticScratchBytes);
    return fail(QEPW_ERR_WORKSPACE, msg);
  }
  void* ptr = nullptr;
  if (need > 0) {
    const int rc = static_scratch(&ptr);
    if (rc != QEPW_OK) return rc;
  }
  return finish(qepw::gemm(precision, op_a, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, ptr,
                           need > 0 ? kStaticScratchBytes : 0, static_cast<hipStream_t>(stream)),
                "qepw_gemm_static");
}

int qepw_column_squared_norms(int precision, const void* v, int64_t ldv, int64_t nrows, int nb,
                              double* out, void* stream) {
  if (nb > 0 && (!v || !out || ldv < nrows))
    return fail(QEPW_ERR_ARG, "qepw_column_squared_norms: bad argument");
  return finish(qepw::column_squared_norms(precision, v, ldv, nrows, nb, out,
                                           static_cast<hipStream_t>(stream)),
                "qepw_column_squared_norms");
}

int qepw_column_inner_products(int precision, const void* l, int64_t ldl, const void* r,
                               int64_t ldr, int64_t nrows, int nb, double* out,
                               void* stream) {
  if (nb > 0 && (!l || !r || !out || ldl < nrows || (ldr < nrows && ldr != 0)))
    return fail(QEPW_ERR_ARG, "qepw_column_inner_products: bad argument");
  return finish(qepw::column_inner_products(precision, l, ldl, r, ldr, nrows, nb, out,
                                            static_cast<hipStream_t>(stream)),
                "qepw_column_inner_products");
}

int qepw_subtract_band_energies(int precision, void* r, int64_t ldr, const void* x,
                                int64_t ldx, const double* e, int64_t nrows, int nb,
                                void* stream) {
  if (nb > 0 && (!r || !x || !e || ldr < nrows || ldx < nrows))
    return fail(QEPW_ERR_ARG, "qepw_subtract_band_energies: bad argument");
  return finish(qepw::subtract_band_energies(precision, r, ldr, x, ldx, e, nrows, nb,
                                             static_cast<hipStream_t>(stream)),
                "qepw_subtract_band_energies");
}

int qepw_qe_precondition(int precision, const void* r, int64_t ldr, const int32_t* cols,
                         const double* e, const double* diag, void* out, int64_t ldo,
                         int64_t nrows, int nb, void* stream) {
  if (nb > 0 && (!r || !e || !diag || !out || ldr < nrows || ldo < nrows))
    return fail(QEPW_ERR_ARG, "qepw_qe_precondition: bad argument");
  return finish(qepw::qe_precondition(precision, r, ldr, cols, e, diag, out, ldo, nrows, nb,
                                      static_cast<hipStream_t>(stream)),
                "qepw_qe_precondition");
}

int qepw_normalize_selected_columns(int precision, const void* src, int64_t lds,
                                    const double* sq, const int32_t* sel, void* dst,
                                    int64_t ldd, int64_t nrows, int nb, void* stream) {
  if (nb > 0 && (!src || !sq || !sel || !dst || lds < nrows || ldd < nrows))
    return fail(QEPW_ERR_ARG, "qepw_normalize_selected_columns: bad argument");
  return finish(qepw::normalize_selected_columns(precision, src, lds, sq, sel, dst, ldd, nrows,
                                                 nb, static_cast<hipStream_t>(stream)),
                "qepw_normalize_selected_columns");
}

int qepw_restart_applied(int precision, const void* x, int64_t ldx, const void* r,
                         int64_t ldr, const double* e, void* hv, int64_t ldh, int64_t nrows,
                         int nb, void* stream) {
  if (nb > 0 && (!x || !r || !e || !hv || ldx < nrows || ldr < nrows || ldh < nrows))
    return fail(QEPW_ERR_ARG, "qepw_restart_applied: bad argument");
  return finish(qepw::restart_applied(precision, x, ldx, r, ldr, e, hv, ldh, nrows, nb,
                                      static_cast<hipStream_t>(stream)),
                "qepw_restart_applied");
}

int qepw_permute_rows(int src_precision, int dst_precision, int direction, const int32_t* perm,
                      const void* src, int64_t lds, void* dst, int64_t ldd, int64_t nrows,
                      int nb, void* stream) {
  if (nb > 0 && (!src || !dst || lds < nrows || ldd < nrows))
    return fail(QEPW_ERR_ARG, "qepw_permute_rows: bad argument");
  if (src == dst && perm) return fail(QEPW_ERR_ARG, "qepw_permute_rows: in-place permutation");
  return finish(qepw::permute_rows(src_precision, dst_precision, direction, perm, src, lds, dst,
                                   ldd, nrows, nb, static_cast<hipStream_t>(stream)),
                "qepw_permute_rows");
}

int qepw_copy_columns(int src_precision, int dst_precision, const void* src, int64_t lds,
                      const int32_t* src_cols, void* dst, int64_t ldd, const int32_t* dst_cols,
                      const double* scale, int64_t nrows, int ncols, void* stream) {
  if (ncols > 0 && nrows > 0 && (!src || !dst || lds < nrows || ldd < nrows))
    return fail(QEPW_ERR_ARG, "qepw_copy_columns: bad argument");
  if (ncols < 0 || nrows < 0) return fail(QEPW_ERR_ARG, "qepw_copy_columns: negative size");
  return finish(qepw::copy_columns(src_precision, dst_precision, src, lds, src_cols, dst, ldd,
                                   dst_cols, scale, nrows, ncols,
                                   static_cast<hipStream_t>(stream)),
                "qepw_copy_columns");
}

int qepw_set_block(int precision, void* dst, int64_t ldd, int64_t nrows, int ncols, double diag,
                   const double* d, void* stream) {
  if (ncols > 0 && nrows > 0 && (!dst || ldd < nrows))
    return fail(QEPW_ERR_ARG, "qepw_set_block: bad argument");
  if (ncols < 0 || nrows < 0) return fail(QEPW_ERR_ARG, "qepw_set_block: negative size");
  return finish(qepw::set_block(precision, dst, ldd, nrows, ncols, diag, d,
                                static_cast<hipStream_t>(stream)),
                "qepw_set_block");
}

int qepw_accumulate_projector_diagonal(const void* coupled, int64_t ldc, const void* beta,
                                       int64_t ldb, int64_t nrows, int nch,
                                       double multiplicity, double* diagonal, void* stream) {
  if (nrows > 0 && nch > 0 && (!coupled || !beta || !diagonal || ldc < nrows || ldb < nrows))
    return fail(QEPW_ERR_ARG, "qepw_accumulate_projector_diagonal: bad argument");
  if (nch <= 0) return QEPW_OK;
  return finish(qepw::row_real_dot(coupled, ldc, beta, ldb, nrows, nch, multiplicity, diagonal,
                                   static_cast<hipStream_t>(stream)),
                "qepw_accumulate_projector_diagonal");
}

}  // extern "C"

extern "C" {

size_t qepw_accumulate_density_workspace(const qepw_plan* plan, int precision, int nb) {
  if (!plan || (precision != QEPW_C128 && precision != QEPW_C64)) return 0;
  return qepw::density_workspace_bytes(plan->p, precision, nb);
}

int qepw_accumulate_density(const qepw_plan* plan, int precision, const void* vectors,
                            int64_t ldv, int nb, const double* weights, double scale,
                            double* out, void* workspace, size_t ws_bytes, void* stream) {
  if (!plan || !out || !workspace) return fail(QEPW_ERR_ARG, "qepw_accumulate_density: NULL argument");
  if (nb < 0) return fail(QEPW_ERR_ARG, "qepw_accumulate_density: negative band count");
  if (nb > 0 && (!vectors || !weights || ldv < plan->p.npw))
    return fail(QEPW_ERR_ARG, "qepw_accumulate_density: bad vectors / weights");
  return finish(qepw::accumulate_density(plan->p, precision, vectors, ldv, nb, weights, scale, out,
                                         workspace, ws_bytes, static_cast<hipStream_t>(stream)),
                "qepw_accumulate_density");
}

int qepw_scale_rows(int precision, const void* src, int64_t lds, const double* s,
                    int64_t sstride, const double* cfac, void* dst, int64_t ldd, int64_t nrows,
                    int nb, void* stream) {
  if (nb > 0 && nrows > 0 && (!src || !dst || lds < nrows || ldd < nrows || sstride < 0))
    return fail(QEPW_ERR_ARG, "qepw_scale_rows: bad argument");
  return finish(qepw::scale_rows(precision, src, lds, s, sstride, cfac, dst, ldd, nrows, nb,
                                 static_cast<hipStream_t>(stream)),
                "qepw_scale_rows");
}

int qepw_column_axpby(int precision, const void* x, int64_t ldx, const double* a, void* y,
                      int64_t ldy, const double* b, int64_t nrows, int ncols, void* stream) {
  if (nrows < 0 || ncols < 0) return fail(QEPW_ERR_ARG, "qepw_column_axpby: negative size");
  if (nrows > 0 && ncols > 0 && (!x || !y || ldx < nrows || ldy < nrows))
    return fail(QEPW_ERR_ARG, "qepw_column_axpby: bad argument");
  return finish(qepw::column_axpby(precision, x, ldx, a, y, ldy, b, nrows, ncols,
                                   static_cast<hipStream_t>(stream)),
                "qepw_column_axpby");
}

int qepw_column_diagonal_expectations(int precision, const void* v, int64_t ldv,
                                      const double* d, int64_t nrows, int nb, double* out,
                                      void* stream) {
  if (nrows < 0 || nb < 0)
    return fail(QEPW_ERR_ARG, "qepw_column_diagonal_expectations: negative size");
  if (nb > 0 && (!v || !d || !out || ldv < nrows))
    return fail(QEPW_ERR_ARG, "qepw_column_diagonal_expectations: bad argument");
  return finish(qepw::column_diagonal_expectations(precision, v, ldv, d, nrows, nb, out,
                                                   static_cast<hipStream_t>(stream)),
                "qepw_column_diagonal_expectations");
}

int qepw_rmm_precondition(int precision, const void* r, int64_t ldr, const int32_t* cols,
                          const double* kinetic, const double* expectations, void* out,
                          int64_t ldo, int64_t nrows, int ncols, void* stream) {
  if (nrows < 0 || ncols < 0) return fail(QEPW_ERR_ARG, "qepw_rmm_precondition: negative size");
  if (nrows > 0 && ncols > 0 &&
      (!r || !kinetic || !expectations || !out || ldr < nrows || ldo < nrows))
    return fail(QEPW_ERR_ARG, "qepw_rmm_precondition: bad argument");
  return finish(qepw::rmm_precondition(precision, r, ldr, cols, kinetic, expectations, out, ldo,
                                       nrows, ncols, static_cast<hipStream_t>(stream)),
                "qepw_rmm_precondition");
}

int qepw_history_combine(int precision, const void* hist, int64_t ldh, int64_t hstride,
                         int count, const int32_t* cols, const double* coef, void* out,
                         int64_t ldo, int64_t nrows, int ncols, void* stream) {
  if (nrows < 0 || ncols < 0 || count < 0)
    return fail(QEPW_ERR_ARG, "qepw_history_combine: negative size");
  if (nrows > 0 && ncols > 0 &&
      (!hist || !coef || !out || count < 1 || ldh < nrows || ldo < nrows || hstride < 0))
    return fail(QEPW_ERR_ARG, "qepw_history_combine: bad argument");
  return finish(qepw::history_combine(precision, hist, ldh, hstride, count, cols, coef, out, ldo,
                                      nrows, ncols, static_cast<hipStream_t>(stream)),
                "qepw_history_combine");
}

int qepw_build_projectors(int precision, const void* beta, int64_t ldb, int nch, const double* q,
                          const double* tau, int natom, void* out, int64_t ldo, int64_t nrows,
                          void* stream) {
  if (natom > 0 && nch > 0 && nrows > 0 &&
      (!beta || !q || !tau || !out || ldb < nrows || ldo < nrows))
    return fail(QEPW_ERR_ARG, "qepw_build_projectors: bad argument");
  return finish(qepw::build_projectors(precision, beta, ldb, nch, q, tau, natom, out, ldo, nrows,
                                       static_cast<hipStream_t>(stream)),
                "qepw_build_projectors");
}

int qepw_projector_basis(const double* q, int64_t npw, const double* table, int64_t ldt,
                         int nq, double dq, const int32_t* chan_radial, const int32_t* chan_l,
                         const int32_t* chan_m, int nch, void* beta, int64_t ldb, void* stream) {
  if (npw < 0 || nch < 0) return fail(QEPW_ERR_ARG, "qepw_projector_basis: negative size");
  if (npw == 0 || nch == 0) return QEPW_OK;
  if (!q || !table || !chan_radial || !chan_l || !chan_m || !beta || ldb < npw || nq < 4 ||
      ldt < nq || !(dq > 0.0))
    return fail(QEPW_ERR_ARG, "qepw_projector_basis: bad argument");
  return finish(qepw::projector_basis(q, npw, table, ldt, nq, dq, chan_radial, chan_l, chan_m,
                                      nch, beta, ldb, static_cast<hipStream_t>(stream)),
                "qepw_projector_basis");
}

int qepw_phase_blocks(int precision, const void* beta, int64_t ldb, int nch,
                      const void* phases, int64_t ldp, int natom, const int32_t* perm, void* out,
                      int64_t ldo, int64_t nrows, void* stream) {
  if (nch < 0 || natom < 0 || nrows < 0)
    return fail(QEPW_ERR_ARG, "qepw_phase_blocks: negative size");
  if (natom > 0 && nch > 0 && nrows > 0 &&
      (!beta || !phases || !out || ldb < nrows || ldp < nrows || ldo < nrows))
    return fail(QEPW_ERR_ARG, "qepw_phase_blocks: bad argument");
  return finish(qepw::phase_blocks(precision, beta, ldb, nch, phases, ldp, natom, perm, out, ldo,
                                   nrows, static_cast<hipStream_t>(stream)),
                "qepw_phase_blocks");
}

int qepw_projector_forces(int precision, const void* dO, int64_t lddo, int64_t dstride,
                          const void* DO, int64_t lddO, const double* f, int nb,
                          const int32_t* off, int natom, double* out, void* stream) {
  if (natom > 0 && (!dO || !DO || !f || !off || !out || nb < 0))
    return fail(QEPW_ERR_ARG, "qepw_projector_forces: bad argument");
  return finish(qepw::projector_forces(precision, dO, lddo, dstride, DO, lddO, f, nb, off, natom,
                                       out, static_cast<hipStream_t>(stream)),
                "qepw_projector_forces");
}

}  // extern "C"

extern "C" {

int qepw_hermitian_complete(void* a, int64_t lda, int first, int active, void* stream) {
  if (!a || first < 0 || active < first || lda < active)
    return fail(QEPW_ERR_ARG, "qepw_hermitian_complete: bad argument");
  return finish(qepw::hermitian_complete(a, lda, first, active, static_cast<hipStream_t>(stream)),
                "qepw_hermitian_complete");
}

int qepw_shift_scale(const void* a, int64_t lda, int n, double sigma, void* x, int64_t ldx,
                     double* bound, void* stream) {
  if (n < 0 || (n > 0 && (!a || !x || !bound || lda < n || ldx < n)))
    return fail(QEPW_ERR_ARG, "qepw_shift_scale: bad argument");
  return finish(qepw::shift_scale(a, lda, n, sigma, x, ldx, bound,
                                  static_cast<hipStream_t>(stream)),
                "qepw_shift_scale");
}

int qepw_hermitize(void* x, int64_t ldx, int n, void* stream) {
  if (n < 0 || (n > 0 && (!x || ldx < n))) return fail(QEPW_ERR_ARG, "qepw_hermitize: bad argument");
  return finish(qepw::hermitize(x, ldx, n, static_cast<hipStream_t>(stream)), "qepw_hermitize");
}

int qepw_sign_check(const void* x, int64_t ldx, const void* x2, int64_t ld2, int n, double* out,
                    void* stream) {
  if (n <= 0 || !x || !x2 || !out || ldx < n || ld2 < n)
    return fail(QEPW_ERR_ARG, "qepw_sign_check: bad argument");
  return finish(qepw::sign_check(x, ldx, x2, ld2, n, out, static_cast<hipStream_t>(stream)),
                "qepw_sign_check");
}

size_t qepw_sign_iteration_workspace(int n) {
  return n > 0 ? qepw::sign_iteration_workspace(n) : 0;
}

int qepw_sign_iteration(void* x, void* x2, void* y, int64_t ld, int n, const double* alphas,
                        int nalpha, int* in_x2, void* workspace, size_t workspace_bytes,
                        void* stream) {
  if (n <= 0 || nalpha < 0 || !x || !x2 || !y || !in_x2 || ld < n || (nalpha > 0 && !alphas))
    return fail(QEPW_ERR_ARG, "qepw_sign_iteration: bad argument");
  if (x == x2 || x == y || x2 == y)
    return fail(QEPW_ERR_ARG, "qepw_sign_iteration: x, x2 and y must be distinct");
  if (workspace_bytes < qepw::sign_iteration_workspace(n))
    return fail(QEPW_ERR_WORKSPACE, "qepw_sign_iteration: workspace too small");
  *in_x2 = 0;
  return finish(qepw::sign_iteration(x, x2, y, ld, n, alphas, nalpha, in_x2, workspace,
                                     workspace_bytes, static_cast<hipStream_t>(stream)),
                "qepw_sign_iteration");
}

int qepw_set_host_lapack(void* zpotrf, void* ztrtri, void* zheevd, void* cheevd) {
  if (!zpotrf || !ztrtri || !zheevd || !cheevd)
    return fail(QEPW_ERR_ARG, "qepw_set_host_lapack: NULL routine");
  qepw::set_host_lapack(zpotrf, ztrtri, zheevd, cheevd);
  return QEPW_OK;
}

size_t qepw_rr_workspace(int maxsub) { return maxsub > 0 ? qepw::rr_workspace_bytes(maxsub) : 0; }

int qepw_rr_standardize(void* h, void* s, void* w, void* a, int maxsub, int first, int active,
                        int* not_pd, void* workspace, size_t workspace_bytes, void* stream) {
  if (!h || !s || !w || !a || !not_pd || maxsub < 1 || first < 0 || first > active ||
      active > maxsub || (active > first && !workspace))
    return fail(QEPW_ERR_ARG, "qepw_rr_standardize: bad argument");
  if (!qepw::host_lapack_ready())
    return fail(QEPW_ERR_UNSUPPORTED,
                "qepw_rr_standardize: no host LAPACK registered (qepw_set_host_lapack)");
  const int rc = qepw::rr_standardize(h, s, w, a, maxsub, first, active, not_pd, workspace,
                                      workspace_bytes, static_cast<hipStream_t>(stream));
  if (rc != QEPW_OK) g_error = "qepw_rr_standardize: " + qepw::rr_message();
  return rc;
}

int qepw_rr_spectral(const qepw_rr_spectral_args* args, double* values,
                     qepw_rr_spectral_result* result) {
  if (!args || !values || !result) return fail(QEPW_ERR_ARG, "qepw_rr_spectral: NULL argument");
  const qepw_rr_spectral_args& g = *args;
  if (g.k < 1 || g.n < g.k || g.n > g.maxsub || g.count_hint < 1 || !g.A || !g.W || !g.X ||
      !g.X2 || !g.Y || !g.u_prev || !g.gauss || !g.coef || !g.rb || !g.work ||
      g.X == g.X2 || g.X == g.Y || g.X2 == g.Y || !(g.gap > 0.0))
    return fail(QEPW_ERR_ARG, "qepw_rr_spectral: bad argument");
  if (g.work_bytes < qepw::rr_workspace_bytes(g.maxsub))
    return fail(QEPW_ERR_WORKSPACE, "qepw_rr_spectral: workspace too small");
  if (!qepw::host_lapack_ready())
    return fail(QEPW_ERR_UNSUPPORTED,
                "qepw_rr_spectral: no host LAPACK registered (qepw_set_host_lapack)");
  const int rc = qepw::rr_spectral(args, values, result);
  if (rc != QEPW_OK) g_error = "qepw_rr_spectral: " + qepw::rr_message();
  return rc;
}

int qepw_cholesky_qr(int precision, const void* block, int64_t ldb, int64_t npw, int nb,
                     double condition, void* out, int64_t ldo, int* ok, void* workspace,
                     size_t workspace_bytes, void* stream) {
  if ((precision != QEPW_C128 && precision != QEPW_C64) || !ok || nb < 0 || npw < 1 ||
      ldb < npw || ldo < npw || (nb > 0 && (!block || !out || !workspace)) || block == out)
    return fail(QEPW_ERR_ARG, "qepw_cholesky_qr: bad argument");
  if (!qepw::host_lapack_ready())
    return fail(QEPW_ERR_UNSUPPORTED,
                "qepw_cholesky_qr: no host LAPACK registered (qepw_set_host_lapack)");
  const int rc = qepw::cholesky_qr(precision, block, ldb, npw, nb, condition, out, ldo, ok,
                                   workspace, workspace_bytes, static_cast<hipStream_t>(stream));
  if (rc != QEPW_OK) g_error = "qepw_cholesky_qr: " + qepw::rr_message();
  return rc;
}

int qepw_set_diagonal(void* a, int64_t lda, int n, const double* d, void* stream) {
  if (n < 0 || (n > 0 && (!a || lda < n)))
    return fail(QEPW_ERR_ARG, "qepw_set_diagonal: bad argument");
  return finish(qepw::set_diagonal(a, lda, n, d, static_cast<hipStream_t>(stream)),
                "qepw_set_diagonal");
}

}  // extern "C"
