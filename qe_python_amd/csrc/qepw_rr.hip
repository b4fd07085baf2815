// Host side of the device Rayleigh-Ritz in native code: the standardisation
// of the reduced pencil and the spectral solve of the standardised matrix,
// each one call that launches its kernels, reads the few small matrices it
// needs back through pinned memory, factorises them with the host LAPACK the
// caller registered (qepw_set_host_lapack) and uploads the factors -- the
// sequence the Python driver (qe_python_amd/eigensolver.py) runs step by
// step, without an interpreter round trip per launch.
//
// Reference: the reduced eigensolve of cegterg, zhegvd through
// generalized_eigh (qepy_pw/_native_fft.pyx:667-691, called from
// qepy_pw/diagonalization.py:699-726); the device form (spectral division)
// is described in qepw_eig.hip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <mutex>
#include <string>
#include <vector>

#include "qepw_internal.h"

namespace qepw {

namespace {

using zcplx = double2;
using ccplx = float2;

// Fortran LAPACK entry points (arguments by reference, character arguments
// of length one)
using potrf_fn = void (*)(char*, int*, zcplx*, int*, int*);
using trtri_fn = void (*)(char*, char*, int*, zcplx*, int*, int*);
using zheevd_fn = void (*)(char*, char*, int*, zcplx*, int*, double*, zcplx*, int*, double*,
                           int*, int*, int*, int*);
using cheevd_fn = void (*)(char*, char*, int*, ccplx*, int*, float*, ccplx*, int*, float*,
                           int*, int*, int*, int*);

struct HostLapack {
  potrf_fn zpotrf = nullptr;
  trtri_fn ztrtri = nullptr;
  zheevd_fn zheevd = nullptr;
  cheevd_fn cheevd = nullptr;
};
HostLapack g_lapack;

thread_local std::string t_message;

// Pinned host staging.  A device-to-host read waits for its stream, so a
// buffer read into is free again on return; an upload records an event, and
// its buffer is not written again before that event has completed.
struct Pinned {
  void* ptr = nullptr;
  size_t bytes = 0;
  bool in_flight = false;
  hipEvent_t done = nullptr;
  // the previous upload from this buffer has been copied
  hipError_t settle() {
    if (!in_flight) return hipSuccess;
    in_flight = false;
    return hipEventSynchronize(done);
  }
  hipError_t reserve(size_t need) {
    hipError_t e = settle();
    if (e != hipSuccess || need <= bytes) return e;
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
    const size_t grow = std::max(need, (size_t)1 << 20);
    e = hipHostMalloc(&ptr, grow, hipHostMallocDefault);
    if (e == hipSuccess) bytes = grow;
    return e;
  }
};

// One call's staging buffers, leased from a process-wide pool (calls from
// several host threads each hold their own set; the sets live until the
// process ends, so nothing is freed while the runtime shuts down).
struct Staging {
  Pinned read, up[2];
};
std::mutex g_staging_mutex;
std::vector<Staging*> g_staging_free;

struct StagingLease {
  Staging* s;
  StagingLease() {
    std::lock_guard<std::mutex> lock(g_staging_mutex);
    if (g_staging_free.empty()) {
      s = new Staging();
    } else {
      s = g_staging_free.back();
      g_staging_free.pop_back();
    }
  }
  ~StagingLease() {
    std::lock_guard<std::mutex> lock(g_staging_mutex);
    g_staging_free.push_back(s);
  }
};

struct Fail {
  int code;
};

void check_hip(hipError_t e, const char* where) {
  if (e != hipSuccess) {
    t_message = std::string(where) + ": " + hipGetErrorString(e);
    throw Fail{QEPW_ERR_HIP};
  }
}

void check_rc(int rc, const char* where) {
  if (rc != QEPW_OK) {
    t_message = std::string(where) + (rc == QEPW_ERR_WORKSPACE ? ": workspace too small"
                                      : rc == QEPW_ERR_ARG      ? ": invalid argument"
                                                                : ": launch failed");
    throw Fail{rc};
  }
}

// device scratch carved from the caller's workspace, 256-byte aligned
struct Carver {
  char* base;
  size_t left;
  void* take(size_t bytes, const char* what) {
    const size_t b = (bytes + 255) & ~(size_t)255;
    if (b > left) {
      t_message = std::string("rayleigh-ritz workspace too small for ") + what;
      throw Fail{QEPW_ERR_WORKSPACE};
    }
    void* p = base;
    base += b;
    left -= b;
    return p;
  }
};

struct Ctx {
  hipStream_t stream;
  void* gemm_ws;
  size_t gemm_ws_bytes;
  Staging* st;
};

void gemm_z(const Ctx& c, int op, int64_t m, int64_t n, int64_t k, double alpha,
            const zcplx* a, int64_t lda, const zcplx* b, int64_t ldb, double beta, zcplx* out,
            int64_t ldc) {
  const double al[2] = {alpha, 0.0}, be[2] = {beta, 0.0};
  check_rc(gemm(QEPW_C128, op, m, n, k, al, a, lda, b, ldb, be, out, ldc, c.gemm_ws,
                c.gemm_ws_bytes, c.stream),
           "rayleigh-ritz gemm");
}

// device (rows x cols, ld) complex128 -> host column-major (rows x cols)
// device bytes -> host, through the call's read buffer (waits for the stream)
void read_raw(const Ctx& c, const void* dev, int64_t ld_bytes, size_t row_bytes, int cols,
              void* host) {
  Pinned& p = c.st->read;
  const size_t bytes = row_bytes * (size_t)cols;
  check_hip(p.reserve(bytes), "hipHostMalloc");
  check_hip(hipMemcpy2DAsync(p.ptr, row_bytes, dev, ld_bytes, row_bytes, cols,
                             hipMemcpyDeviceToHost, c.stream),
            "rayleigh-ritz read");
  check_hip(hipStreamSynchronize(c.stream), "rayleigh-ritz wait");
  std::memcpy(host, p.ptr, bytes);
}

// device (rows x cols, ld) complex128 -> host column-major (rows x cols)
void read_block(const Ctx& c, const zcplx* dev, int64_t ld, int rows, int cols, zcplx* host) {
  read_raw(c, dev, ld * sizeof(zcplx), rows * sizeof(zcplx), cols, host);
}

void read_doubles(const Ctx& c, const double* dev, size_t count, double* host) {
  read_raw(c, dev, count * sizeof(double), count * sizeof(double), 1, host);
}

// host bytes -> device, through upload slot s
void upload_raw(const Ctx& c, int s, const void* host, size_t bytes, void* dev) {
  Pinned& p = c.st->up[s];
  check_hip(p.reserve(bytes), "hipHostMalloc");
  if (!p.done) check_hip(hipEventCreateWithFlags(&p.done, hipEventDisableTiming), "hipEventCreate");
  std::memcpy(p.ptr, host, bytes);
  check_hip(hipMemcpyAsync(dev, p.ptr, bytes, hipMemcpyHostToDevice, c.stream),
            "rayleigh-ritz upload");
  check_hip(hipEventRecord(p.done, c.stream), "hipEventRecord");
  p.in_flight = true;
}

// host column-major (rows x cols) -> device (ld = rows)
void upload_block(const Ctx& c, int s, const zcplx* host, int rows, int cols, zcplx* dev) {
  upload_raw(c, s, host, (size_t)rows * cols * sizeof(zcplx), dev);
}

// lower Cholesky factor L of the Hermitian part of column-major m (n x n,
// ld n) and L^-1 (lower, strict upper zero) in place; false if m is not
// positive definite
bool cholesky_inverse(std::vector<zcplx>& m, int n) {
  for (int j = 0; j < n; ++j)
    for (int i = j; i < n; ++i) {
      const zcplx a = m[i + (size_t)j * n], b = m[j + (size_t)i * n];
      // 0.5 (m + m^H), lower triangle
      m[i + (size_t)j * n] = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
    }
  char lo = 'L', nd = 'N';
  int nn = n, ld = n, info = 0;
  g_lapack.zpotrf(&lo, &nn, m.data(), &ld, &info);
  if (info != 0) return false;
  g_lapack.ztrtri(&lo, &nd, &nn, m.data(), &ld, &info);
  if (info != 0) return false;
  for (int j = 1; j < n; ++j)
    for (int i = 0; i < j; ++i) m[i + (size_t)j * n] = make_double2(0.0, 0.0);
  return true;
}

// all eigenpairs of the Hermitian part of column-major b (n x n), ascending;
// vectors overwrite b; in complex64 when single
bool hermitian_eigh(std::vector<zcplx>& b, int n, bool single, std::vector<double>& w) {
  for (int j = 0; j < n; ++j)
    for (int i = j; i < n; ++i) {
      const zcplx a = b[i + (size_t)j * n], c = b[j + (size_t)i * n];
      const zcplx h = make_double2(0.5 * (a.x + c.x), 0.5 * (a.y - c.y));
      b[i + (size_t)j * n] = h;
      b[j + (size_t)i * n] = make_double2(h.x, -h.y);
    }
  char jv = 'V', lo = 'L';
  int nn = n, ld = n, info = 0;
  int lwork = 2 * n + n * n, lrwork = 1 + 5 * n + 2 * n * n, liwork = 3 + 5 * n;
  std::vector<int> iwork(liwork);
  w.assign(n, 0.0);
  if (!single) {
    std::vector<zcplx> work(lwork);
    std::vector<double> rwork(lrwork);
    g_lapack.zheevd(&jv, &lo, &nn, b.data(), &ld, w.data(), work.data(), &lwork, rwork.data(),
                    &lrwork, iwork.data(), &liwork, &info);
    return info == 0;
  }
  std::vector<ccplx> a32((size_t)n * n), work(lwork);
  std::vector<float> w32(n), rwork(lrwork);
  for (size_t i = 0; i < a32.size(); ++i)
    a32[i] = make_float2((float)b[i].x, (float)b[i].y);
  g_lapack.cheevd(&jv, &lo, &nn, a32.data(), &ld, w32.data(), work.data(), &lwork, rwork.data(),
                  &lrwork, iwork.data(), &liwork, &info);
  if (info != 0) return false;
  for (int i = 0; i < n; ++i) w[i] = (double)w32[i];
  for (size_t i = 0; i < a32.size(); ++i) b[i] = make_double2(a32[i].x, a32[i].y);
  return true;
}

// Chen-Chow scaling factors of the Newton-Schulz sign iteration from a lower
// bound of |eigenvalues| of X0 (all |x| <= 1)
std::vector<double> scaled_ns_alphas(double lower, double tol = 1e-15, int limit = 60) {
  std::vector<double> alphas;
  double ell = std::max(std::min(lower, 1.0), 1e-16);
  while (1.0 - ell > tol && (int)alphas.size() < limit) {
    const double a = std::sqrt(3.0 / (1.0 + ell + ell * ell));
    alphas.push_back(a);
    ell = 0.5 * a * ell * (3.0 - a * a * ell * ell);
  }
  return alphas;
}

inline zcplx* blk(void* base, int64_t ld, int r0, int c0) {
  return static_cast<zcplx*>(base) + r0 + (int64_t)c0 * ld;
}

}  // namespace

void set_host_lapack(void* zpotrf, void* ztrtri, void* zheevd, void* cheevd) {
  g_lapack.zpotrf = reinterpret_cast<potrf_fn>(zpotrf);
  g_lapack.ztrtri = reinterpret_cast<trtri_fn>(ztrtri);
  g_lapack.zheevd = reinterpret_cast<zheevd_fn>(zheevd);
  g_lapack.cheevd = reinterpret_cast<cheevd_fn>(cheevd);
}

bool host_lapack_ready() {
  return g_lapack.zpotrf && g_lapack.ztrtri && g_lapack.zheevd && g_lapack.cheevd;
}

const std::string& rr_message() { return t_message; }

size_t rr_workspace_bytes(int maxsub) {
  // Z, M, T, R, the uploaded factor; Omega, Q0, Q, AQ, B, the uploaded
  // rotations, U; GEMM split partials
  const size_t sq = (size_t)maxsub * maxsub * sizeof(zcplx) + 256;
  return 12 * sq + ((size_t)32 << 20);
}

// Standardisation of rows first:active (see DeviceRayleighRitz._standardize):
// W = L^-1 grows by the new block, A = W H W^H gains its new rows.
int rr_standardize(void* H, void* S, void* W, void* A, int maxsub, int first, int active,
                   int* not_pd, void* work, size_t work_bytes, hipStream_t stream) {
  *not_pd = 0;
  try {
    const int64_t ld = maxsub;
    const int nnew = active - first;
    if (nnew <= 0) return QEPW_OK;
    Carver cv{static_cast<char*>(work), work_bytes};
    StagingLease lease;
    Ctx c{stream, nullptr, 0, lease.s};
    const size_t z = sizeof(zcplx);
    zcplx* Z = static_cast<zcplx*>(cv.take((size_t)nnew * std::max(1, first) * z, "Z"));
    zcplx* M = static_cast<zcplx*>(cv.take((size_t)nnew * nnew * z, "M"));
    zcplx* T = static_cast<zcplx*>(cv.take((size_t)std::max(1, first) * nnew * z, "T"));
    zcplx* R = static_cast<zcplx*>(cv.take((size_t)active * nnew * z, "R"));
    zcplx* Wd = static_cast<zcplx*>(cv.take((size_t)nnew * nnew * z, "W22"));
    c.gemm_ws_bytes = cv.left > 256 ? cv.left - 256 : 0;
    c.gemm_ws = cv.take(c.gemm_ws_bytes, "gemm");
    std::vector<zcplx> m((size_t)nnew * nnew);
    if (first == 0) {
      read_block(c, blk(S, ld, 0, 0), ld, active, active, m.data());
    } else {
      // Z = W11 S12 (first x nnew); M = S22 - Z^H Z
      gemm_z(c, QEPW_OP_N, first, nnew, first, 1.0, blk(W, ld, 0, 0), ld,
             blk(S, ld, 0, first), ld, 0.0, Z, first);
      check_rc(copy_columns(QEPW_C128, QEPW_C128, blk(S, ld, first, first), ld, nullptr, M, nnew,
                            nullptr, nullptr, nnew, nnew, stream),
               "rayleigh-ritz copy");
      gemm_z(c, QEPW_OP_C, nnew, nnew, first, -1.0, Z, first, Z, first, 1.0, M, nnew);
      read_block(c, M, nnew, nnew, nnew, m.data());
    }
    if (!cholesky_inverse(m, nnew)) {
      *not_pd = 1;
      return QEPW_OK;
    }
    upload_block(c, 0, m.data(), nnew, nnew, Wd);
    check_rc(copy_columns(QEPW_C128, QEPW_C128, Wd, nnew, nullptr, blk(W, ld, first, first), ld,
                          nullptr, nullptr, nnew, nnew, stream),
             "rayleigh-ritz copy");
    if (first > 0) {
      // the strictly upper part of the new columns stays zero
      check_rc(set_block(QEPW_C128, blk(W, ld, 0, first), ld, first, nnew, 0.0, nullptr, stream),
               "rayleigh-ritz zero");
      // W21 = -W22 (Z^H W11)
      gemm_z(c, QEPW_OP_C, nnew, first, first, 1.0, Z, first, blk(W, ld, 0, 0), ld, 0.0, T,
             nnew);
      gemm_z(c, QEPW_OP_N, nnew, first, nnew, -1.0, Wd, nnew, T, nnew, 0.0,
             blk(W, ld, first, 0), ld);
    }
    // new rows of A: A[first:, :] = W[first:, :] H W^H
    gemm_z(c, QEPW_OP_N, nnew, active, active, 1.0, blk(W, ld, first, 0), ld, blk(H, ld, 0, 0),
           ld, 0.0, R, nnew);
    gemm_z(c, QEPW_OP_N | QEPW_OP_BC, nnew, active, active, 1.0, R, nnew, blk(W, ld, 0, 0), ld,
           0.0, blk(A, ld, first, 0), ld);
    check_rc(hermitian_complete(A, ld, first, active, stream), "rayleigh-ritz hermitian");
    return QEPW_OK;
  } catch (const Fail& f) {
    return f.code;
  } catch (const std::exception& e) {
    // host allocations (staging, LAPACK work arrays)
    t_message = std::string("host allocation: ") + e.what();
    return QEPW_ERR_HIP;
  }
}

// Spectral solve (see DeviceRayleighRitz._spectral), given the shift and the
// gap around it: sign iteration, range sample of the projector below the
// shift, Rayleigh-Ritz on it.  status: 0 solved, 1 sign iteration not
// converged, 2 fewer than k eigenvalues below the shift, 3 sample Gram
// matrix not positive definite, 4 host eigensolve failed.
int rr_spectral(const qepw_rr_spectral_args* a, double* values, qepw_rr_spectral_result* out) {
  try {
    const int n = a->n, k = a->k;
    const int64_t ld = a->maxsub;
    hipStream_t stream = static_cast<hipStream_t>(a->stream);
    Carver cv{static_cast<char*>(a->work), a->work_bytes};
    StagingLease lease;
    Ctx c{stream, nullptr, 0, lease.s};
    const size_t sq = (size_t)a->maxsub * a->maxsub * sizeof(zcplx);
    zcplx* Om = static_cast<zcplx*>(cv.take(sq, "Omega"));
    zcplx* Q0 = static_cast<zcplx*>(cv.take(sq, "Q0"));
    zcplx* Q = static_cast<zcplx*>(cv.take(sq, "Q"));
    zcplx* AQ = static_cast<zcplx*>(cv.take(sq, "AQ"));
    zcplx* B = static_cast<zcplx*>(cv.take(sq, "B"));
    zcplx* Tt = static_cast<zcplx*>(cv.take(sq, "T"));
    zcplx* Yd = static_cast<zcplx*>(cv.take(sq, "Y"));
    zcplx* U = static_cast<zcplx*>(cv.take(sq, "U"));
    c.gemm_ws_bytes = cv.left > 256 ? cv.left - 256 : 0;
    c.gemm_ws = cv.take(c.gemm_ws_bytes, "gemm");
    double* rb = a->rb;
    const int head = 2 * a->maxsub + 4;
    double* sig = rb;
    double* bound = rb + a->maxsub + 2;
    out->status = 0;
    out->ns_steps = 0;
    out->slow_shift = 0;
    out->m_below = 0;

    zcplx* X = static_cast<zcplx*>(a->X);
    zcplx* X2 = static_cast<zcplx*>(a->X2);
    zcplx* Y = static_cast<zcplx*>(a->Y);
    check_rc(shift_scale(a->A, ld, n, a->sigma, X, ld, bound, stream), "qepw_shift_scale");
    // the scaling schedule needs ||A - sigma I||_1 only as an estimate:
    // twice the previous solve's value, or the exact one read back
    double bound_est = 2.0 * a->bound_prev;
    if (!(a->bound_prev > 0.0)) read_doubles(c, bound, 1, &bound_est);
    out->bound = bound_est;
    std::vector<double> alphas = scaled_ns_alphas(0.25 * a->gap / bound_est);
    alphas.push_back(1.0);
    alphas.push_back(1.0);
    auto sign_steps = [&](const std::vector<double>& sched) {
      int in_x2 = 0;
      check_rc(sign_iteration(X, X2, Y, ld, n, sched.data(), (int)sched.size(), &in_x2, nullptr,
                              0, stream),
               "qepw_sign_iteration");
      if (in_x2) std::swap(X, X2);
      out->ns_steps += (int)sched.size();
      check_rc(sign_check(X, ld, Y, ld, n, sig, stream), "qepw_sign_check");
    };
    // Q0 = (I - X) Omega / 2 for Omega = [previous eigenvectors, Gaussian
    // columns] and its Gram matrix after the sign check in rb
    auto sampled_range = [&](int count) {
      int start = 0;
      if (a->have_prev) {
        start = std::min(k, count);
        check_rc(copy_columns(QEPW_C128, QEPW_C128, a->u_prev, ld, nullptr, Om, ld, nullptr,
                              nullptr, n, start, stream),
                 "rayleigh-ritz copy");
      }
      if (start < count)
        check_rc(copy_columns(QEPW_C128, QEPW_C128, blk(a->gauss, ld, 0, start), ld, nullptr,
                              blk(Om, ld, 0, start), ld, nullptr, nullptr, n, count - start,
                              stream),
                 "rayleigh-ritz copy");
      check_rc(copy_columns(QEPW_C128, QEPW_C128, Om, ld, nullptr, Q0, n, nullptr, nullptr, n,
                            count, stream),
               "rayleigh-ritz copy");
      gemm_z(c, QEPW_OP_N, n, count, n, -0.5, X, ld, Om, ld, 0.5, Q0, n);
      gemm_z(c, QEPW_OP_C, count, count, n, 1.0, Q0, n, Q0, n, 0.0,
             reinterpret_cast<zcplx*>(rb + head), count);
    };
    std::vector<double> host;
    auto unpack = [&](int count) {
      host.resize((size_t)head + 2 * (size_t)count * count);
      read_doubles(c, rb, host.size(), host.data());
    };

    sign_steps(alphas);
    int count = std::min(n, a->count_hint);
    sampled_range(count);
    unpack(count);
    double res2 = host[0], trace = host[1];
    out->bound = host[a->maxsub + 2];
    bool have_g = true;
    out->slow_shift = res2 >= 1e-20 * n ? 1 : 0;
    for (int attempt = 0; res2 >= 1e-20 * n && attempt < 2; ++attempt) {
      sign_steps(std::vector<double>(12, 1.0));
      double s2[2];
      read_doubles(c, sig, 2, s2);
      res2 = s2[0];
      trace = s2[1];
      have_g = false;
    }
    if (res2 >= 1e-20 * n) {
      out->status = 1;
      return QEPW_OK;
    }
    const int m_below = (int)std::lround(0.5 * (n - trace));
    out->m_below = m_below;
    if (m_below < k) {
      out->status = 2;
      return QEPW_OK;
    }
    if (!have_g || m_below > count) {
      count = std::max(count, m_below);
      sampled_range(count);
      unpack(count);
    }
    const int mp = m_below;
    // the leading mp x mp block of the Gram matrix (column-major, ld count)
    const zcplx* g = reinterpret_cast<const zcplx*>(host.data() + head);
    std::vector<zcplx> lg((size_t)mp * mp);
    for (int j = 0; j < mp; ++j)
      for (int i = 0; i < mp; ++i) lg[i + (size_t)j * mp] = g[i + (size_t)j * count];
    if (!cholesky_inverse(lg, mp)) {
      out->status = 3;
      return QEPW_OK;
    }
    // t = L^-H (upper), uploaded column-major
    std::vector<zcplx> t((size_t)mp * mp);
    for (int j = 0; j < mp; ++j)
      for (int i = 0; i < mp; ++i) {
        const zcplx v = lg[j + (size_t)i * mp];
        t[i + (size_t)j * mp] = make_double2(v.x, -v.y);
      }
    upload_block(c, 0, t.data(), mp, mp, Tt);
    gemm_z(c, QEPW_OP_N, n, mp, mp, 1.0, Q0, n, Tt, mp, 0.0, Q, n);
    gemm_z(c, QEPW_OP_N, n, mp, n, 1.0, static_cast<const zcplx*>(a->A), ld, Q, n, 0.0, AQ, n);
    gemm_z(c, QEPW_OP_C, mp, mp, n, 1.0, Q, n, AQ, n, 0.0, B, mp);
    std::vector<zcplx> bt((size_t)mp * mp);
    read_block(c, B, mp, mp, mp, bt.data());
    std::vector<double> w;
    if (!hermitian_eigh(bt, mp, a->host_single != 0, w) || mp < k) {
      out->status = 4;
      return QEPW_OK;
    }
    // U = Q y[:, :k]; coef = W^H U; the previous eigenvectors become U
    upload_block(c, 1, bt.data(), mp, k, Yd);
    gemm_z(c, QEPW_OP_N, n, k, mp, 1.0, Q, n, Yd, mp, 0.0, U, n);
    gemm_z(c, QEPW_OP_C, n, k, n, 1.0, static_cast<const zcplx*>(a->W), ld, U, n, 0.0,
           static_cast<zcplx*>(a->coef), n);
    check_rc(set_block(QEPW_C128, a->u_prev, ld, ld, k, 0.0, nullptr, stream),
             "rayleigh-ritz zero");
    check_rc(copy_columns(QEPW_C128, QEPW_C128, U, n, nullptr, a->u_prev, ld, nullptr, nullptr, n,
                          k, stream),
             "rayleigh-ritz copy");
    for (int i = 0; i < mp; ++i) values[i] = w[i];
    return QEPW_OK;
  } catch (const Fail& f) {
    return f.code;
  } catch (const std::exception& e) {
    // host allocations (staging, LAPACK work arrays)
    t_message = std::string("host allocation: ") + e.what();
    return QEPW_ERR_HIP;
  }
}

// Cholesky QR of an (npw x nb) block of the given precision (see
// _cholesky_orthonormalize): Gram = B^H B on the device, L L^H = Gram in FP64
// on the host, out = B L^-H.  *ok = 0 (out not written) when the Gram
// matrix is not positive definite or its squared condition estimate
// (max/min of diag L)^2 exceeds ``condition``.
int cholesky_qr(int precision, const void* block, int64_t ldb, int64_t npw, int nb,
                double condition, void* out, int64_t ldo, int* ok, void* work,
                size_t work_bytes, hipStream_t stream) {
  *ok = 0;
  try {
    if (nb <= 0) {
      *ok = 1;
      return QEPW_OK;
    }
    const bool single = precision == QEPW_C64;
    const size_t esz = single ? sizeof(ccplx) : sizeof(zcplx);
    Carver cv{static_cast<char*>(work), work_bytes};
    StagingLease lease;
    Ctx c{stream, nullptr, 0, lease.s};
    void* gram = cv.take((size_t)nb * nb * esz, "Gram");
    void* tdev = cv.take((size_t)nb * nb * esz, "L^-H");
    c.gemm_ws_bytes = cv.left > 256 ? cv.left - 256 : 0;
    c.gemm_ws = cv.take(c.gemm_ws_bytes, "gemm");
    const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
    check_rc(gemm(precision, QEPW_OP_C, nb, nb, npw, one, block, ldb, block, ldb, zero, gram, nb,
                  c.gemm_ws, c.gemm_ws_bytes, stream),
             "cholesky qr gemm");
    std::vector<zcplx> m((size_t)nb * nb);
    if (single) {
      std::vector<ccplx> g32((size_t)nb * nb);
      read_raw(c, gram, (int64_t)g32.size() * sizeof(ccplx), g32.size() * sizeof(ccplx), 1,
               g32.data());
      for (size_t i = 0; i < m.size(); ++i) m[i] = make_double2(g32[i].x, g32[i].y);
    } else {
      read_block(c, static_cast<zcplx*>(gram), nb, nb, nb, m.data());
    }
    // Hermitian part, lower triangle, and its Cholesky factor
    for (int j = 0; j < nb; ++j)
      for (int i = j; i < nb; ++i) {
        const zcplx a = m[i + (size_t)j * nb], b = m[j + (size_t)i * nb];
        m[i + (size_t)j * nb] = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
      }
    char lo = 'L', nd = 'N';
    int nn = nb, ld = nb, info = 0;
    g_lapack.zpotrf(&lo, &nn, m.data(), &ld, &info);
    if (info != 0) return QEPW_OK;
    double dmin = INFINITY, dmax = 0.0;
    for (int i = 0; i < nb; ++i) {
      const zcplx d = m[i + (size_t)i * nb];
      const double a = std::hypot(d.x, d.y);
      dmin = std::min(dmin, a);
      dmax = std::max(dmax, a);
    }
    if (!(dmin > 0.0) || (dmax / dmin) * (dmax / dmin) > condition) return QEPW_OK;
    g_lapack.ztrtri(&lo, &nd, &nn, m.data(), &ld, &info);
    if (info != 0) return QEPW_OK;
    // t = L^-H (upper triangular), column-major, in the block's precision
    std::vector<zcplx> t((size_t)nb * nb, make_double2(0.0, 0.0));
    for (int j = 0; j < nb; ++j)
      for (int i = 0; i <= j; ++i) {
        const zcplx v = m[j + (size_t)i * nb];
        t[i + (size_t)j * nb] = make_double2(v.x, -v.y);
      }
    if (single) {
      std::vector<ccplx> t32(t.size());
      for (size_t i = 0; i < t.size(); ++i) t32[i] = make_float2((float)t[i].x, (float)t[i].y);
      upload_raw(c, 0, t32.data(), t32.size() * sizeof(ccplx), tdev);
    } else {
      upload_block(c, 0, t.data(), nb, nb, static_cast<zcplx*>(tdev));
    }
    check_rc(gemm(precision, QEPW_OP_N, npw, nb, nb, one, block, ldb, tdev, nb, zero, out, ldo,
                  c.gemm_ws, c.gemm_ws_bytes, stream),
             "cholesky qr gemm");
    *ok = 1;
    return QEPW_OK;
  } catch (const Fail& f) {
    return f.code;
  } catch (const std::exception& e) {
    // host allocations (staging, LAPACK work arrays)
    t_message = std::string("host allocation: ") + e.what();
    return QEPW_ERR_HIP;
  }
}

}  // namespace qepw
