// Bandwidth-bound Davidson kernels over (npw, nb) column-major blocks.
//
// Each reproduces one elementwise / reduction kernel of the reference's
// native extension (qepy_pw/_native_fft.pyx); the formula is cited at each
// kernel.  Reductions accumulate in FP64 for both storage precisions and use
// a fixed per-column tree, so they are deterministic.
#include "qepw_internal.h"

namespace qepw {

constexpr int kRedThreads = 256;

template <typename T>
__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int s = kRedThreads / 2; s > 0; s >>= 1) {
    if (tid < s) sh[tid] += sh[tid + s];
    __syncthreads();
  }
  return sh[0];
}

// Fixed-order tree sum over a THREADS-wide workgroup.
template <int THREADS>
__device__ __forceinline__ double tree_sum(double v, double* sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int s = THREADS / 2; s > 0; s >>= 1) {
    if (tid < s) sh[tid] += sh[tid + s];
    __syncthreads();
  }
  return sh[0];
}

// Blocks of at most this many columns reduce each column with a 1024-wide
// workgroup instead of 256 (a single column of npw ~ 24k is otherwise a
// ~30 us serial loop: the band-by-band CG's inner products).
constexpr int kWideColumns = 64;

// column_squared_norms (_native_fft.pyx:1153-1191): out[b] = sum_g |v_gb|^2
template <typename T, int THREADS>
__global__ __launch_bounds__(THREADS) void k_col_sqnorm(const cplx<T>* __restrict__ v,
                                                        int64_t ldv, int64_t nrows,
                                                        double* __restrict__ out) {
  __shared__ double sh[THREADS];
  const int64_t b = blockIdx.x;
  const cplx<T>* col = v + b * ldv;
  double acc = 0.0;
  for (int64_t g = threadIdx.x; g < nrows; g += THREADS) {
    const cplx<T> z = col[g];
    acc += (double)z.x * (double)z.x + (double)z.y * (double)z.y;
  }
  const double s = tree_sum<THREADS>(acc, sh);
  if (threadIdx.x == 0) out[b] = s;
}

// column_inner_products (_native_fft.pyx:1241-1291): out[b] = <l_b | r_b>
template <typename T, int THREADS>
__global__ __launch_bounds__(THREADS) void k_col_inner(const cplx<T>* __restrict__ l,
                                                       int64_t ldl,
                                                       const cplx<T>* __restrict__ r,
                                                       int64_t ldr, int64_t nrows,
                                                       double* __restrict__ out) {
  __shared__ double sh[THREADS];
  const int64_t b = blockIdx.x;
  double ar = 0.0, ai = 0.0;
  for (int64_t g = threadIdx.x; g < nrows; g += THREADS) {
    const cplx<T> x = l[b * ldl + g];
    const cplx<T> y = r[b * ldr + g];
    ar += (double)x.x * y.x + (double)x.y * y.y;
    ai += (double)x.x * y.y - (double)x.y * y.x;
  }
  const double sr = tree_sum<THREADS>(ar, sh);
  __syncthreads();  // every thread has read sh[0] before it is reused
  const double si = tree_sum<THREADS>(ai, sh);
  if (threadIdx.x == 0) { out[2 * b] = sr; out[2 * b + 1] = si; }
}

// subtract_band_energies (_native_fft.pyx:1194-1238): r[:,b] -= e[b] x[:,b]
template <typename T>
__global__ void k_sub_energies(cplx<T>* __restrict__ r, int64_t ldr,
                               const cplx<T>* __restrict__ x, int64_t ldx,
                               const double* __restrict__ e, int64_t nrows) {
  const int64_t b = blockIdx.y;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nrows) return;
  const T eb = (T)e[b];
  const cplx<T> xv = x[b * ldx + g];
  cplx<T> rv = r[b * ldr + g];
  rv.x -= eb * xv.x;
  rv.y -= eb * xv.y;
  r[b * ldr + g] = rv;
}

// qe_precondition (_native_fft.pyx:852-899), QE g_psi in Rydberg units.
template <typename T>
__global__ void k_precondition(const cplx<T>* __restrict__ r, int64_t ldr,
                               const int* __restrict__ cols, const double* __restrict__ e,
                               const double* __restrict__ diag, cplx<T>* __restrict__ out,
                               int64_t ldo, int64_t nrows) {
  const int64_t j = blockIdx.y;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nrows) return;
  const int64_t c = cols ? cols[j] : j;
  const double x = 2.0 * (diag[g] - e[j]);
  const double d = 0.5 * (1.0 + x + sqrt(1.0 + (x - 1.0) * (x - 1.0)));
  const double f = 2.0 / d;
  const cplx<T> v = r[c * ldr + g];
  out[j * ldo + g] = mk<T>((T)(f * (double)v.x), (T)(f * (double)v.y));
}

// normalize_selected_columns (_native_fft.pyx:1495-1530)
template <typename T>
__global__ void k_normalize_sel(const cplx<T>* __restrict__ src, int64_t lds,
                                const double* __restrict__ sq, const int* __restrict__ sel,
                                cplx<T>* __restrict__ dst, int64_t ldd, int64_t nrows) {
  const int64_t j = blockIdx.y;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nrows) return;
  const int64_t c = sel[j];
  const double inv = 1.0 / sqrt(sq[c]);
  const cplx<T> v = src[c * lds + g];
  dst[j * ldd + g] = mk<T>((T)((double)v.x * inv), (T)((double)v.y * inv));
}

// davidson restart (diagonalization.py:843-868): hv = x e + r
template <typename T>
__global__ void k_restart(const cplx<T>* __restrict__ x, int64_t ldx,
                          const cplx<T>* __restrict__ r, int64_t ldr,
                          const double* __restrict__ e, cplx<T>* __restrict__ hv, int64_t ldh,
                          int64_t nrows) {
  const int64_t b = blockIdx.y;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nrows) return;
  const T eb = (T)e[b];
  const cplx<T> xv = x[b * ldx + g];
  const cplx<T> rv = r[b * ldr + g];
  hv[b * ldh + g] = mk<T>(xv.x * eb + rv.x, xv.y * eb + rv.y);
}

template <typename TS, typename TD>
__global__ void k_permute(const int* __restrict__ perm, int direction,
                          const cplx<TS>* __restrict__ src, int64_t lds,
                          cplx<TD>* __restrict__ dst, int64_t ldd, int64_t nrows) {
  const int64_t b = blockIdx.y;
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= nrows) return;
  const int64_t p = perm ? perm[d] : d;
  if (direction == 0) {
    const cplx<TS> v = src[b * lds + p];
    dst[b * ldd + d] = mk<TD>((TD)v.x, (TD)v.y);
  } else {
    const cplx<TS> v = src[b * lds + d];
    dst[b * ldd + p] = mk<TD>((TD)v.x, (TD)v.y);
  }
}

// accumulate_projector_diagonal (_native_fft.pyx:734-779)
__global__ void k_row_real_dot(const double2* __restrict__ a, int64_t lda,
                               const double2* __restrict__ b, int64_t ldb, int64_t nrows,
                               int ncols, double scale, double* __restrict__ acc) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nrows) return;
  double s = 0.0;
  for (int i = 0; i < ncols; ++i) {
    const double2 x = a[i * lda + g];
    const double2 y = b[i * ldb + g];
    s += x.x * y.x + x.y * y.y;
  }
  acc[g] += scale * s;
}

// dst[g, b] = s[g] c[b] src[g, b] (either factor may be absent): the
// (k+G)_c psi factor of the projector derivative (pw/scf.py:1580-1590) and
// the occupation-weighted H psi of the band-energy gradient
template <typename T>
__global__ void k_scale_rows(const cplx<T>* __restrict__ src, int64_t lds,
                             const double* __restrict__ s, int64_t sstride,
                             const double* __restrict__ cfac,
                             cplx<T>* __restrict__ dst, int64_t ldd, int64_t nrows) {
  const int64_t b = blockIdx.y;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nrows) return;
  const T f = (T)((s ? s[g * sstride] : 1.0) * (cfac ? cfac[b] : 1.0));
  const cplx<T> v = src[b * lds + g];
  dst[b * ldd + g] = mk<T>(f * v.x, f * v.y);
}

// B[g, a*nch + i] = beta[g, i] * exp(-i q_g . tau_a): the atom-packed
// projector of pw/scf.py:657-686 recomputed from positions on the device.
template <typename T>
__global__ void k_build_projectors(const cplx<T>* __restrict__ beta, int64_t ldb, int nch,
                                   const double* __restrict__ q, const double* __restrict__ tau,
                                   int natom, cplx<T>* __restrict__ out, int64_t ldo,
                                   int64_t nrows) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int a = blockIdx.y;
  if (g >= nrows || a >= natom) return;
  const double arg = -(q[3 * g] * tau[3 * a] + q[3 * g + 1] * tau[3 * a + 1] +
                       q[3 * g + 2] * tau[3 * a + 2]);
  double sn, cs;
  sincos(arg, &sn, &cs);
  for (int i = 0; i < nch; ++i) {
    const cplx<T> bv = beta[(int64_t)i * ldb + g];
    const double re = (double)bv.x * cs - (double)bv.y * sn;
    const double im = (double)bv.x * sn + (double)bv.y * cs;
    out[(int64_t)(a * nch + i) * ldo + g] = mk<T>((T)re, (T)im);
  }
}

// out[d, a*nch + i] = beta[r, i] phases[r, a] with r = perm[d] (or d): the
// atom blocks of a factorized projector term (reference
// qepy_pw/diagonalization.py:21-28, FactorizedProjectorTerm) packed on the
// device, optionally gathered into a plan's device row order.
template <typename T>
__global__ void k_phase_blocks(const cplx<T>* __restrict__ beta, int64_t ldb, int nch,
                               const cplx<T>* __restrict__ phases, int64_t ldp, int natom,
                               const int* __restrict__ perm, cplx<T>* __restrict__ out,
                               int64_t ldo, int64_t nrows) {
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int a = blockIdx.y;
  if (d >= nrows || a >= natom) return;
  const int64_t r = perm ? perm[d] : d;
  const cplx<T> ph = phases[(int64_t)a * ldp + r];
  for (int i = 0; i < nch; ++i) {
    const cplx<T> bv = beta[(int64_t)i * ldb + r];
    const double re = (double)bv.x * ph.x - (double)bv.y * ph.y;
    const double im = (double)bv.x * ph.y + (double)bv.y * ph.x;
    out[(int64_t)(a * nch + i) * ldo + d] = mk<T>((T)re, (T)im);
  }
}

// out[a*3 + c] = 2 sum_n f_n sum_{i in atom a} Re(conj(dO_c[i, n]) DO[i, n]),
// the force-branch contraction of pw/scf.py:1591-1599 (dE/dtau = -F).
// dO: three (P, nb) blocks with column stride lddo and block stride
// dstride; DO = D O; atom a owns channels [off[a], off[a+1]).
template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_projector_forces(
    const cplx<T>* __restrict__ dO, int64_t lddo, int64_t dstride,
    const cplx<T>* __restrict__ DO, int64_t lddO, const double* __restrict__ f, int nb,
    const int* __restrict__ off, double* __restrict__ out) {
  __shared__ double sh[kRedThreads];
  const int a = blockIdx.x, c = blockIdx.y;
  const int i0 = off[a], i1 = off[a + 1];
  const int nch = i1 - i0;
  double acc = 0.0;
  for (int e = threadIdx.x; e < nch * nb; e += kRedThreads) {
    const int n = e / nch, i = i0 + (e - n * nch);
    const cplx<T> d = dO[c * dstride + (int64_t)n * lddo + i];
    const cplx<T> o = DO[(int64_t)n * lddO + i];
    acc += f[n] * ((double)d.x * o.x + (double)d.y * o.y);
  }
  const double s = block_sum<T>(acc, sh);
  if (threadIdx.x == 0) out[a * 3 + c] = 2.0 * s;
}

// dst[:, dcol(j)] = s_j src[:, scol(j)] for j < ncols (column maps and scale
// optional; the precisions of src and dst may differ): the column gathers,
// scatters, widenings and per-column scalings of the Davidson bookkeeping.
template <typename TS, typename TD>
__global__ void k_copy_columns(const cplx<TS>* __restrict__ src, int64_t lds,
                               const int* __restrict__ scol, cplx<TD>* __restrict__ dst,
                               int64_t ldd, const int* __restrict__ dcol,
                               const double* __restrict__ scale, int64_t nrows) {
  const int j = blockIdx.y;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nrows) return;
  const int64_t sc = scol ? scol[j] : j;
  const int64_t dc = dcol ? dcol[j] : j;
  const cplx<TS> v = src[sc * lds + g];
  const double s = scale ? scale[j] : 1.0;
  dst[dc * ldd + g] = mk<TD>((TD)((double)v.x * s), (TD)((double)v.y * s));
}

// dst[i, j] = (i == j) ? (d ? d[i] : diag) : 0 over an nrows x ncols block.
template <typename T>
__global__ void k_set_block(cplx<T>* __restrict__ dst, int64_t ldd, int64_t nrows, double diag,
                            const double* __restrict__ d) {
  const int64_t j = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  const double v = (i == j) ? (d ? d[i] : diag) : 0.0;
  dst[j * ldd + i] = mk<T>((T)v, T(0));
}

static inline dim3 rows_grid(int64_t nrows, int nb) {
  return dim3((unsigned)((nrows + 255) / 256), (unsigned)nb);
}

static inline int status() { return hipGetLastError() == hipSuccess ? QEPW_OK : QEPW_ERR_HIP; }

#define QEPW_DISPATCH(prec, CALL_D, CALL_F)          \
  do {                                               \
    if ((prec) == QEPW_C128) { CALL_D; }             \
    else if ((prec) == QEPW_C64) { CALL_F; }         \
    else return QEPW_ERR_ARG;                        \
  } while (0)

int column_squared_norms(int precision, const void* v, int64_t ldv, int64_t nrows, int nb,
                         double* out, hipStream_t stream) {
  if (nb <= 0) return QEPW_OK;
  QEPW_DISPATCH(precision,
      if (nb <= kWideColumns)
        hipLaunchKernelGGL((k_col_sqnorm<double, 1024>), dim3(nb), dim3(1024), 0, stream,
                           static_cast<const double2*>(v), ldv, nrows, out);
      else
        hipLaunchKernelGGL((k_col_sqnorm<double, kRedThreads>), dim3(nb), dim3(kRedThreads), 0,
                           stream, static_cast<const double2*>(v), ldv, nrows, out),
      if (nb <= kWideColumns)
        hipLaunchKernelGGL((k_col_sqnorm<float, 1024>), dim3(nb), dim3(1024), 0, stream,
                           static_cast<const float2*>(v), ldv, nrows, out);
      else
        hipLaunchKernelGGL((k_col_sqnorm<float, kRedThreads>), dim3(nb), dim3(kRedThreads), 0,
                           stream, static_cast<const float2*>(v), ldv, nrows, out));
  return status();
}

int column_inner_products(int precision, const void* l, int64_t ldl, const void* r,
                          int64_t ldr, int64_t nrows, int nb, double* out,
                          hipStream_t stream) {
  if (nb <= 0) return QEPW_OK;
  QEPW_DISPATCH(precision,
      if (nb <= kWideColumns)
        hipLaunchKernelGGL((k_col_inner<double, 1024>), dim3(nb), dim3(1024), 0, stream,
                           static_cast<const double2*>(l), ldl,
                           static_cast<const double2*>(r), ldr, nrows, out);
      else
        hipLaunchKernelGGL((k_col_inner<double, kRedThreads>), dim3(nb), dim3(kRedThreads), 0,
                           stream, static_cast<const double2*>(l), ldl,
                           static_cast<const double2*>(r), ldr, nrows, out),
      if (nb <= kWideColumns)
        hipLaunchKernelGGL((k_col_inner<float, 1024>), dim3(nb), dim3(1024), 0, stream,
                           static_cast<const float2*>(l), ldl,
                           static_cast<const float2*>(r), ldr, nrows, out);
      else
        hipLaunchKernelGGL((k_col_inner<float, kRedThreads>), dim3(nb), dim3(kRedThreads), 0,
                           stream, static_cast<const float2*>(l), ldl,
                           static_cast<const float2*>(r), ldr, nrows, out));
  return status();
}

int subtract_band_energies(int precision, void* r, int64_t ldr, const void* x, int64_t ldx,
                           const double* e, int64_t nrows, int nb, hipStream_t stream) {
  if (nb <= 0 || nrows <= 0) return QEPW_OK;
  QEPW_DISPATCH(precision,
      hipLaunchKernelGGL(k_sub_energies<double>, rows_grid(nrows, nb), dim3(256), 0, stream,
                         static_cast<double2*>(r), ldr, static_cast<const double2*>(x), ldx,
                         e, nrows),
      hipLaunchKernelGGL(k_sub_energies<float>, rows_grid(nrows, nb), dim3(256), 0, stream,
                         static_cast<float2*>(r), ldr, static_cast<const float2*>(x), ldx,
                         e, nrows));
  return status();
}

int qe_precondition(int precision, const void* r, int64_t ldr, const int* cols,
                    const double* e, const double* diag, void* out, int64_t ldo,
                    int64_t nrows, int nb, hipStream_t stream) {
  if (nb <= 0 || nrows <= 0) return QEPW_OK;
  QEPW_DISPATCH(precision,
      hipLaunchKernelGGL(k_precondition<double>, rows_grid(nrows, nb), dim3(256), 0, stream,
                         static_cast<const double2*>(r), ldr, cols, e, diag,
                         static_cast<double2*>(out), ldo, nrows),
      hipLaunchKernelGGL(k_precondition<float>, rows_grid(nrows, nb), dim3(256), 0, stream,
                         static_cast<const float2*>(r), ldr, cols, e, diag,
                         static_cast<float2*>(out), ldo, nrows));
  return status();
}

int normalize_selected_columns(int precision, const void* src, int64_t lds,
                               const double* sq, const int* sel, void* dst, int64_t ldd,
                               int64_t nrows, int nb, hipStream_t stream) {
  if (nb <= 0 || nrows <= 0) return QEPW_OK;
  QEPW_DISPATCH(precision,
      hipLaunchKernelGGL(k_normalize_sel<double>, rows_grid(nrows, nb), dim3(256), 0, stream,
                         static_cast<const double2*>(src), lds, sq, sel,
                         static_cast<double2*>(dst), ldd, nrows),
      hipLaunchKernelGGL(k_normalize_sel<float>, rows_grid(nrows, nb), dim3(256), 0, stream,
                         static_cast<const float2*>(src), lds, sq, sel,
                         static_cast<float2*>(dst), ldd, nrows));
  return status();
}

int restart_applied(int precision, const void* x, int64_t ldx, const void* r, int64_t ldr,
                    const double* e, void* hv, int64_t ldh, int64_t nrows, int nb,
                    hipStream_t stream) {
  if (nb <= 0 || nrows <= 0) return QEPW_OK;
  QEPW_DISPATCH(precision,
      hipLaunchKernelGGL(k_restart<double>, rows_grid(nrows, nb), dim3(256), 0, stream,
                         static_cast<const double2*>(x), ldx,
                         static_cast<const double2*>(r), ldr, e,
                         static_cast<double2*>(hv), ldh, nrows),
      hipLaunchKernelGGL(k_restart<float>, rows_grid(nrows, nb), dim3(256), 0, stream,
                         static_cast<const float2*>(x), ldx,
                         static_cast<const float2*>(r), ldr, e,
                         static_cast<float2*>(hv), ldh, nrows));
  return status();
}

int permute_rows(int sp, int dp, int direction, const int* perm, const void* src,
                 int64_t lds, void* dst, int64_t ldd, int64_t nrows, int nb,
                 hipStream_t stream) {
  if (nb <= 0 || nrows <= 0) return QEPW_OK;
  if (direction != 0 && direction != 1) return QEPW_ERR_ARG;
  const dim3 g = rows_grid(nrows, nb);
  if (sp == QEPW_C128 && dp == QEPW_C128)
    hipLaunchKernelGGL((k_permute<double, double>), g, dim3(256), 0, stream, perm, direction,
                       static_cast<const double2*>(src), lds, static_cast<double2*>(dst), ldd,
                       nrows);
  else if (sp == QEPW_C128 && dp == QEPW_C64)
    hipLaunchKernelGGL((k_permute<double, float>), g, dim3(256), 0, stream, perm, direction,
                       static_cast<const double2*>(src), lds, static_cast<float2*>(dst), ldd,
                       nrows);
  else if (sp == QEPW_C64 && dp == QEPW_C128)
    hipLaunchKernelGGL((k_permute<float, double>), g, dim3(256), 0, stream, perm, direction,
                       static_cast<const float2*>(src), lds, static_cast<double2*>(dst), ldd,
                       nrows);
  else if (sp == QEPW_C64 && dp == QEPW_C64)
    hipLaunchKernelGGL((k_permute<float, float>), g, dim3(256), 0, stream, perm, direction,
                       static_cast<const float2*>(src), lds, static_cast<float2*>(dst), ldd,
                       nrows);
  else
    return QEPW_ERR_ARG;
  return status();
}

int row_real_dot(const void* a, int64_t lda, const void* b, int64_t ldb, int64_t nrows,
                 int ncols, double scale, double* acc, hipStream_t stream) {
  if (nrows <= 0) return QEPW_OK;
  hipLaunchKernelGGL(k_row_real_dot, dim3((unsigned)((nrows + 255) / 256)), dim3(256), 0,
                     stream, static_cast<const double2*>(a), lda,
                     static_cast<const double2*>(b), ldb, nrows, ncols, scale, acc);
  return status();
}

}  // namespace qepw

namespace qepw {

// y[:, j] = a[j] x[:, j] + b[j] y[:, j] (a, b: device double, NULL for 1;
// y is not read where b[j] == 0): the per-column vector updates of the
// iterative solvers -- CG's line rotation and conjugation
// (qepy_pw/diagonalization.py:1105-1160), the bpcg steps and their rollback
// (1225-1310), the RMM-DIIS step (1721-1729).
template <typename T>
__global__ void k_col_axpby(const cplx<T>* __restrict__ x, int64_t ldx,
                            const double* __restrict__ a, cplx<T>* __restrict__ y, int64_t ldy,
                            const double* __restrict__ bcoef, int64_t nrows) {
  const int64_t j = blockIdx.y;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nrows) return;
  const double fa = a ? a[j] : 1.0;
  const double fb = bcoef ? bcoef[j] : 1.0;
  const cplx<T> xv = x[j * ldx + g];
  double re = fa * (double)xv.x, im = fa * (double)xv.y;
  if (fb != 0.0) {
    const cplx<T> yv = y[j * ldy + g];
    re += fb * (double)yv.x;
    im += fb * (double)yv.y;
  }
  y[j * ldy + g] = mk<T>((T)re, (T)im);
}

// column_diagonal_expectations (_native_fft.pyx:1294-1343):
// out[b] = sum_g d[g] |v_gb|^2, fixed-order block tree
template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_col_diag_expect(const cplx<T>* __restrict__ v,
                                                                 int64_t ldv,
                                                                 const double* __restrict__ d,
                                                                 int64_t nrows,
                                                                 double* __restrict__ out) {
  __shared__ double sh[kRedThreads];
  const int64_t b = blockIdx.x;
  const cplx<T>* col = v + b * ldv;
  double acc = 0.0;
  for (int64_t g = threadIdx.x; g < nrows; g += kRedThreads) {
    const cplx<T> z = col[g];
    acc += d[g] * ((double)z.x * (double)z.x + (double)z.y * (double)z.y);
  }
  const double s = block_sum<T>(acc, sh);
  if (threadIdx.x == 0) out[b] = s;
}

// rmm_kinetic_precondition (_native_fft.pyx:1346-1408): the
// Kresse-Furthmueller diagonal K(x) = -4/(3 <T>) p(x) / (p(x) + 16 x^4),
// p(x) = 27 + 18x + 12x^2 + 8x^3, x = T_g / (1.5 <T>), applied to column
// cols[j] (or j) of r.
template <typename T>
__global__ void k_rmm_precondition(const cplx<T>* __restrict__ r, int64_t ldr,
                                   const int* __restrict__ cols,
                                   const double* __restrict__ kinetic,
                                   const double* __restrict__ expect, cplx<T>* __restrict__ out,
                                   int64_t ldo, int64_t nrows) {
  const int64_t j = blockIdx.y;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nrows) return;
  const int64_t c = cols ? cols[j] : j;
  const double e = expect[j];
  const double x = kinetic[g] / (1.5 * e);
  const double x2 = x * x;
  const double num = 27.0 + x * (18.0 + x * (12.0 + 8.0 * x));
  const double den = num + 16.0 * x2 * x2;
  const double f = (-4.0 / 3.0) * num / (e * den);
  const cplx<T> v = r[c * ldr + g];
  out[j * ldo + g] = mk<T>((T)(f * (double)v.x), (T)(f * (double)v.y));
}

// out[:, j] = sum_i coef[j, i] hist_i[:, cols[j]] with hist_i = hist + i *
// hstride and coef (ncols, count) complex row-major: the DIIS combination of
// each band's own history (qepy_pw/diagonalization.py:1655-1662).
template <typename T>
__global__ void k_history_combine(const cplx<T>* __restrict__ hist, int64_t ldh,
                                  int64_t hstride, int count, const int* __restrict__ cols,
                                  const double* __restrict__ coef, cplx<T>* __restrict__ out,
                                  int64_t ldo, int64_t nrows) {
  const int64_t j = blockIdx.y;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nrows) return;
  const int64_t c = cols ? cols[j] : j;
  double re = 0.0, im = 0.0;
  for (int i = 0; i < count; ++i) {
    const cplx<T> h = hist[i * hstride + c * ldh + g];
    const double wr = coef[2 * (j * count + i)], wi = coef[2 * (j * count + i) + 1];
    re += wr * (double)h.x - wi * (double)h.y;
    im += wr * (double)h.y + wi * (double)h.x;
  }
  out[j * ldo + g] = mk<T>((T)re, (T)im);
}

int column_axpby(int precision, const void* x, int64_t ldx, const double* a, void* y,
                 int64_t ldy, const double* b, int64_t nrows, int ncols, hipStream_t stream) {
  if (ncols <= 0 || nrows <= 0) return QEPW_OK;
  QEPW_DISPATCH(precision,
      hipLaunchKernelGGL(k_col_axpby<double>, rows_grid(nrows, ncols), dim3(256), 0, stream,
                         static_cast<const double2*>(x), ldx, a, static_cast<double2*>(y), ldy,
                         b, nrows),
      hipLaunchKernelGGL(k_col_axpby<float>, rows_grid(nrows, ncols), dim3(256), 0, stream,
                         static_cast<const float2*>(x), ldx, a, static_cast<float2*>(y), ldy,
                         b, nrows));
  return status();
}

int column_diagonal_expectations(int precision, const void* v, int64_t ldv, const double* d,
                                 int64_t nrows, int nb, double* out, hipStream_t stream) {
  if (nb <= 0) return QEPW_OK;
  QEPW_DISPATCH(precision,
      hipLaunchKernelGGL(k_col_diag_expect<double>, dim3(nb), dim3(kRedThreads), 0, stream,
                         static_cast<const double2*>(v), ldv, d, nrows, out),
      hipLaunchKernelGGL(k_col_diag_expect<float>, dim3(nb), dim3(kRedThreads), 0, stream,
                         static_cast<const float2*>(v), ldv, d, nrows, out));
  return status();
}

int rmm_precondition(int precision, const void* r, int64_t ldr, const int* cols,
                     const double* kinetic, const double* expect, void* out, int64_t ldo,
                     int64_t nrows, int ncols, hipStream_t stream) {
  if (ncols <= 0 || nrows <= 0) return QEPW_OK;
  QEPW_DISPATCH(precision,
      hipLaunchKernelGGL(k_rmm_precondition<double>, rows_grid(nrows, ncols), dim3(256), 0,
                         stream, static_cast<const double2*>(r), ldr, cols, kinetic, expect,
                         static_cast<double2*>(out), ldo, nrows),
      hipLaunchKernelGGL(k_rmm_precondition<float>, rows_grid(nrows, ncols), dim3(256), 0,
                         stream, static_cast<const float2*>(r), ldr, cols, kinetic, expect,
                         static_cast<float2*>(out), ldo, nrows));
  return status();
}

int history_combine(int precision, const void* hist, int64_t ldh, int64_t hstride, int count,
                    const int* cols, const double* coef, void* out, int64_t ldo, int64_t nrows,
                    int ncols, hipStream_t stream) {
  if (ncols <= 0 || nrows <= 0) return QEPW_OK;
  QEPW_DISPATCH(precision,
      hipLaunchKernelGGL(k_history_combine<double>, rows_grid(nrows, ncols), dim3(256), 0,
                         stream, static_cast<const double2*>(hist), ldh, hstride, count, cols,
                         coef, static_cast<double2*>(out), ldo, nrows),
      hipLaunchKernelGGL(k_history_combine<float>, rows_grid(nrows, ncols), dim3(256), 0,
                         stream, static_cast<const float2*>(hist), ldh, hstride, count, cols,
                         coef, static_cast<float2*>(out), ldo, nrows));
  return status();
}

int scale_rows(int precision, const void* src, int64_t lds, const double* s, int64_t sstride,
               const double* cfac, void* dst, int64_t ldd, int64_t nrows, int nb,
               hipStream_t stream) {
  if (nb <= 0 || nrows <= 0) return QEPW_OK;
  QEPW_DISPATCH(precision,
      hipLaunchKernelGGL(k_scale_rows<double>, rows_grid(nrows, nb), dim3(256), 0, stream,
                         static_cast<const double2*>(src), lds, s, sstride, cfac,
                         static_cast<double2*>(dst), ldd, nrows),
      hipLaunchKernelGGL(k_scale_rows<float>, rows_grid(nrows, nb), dim3(256), 0, stream,
                         static_cast<const float2*>(src), lds, s, sstride, cfac,
                         static_cast<float2*>(dst), ldd, nrows));
  return status();
}

// Real spherical harmonic of QE ylmr2 order (m = 0, then cos and sin of
// m = 1..l; Condon-Shortley phase) at the unit direction (x, y, z), l <= 3.
__device__ __forceinline__ double qe_real_harmonic(int l, int m, double x, double y, double z) {
  const double pi = 3.14159265358979323846;
  switch (l) {
    case 0:
      return 1.0 / sqrt(4.0 * pi);
    case 1: {
      const double c = sqrt(3.0 / (4.0 * pi));
      return m == 0 ? c * z : (m == 1 ? -c * x : -c * y);
    }
    case 2: {
      const double a = sqrt(15.0 / (4.0 * pi));
      switch (m) {
        case 0: return sqrt(5.0 / (16.0 * pi)) * (3.0 * z * z - 1.0);
        case 1: return -a * x * z;
        case 2: return -a * y * z;
        case 3: return sqrt(15.0 / (16.0 * pi)) * (x * x - y * y);
        default: return a * x * y;
      }
    }
    default: {
      switch (m) {
        case 0: return 0.25 * sqrt(7.0 / pi) * (5.0 * z * z * z - 3.0 * z);
        case 1: return -0.125 * sqrt(42.0 / pi) * (5.0 * z * z - 1.0) * x;
        case 2: return -0.125 * sqrt(42.0 / pi) * (5.0 * z * z - 1.0) * y;
        case 3: return 0.25 * sqrt(105.0 / pi) * z * (x * x - y * y);
        case 4: return 0.5 * sqrt(105.0 / pi) * x * y * z;
        case 5: return -0.125 * sqrt(70.0 / pi) * (x * x * x - 3.0 * x * y * y);
        default: return -0.125 * sqrt(70.0 / pi) * (3.0 * x * x * y - y * y * y);
      }
    }
  }
}

// beta[g, j] = table_{p_j}(|q_g|) Y_{l_j m_j}(q_g) (-i)^{l_j}: the per-k
// assembly of the species projectors (QE init_us_2; reference
// qepy_pw/upf.py:483-541) with QE's four-point cubic interpolation of the
// radial table (upf.py:880-912): samples i0 .. i0+3, i0 = floor(|q| / dq).
// The +z direction stands for q = 0.
__global__ void k_projector_basis(const double* __restrict__ q, int64_t npw,
                                  const double* __restrict__ table, int64_t ldt, int nq,
                                  double dq, const int* __restrict__ chan_radial,
                                  const int* __restrict__ chan_l, const int* __restrict__ chan_m,
                                  double2* __restrict__ beta, int64_t ldb) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (g >= npw) return;
  const double qx = q[3 * g], qy = q[3 * g + 1], qz = q[3 * g + 2];
  const double qn = sqrt(qx * qx + qy * qy + qz * qz);
  double x = 0.0, y = 0.0, z = 1.0;
  if (qn > 1.0e-14) {
    x = qx / qn;
    y = qy / qn;
    z = qz / qn;
  }
  const double scaled = qn / dq;
  int lower = (int)floor(scaled);
  // the host checks that the table covers every |q|; the clamp only keeps
  // the reads inside it
  lower = lower < 0 ? 0 : (lower > nq - 4 ? nq - 4 : lower);
  const double px = scaled - lower;
  const double ux = 1.0 - px, vx = 2.0 - px, wx = 3.0 - px;
  const double* t = table + (int64_t)chan_radial[j] * ldt + lower;
  double radial = t[0] * (ux * vx * wx / 6.0);
  radial += t[1] * (px * vx * wx / 2.0);
  radial += t[2] * (-px * ux * wx / 2.0);
  radial += t[3] * (px * ux * vx / 6.0);
  const int l = chan_l[j];
  const double v = radial * qe_real_harmonic(l, chan_m[j], x, y, z);
  double2 out;
  switch (l & 3) {  // (-i)^l
    case 0: out = make_double2(v, 0.0); break;
    case 1: out = make_double2(0.0, -v); break;
    case 2: out = make_double2(-v, 0.0); break;
    default: out = make_double2(0.0, v); break;
  }
  beta[g + (int64_t)j * ldb] = out;
}

int projector_basis(const double* q, int64_t npw, const double* table, int64_t ldt, int nq,
                    double dq, const int* chan_radial, const int* chan_l, const int* chan_m,
                    int nch, void* beta, int64_t ldb, hipStream_t stream) {
  if (npw <= 0 || nch <= 0) return QEPW_OK;
  const dim3 grid((unsigned)((npw + 255) / 256), (unsigned)nch);
  hipLaunchKernelGGL(k_projector_basis, grid, dim3(256), 0, stream, q, npw, table, ldt, nq, dq,
                     chan_radial, chan_l, chan_m, static_cast<double2*>(beta), ldb);
  return hipGetLastError() == hipSuccess ? QEPW_OK : QEPW_ERR_HIP;
}

int build_projectors(int precision, const void* beta, int64_t ldb, int nch, const double* q,
                     const double* tau, int natom, void* out, int64_t ldo, int64_t nrows,
                     hipStream_t stream) {
  if (natom <= 0 || nrows <= 0 || nch <= 0) return QEPW_OK;
  if (natom > 65535) return QEPW_ERR_UNSUPPORTED;
  QEPW_DISPATCH(precision,
      hipLaunchKernelGGL(k_build_projectors<double>, rows_grid(nrows, natom), dim3(256), 0,
                         stream, static_cast<const double2*>(beta), ldb, nch, q, tau, natom,
                         static_cast<double2*>(out), ldo, nrows),
      hipLaunchKernelGGL(k_build_projectors<float>, rows_grid(nrows, natom), dim3(256), 0,
                         stream, static_cast<const float2*>(beta), ldb, nch, q, tau, natom,
                         static_cast<float2*>(out), ldo, nrows));
  return status();
}

int phase_blocks(int precision, const void* beta, int64_t ldb, int nch, const void* phases,
                 int64_t ldp, int natom, const int* perm, void* out, int64_t ldo, int64_t nrows,
                 hipStream_t stream) {
  if (natom <= 0 || nrows <= 0 || nch <= 0) return QEPW_OK;
  if (natom > 65535) return QEPW_ERR_UNSUPPORTED;
  QEPW_DISPATCH(precision,
      hipLaunchKernelGGL(k_phase_blocks<double>, rows_grid(nrows, natom), dim3(256), 0, stream,
                         static_cast<const double2*>(beta), ldb, nch,
                         static_cast<const double2*>(phases), ldp, natom, perm,
                         static_cast<double2*>(out), ldo, nrows),
      hipLaunchKernelGGL(k_phase_blocks<float>, rows_grid(nrows, natom), dim3(256), 0, stream,
                         static_cast<const float2*>(beta), ldb, nch,
                         static_cast<const float2*>(phases), ldp, natom, perm,
                         static_cast<float2*>(out), ldo, nrows));
  return status();
}

int projector_forces(int precision, const void* dO, int64_t lddo, int64_t dstride,
                     const void* DO, int64_t lddO, const double* f, int nb, const int* off,
                     int natom, double* out, hipStream_t stream) {
  if (natom <= 0) return QEPW_OK;
  if (natom > 0x7fffffff / 3) return QEPW_ERR_UNSUPPORTED;
  QEPW_DISPATCH(precision,
      hipLaunchKernelGGL(k_projector_forces<double>, dim3(natom, 3), dim3(kRedThreads), 0,
                         stream, static_cast<const double2*>(dO), lddo, dstride,
                         static_cast<const double2*>(DO), lddO, f, nb, off, out),
      hipLaunchKernelGGL(k_projector_forces<float>, dim3(natom, 3), dim3(kRedThreads), 0,
                         stream, static_cast<const float2*>(dO), lddo, dstride,
                         static_cast<const float2*>(DO), lddO, f, nb, off, out));
  return status();
}

int copy_columns(int src_precision, int dst_precision, const void* src, int64_t lds,
                 const int* scol, void* dst, int64_t ldd, const int* dcol, const double* scale,
                 int64_t nrows, int ncols, hipStream_t stream) {
  if (ncols <= 0 || nrows <= 0) return QEPW_OK;
  if (ncols > 65535) return QEPW_ERR_UNSUPPORTED;
  const dim3 g = rows_grid(nrows, ncols);
#define QEPW_COPY_COLS(TS, TD)                                                              \
  hipLaunchKernelGGL((k_copy_columns<TS, TD>), g, dim3(256), 0, stream,                    \
                     static_cast<const cplx<TS>*>(src), lds, scol, static_cast<cplx<TD>*>(dst), \
                     ldd, dcol, scale, nrows)
  if (src_precision == QEPW_C128 && dst_precision == QEPW_C128) QEPW_COPY_COLS(double, double);
  else if (src_precision == QEPW_C128 && dst_precision == QEPW_C64) QEPW_COPY_COLS(double, float);
  else if (src_precision == QEPW_C64 && dst_precision == QEPW_C128) QEPW_COPY_COLS(float, double);
  else if (src_precision == QEPW_C64 && dst_precision == QEPW_C64) QEPW_COPY_COLS(float, float);
  else return QEPW_ERR_ARG;
#undef QEPW_COPY_COLS
  return status();
}

int set_block(int precision, void* dst, int64_t ldd, int64_t nrows, int ncols, double diag,
              const double* d, hipStream_t stream) {
  if (ncols <= 0 || nrows <= 0) return QEPW_OK;
  if (ncols > 65535) return QEPW_ERR_UNSUPPORTED;
  QEPW_DISPATCH(precision,
      hipLaunchKernelGGL(k_set_block<double>, rows_grid(nrows, ncols), dim3(256), 0, stream,
                         static_cast<double2*>(dst), ldd, nrows, diag, d),
      hipLaunchKernelGGL(k_set_block<float>, rows_grid(nrows, ncols), dim3(256), 0, stream,
                         static_cast<float2*>(dst), ldd, nrows, diag, d));
  return status();
}

}  // namespace qepw
