// Small dense kernels of the device Rayleigh-Ritz (FP64, n <= a few hundred).
//
// The reduced Davidson problem H c = e S c (reference: zhegvd via
// generalized_eigh, qepy_pw/_native_fft.pyx:667-691) is solved on the device
// by a spectral division: with W = L^-1 (S = L L^H) kept up to date
// incrementally, A = W H W^H, the Newton-Schulz iteration
// X <- (a/2) X (3 I - a^2 X^2) drives X0 = (A - sigma I)/s to sign(A - sigma I),
// P = (I - X)/2 projects on the eigenvectors below sigma, and a Rayleigh-Ritz
// over range(P) gives the lowest pairs.  These kernels are the non-GEMM
// pieces; all matrices are column-major complex128 with leading dimension ld.
#include "qepw_internal.h"

namespace qepw {

constexpr int kEigThreads = 256;

// Rows [first, active) of A hold new values for columns [0, active):
// mirror them into columns [first, active) of rows [0, first) and
// symmetrise the new diagonal block (update_projected,
// qepy_pw/diagonalization.py:658-697).
__global__ void k_hermitian_complete(double2* __restrict__ a, int64_t lda, int first,
                                     int active) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nnew = active - first;
  if (idx >= (int64_t)nnew * active) return;
  const int r = first + (int)(idx % nnew);
  const int c = (int)(idx / nnew);
  if (c < first) {
    const double2 v = a[r + c * lda];
    a[c + r * lda] = make_double2(v.x, -v.y);
  } else if (c < r) {
    const double2 u = a[r + c * lda];
    const double2 w = a[c + r * lda];
    const double2 m = make_double2(0.5 * (u.x + w.x), 0.5 * (u.y - w.y));
    a[r + c * lda] = m;
    a[c + r * lda] = make_double2(m.x, -m.y);
  } else if (c == r) {
    a[r + c * lda].y = 0.0;
  }
}

// Sum over one column of a wave64: lane 0 ends with the total.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// colsum[j] = sum_i |A_ij - sigma delta_ij|: one wave per column (A is
// Hermitian, so these are also the row sums)
__global__ __launch_bounds__(kEigThreads) void k_shift_colsums(const double2* __restrict__ a,
                                                               int64_t lda, int n, double sigma,
                                                               double* __restrict__ colsum) {
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * (kEigThreads / 64) + (threadIdx.x >> 6);
  if (j >= n) return;
  double s = 0.0;
  for (int i = lane; i < n; i += 64) {
    double2 v = a[i + (int64_t)j * lda];
    if (i == j) v.x -= sigma;
    s += sqrt(v.x * v.x + v.y * v.y);
  }
  s = wave_sum(s);
  if (lane == 0) colsum[j] = s;
}

// X = (A - sigma I) / bound[0], bound[0] = max_j colsum[j] (>= ||A - sigma I||_2)
// with colsum = bound + 1; every block forms the maximum itself
__global__ __launch_bounds__(256) void k_shift_scale(const double2* __restrict__ a,
                                                     int64_t lda, int n, double sigma,
                                                     double* __restrict__ bound,
                                                     double2* __restrict__ x, int64_t ldx) {
  __shared__ double sh[256];
  double best = 0.0;
  for (int j = threadIdx.x; j < n; j += 256) best = fmax(best, bound[1 + j]);
  sh[threadIdx.x] = best;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  const double b = sh[0] > 0.0 ? sh[0] : 1.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) bound[0] = b;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  const double inv = 1.0 / b;
  double2 v = a[i + (int64_t)j * lda];
  if (i == j) v.x -= sigma;
  x[i + (int64_t)j * ldx] = make_double2(v.x * inv, v.y * inv);
}

// X = (X + X^H) / 2 in place
__global__ void k_hermitize(double2* __restrict__ x, int64_t ldx, int n) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  if (i < j) return;
  if (i == j) {
    x[i + (int64_t)j * ldx].y = 0.0;
    return;
  }
  const double2 u = x[i + (int64_t)j * ldx];
  const double2 w = x[j + (int64_t)i * ldx];
  const double2 m = make_double2(0.5 * (u.x + w.x), 0.5 * (u.y - w.y));
  x[i + (int64_t)j * ldx] = m;
  x[j + (int64_t)i * ldx] = make_double2(m.x, -m.y);
}

// part[j] = sum_i |X2_ij - delta_ij|^2: one wave per column
__global__ __launch_bounds__(kEigThreads) void k_sign_colsums(const double2* __restrict__ x2,
                                                              int64_t ld2, int n,
                                                              double* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * (kEigThreads / 64) + (threadIdx.x >> 6);
  if (j >= n) return;
  double r = 0.0;
  for (int i = lane; i < n; i += 64) {
    double2 v = x2[i + (int64_t)j * ld2];
    if (i == j) v.x -= 1.0;
    r += v.x * v.x + v.y * v.y;
  }
  r = wave_sum(r);
  if (lane == 0) part[j] = r;
}

// out[0] = sum_j part[j] = ||X2 - I||_F^2, out[1] = Re trace(X) with
// part = out + 2; one block, fixed order
__global__ __launch_bounds__(kEigThreads) void k_sign_final(const double2* __restrict__ x,
                                                            int64_t ldx, int n,
                                                            double* __restrict__ out) {
  __shared__ double sh[2][kEigThreads];
  double r = 0.0, t = 0.0;
  for (int j = threadIdx.x; j < n; j += kEigThreads) {
    r += out[2 + j];
    t += x[j + (int64_t)j * ldx].x;
  }
  sh[0][threadIdx.x] = r;
  sh[1][threadIdx.x] = t;
  __syncthreads();
  for (int s = kEigThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sh[0][0];
    out[1] = sh[1][0];
  }
}

// A[:n, :n] = diag(d) (identity when d is null)
__global__ void k_set_diagonal(double2* __restrict__ a, int64_t lda, int n,
                               const double* __restrict__ d) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  a[i + (int64_t)j * lda] = make_double2(i == j ? (d ? d[i] : 1.0) : 0.0, 0.0);
}

static inline unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }
static inline int status() { return hipGetLastError() == hipSuccess ? QEPW_OK : QEPW_ERR_HIP; }

int hermitian_complete(void* a, int64_t lda, int first, int active, hipStream_t stream) {
  if (active <= first) return QEPW_OK;
  const int64_t work = (int64_t)(active - first) * active;
  hipLaunchKernelGGL(k_hermitian_complete, dim3(blocks_for(work)), dim3(256), 0, stream,
                     static_cast<double2*>(a), lda, first, active);
  return status();
}

int shift_scale(const void* a, int64_t lda, int n, double sigma, void* x, int64_t ldx,
                double* bound, hipStream_t stream) {
  if (n <= 0) return QEPW_OK;
  const unsigned col_blocks = (unsigned)((n + kEigThreads / 64 - 1) / (kEigThreads / 64));
  hipLaunchKernelGGL(k_shift_colsums, dim3(col_blocks), dim3(kEigThreads), 0, stream,
                     static_cast<const double2*>(a), lda, n, sigma, bound + 1);
  hipLaunchKernelGGL(k_shift_scale, dim3(blocks_for((int64_t)n * n)), dim3(256), 0, stream,
                     static_cast<const double2*>(a), lda, n, sigma, bound,
                     static_cast<double2*>(x), ldx);
  return status();
}

int hermitize(void* x, int64_t ldx, int n, hipStream_t stream) {
  if (n <= 0) return QEPW_OK;
  hipLaunchKernelGGL(k_hermitize, dim3(blocks_for((int64_t)n * n)), dim3(256), 0, stream,
                     static_cast<double2*>(x), ldx, n);
  return status();
}

int sign_check(const void* x, int64_t ldx, const void* x2, int64_t ld2, int n, double* out,
               hipStream_t stream) {
  if (n <= 0) return QEPW_OK;
  const unsigned col_blocks = (unsigned)((n + kEigThreads / 64 - 1) / (kEigThreads / 64));
  hipLaunchKernelGGL(k_sign_colsums, dim3(col_blocks), dim3(kEigThreads), 0, stream,
                     static_cast<const double2*>(x2), ld2, n, out + 2);
  hipLaunchKernelGGL(k_sign_final, dim3(1), dim3(kEigThreads), 0, stream,
                     static_cast<const double2*>(x), ldx, n, out);
  return status();
}

size_t sign_iteration_workspace(int) { return 0; }

// Scaled Newton-Schulz steps X <- (a/2) X (3 I - a^2 X^2), one per entry of
// ``alphas``: Y = X X, X' = 1.5 a X - 0.5 a^3 X Y, each a Hermitian product
// (herm_product: lower-triangle tiles, mirrored, so every iterate is exactly
// Hermitian).  The iterate ping-pongs between x and x2; *in_x2 tells which
// holds the result, and y is left holding its square for the caller's
// convergence check.
int sign_iteration(void* x, void* x2, void* y, int64_t ld, int n, const double* alphas,
                   int nalpha, int* in_x2, void*, size_t, hipStream_t stream) {
  double2* cur = static_cast<double2*>(x);
  double2* nxt = static_cast<double2*>(x2);
  for (int i = 0; i < nalpha; ++i) {
    const double a = alphas[i];
    int rc = herm_product(n, cur, cur, cur, y, ld, 1.0, 0.0, stream);
    if (rc) return rc;
    rc = herm_product(n, cur, y, cur, nxt, ld, -0.5 * a * a * a, 1.5 * a, stream);
    if (rc) return rc;
    double2* t = cur;
    cur = nxt;
    nxt = t;
  }
  *in_x2 = cur == static_cast<double2*>(x2) ? 1 : 0;
  return herm_product(n, cur, cur, cur, y, ld, 1.0, 0.0, stream);
}

int set_diagonal(void* a, int64_t lda, int n, const double* d, hipStream_t stream) {
  if (n <= 0) return QEPW_OK;
  hipLaunchKernelGGL(k_set_diagonal, dim3(blocks_for((int64_t)n * n)), dim3(256), 0, stream,
                     static_cast<double2*>(a), lda, n, d);
  return status();
}

}  // namespace qepw
