// Complex GEMM on CDNA4 matrix cores: C = alpha op(A) op(B) + beta C.
//
// Every complex product is four real MFMA products accumulated into separate
// real / imaginary accumulators (no 3M trick, which loses accuracy):
//   op N : re += Ar Br - Ai Bi ;  im += Ar Bi + Ai Br
//   op C : re += Ar Br + Ai Bi ;  im += Ar Bi - Ai Br
// (op(B) = B^H is conjugated as it is staged into LDS).
// using v_mfma_f64_16x16x4f64 (FP64) or v_mfma_f32_16x16x4f32 (FP32, FP32
// accumulation).  A 256-thread workgroup owns a TILE x TILE output tile
// (TILE = 64, or 32 for small outputs such as the reduced problem's square
// products, where 64x64 tiles leave most CUs idle); each of its four waves
// owns a quarter (2x2 or 1x1 MFMA blocks of 16x16) and walks K in steps of
// 16 staged through LDS.  Tall-skinny A^H B products (K = npw up
// to ~24k, M,N <= 512) are split over K: every split writes its partial tile
// to the workspace and a second kernel sums the splits in fixed order, so
// results are bitwise reproducible run to run.
//
// MFMA 16x16x4 operand layout (lane l of a wave64): A element (row l%16,
// k l/16); B element (k l/16, col l%16).  Accumulator v of lane l holds
// column l%16 and row 4*(l/16)+v for the f32 instruction but row l/16+4v
// for the f64 one (measured on gfx950 with an in-bounds probe).
#include <type_traits>

#include "qepw_internal.h"

namespace qepw {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
template <> struct Mfma<double> {
  using acc = v4d;
  static __device__ __forceinline__ acc op(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct Mfma<float> {
  using acc = v4f;
  static __device__ __forceinline__ acc op(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

constexpr int kBK = 16;
constexpr int kGemmThreads = 256;
constexpr int kPad = 1;

// TM x TN output tile per 256-thread workgroup; its four waves form a
// WAVES_M x (4 / WAVES_M) grid, each owning a (TM / WAVES_M) x (TN * WAVES_M / 4)
// block of 16x16 MFMA tiles.  Rectangular tiles let a tall-skinny product
// (npw x 128 outputs) cover all its columns in one workgroup, so the tall
// operand is streamed from HBM once.
template <typename T, bool CONJ_A, bool CONJ_B, int TM, int TN, int WAVES_M>
__global__ __launch_bounds__(kGemmThreads) void k_cgemm(
    int64_t M, int64_t N, int64_t K, int64_t kchunk, const cplx<T>* __restrict__ A,
    int64_t lda, const cplx<T>* __restrict__ B, int64_t ldb, cplx<T>* __restrict__ C,
    int64_t ldc, cplx<T>* __restrict__ partial, T alpha_r, T alpha_i, T beta_r,
    T beta_i) {
  using Acc = typename Mfma<T>::acc;
  constexpr int WAVES_N = 4 / WAVES_M;
  constexpr int WTM = TM / WAVES_M;  // wave tile
  constexpr int WTN = TN / WAVES_N;
  constexpr int NBM = WTM / 16;      // MFMA blocks per wave and dimension
  constexpr int NBN = WTN / 16;
  static_assert(NBM >= 1 && NBN >= 1 && WAVES_M * WAVES_N == 4, "tile shape");
  __shared__ cplx<T> As[kBK][TM + kPad];
  __shared__ cplx<T> Bs[kBK][TN + kPad];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave % WAVES_M) * WTM;
  const int wn = (wave / WAVES_M) * WTN;
  const int64_t m0 = (int64_t)blockIdx.x * TM;
  const int64_t n0 = (int64_t)blockIdx.y * TN;
  const int64_t kbeg = (int64_t)blockIdx.z * kchunk;
  const int64_t kend = (kbeg + kchunk) < K ? (kbeg + kchunk) : K;
  const cplx<T> zero = mk<T>(T(0), T(0));

  Acc re[NBM][NBN], im[NBM][NBN];
#pragma unroll
  for (int i = 0; i < NBM; ++i)
#pragma unroll
    for (int j = 0; j < NBN; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) { re[i][j][v] = T(0); im[i][j][v] = T(0); }

  // Software pipeline: the operands of step k0 + kBK are fetched into
  // registers while the MFMAs of step k0 run from LDS, so the global-load
  // latency overlaps the matrix work instead of preceding it.
  constexpr int kPerA = kBK * TM / kGemmThreads;  // elements per thread
  constexpr int kPerB = kBK * TN / kGemmThreads;
  static_assert(kPerA >= 1 && kPerB >= 1, "tile too small for the workgroup");
  cplx<T> ra[kPerA], rb[kPerB];
  // element e of the kBK x TM / kBK x TN operand tiles of step K0
  auto load_a = [=](int e, int64_t k0) -> cplx<T> {
    const int kk = CONJ_A ? (e & (kBK - 1)) : (e / TM);
    const int mm = CONJ_A ? (e / kBK) : (e % TM);
    const int64_t ka = k0 + kk, m = m0 + mm;
    if (ka >= kend || m >= M) return zero;
    return CONJ_A ? A[ka + m * lda] : A[m + ka * lda];
  };
  auto load_b = [=](int e, int64_t k0) -> cplx<T> {
    const int kb = CONJ_B ? (e / TN) : (e & (kBK - 1));
    const int nb = CONJ_B ? (e % TN) : (e / kBK);
    const int64_t kq = k0 + kb, n = n0 + nb;
    if (kq >= kend || n >= N) return zero;
    if (CONJ_B) {  // op(B) = B^H, B stored n x k
      cplx<T> w = B[n + kq * ldb];
      w.y = -w.y;
      return w;
    }
    return B[kq + n * ldb];
  };

  if (kbeg < kend) {
#pragma unroll
    for (int p = 0; p < kPerA; ++p) ra[p] = load_a(tid + p * kGemmThreads, kbeg);
#pragma unroll
    for (int p = 0; p < kPerB; ++p) rb[p] = load_b(tid + p * kGemmThreads, kbeg);
  }
  for (int64_t k0 = kbeg; k0 < kend; k0 += kBK) {
#pragma unroll
    for (int p = 0; p < kPerA; ++p) {
      const int e = tid + p * kGemmThreads;
      if (CONJ_A) As[e & (kBK - 1)][e / kBK] = ra[p];
      else As[e / TM][e % TM] = ra[p];
    }
#pragma unroll
    for (int p = 0; p < kPerB; ++p) {
      const int e = tid + p * kGemmThreads;
      if (CONJ_B) Bs[e / TN][e % TN] = rb[p];
      else Bs[e & (kBK - 1)][e / kBK] = rb[p];
    }
    __syncthreads();
    if (k0 + kBK < kend) {
#pragma unroll
      for (int p = 0; p < kPerA; ++p) ra[p] = load_a(tid + p * kGemmThreads, k0 + kBK);
#pragma unroll
      for (int p = 0; p < kPerB; ++p) rb[p] = load_b(tid + p * kGemmThreads, k0 + kBK);
    }
#pragma unroll
    for (int ks = 0; ks < kBK; ks += 4) {
      const int kk = ks + (lane >> 4);
      cplx<T> a[NBM], b[NBN];
#pragma unroll
      for (int i = 0; i < NBM; ++i) a[i] = As[kk][wm + 16 * i + (lane & 15)];
#pragma unroll
      for (int j = 0; j < NBN; ++j) b[j] = Bs[kk][wn + 16 * j + (lane & 15)];
#pragma unroll
      for (int i = 0; i < NBM; ++i) {
        const T ai_re = CONJ_A ? a[i].y : -a[i].y;
        const T ai_im = CONJ_A ? -a[i].y : a[i].y;
#pragma unroll
        for (int j = 0; j < NBN; ++j) {
          re[i][j] = Mfma<T>::op(a[i].x, b[j].x, re[i][j]);
          re[i][j] = Mfma<T>::op(ai_re, b[j].y, re[i][j]);
          im[i][j] = Mfma<T>::op(a[i].x, b[j].y, im[i][j]);
          im[i][j] = Mfma<T>::op(ai_im, b[j].x, im[i][j]);
        }
      }
    }
    __syncthreads();
  }

  const bool direct = gridDim.z == 1;
#pragma unroll
  for (int i = 0; i < NBM; ++i)
#pragma unroll
    for (int j = 0; j < NBN; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t row = m0 + wm + 16 * i + Mfma<T>::row(lane, v);
        const int64_t col = n0 + wn + 16 * j + (lane & 15);
        if (row >= M || col >= N) continue;
        const T sr = re[i][j][v], si = im[i][j][v];
        if (direct) {
          cplx<T> out;
          out.x = alpha_r * sr - alpha_i * si;
          out.y = alpha_r * si + alpha_i * sr;
          if (beta_r != T(0) || beta_i != T(0)) {
            const cplx<T> c = C[row + col * ldc];
            out.x += beta_r * c.x - beta_i * c.y;
            out.y += beta_r * c.y + beta_i * c.x;
          }
          C[row + col * ldc] = out;
        } else {
          partial[((int64_t)blockIdx.z * N + col) * M + row] = mk<T>(sr, si);
        }
      }
}

// Sums the split-K partial tiles: a block owns 64 consecutive outputs and
// each of its four waves adds every fourth split (eight loads in flight per
// lane); the four partial sums are combined in LDS in wave order, so the
// summation order is fixed and results are reproducible.
template <typename T>
__global__ __launch_bounds__(256) void k_reduce_splits(
    int64_t M, int64_t N, int splits, const cplx<T>* __restrict__ partial,
    cplx<T>* __restrict__ C, int64_t ldc, T alpha_r, T alpha_i, T beta_r, T beta_i) {
  __shared__ T sr_part[4][64], si_part[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t total = M * N;
  const int64_t idx = (int64_t)blockIdx.x * 64 + lane;
  T sr = T(0), si = T(0);
  if (idx < total) {
    const cplx<T>* p = partial + idx;
    int z = wave;
    for (; z + 28 < splits; z += 32) {
      cplx<T> v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = p[(int64_t)(z + 4 * u) * total];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        sr += v[u].x;
        si += v[u].y;
      }
    }
    for (; z < splits; z += 4) {
      const cplx<T> v = p[(int64_t)z * total];
      sr += v.x;
      si += v.y;
    }
  }
  sr_part[wave][lane] = sr;
  si_part[wave][lane] = si;
  __syncthreads();
  if (wave != 0 || idx >= total) return;
  sr = sr_part[0][lane] + sr_part[1][lane] + sr_part[2][lane] + sr_part[3][lane];
  si = si_part[0][lane] + si_part[1][lane] + si_part[2][lane] + si_part[3][lane];
  const int64_t col = idx / M, row = idx - col * M;
  cplx<T> out;
  out.x = alpha_r * sr - alpha_i * si;
  out.y = alpha_r * si + alpha_i * sr;
  if (beta_r != T(0) || beta_i != T(0)) {
    const cplx<T> c = C[row + col * ldc];
    out.x += beta_r * c.x - beta_i * c.y;
    out.y += beta_r * c.y + beta_i * c.x;
  }
  C[row + col * ldc] = out;
}

// One-column products (n = 1: a single band's projections and projector
// sums) as matrix-vector passes instead of MFMA tiles with one live column.
// Both write split-K partials part[s * m + i] that k_reduce_splits sums in a
// fixed order and scales by alpha, beta.
//   k_gemv_n: sum_{j in split s} A[i + j lda] b_j, one thread per row i;
//   k_gemv_c: sum_{g in split s} conj(A[g + i lda]) b_g, one workgroup per
//             (i, s) with a fixed-order tree sum.
// b_j = B[j bstride], conjugated when CB.
template <typename T, bool CB>
__global__ __launch_bounds__(256) void k_gemv_n(int64_t m, int64_t k, int64_t kchunk,
                                                const cplx<T>* __restrict__ A, int64_t lda,
                                                const cplx<T>* __restrict__ B, int64_t bstride,
                                                cplx<T>* __restrict__ part) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  const int64_t j0 = (int64_t)blockIdx.y * kchunk;
  const int64_t j1 = j0 + kchunk < k ? j0 + kchunk : k;
  T sr = T(0), si = T(0);
#pragma unroll 8
  for (int64_t j = j0; j < j1; ++j) {
    const cplx<T> a = A[i + j * lda];
    const cplx<T> b = B[j * bstride];
    const T bi = CB ? -b.y : b.y;
    sr += a.x * b.x - a.y * bi;
    si += a.x * bi + a.y * b.x;
  }
  part[(int64_t)blockIdx.y * m + i] = mk<T>(sr, si);
}

template <typename T, bool CB>
__global__ __launch_bounds__(256) void k_gemv_c(int64_t m, int64_t k, int64_t kchunk,
                                                const cplx<T>* __restrict__ A, int64_t lda,
                                                const cplx<T>* __restrict__ B, int64_t bstride,
                                                cplx<T>* __restrict__ part) {
  __shared__ T shr[256], shi[256];
  const int64_t i = blockIdx.x;
  const int64_t g0 = (int64_t)blockIdx.y * kchunk;
  const int64_t g1 = g0 + kchunk < k ? g0 + kchunk : k;
  const cplx<T>* col = A + i * lda;
  T sr = T(0), si = T(0);
  for (int64_t g = g0 + threadIdx.x; g < g1; g += 256) {
    const cplx<T> a = col[g];
    const cplx<T> b = B[g * bstride];
    const T bi = CB ? -b.y : b.y;
    sr += a.x * b.x + a.y * bi;
    si += a.x * bi - a.y * b.x;
  }
  const int tid = threadIdx.x;
  shr[tid] = sr;
  shi[tid] = si;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) {
      shr[tid] += shr[tid + w];
      shi[tid] += shi[tid + w];
    }
    __syncthreads();
  }
  if (tid == 0) part[(int64_t)blockIdx.y * m + i] = mk<T>(shr[0], shi[0]);
}

// Splits of a one-column product: about 1024 workgroups in all, each split of
// K at least 8 columns of A (k_gemv_n) or 1024 rows (k_gemv_c).
struct GemvPlan {
  int64_t splits;
  int64_t kchunk;
};

static GemvPlan gemv_plan(bool ca, int64_t m, int64_t k) {
  const int64_t blocks = ca ? m : (m + 255) / 256;
  const int64_t most = ca ? (k + 1023) / 1024 : (k + 7) / 8;
  int64_t splits = (1024 + blocks - 1) / blocks;
  if (splits > most) splits = most;
  if (splits > 65535) splits = 65535;
  if (splits < 1) splits = 1;
  const int64_t kchunk = (k + splits - 1) / splits;
  return {(k + kchunk - 1) / kchunk, kchunk};
}

// Tile variants of k_cgemm: (TM, TN, WAVES_M).
struct GemmVariant {
  int tm, tn, waves_m;
};
constexpr GemmVariant kVariants[] = {
    {32, 32, 2},    // 0: small outputs (the reduced problem), wave 16x16
    {64, 64, 2},    // 1: square-ish outputs, wave 32x32
    {32, 128, 1},   // 2: tall outputs with up to 128 columns: A streamed once
    {64, 128, 2},   // 3: tall outputs, wave 32x64
    {128, 64, 2},   // 4: tall-row outputs (A^H B with m > n), wave 64x32
    {128, 32, 4},   // 5: tall outputs with 32 columns
    {32, 64, 1},    // 6: tall outputs with up to 64 columns, wave 32x16
};
constexpr int kNumVariants = (int)(sizeof(kVariants) / sizeof(kVariants[0]));

struct SplitPlan {
  int64_t splits;
  int64_t kchunk;
  int variant;
};

// Tuned choices (qepw_gemm_set_rule): the first rule whose ranges hold the
// shape wins over the built-in heuristic.
struct GemmRule {
  int precision, op;
  int64_t m_lo, m_hi, n_lo, n_hi, k_lo, k_hi;
  int variant;
  int64_t splits;
};
constexpr int kMaxRules = 256;
GemmRule g_rules[kMaxRules];
int g_nrules = 0;

static int64_t tiles_of(int variant, int64_t m, int64_t n) {
  const GemmVariant& v = kVariants[variant];
  return ((m + v.tm - 1) / v.tm) * ((n + v.tn - 1) / v.tn);
}

static SplitPlan finish_plan(int variant, int64_t splits, int64_t k) {
  if (splits < 1) splits = 1;
  int64_t kchunk = (k + splits - 1) / splits;
  kchunk = ((kchunk + kBK - 1) / kBK) * kBK;
  if (kchunk < kBK) kchunk = kBK;
  splits = (k + kchunk - 1) / kchunk;
  if (splits < 1) splits = 1;
  return {splits, kchunk, variant};
}

static SplitPlan split_plan(int precision, int op_a, int64_t m, int64_t n, int64_t k) {
  for (int r = 0; r < g_nrules; ++r) {
    const GemmRule& g = g_rules[r];
    if (g.precision == precision && g.op == op_a && m >= g.m_lo && m <= g.m_hi &&
        n >= g.n_lo && n <= g.n_hi && k >= g.k_lo && k <= g.k_hi)
      return finish_plan(g.variant, g.splits, k);
  }
  const bool ca = (op_a & QEPW_OP_C) != 0;
  const bool tall = ca && k > 2048;
  int variant;
  if (tall) {
    // A^H B over npw: parallelism from splitting K
    variant = 1;
  } else if (!ca && m > 4096 && n <= 128) {
    // npw x n outputs (Ritz vectors, residuals, projector sums): all n
    // columns in one workgroup, so the npw x k operand is read once
    variant = n > 64 ? 2 : (n > 32 ? 6 : 0);
  } else {
    const int64_t tiles64 = ((m + 63) / 64) * ((n + 63) / 64);
    variant = tiles64 < 256 ? 0 : 1;
  }
  const int64_t tiles = tiles_of(variant, m, n);
  int64_t splits = 1;
  // split K when the output tiles alone cannot fill the 256 CUs: the tall
  // A^H B products over npw (about two workgroups per CU and at least 256 of
  // K per split: shorter splits spend more on writing and summing their
  // partial tiles than on the product) and the small square products of the
  // reduced eigensolve (small partial tiles: four workgroups per CU, at
  // least 128 of K per split)
  if (tall || (ca && k > 512) || (tiles < 256 && k >= 256)) {
    const int64_t want = ((tall ? 512 : 1024) + tiles - 1) / tiles;
    const int64_t most = k / (tall ? 256 : 128);
    splits = want < most ? want : most;
  }
  return finish_plan(variant, splits, k);
}

int gemm_set_rule(int precision, int op_a, const int64_t* ranges, int variant, int64_t splits) {
  if (precision != QEPW_C128 && precision != QEPW_C64) return QEPW_ERR_ARG;
  if (op_a & ~(QEPW_OP_C | QEPW_OP_BC)) return QEPW_ERR_ARG;
  if (variant < 0 || variant >= kNumVariants || splits < 1 || splits > 65535 || !ranges)
    return QEPW_ERR_ARG;
  if (g_nrules >= kMaxRules) return QEPW_ERR_UNSUPPORTED;
  g_rules[g_nrules++] = {precision, op_a,      ranges[0], ranges[1], ranges[2],
                         ranges[3], ranges[4], ranges[5], variant,   splits};
  return QEPW_OK;
}

void gemm_clear_rules() { g_nrules = 0; }

int gemm_variant_count() { return kNumVariants; }

int gemm_plan(int precision, int op_a, int64_t m, int64_t n, int64_t k, int64_t* out) {
  const SplitPlan sp = split_plan(precision, op_a & (QEPW_OP_C | QEPW_OP_BC), m, n, k);
  out[0] = sp.variant;
  out[1] = sp.splits;
  out[2] = kVariants[sp.variant].tm;
  out[3] = kVariants[sp.variant].tn;
  return QEPW_OK;
}

// Products of commuting Hermitian matrices (the Newton-Schulz steps of the
// reduced eigensolve): C = alpha X Y + beta D with X, Y, D Hermitian
// polynomials of one matrix, so that C is Hermitian.  Only the lower-triangle
// 16x16 tiles are computed, one per workgroup, and each is written together
// with its conjugate transpose (the diagonal made real): C is exactly
// Hermitian, which the conjugated loads of Y below rely on.  Y's element
// (k, j) is read as conj(Y[j, k]), a contiguous run like X's (i, k).  The four
// waves of a workgroup split K, each fetching its next kHermUnroll MFMA
// steps while the current ones run, and their partial tiles are summed in LDS
// in wave order, so results are reproducible.  Two steps in flight measured
// fastest (n = 264 / 432 / 512: 7.1 / 15.2 / 21.4 us against 8.4 / 18.3 /
// 26.3 us with eight; four and six in between).
constexpr int kHermUnroll = 2;

__global__ __launch_bounds__(256) void k_herm_product(
    int n, const double2* __restrict__ X, const double2* __restrict__ Y,
    const double2* __restrict__ D, double2* __restrict__ C, int64_t ld, double alpha,
    double beta) {
  constexpr int U = kHermUnroll;
  __shared__ double2 part[4][256];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // tile (ti, tj), ti >= tj, of linear index b = ti (ti + 1) / 2 + tj
  const int b = blockIdx.x;
  int ti = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= b) ++ti;
  while (ti * (ti + 1) / 2 > b) --ti;
  const int tj = b - ti * (ti + 1) / 2;
  const int m0 = 16 * ti, n0 = 16 * tj;
  // this wave's K steps (4 columns of K each): [s0, s1)
  const int steps = (n + 3) >> 2;
  const int per = (steps + 3) >> 2;
  const int s0 = min(steps, wave * per);
  const int s1 = min(steps, s0 + per);
  const int r = lane & 15, kq = lane >> 4;
  const int row = m0 + r, col = n0 + r;
  const bool rok = row < n, cok = col < n;
  v4d re = {0.0, 0.0, 0.0, 0.0}, im = {0.0, 0.0, 0.0, 0.0};
  // operands are loaded unconditionally from an in-bounds element (the
  // matrix corner when out of range) and zeroed by register selects, which
  // keeps the loads on the global path and the operands in registers
  const double2* xr = X + (rok ? row : 0);
  const double2* yc = Y + (cok ? col : 0);
  double ar[U], ai[U], wr[U], wi[U], nar[U], nai[U], nwr[U], nwi[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int k = 4 * (s0 + u) + kq;
    const bool ok = s0 + u < s1 && k < n;
    const int64_t off = ok ? (int64_t)k * ld : 0;
    const double2 xa = xr[off], yw = yc[off];
    const bool ua = ok && rok, uw = ok && cok;
    ar[u] = ua ? xa.x : 0.0;
    ai[u] = ua ? xa.y : 0.0;
    wr[u] = uw ? yw.x : 0.0;
    wi[u] = uw ? yw.y : 0.0;
  }
  for (int s = s0; s < s1; s += U) {
    const int sn = s + U;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = 4 * (sn + u) + kq;
      const bool ok = sn + u < s1 && k < n;
      const int64_t off = ok ? (int64_t)k * ld : 0;
      const double2 xa = xr[off], yw = yc[off];
      const bool ua = ok && rok, uw = ok && cok;
      nar[u] = ua ? xa.x : 0.0;
      nai[u] = ua ? xa.y : 0.0;
      nwr[u] = uw ? yw.x : 0.0;
      nwi[u] = uw ? yw.y : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      // (X Y)_ij += X_ik conj(Y_jk)
      re = Mfma<double>::op(ar[u], wr[u], re);
      re = Mfma<double>::op(ai[u], wi[u], re);
      im = Mfma<double>::op(ai[u], wr[u], im);
      im = Mfma<double>::op(-ar[u], wi[u], im);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      ar[u] = nar[u];
      ai[u] = nai[u];
      wr[u] = nwr[u];
      wi[u] = nwi[u];
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v)
    part[wave][Mfma<double>::row(lane, v) + 16 * r] = make_double2(re[v], im[v]);
  __syncthreads();
  const int er = tid & 15, ec = tid >> 4;
  const int gi = m0 + er, gj = n0 + ec;
  if (gi >= n || gj >= n || gi < gj) return;
  double sr = 0.0, si = 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    sr += part[q][tid].x;
    si += part[q][tid].y;
  }
  double2 out = make_double2(alpha * sr, alpha * si);
  if (beta != 0.0) {
    const double2 d = D[gi + (int64_t)gj * ld];
    out.x += beta * d.x;
    out.y += beta * d.y;
  }
  if (gi == gj) {
    out.y = 0.0;
    C[gi + (int64_t)gj * ld] = out;
  } else {
    C[gi + (int64_t)gj * ld] = out;
    C[gj + (int64_t)gi * ld] = make_double2(out.x, -out.y);
  }
}

int herm_product(int n, const void* x, const void* y, const void* d, void* c, int64_t ld,
                 double alpha, double beta, hipStream_t stream) {
  if (n <= 0) return QEPW_OK;
  if (ld < n) return QEPW_ERR_ARG;
  const int nt = (n + 15) / 16;
  hipLaunchKernelGGL(k_herm_product, dim3((unsigned)(nt * (nt + 1) / 2)), dim3(256), 0,
                     stream, n, static_cast<const double2*>(x), static_cast<const double2*>(y),
                     static_cast<const double2*>(d), static_cast<double2*>(c), ld, alpha, beta);
  return hipGetLastError() == hipSuccess ? QEPW_OK : QEPW_ERR_HIP;
}

size_t gemm_workspace_bytes(int precision, int op, int64_t m, int64_t n, int64_t k) {
  const size_t elem = precision == QEPW_C128 ? sizeof(double2) : sizeof(float2);
  if (n == 1 && k > 0 && m > 0)
    return (size_t)gemv_plan((op & QEPW_OP_C) != 0, m, k).splits * (size_t)m * elem;
  const SplitPlan sp = split_plan(precision, op & (QEPW_OP_C | QEPW_OP_BC), m, n, k);
  if (sp.splits <= 1) return 0;
  const size_t c = precision == QEPW_C128 ? sizeof(double2) : sizeof(float2);
  return (size_t)sp.splits * (size_t)m * (size_t)n * c;
}

template <typename T>
static int gemv_t(int op, int64_t m, int64_t k, const double* alpha, const void* a, int64_t lda,
                  const void* b, int64_t ldb, const double* beta, void* c, int64_t ldc,
                  void* ws, size_t ws_bytes, hipStream_t stream) {
  const bool ca = (op & QEPW_OP_C) != 0, cb = (op & QEPW_OP_BC) != 0;
  const GemvPlan gp = gemv_plan(ca, m, k);
  if (ws_bytes < (size_t)gp.splits * (size_t)m * sizeof(cplx<T>)) return QEPW_ERR_WORKSPACE;
  // the single column of op(B): B[j] (k x 1), or conj(B[j ldb]) for its (1 x k) adjoint
  const int64_t bstride = cb ? ldb : 1;
  const cplx<T>* pa = static_cast<const cplx<T>*>(a);
  const cplx<T>* pb = static_cast<const cplx<T>*>(b);
  cplx<T>* part = static_cast<cplx<T>*>(ws);
  if (ca) {
    const dim3 grid((unsigned)m, (unsigned)gp.splits);
    if (cb)
      hipLaunchKernelGGL((k_gemv_c<T, true>), grid, dim3(256), 0, stream, m, k, gp.kchunk, pa,
                         lda, pb, bstride, part);
    else
      hipLaunchKernelGGL((k_gemv_c<T, false>), grid, dim3(256), 0, stream, m, k, gp.kchunk, pa,
                         lda, pb, bstride, part);
  } else {
    const dim3 grid((unsigned)((m + 255) / 256), (unsigned)gp.splits);
    if (cb)
      hipLaunchKernelGGL((k_gemv_n<T, true>), grid, dim3(256), 0, stream, m, k, gp.kchunk, pa,
                         lda, pb, bstride, part);
    else
      hipLaunchKernelGGL((k_gemv_n<T, false>), grid, dim3(256), 0, stream, m, k, gp.kchunk, pa,
                         lda, pb, bstride, part);
  }
  hipLaunchKernelGGL(k_reduce_splits<T>, dim3((unsigned)((m + 63) / 64)), dim3(256), 0, stream,
                     m, (int64_t)1, (int)gp.splits, part, static_cast<cplx<T>*>(c), ldc,
                     (T)alpha[0], (T)alpha[1], (T)beta[0], (T)beta[1]);
  return hipGetLastError() == hipSuccess ? QEPW_OK : QEPW_ERR_HIP;
}

template <typename T>
static int gemm_t(int op, int64_t m, int64_t n, int64_t k, const double* alpha,
                  const void* a, int64_t lda, const void* b, int64_t ldb, const double* beta,
                  void* c, int64_t ldc, void* ws, size_t ws_bytes, hipStream_t stream) {
  if (n == 1 && k > 0)
    return gemv_t<T>(op, m, k, alpha, a, lda, b, ldb, beta, c, ldc, ws, ws_bytes, stream);
  const SplitPlan sp = split_plan(std::is_same<T, double>::value ? QEPW_C128 : QEPW_C64,
                                  op & (QEPW_OP_C | QEPW_OP_BC), m, n, k);
  if (sp.splits > 65535) return QEPW_ERR_UNSUPPORTED;
  const size_t need = sp.splits > 1 ? (size_t)sp.splits * m * n * sizeof(cplx<T>) : 0;
  if (ws_bytes < need) return QEPW_ERR_WORKSPACE;
  const GemmVariant gv = kVariants[sp.variant];
  if ((m + gv.tm - 1) / gv.tm > 0x7fffffff || (n + gv.tn - 1) / gv.tn > 65535)
    return QEPW_ERR_UNSUPPORTED;
  const dim3 grid((unsigned)((m + gv.tm - 1) / gv.tm), (unsigned)((n + gv.tn - 1) / gv.tn),
                  (unsigned)sp.splits);
  const T ar = (T)alpha[0], ai = (T)alpha[1], br = (T)beta[0], bi = (T)beta[1];
  cplx<T>* part = static_cast<cplx<T>*>(ws);
  const cplx<T>* pa = static_cast<const cplx<T>*>(a);
  const cplx<T>* pb = static_cast<const cplx<T>*>(b);
  cplx<T>* pc = static_cast<cplx<T>*>(c);
#define QEPW_GEMM_TILE(CA, CB, TM, TN, WM)                                                \
  hipLaunchKernelGGL((k_cgemm<T, CA, CB, TM, TN, WM>), grid, dim3(kGemmThreads), 0, stream, \
                     m, n, k, sp.kchunk, pa, lda, pb, ldb, pc, ldc, part, ar, ai, br, bi)
#define QEPW_GEMM_LAUNCH(CA, CB)                                   \
  do {                                                             \
    switch (sp.variant) {                                          \
      case 0: QEPW_GEMM_TILE(CA, CB, 32, 32, 2); break;            \
      case 1: QEPW_GEMM_TILE(CA, CB, 64, 64, 2); break;            \
      case 2: QEPW_GEMM_TILE(CA, CB, 32, 128, 1); break;           \
      case 3: QEPW_GEMM_TILE(CA, CB, 64, 128, 2); break;           \
      case 4: QEPW_GEMM_TILE(CA, CB, 128, 64, 2); break;           \
      case 5: QEPW_GEMM_TILE(CA, CB, 128, 32, 4); break;           \
      default: QEPW_GEMM_TILE(CA, CB, 32, 64, 1); break;           \
    }                                                              \
  } while (0)
  const bool ca = (op & QEPW_OP_C) != 0, cb = (op & QEPW_OP_BC) != 0;
  if (ca && cb) QEPW_GEMM_LAUNCH(true, true);
  else if (ca) QEPW_GEMM_LAUNCH(true, false);
  else if (cb) QEPW_GEMM_LAUNCH(false, true);
  else QEPW_GEMM_LAUNCH(false, false);
#undef QEPW_GEMM_LAUNCH
#undef QEPW_GEMM_TILE
  if (sp.splits > 1) {
    const int64_t total = m * n;
    hipLaunchKernelGGL(k_reduce_splits<T>, dim3((unsigned)((total + 63) / 64)), dim3(256),
                       0, stream, m, n, (int)sp.splits, part, static_cast<cplx<T>*>(c), ldc,
                       ar, ai, br, bi);
  }
  return hipGetLastError() == hipSuccess ? QEPW_OK : QEPW_ERR_HIP;
}

int gemm(int precision, int op, int64_t m, int64_t n, int64_t k, const double* alpha,
         const void* a, int64_t lda, const void* b, int64_t ldb, const double* beta,
         void* c, int64_t ldc, void* ws, size_t ws_bytes, hipStream_t stream) {
  if (m < 0 || n < 0 || k < 0) return QEPW_ERR_ARG;
  if (op & ~(QEPW_OP_C | QEPW_OP_BC)) return QEPW_ERR_ARG;
  if (m == 0 || n == 0) return QEPW_OK;
  const bool ca = (op & QEPW_OP_C) != 0, cb = (op & QEPW_OP_BC) != 0;
  if (ldc < m || ldb < (cb ? n : (k > 0 ? k : 1))) return QEPW_ERR_ARG;
  if (!ca ? lda < m : lda < (k > 0 ? k : 1)) return QEPW_ERR_ARG;
  if (precision == QEPW_C128)
    return gemm_t<double>(op, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, ws, ws_bytes,
                          stream);
  if (precision == QEPW_C64)
    return gemm_t<float>(op, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, ws, ws_bytes,
                         stream);
  return QEPW_ERR_ARG;
}

}  // namespace qepw
