// Sphere-pruned local-potential pipeline: V_loc(r) psi + T psi on gfx950.
//
// Per band the reference computes (qepy_pw/_native_fft.pyx:2338-2455,
// apply_serial; stick/plane form at :58-135, apply_serial_spatial):
//   grid = 0; grid[slot_G] = psi_G; backward 3-D FFT; grid *= V(r);
//   forward 3-D FFT; out_G = grid[slot_G] / N + T_G psi_G.
// Here the 3-D FFT is split into three launches over all bands at once:
//   A  k_sticks_inverse : gather the rows of ``group`` sticks (contiguous in
//                         the device order) into LDS lines, backward z-FFT,
//                         write the (n3, nst) stick buffer;
//   B  k_planes         : one workgroup per (z-plane, band): load the plane's
//                         active sticks into an n1 x n2 LDS plane, backward
//                         y-FFT on active x columns only, backward x-FFT,
//                         multiply by V(z, x, y), forward x-FFT, forward
//                         y-FFT on active columns, store the active sticks;
//   C  k_sticks_forward : forward z-FFT, gather the sphere rows, scale 1/N and
//                         add T_G psi_G.
// Only active sticks (19% of the 72^3 grid at 30 Ry) travel through HBM.
#include <algorithm>
#include <vector>

#include "qepw_internal.h"
#include "qepw_fft_lds.h"
#include "qepw_regfft.h"

namespace qepw {

constexpr int kStickPts = 16;

template <typename T>
__global__ __launch_bounds__(kStickThreads) void k_sticks_inverse(
    const cplx<T>* __restrict__ psi, int64_t ldpsi, cplx<T>* __restrict__ work,
    int64_t band_stride, int nst, int n3, const int* __restrict__ off,
    const int* __restrict__ gpos, int group, LinePlan pz,
    const cplx<T>* __restrict__ tw) {
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* L = reinterpret_cast<cplx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int s0 = blockIdx.x * group;
  const int ns = min(group, nst - s0);
  const int64_t b = blockIdx.y;
  const int len = ns * n3;
  const int lsd = n3 + 1;  // padded line stride: lines start on different LDS banks
  const cplx<T> zero = mk<T>(T(0), T(0));
  for (int i = tid; i < ns * lsd; i += kStickThreads) L[i] = zero;
  __syncthreads();
  const int d0 = off[s0], d1 = off[s0 + ns];
  const int shift = s0 * n3;
  const cplx<T>* src = psi + b * ldpsi;
  for (int d = d0 + tid; d < d1; d += kStickThreads) {
    const int q = gpos[d] - shift;  // l * n3 + z
    L[q + q / n3] = src[d];
  }
  __syncthreads();
  lds_fft<T, kStickThreads, kStickPts>(L, ns, lsd, 1, nullptr, pz, tw, +1, tid);
  cplx<T>* w = work + b * band_stride + s0;
  for (int i = tid; i < len; i += kStickThreads) {
    const int z = i / ns;
    const int l = i - z * ns;
    w[(int64_t)z * nst + l] = L[l * lsd + z];
  }
}

template <typename T>
__global__ __launch_bounds__(kStickThreads) void k_sticks_forward(
    const cplx<T>* __restrict__ work, int64_t band_stride, const cplx<T>* __restrict__ psi,
    int64_t ldpsi, const T* __restrict__ kinetic, cplx<T>* __restrict__ out,
    int64_t ldout, int nst, int n3, const int* __restrict__ off,
    const int* __restrict__ gpos, int group, LinePlan pz,
    const cplx<T>* __restrict__ tw, T inv_n) {
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* L = reinterpret_cast<cplx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int s0 = blockIdx.x * group;
  const int ns = min(group, nst - s0);
  const int64_t b = blockIdx.y;
  const int len = ns * n3;
  const int lsd = n3 + 1;  // padded line stride (see k_sticks_inverse)
  const cplx<T>* w = work + b * band_stride + s0;
  for (int i = tid; i < len; i += kStickThreads) {
    const int z = i / ns;
    const int l = i - z * ns;
    L[l * lsd + z] = w[(int64_t)z * nst + l];
  }
  __syncthreads();
  lds_fft<T, kStickThreads, kStickPts>(L, ns, lsd, 1, nullptr, pz, tw, -1, tid);
  const int d0 = off[s0], d1 = off[s0 + ns];
  const int shift = s0 * n3;
  const cplx<T>* src = psi + b * ldpsi;
  cplx<T>* dst = out + b * ldout;
  for (int d = d0 + tid; d < d1; d += kStickThreads) {
    const int q = gpos[d] - shift;
    cplx<T> v = cscale(L[q + q / n3], inv_n);
    if (kinetic) {
      const T k = kinetic[d];
      const cplx<T> p = src[d];
      v.x += k * p.x;
      v.y += k * p.y;
    }
    dst[d] = v;
  }
}

enum PlaneMode { kPlaneApply = 0, kPlaneReal = 1 };

template <typename T, int MODE>
__global__ __launch_bounds__(kPlaneThreads) void k_planes(
    cplx<T>* __restrict__ work, int64_t band_stride, int nst,
    const int* __restrict__ plane_of, const int* __restrict__ xcols, int nxcols,
    const T* __restrict__ V, T* __restrict__ real_out, int n1, int n2,
    LinePlan p1, LinePlan p2, const cplx<T>* __restrict__ tw1,
    const cplx<T>* __restrict__ tw2) {
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* P = reinterpret_cast<cplx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int z = blockIdx.x;
  const int64_t b = blockIdx.y;
  const int np = n1 * n2;
  const cplx<T> zero = mk<T>(T(0), T(0));
  cplx<T>* w = work + b * band_stride + (int64_t)z * nst;
  for (int i = tid; i < np; i += kPlaneThreads) P[i] = zero;
  __syncthreads();
  for (int s = tid; s < nst; s += kPlaneThreads) P[plane_of[s]] = w[s];
  __syncthreads();
  // backward along y (contiguous) for the active x columns
  lds_fft<T, kPlaneThreads, kPlanePoints>(P, nxcols, n2, 1, xcols, p2, tw2, +1, tid);
  // backward along x (stride n2) for every y row
  lds_fft<T, kPlaneThreads, kPlanePoints>(P, n2, 1, n2, nullptr, p1, tw1, +1, tid);
  // V and the real output use the (n3, n2, n1) layout: element (x, y) of
  // plane z at z*np + y*n1 + x; the LDS plane holds (x, y) at x*n2 + y.
  if (MODE == kPlaneReal) {
    T* o = real_out + (int64_t)z * np;
    for (int i = tid; i < np; i += kPlaneThreads) {
      const int y = i / n1, x = i - y * n1;
      o[i] = P[x * n2 + y].x;
    }
    return;
  }
  const T* v = V + (int64_t)z * np;
  for (int i = tid; i < np; i += kPlaneThreads) {
    const int x = i / n2, y = i - x * n2;
    P[i] = cscale(P[i], v[y * n1 + x]);
  }
  __syncthreads();
  lds_fft<T, kPlaneThreads, kPlanePoints>(P, n2, 1, n2, nullptr, p1, tw1, -1, tid);
  lds_fft<T, kPlaneThreads, kPlanePoints>(P, nxcols, n2, 1, xcols, p2, tw2, -1, tid);
  for (int s = tid; s < nst; s += kPlaneThreads) w[s] = P[plane_of[s]];
}

template <typename T> const cplx<T>* twiddles(const Plan& p, int axis);

// Split plane stage for planes that no single workgroup can hold (large
// FP64 grids): three passes through a per-band buffer T[z][c][y] of the
// active columns' full y lines (n3 x nxc x n2 elements per band):
//   k_split_cols_inverse : scatter each column's sticks into its y line,
//                          backward y-FFT, store the line;
//   k_split_rows         : gather row y from the columns, backward x-FFT,
//                          times V (or, for the real-space potential, store
//                          the real part and stop), forward x-FFT, scatter
//                          back to the columns;
//   k_split_cols_forward : forward y-FFT of each column, gather its sticks.
// Every 1-D FFT is the Stockham LDS FFT of a group of lines.
constexpr int kSplitThreads = 256;

template <typename T>
__global__ __launch_bounds__(kSplitThreads) void k_split_cols_inverse(
    const cplx<T>* __restrict__ work, int64_t band_stride, int nst,
    const int* __restrict__ colstart, const int* __restrict__ stick_q, int nxc, int n2,
    cplx<T>* __restrict__ tbuf, int64_t tband, int group, LinePlan p2,
    const cplx<T>* __restrict__ tw2) {
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* Ls = reinterpret_cast<cplx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * group;
  const int ng = min(group, nxc - c0);
  const int z = blockIdx.y;
  const int64_t b = blockIdx.z;
  const cplx<T> zero = mk<T>(T(0), T(0));
  for (int i = tid; i < ng * n2; i += kSplitThreads) Ls[i] = zero;
  __syncthreads();
  const cplx<T>* src = work + b * band_stride + (int64_t)z * nst;
  const int s0 = colstart[c0], s1 = colstart[c0 + ng];
  for (int st = s0 + tid; st < s1; st += kSplitThreads) {
    const int q = stick_q[st];  // c * n2 + y
    Ls[q - c0 * n2] = src[st];
  }
  __syncthreads();
  lds_fft<T, kSplitThreads, kPlanePoints>(Ls, ng, n2, 1, nullptr, p2, tw2, +1, tid);
  cplx<T>* dst = tbuf + b * tband + ((int64_t)z * nxc + c0) * n2;
  for (int i = tid; i < ng * n2; i += kSplitThreads) dst[i] = Ls[i];
}

template <typename T, int MODE>
__global__ __launch_bounds__(kSplitThreads) void k_split_rows(
    cplx<T>* __restrict__ tbuf, int64_t tband, const int* __restrict__ xcols, int nxc, int n1,
    int n2, const T* __restrict__ V, T* __restrict__ real_out, int group, LinePlan p1,
    const cplx<T>* __restrict__ tw1) {
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* Ls = reinterpret_cast<cplx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int y0 = blockIdx.x * group;
  const int nr = min(group, n2 - y0);
  const int z = blockIdx.y;
  const int64_t b = blockIdx.z;
  const cplx<T> zero = mk<T>(T(0), T(0));
  for (int i = tid; i < nr * n1; i += kSplitThreads) Ls[i] = zero;
  __syncthreads();
  cplx<T>* base = tbuf + b * tband + (int64_t)z * nxc * n2 + y0;
  for (int i = tid; i < nr * nxc; i += kSplitThreads) {
    const int r = i % nr, c = i / nr;
    Ls[r * n1 + xcols[c]] = base[(int64_t)c * n2 + r];
  }
  __syncthreads();
  lds_fft<T, kSplitThreads, kPlanePoints>(Ls, nr, n1, 1, nullptr, p1, tw1, +1, tid);
  // V and the real output use the (n3, n2, n1) layout: (x, y) of plane z at
  // (z * n2 + y) * n1 + x
  const int64_t row0 = ((int64_t)z * n2 + y0) * n1;
  if (MODE == kPlaneReal) {
    for (int i = tid; i < nr * n1; i += kSplitThreads) real_out[row0 + i] = Ls[i].x;
    return;
  }
  for (int i = tid; i < nr * n1; i += kSplitThreads) Ls[i] = cscale(Ls[i], V[row0 + i]);
  __syncthreads();
  lds_fft<T, kSplitThreads, kPlanePoints>(Ls, nr, n1, 1, nullptr, p1, tw1, -1, tid);
  for (int i = tid; i < nr * nxc; i += kSplitThreads) {
    const int r = i % nr, c = i / nr;
    base[(int64_t)c * n2 + r] = Ls[r * n1 + xcols[c]];
  }
}

template <typename T>
__global__ __launch_bounds__(kSplitThreads) void k_split_cols_forward(
    const cplx<T>* __restrict__ tbuf, int64_t tband, const int* __restrict__ colstart,
    const int* __restrict__ stick_q, int nxc, int n2, cplx<T>* __restrict__ work,
    int64_t band_stride, int nst, int group, LinePlan p2, const cplx<T>* __restrict__ tw2) {
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* Ls = reinterpret_cast<cplx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * group;
  const int ng = min(group, nxc - c0);
  const int z = blockIdx.y;
  const int64_t b = blockIdx.z;
  const cplx<T>* src = tbuf + b * tband + ((int64_t)z * nxc + c0) * n2;
  for (int i = tid; i < ng * n2; i += kSplitThreads) Ls[i] = src[i];
  __syncthreads();
  lds_fft<T, kSplitThreads, kPlanePoints>(Ls, ng, n2, 1, nullptr, p2, tw2, -1, tid);
  cplx<T>* dst = work + b * band_stride + (int64_t)z * nst;
  const int s0 = colstart[c0], s1 = colstart[c0 + ng];
  for (int st = s0 + tid; st < s1; st += kSplitThreads) {
    const int q = stick_q[st];
    dst[st] = Ls[q - c0 * n2];
  }
}

// Lines per workgroup of the split stage for lines of plan lp (0: none fits).
template <typename T>
static int split_group(const LinePlan& lp) {
  const int cap = lds_fft_capacity(lp, kSplitThreads, kPlanePoints) / lp.n;
  const int by_lds = kMaxLdsBytes / (lp.n * (int)sizeof(cplx<T>));
  int g = cap < by_lds ? cap : by_lds;
  if (g > 16) g = 16;
  return g > 0 ? g : 0;
}

template <typename T>
static bool split_available(const Plan& p, const StickTables& st) {
  return st.d_colstart && st.d_stickq && split_group<T>(p.line[0]) > 0 &&
         split_group<T>(p.line[1]) > 0 && p.n1 <= kMaxLine && p.n2 <= kMaxLine;
}

// One band's T buffer of the split stage, in elements.
static int64_t split_band_elems(const Plan& p, const StickTables& st) {
  return (int64_t)p.n3 * st.nxcols * p.n2;
}

template <typename T>
static void launch_split(bool real_mode, const Plan& p, const StickTables& st, cplx<T>* work,
                         int64_t bstride, cplx<T>* tbuf, const T* V, T* real_out, int nb,
                         hipStream_t stream) {
  const int gc = split_group<T>(p.line[1]);
  const int gr = split_group<T>(p.line[0]);
  const size_t lc = (size_t)gc * p.n2 * sizeof(cplx<T>);
  const size_t lr = (size_t)gr * p.n1 * sizeof(cplx<T>);
  const int64_t tband = split_band_elems(p, st);
  hipFuncSetAttribute(reinterpret_cast<const void*>(&k_split_cols_inverse<T>),
                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lc);
  hipFuncSetAttribute(reinterpret_cast<const void*>(&k_split_cols_forward<T>),
                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lc);
  hipFuncSetAttribute(reinterpret_cast<const void*>(&k_split_rows<T, kPlaneApply>),
                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lr);
  hipFuncSetAttribute(reinterpret_cast<const void*>(&k_split_rows<T, kPlaneReal>),
                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lr);
  const dim3 gcols((unsigned)((st.nxcols + gc - 1) / gc), (unsigned)p.n3, (unsigned)nb);
  const dim3 grows((unsigned)((p.n2 + gr - 1) / gr), (unsigned)p.n3, (unsigned)nb);
  hipLaunchKernelGGL(k_split_cols_inverse<T>, gcols, dim3(kSplitThreads), lc, stream, work,
                     bstride, st.nst, st.d_colstart, st.d_stickq, st.nxcols, p.n2, tbuf, tband,
                     gc, p.line[1], twiddles<T>(p, 1));
  if (real_mode) {
    hipLaunchKernelGGL((k_split_rows<T, kPlaneReal>), grows, dim3(kSplitThreads), lr, stream,
                       tbuf, tband, st.d_xcols, st.nxcols, p.n1, p.n2, V, real_out, gr,
                       p.line[0], twiddles<T>(p, 0));
    return;
  }
  hipLaunchKernelGGL((k_split_rows<T, kPlaneApply>), grows, dim3(kSplitThreads), lr, stream,
                     tbuf, tband, st.d_xcols, st.nxcols, p.n1, p.n2, V, real_out, gr, p.line[0],
                     twiddles<T>(p, 0));
  hipLaunchKernelGGL(k_split_cols_forward<T>, gcols, dim3(kSplitThreads), lc, stream, tbuf,
                     tband, st.d_colstart, st.d_stickq, st.nxcols, p.n2, work, bstride, st.nst,
                     gc, p.line[1], twiddles<T>(p, 1));
}

// Fast xy-plane kernel for square planes (n1 = n2 = N): the LDS holds only
// the nxc active x columns of the plane, Q[c * NP + y], plus one all-zero
// column; every 1-D FFT is a register "team" FFT (qepw_regfft.h).
//
// The backward transforms are decimation-in-frequency team FFTs, whose input
// slot s of lane t is position t + M s: the sphere's x and y extents leave
// whole slots empty for every lane (with the usual grids only s = 0, 1, L-2,
// L-1 hold sticks), so with KEEP naming the occupied slots the empty inputs
// are literal zeros that the compiler folds away (this file is built with
// -fno-signed-zeros for that), and the loads of them disappear.  The forward
// transforms are decimation-in-time team FFTs, whose output slot s of lane t
// is again position t + M s: only the kept slots are computed and stored.
// The row pass chains backward x-FFT (DIF), V(r) multiply and forward x-FFT
// (DIT) in registers -- the DIT consumes the DIF's output order -- at two
// team exchanges per row.
//
// LDS holds the plane and nothing else, so that more workgroups share a CU
// (FP64 at the 72^3 workload: 42 KB, three per CU): a column FFT exchanges
// through its own column (its values are in registers by then), and the row
// pass reads a batch of rows into registers, after which the plane's storage
// is free and each team exchanges through a line of it; the outputs wait in
// registers until every team is done, then go back to their rows.
constexpr int kTeamThreads = 256;
constexpr int kPlaneTeamThreads = 256;

// Column stride padding of the plane kernel's LDS plane (one element in
// both precisions: strides of N + 2, 3, 5 and 9 FP32 elements measured the
// same at the 72^3 workload).
template <typename T> constexpr int plane_pad() { return 1; }

template <int N> struct TeamGeom {
  static constexpr int L = team_L(N);
  static constexpr int M = N / L;
  static constexpr int per_wave = 64 / M;
  static constexpr int teams = (kTeamThreads / 64) * per_wave;
  // input slots kept by the pruned kernel: the first two and the last two
  static constexpr unsigned all = (1u << L) - 1u;
  static constexpr unsigned pruned = L >= 6 ? (0x3u | (0x3u << (L - 2))) : all;
  // the plane kernel's workgroup: kPlaneTeamThreads threads, plane_teams
  // teams -- enough for the usual sphere's active columns in one round of
  // each column pass and the rows in two rounds
  static constexpr int plane_teams = (kPlaneTeamThreads / 64) * per_wave;
  // rows a team holds in registers per batch of the plane kernel's row pass:
  // all of its rows when there are at most three, and then the row pass
  // exchanges through the plane's own storage; else one, with exchange lines
  // of their own after the plane
  static constexpr int rows = (N + plane_teams - 1) / plane_teams;
  static constexpr bool rows_in_place = rows <= 3;
  static constexpr int row_batch = rows_in_place ? rows : 1;
};

template <typename T, int N, bool PRUNE>
__device__ __forceinline__ void planes_team_body(
    int z, int64_t b, cplx<T>* __restrict__ work, int64_t band_stride, int nst,
    const int* __restrict__ colpos, const int* __restrict__ colmap, int nxc,
    const T* __restrict__ V, const cplx<T>* __restrict__ tw) {
  using G = TeamGeom<N>;
  constexpr int L = G::L, M = G::M, RB = G::row_batch;
  constexpr unsigned KEEP = PRUNE ? G::pruned : G::all;
  extern __shared__ unsigned char smem_raw[];
  // Q[c * NP + y]: active column c, padded stride NP = N + 1 so that the
  // columns start on different LDS banks.  Column nxc stays zero through the
  // column passes: the row pass reads the inactive x positions from it and
  // writes their (discarded) outputs to it.
  constexpr int NP = N + plane_pad<T>();
  cplx<T>* Q = reinterpret_cast<cplx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int team_in_wave = lane / M;
  const int t = lane - team_in_wave * M;
  const bool in_team = team_in_wave < G::per_wave;
  const int team = wave * G::per_wave + team_in_wave;
  const int nq = (nxc + 1) * NP;
  // the team's exchange line of the row pass: in the plane's storage when
  // one batch holds every row, else after it
  cplx<T>* E = Q + (G::rows_in_place ? 0 : nq) + team * NP;
  const cplx<T> zero = mk<T>(T(0), T(0));
  cplx<T>* w = work + b * band_stride + (int64_t)z * nst;

  // per-lane tables, the same for every row of the plane: twiddles, and the
  // Q offsets of this lane's row slots x = t + M s (lanes t < M, kept s),
  // which are both the backward inputs and the forward outputs
  TeamTwiddles<T, N> twr;
  twr.load(t, tw);
  int cq[L];
#pragma unroll
  for (int s = 0; s < L; ++s) {
    const int c = (t < M && ((KEEP >> s) & 1u)) ? colmap[t + M * s] : -1;
    cq[s] = (c < 0 ? nxc : c) * NP;
  }

  // the zero column (ordered before the row pass by the barrier after the
  // column pass, which does not touch it)
  for (int i = tid; i < NP; i += kPlaneTeamThreads) Q[nxc * NP + i] = zero;

  // backward y-FFT of every active column, its inputs read straight from the
  // stick buffer (colpos: the stick at (c, y), or -1), in place in LDS: the
  // column is the team's exchange line
  if (in_team) {
    for (int c = team; c < nxc; c += G::plane_teams) {
      cplx<T>* col = Q + c * NP;
      const int* pos = colpos + c * N;
      cplx<T> bb[L], a[M];
      if (t < M) {
        // every position first, then every stick load (from element 0 where
        // the slot holds no stick, replaced by zero): two memory round trips
        // per column instead of two per slot
        int st[L];
#pragma unroll
        for (int s = 0; s < L; ++s) st[s] = ((KEEP >> s) & 1u) ? pos[t + M * s] : -1;
#pragma unroll
        for (int s = 0; s < L; ++s) {
          bb[s] = zero;
          if ((KEEP >> s) & 1u) {
            const cplx<T> v = w[st[s] >= 0 ? st[s] : 0];
            bb[s] = st[s] >= 0 ? v : zero;
          }
        }
      }
      dif_front<T, N, +1>(t, col, twr, bb);
      wave_lds_sync();
      dif_back<T, N, +1>(t, col, a);
      wave_lds_sync();
      if (t < L) {
#pragma unroll
        for (int k = 0; k < M; ++k) col[t + L * k] = a[k];
      }
    }
  }
  __syncthreads();

  // rows, RB per team and batch: backward x-FFT (inputs only at active
  // columns), V, forward x-FFT (outputs only at active columns)
  const T* vz = V + (int64_t)z * N * N;
  for (int y0 = 0; y0 < N; y0 += RB * G::plane_teams) {
    cplx<T> rows[RB][L];
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int y = y0 + team + r * G::plane_teams;
      if (in_team && y < N && t < M) {
#pragma unroll
        for (int s = 0; s < L; ++s) rows[r][s] = ((KEEP >> s) & 1u) ? Q[cq[s] + y] : zero;
      }
    }
    __syncthreads();
    if (in_team) {
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const int y = y0 + team + r * G::plane_teams;
        if (y < N) {
          // V(r) at this lane's backward outputs x = t + L k, fetched first
          T vv[M];
#pragma unroll
          for (int k = 0; k < M; ++k) vv[k] = t < L ? vz[y * N + t + L * k] : T(0);
          cplx<T> a[M];
          // consecutive rows reuse the line: a DIF after a DIT writes only
          // the slots the same lane read last
          dif_front<T, N, +1>(t, E, twr, rows[r]);
          wave_lds_sync();
          dif_back<T, N, +1>(t, E, a);
          if (t < L) {
#pragma unroll
            for (int k = 0; k < M; ++k) a[k] = cscale(a[k], vv[k]);
          }
          dit_front<T, N, -1>(t, E, twr, a);
          wave_lds_sync();
          dit_back<T, N, -1>(t, E, rows[r]);
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int y = y0 + team + r * G::plane_teams;
      if (in_team && y < N && t < M) {
#pragma unroll
        for (int s = 0; s < L; ++s)
          if ((KEEP >> s) & 1u) Q[cq[s] + y] = rows[r][s];
      }
    }
    __syncthreads();
  }

  // forward y-FFT of every active column, in place; only the kept y slots,
  // which hold every stick, are computed, and the sticks among them go
  // straight to the stick buffer
  if (in_team) {
    for (int c = team; c < nxc; c += G::plane_teams) {
      cplx<T>* col = Q + c * NP;
      const int* pos = colpos + c * N;
      cplx<T> a[M], bb[L];
      // the stick positions are fetched before the transform, so that the
      // stores below wait for nothing
      int st[L];
#pragma unroll
      for (int k = 0; k < L; ++k) st[k] = ((KEEP >> k) & 1u) ? pos[t + M * k] : -1;
      if (t < L) {
#pragma unroll
        for (int s = 0; s < M; ++s) a[s] = col[t + L * s];
      }
      wave_lds_sync();
      dit_front<T, N, -1>(t, col, twr, a);
      wave_lds_sync();
      dit_back<T, N, -1>(t, col, bb);
      if (t < M) {
        // every position is consumed before the first store: one wait for
        // the loads above, none between the stores (stores count in vmcnt)
#pragma unroll
        for (int k = 0; k < L; ++k)
          if ((KEEP >> k) & 1u) asm volatile("" ::"v"(st[k]));
#pragma unroll
        for (int k = 0; k < L; ++k)
          if (((KEEP >> k) & 1u) && st[k] >= 0) w[st[k]] = bb[k];
      }
    }
  }
}

// at least three waves per SIMD: the FP64 plane's LDS admits three workgroups
// per CU, and its registers are held to that
template <typename T, int N, bool PRUNE>
__global__ __launch_bounds__(kPlaneTeamThreads, 3) void k_planes_team(
    cplx<T>* __restrict__ work, int64_t band_stride, int nst,
    const int* __restrict__ colpos, const int* __restrict__ colmap, int nxc,
    const T* __restrict__ V, const cplx<T>* __restrict__ tw) {
  planes_team_body<T, N, PRUNE>(blockIdx.x, blockIdx.y, work, band_stride, nst, colpos,
                                colmap, nxc, V, tw);
}

// Slots of the team FFT of length n that hold any active x column or stick y
// (bit s: positions t + M s, t < M); 0 when n has no team FFT.
unsigned team_slot_mask(int n, const std::vector<int>& xs, const std::vector<int>& ys) {
  const int l = team_L(n);
  if (l <= 0) return 0u;
  const int m = n / l;
  unsigned mask = 0u;
  for (int x : xs) mask |= 1u << (x / m);
  for (int y : ys) mask |= 1u << (y / m);
  return mask;
}

// z-stage kernels with register team FFTs for the lengths in QEPW_TEAM_SIZES:
// one line per team (a workgroup holds TeamGeom<N>::teams sticks), the line
// in LDS doubling as the team's exchange buffer; two workgroup barriers in
// all instead of two per radix pass.
template <typename T, int N>
__device__ __forceinline__ void sticks_inverse_body(
    int gx, int64_t b, const cplx<T>* __restrict__ psi, int64_t ldpsi,
    cplx<T>* __restrict__ work, int64_t band_stride, int nst, const int* __restrict__ off,
    const int* __restrict__ gpos, const cplx<T>* __restrict__ tw) {
  using G = TeamGeom<N>;
  constexpr int NP = N + 1;
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* Ls = reinterpret_cast<cplx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int team_in_wave = lane / G::M;
  const int t = lane - team_in_wave * G::M;
  const int team = wave * G::per_wave + team_in_wave;
  const int s0 = gx * G::teams;
  const int ns = min(G::teams, nst - s0);
  const cplx<T> zero = mk<T>(T(0), T(0));
  TeamTwiddles<T, N> twr;
  twr.load(t, tw);
  for (int i = tid; i < ns * NP; i += kTeamThreads) Ls[i] = zero;
  __syncthreads();
  const int d0 = off[s0], d1 = off[s0 + ns];
  const int shift = s0 * N;
  const cplx<T>* src = psi + b * ldpsi;
  for (int d = d0 + tid; d < d1; d += kTeamThreads) {
    const int q = gpos[d] - shift;  // l * N + z
    Ls[q + q / N] = src[d];
  }
  __syncthreads();
  if (team_in_wave < G::per_wave && team < ns)
    team_fft_inplace<T, N, +1>(t, Ls + team * NP, twr);
  __syncthreads();
  cplx<T>* w = work + b * band_stride + s0;
  for (int i = tid; i < ns * N; i += kTeamThreads) {
    const int z = i / ns;
    const int l = i - z * ns;
    w[(int64_t)z * nst + l] = Ls[l * NP + z];
  }
}

template <typename T, int N>
__global__ __launch_bounds__(kTeamThreads) void k_sticks_inverse_team(
    const cplx<T>* __restrict__ psi, int64_t ldpsi, cplx<T>* __restrict__ work,
    int64_t band_stride, int nst, const int* __restrict__ off,
    const int* __restrict__ gpos, const cplx<T>* __restrict__ tw) {
  sticks_inverse_body<T, N>(blockIdx.x, blockIdx.y, psi, ldpsi, work, band_stride, nst, off,
                            gpos, tw);
}

template <typename T, int N>
__device__ __forceinline__ void sticks_forward_body(
    int gx, int64_t b, const cplx<T>* __restrict__ work, int64_t band_stride,
    const cplx<T>* __restrict__ psi, int64_t ldpsi, const T* __restrict__ kinetic,
    cplx<T>* __restrict__ out, int64_t ldout, int nst, const int* __restrict__ off,
    const int* __restrict__ gpos, const cplx<T>* __restrict__ tw, T inv_n) {
  using G = TeamGeom<N>;
  constexpr int NP = N + 1;
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* Ls = reinterpret_cast<cplx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int team_in_wave = lane / G::M;
  const int t = lane - team_in_wave * G::M;
  const int team = wave * G::per_wave + team_in_wave;
  const int s0 = gx * G::teams;
  const int ns = min(G::teams, nst - s0);
  TeamTwiddles<T, N> twr;
  twr.load(t, tw);
  const cplx<T>* w = work + b * band_stride + s0;
  for (int i = tid; i < ns * N; i += kTeamThreads) {
    const int z = i / ns;
    const int l = i - z * ns;
    Ls[l * NP + z] = w[(int64_t)z * nst + l];
  }
  __syncthreads();
  if (team_in_wave < G::per_wave && team < ns)
    team_fft_inplace<T, N, -1>(t, Ls + team * NP, twr);
  __syncthreads();
  const int d0 = off[s0], d1 = off[s0 + ns];
  const int shift = s0 * N;
  const cplx<T>* src = psi + b * ldpsi;
  cplx<T>* dst = out + b * ldout;
  for (int d = d0 + tid; d < d1; d += kTeamThreads) {
    const int q = gpos[d] - shift;
    cplx<T> v = cscale(Ls[q + q / N], inv_n);
    if (kinetic) {
      const T k = kinetic[d];
      const cplx<T> p = src[d];
      v.x += k * p.x;
      v.y += k * p.y;
    }
    dst[d] = v;
  }
}

template <typename T, int N>
__global__ __launch_bounds__(kTeamThreads) void k_sticks_forward_team(
    const cplx<T>* __restrict__ work, int64_t band_stride, const cplx<T>* __restrict__ psi,
    int64_t ldpsi, const T* __restrict__ kinetic, cplx<T>* __restrict__ out, int64_t ldout,
    int nst, const int* __restrict__ off, const int* __restrict__ gpos,
    const cplx<T>* __restrict__ tw, T inv_n) {
  sticks_forward_body<T, N>(blockIdx.x, blockIdx.y, work, band_stride, psi, ldpsi, kinetic,
                            out, ldout, nst, off, gpos, tw, inv_n);
}

// ---- one launch per stage for many problems --------------------------------
// The h_psi FFT phase of up to kMaxFftGroup independent problems (k-points,
// cells: each its own plan, bands, potential and workspace, in one precision)
// runs as three launches in all.  Every problem's workgroups of a stage are
// laid end to end along x (start[p] .. start[p + 1]); a workgroup finds its
// problem from that prefix, then runs the same register team FFT body as the
// one-problem kernels for its grid length, chosen at run time among
// QEPW_GROUP_SIZES.
#define QEPW_GROUP_SIZES(X) X(36) X(48) X(54) X(60) X(64) X(72) X(80) X(96)
constexpr int kMaxFftGroup = 24;

template <typename T> struct StickGroupArgs {
  int count;
  int start[kMaxFftGroup + 1];
  int groups[kMaxFftGroup];        // workgroups per band
  int n[kMaxFftGroup];             // z length
  int nst[kMaxFftGroup];
  const cplx<T>* psi[kMaxFftGroup];
  int64_t ldpsi[kMaxFftGroup];
  cplx<T>* work[kMaxFftGroup];
  int64_t bstride[kMaxFftGroup];
  const int* off[kMaxFftGroup];
  const int* gpos[kMaxFftGroup];
  const cplx<T>* tw[kMaxFftGroup];
  const T* kin[kMaxFftGroup];       // forward stage
  cplx<T>* out[kMaxFftGroup];
  int64_t ldout[kMaxFftGroup];
  T inv_n[kMaxFftGroup];
};

template <typename T> struct PlaneGroupArgs {
  int count;
  int start[kMaxFftGroup + 1];
  int n[kMaxFftGroup];             // plane side (n1 = n2)
  int n3[kMaxFftGroup];
  int prune[kMaxFftGroup];
  int nst[kMaxFftGroup];
  int nxc[kMaxFftGroup];
  cplx<T>* work[kMaxFftGroup];
  int64_t bstride[kMaxFftGroup];
  const int* colpos[kMaxFftGroup];
  const int* colmap[kMaxFftGroup];
  const T* V[kMaxFftGroup];
  const cplx<T>* tw[kMaxFftGroup];
};

__device__ __forceinline__ int group_of(const int* start, int count, int bx) {
  int p = 0;
  while (p + 1 < count && start[p + 1] <= bx) ++p;
  return p;
}

template <typename T, bool FORWARD>
__global__ __launch_bounds__(kTeamThreads) void k_sticks_group(const StickGroupArgs<T> g) {
  const int p = group_of(g.start, g.count, blockIdx.x);
  const int l = blockIdx.x - g.start[p];
  const int gx = l % g.groups[p];
  const int64_t b = l / g.groups[p];
  switch (g.n[p]) {
#define QEPW_STICK_CASE(NN)                                                                     \
  case NN:                                                                                      \
    if (FORWARD)                                                                                \
      sticks_forward_body<T, NN>(gx, b, g.work[p], g.bstride[p], g.psi[p], g.ldpsi[p],          \
                                 g.kin[p], g.out[p], g.ldout[p], g.nst[p], g.off[p], g.gpos[p], \
                                 g.tw[p], g.inv_n[p]);                                          \
    else                                                                                        \
      sticks_inverse_body<T, NN>(gx, b, g.psi[p], g.ldpsi[p], g.work[p], g.bstride[p],          \
                                 g.nst[p], g.off[p], g.gpos[p], g.tw[p]);                       \
    break;
    QEPW_GROUP_SIZES(QEPW_STICK_CASE)
#undef QEPW_STICK_CASE
    default: break;
  }
}

template <typename T>
__global__ __launch_bounds__(kPlaneTeamThreads, 3) void k_planes_group(const PlaneGroupArgs<T> g) {
  const int p = group_of(g.start, g.count, blockIdx.x);
  const int l = blockIdx.x - g.start[p];
  const int z = l % g.n3[p];
  const int64_t b = l / g.n3[p];
  switch (g.n[p]) {
#define QEPW_PLANE_CASE(NN)                                                                     \
  case NN:                                                                                      \
    if (g.prune[p])                                                                             \
      planes_team_body<T, NN, true>(z, b, g.work[p], g.bstride[p], g.nst[p], g.colpos[p],       \
                                    g.colmap[p], g.nxc[p], g.V[p], g.tw[p]);                    \
    else                                                                                        \
      planes_team_body<T, NN, false>(z, b, g.work[p], g.bstride[p], g.nst[p], g.colpos[p],      \
                                     g.colmap[p], g.nxc[p], g.V[p], g.tw[p]);                   \
    break;
    QEPW_GROUP_SIZES(QEPW_PLANE_CASE)
#undef QEPW_PLANE_CASE
    default: break;
  }
}

// Square plane sizes with an instantiated fast kernel.
#define QEPW_TEAM_SIZES(X) \
  X(16) X(18) X(20) X(24) X(25) X(27) X(30) X(32) X(36) X(40) X(45) X(48) X(50) X(54) \
  X(60) X(64) X(72) X(75) X(80) X(81) X(90) X(96) X(100) X(108) X(120)

// Teams per workgroup of the team FFT of length n (0: no team kernel).
static int team_teams(int n) {
  int teams = 0;
#define QEPW_TEAMS_N(NN) if (n == NN) teams = TeamGeom<NN>::teams;
  QEPW_TEAM_SIZES(QEPW_TEAMS_N)
#undef QEPW_TEAMS_N
  return teams;
}

// Whether the pruned plane kernel (only the first two and last two slots
// occupied) serves a plan whose occupied slots are ``mask``.
static bool team_prunable(int n, unsigned mask) {
  bool ok = false;
#define QEPW_PRUNE_N(NN) if (n == NN) ok = (mask & ~TeamGeom<NN>::pruned) == 0u;
  QEPW_TEAM_SIZES(QEPW_PRUNE_N)
#undef QEPW_PRUNE_N
  return ok;
}

// LDS of the plane kernel (k_planes_team / k_planes_group): the active
// columns and the zero column, holding an exchange line per team when the
// row pass exchanges in place, else followed by them.
template <typename T>
static size_t plane_team_lds_bytes(int n, int nxc) {
  size_t teams = 0;
  bool in_place = false;
#define QEPW_TEAMS_OF(NN)                  \
  if (n == NN) {                           \
    teams = TeamGeom<NN>::plane_teams;     \
    in_place = TeamGeom<NN>::rows_in_place; \
  }
  QEPW_TEAM_SIZES(QEPW_TEAMS_OF)
#undef QEPW_TEAMS_OF
  if (teams == 0) return 0;
  const size_t lines = in_place ? std::max((size_t)nxc + 1, teams) : (size_t)nxc + 1 + teams;
  return lines * (size_t)(n + plane_pad<T>()) * sizeof(cplx<T>);
}

template <typename T>
static size_t team_lds_bytes(int n, int nxc) {
  size_t teams = 0;
#define QEPW_TEAMS_OF(NN) if (n == NN) teams = TeamGeom<NN>::teams;
  QEPW_TEAM_SIZES(QEPW_TEAMS_OF)
#undef QEPW_TEAMS_OF
  // no team kernel for this length: the generic kernels run
  if (teams == 0) return 0;
  // active columns, the zero column of k_planes_team, one scratch line per team
  return ((size_t)nxc + 1 + teams) * (size_t)(n + 1) * sizeof(cplx<T>);
}

// Launches the fast plane kernel if one exists for this plane; returns false
// otherwise (the generic kernel is used).
template <typename T>
static bool launch_planes_team(const Plan& p, const StickTables& st, cplx<T>* work,
                               int64_t bstride, const T* V, const cplx<T>* tw, int nb,
                               hipStream_t stream) {
  if (p.n1 != p.n2 || !st.d_colpos || !st.d_colmap) return false;
  const size_t lds = plane_team_lds_bytes<T>(p.n1, st.nxcols);
  if (lds == 0 || lds > (size_t)kMaxLdsBytes) return false;
  const dim3 grid(p.n3, nb);
#define QEPW_LAUNCH_TEAM_K(NN, PR)                                                           \
  do {                                                                                       \
    hipFuncSetAttribute(reinterpret_cast<const void*>(&k_planes_team<T, NN, PR>),           \
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);               \
    hipLaunchKernelGGL((k_planes_team<T, NN, PR>), grid, dim3(kPlaneTeamThreads), lds, stream, \
                       work, bstride, st.nst, st.d_colpos, st.d_colmap, st.nxcols, V, tw);  \
  } while (0)
#define QEPW_LAUNCH_TEAM(NN)                                                                 \
  if (p.n1 == NN) {                                                                          \
    if ((st.slot_mask & ~TeamGeom<NN>::pruned) == 0u)                                        \
      QEPW_LAUNCH_TEAM_K(NN, true);                                                          \
    else                                                                                     \
      QEPW_LAUNCH_TEAM_K(NN, false);                                                         \
    return true;                                                                             \
  }
  QEPW_TEAM_SIZES(QEPW_LAUNCH_TEAM)
#undef QEPW_LAUNCH_TEAM
#undef QEPW_LAUNCH_TEAM_K
  return false;
}

// Launches the team z-stage kernel for n3 if one exists; false otherwise.
template <typename T>
static bool launch_sticks_team(bool inverse, const Plan& p, const StickTables& st,
                               const cplx<T>* psi, int64_t ldpsi, cplx<T>* work,
                               int64_t bstride, const T* kinetic, cplx<T>* out, int64_t ldout,
                               const cplx<T>* tw, T inv_n, int nb, hipStream_t stream) {
#define QEPW_LAUNCH_STICKS(NN)                                                               \
  if (p.n3 == NN) {                                                                          \
    using G = TeamGeom<NN>;                                                                  \
    const size_t lds = (size_t)G::teams * (NN + 1) * sizeof(cplx<T>);                        \
    const dim3 grid((unsigned)((st.nst + G::teams - 1) / G::teams), (unsigned)nb);           \
    if (inverse)                                                                             \
      hipLaunchKernelGGL((k_sticks_inverse_team<T, NN>), grid, dim3(kTeamThreads), lds,      \
                         stream, psi, ldpsi, work, bstride, st.nst, st.d_off, st.d_gpos, tw); \
    else                                                                                     \
      hipLaunchKernelGGL((k_sticks_forward_team<T, NN>), grid, dim3(kTeamThreads), lds,      \
                         stream, work, bstride, psi, ldpsi, kinetic, out, ldout, st.nst,     \
                         st.d_off, st.d_gpos, tw, inv_n);                                    \
    return true;                                                                             \
  }
  QEPW_TEAM_SIZES(QEPW_LAUNCH_STICKS)
#undef QEPW_LAUNCH_STICKS
  return false;
}

template <typename T> const cplx<T>* twiddles(const Plan& p, int axis);
template <> const double2* twiddles<double>(const Plan& p, int axis) { return p.d_tw64[axis]; }
template <> const float2* twiddles<float>(const Plan& p, int axis) { return p.d_tw32[axis]; }

template <typename T>
static int check_fft_limits(const Plan& p, const StickTables& st) {
  const int np = p.n1 * p.n2;
  if (np * (int)sizeof(cplx<T>) > kMaxLdsBytes) return QEPW_ERR_UNSUPPORTED;
  if (np > lds_fft_capacity(p.line[0], kPlaneThreads, kPlanePoints)) return QEPW_ERR_UNSUPPORTED;
  if (np > lds_fft_capacity(p.line[1], kPlaneThreads, kPlanePoints)) return QEPW_ERR_UNSUPPORTED;
  if (st.group * p.n3 > lds_fft_capacity(p.line[2], kStickThreads, kStickPts))
    return QEPW_ERR_UNSUPPORTED;
  return QEPW_OK;
}

int stick_group_for(int n3, int nst) {
  int cap = (kStickThreads * 15) / (n3 > 0 ? n3 : 1);
  int g = cap < 32 ? cap : 32;
  if (g < 1) g = 1;
  if (g > nst && nst > 0) g = nst;
  return g;
}

template <typename T>
static void set_lds_attrs(size_t stick_bytes, size_t plane_bytes) {
  hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sticks_inverse<T>),
                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)stick_bytes);
  hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sticks_forward<T>),
                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)stick_bytes);
  if (plane_bytes == 0 || plane_bytes > (size_t)kMaxLdsBytes) return;
  hipFuncSetAttribute(reinterpret_cast<const void*>(&k_planes<T, kPlaneApply>),
                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)plane_bytes);
  hipFuncSetAttribute(reinterpret_cast<const void*>(&k_planes<T, kPlaneReal>),
                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)plane_bytes);
}

template <typename T>
static bool team_available(const Plan& p, const StickTables& st) {
  if (p.n1 != p.n2 || !st.d_colpos || !st.d_colmap) return false;
  const size_t lds = plane_team_lds_bytes<T>(p.n1, st.nxcols);
  return lds > 0 && lds <= (size_t)kMaxLdsBytes;
}

// How the xy-plane stage of a plan runs.
enum PlanePath { kPathNone = 0, kPathTeam, kPathGeneric, kPathSplit };

template <typename T>
static bool generic_plane_fits(const Plan& p) {
  const int np = p.n1 * p.n2;
  return np * (int)sizeof(cplx<T>) <= kMaxLdsBytes &&
         np <= lds_fft_capacity(p.line[0], kPlaneThreads, kPlanePoints) &&
         np <= lds_fft_capacity(p.line[1], kPlaneThreads, kPlanePoints);
}

template <typename T>
static int plane_path(const Plan& p, const StickTables& st) {
  if (st.group * p.n3 > lds_fft_capacity(p.line[2], kStickThreads, kStickPts)) return kPathNone;
  if (team_available<T>(p, st)) return kPathTeam;
  if (generic_plane_fits<T>(p)) return kPathGeneric;
  if (split_available<T>(p, st)) return kPathSplit;
  return kPathNone;
}

// Workspace of one band: the stick buffer, plus the split stage's column
// lines when that stage runs.
template <typename T>
static size_t band_bytes_t(const Plan& p, const StickTables& st) {
  size_t e = (size_t)p.n3 * (size_t)st.nst;
  if (plane_path<T>(p, st) == kPathSplit) e += (size_t)split_band_elems(p, st);
  return e * sizeof(cplx<T>);
}

size_t apply_local_band_bytes(const Plan& p, int precision) {
  return precision == QEPW_C128 ? band_bytes_t<double>(p, p.sphere)
                                : band_bytes_t<float>(p, p.sphere);
}

template <typename T>
int apply_local_t(const Plan& p, const void* vectors, int64_t ldv, int nb,
                  const void* potential, const void* kinetic, void* result,
                  int64_t ldr, void* workspace, size_t ws_bytes, hipStream_t stream) {
  const StickTables& st = p.sphere;
  const int path = plane_path<T>(p, st);
  if (path == kPathNone) return QEPW_ERR_UNSUPPORTED;
  if (nb <= 0) return QEPW_OK;
  const size_t band_bytes = band_bytes_t<T>(p, st);
  int chunk = (int)(ws_bytes / band_bytes);
  if (chunk < 1) return QEPW_ERR_WORKSPACE;
  if (chunk > 65535) chunk = 65535;
  const size_t stick_lds = (size_t)st.group * (p.n3 + 1) * sizeof(cplx<T>);
  const size_t plane_lds = (size_t)p.n1 * p.n2 * sizeof(cplx<T>);
  set_lds_attrs<T>(stick_lds, path == kPathGeneric ? plane_lds : 0);
  const int64_t bstride = (int64_t)p.n3 * st.nst;
  const T inv_n = T(1.0 / ((double)p.n1 * p.n2 * p.n3));
  cplx<T>* work = static_cast<cplx<T>*>(workspace);
  // the split stage's column lines follow the chunk's stick buffers
  cplx<T>* tbuf = work + (int64_t)chunk * bstride;
  for (int b0 = 0; b0 < nb; b0 += chunk) {
    const int nbc = (nb - b0) < chunk ? (nb - b0) : chunk;
    const cplx<T>* src = static_cast<const cplx<T>*>(vectors) + (int64_t)b0 * ldv;
    cplx<T>* dst = static_cast<cplx<T>*>(result) + (int64_t)b0 * ldr;
    dim3 gs(st.ngroups, nbc), gp(p.n3, nbc);
    if (!launch_sticks_team<T>(true, p, st, src, ldv, work, bstride, nullptr, nullptr, 0,
                               twiddles<T>(p, 2), inv_n, nbc, stream))
      hipLaunchKernelGGL(k_sticks_inverse<T>, gs, dim3(kStickThreads), stick_lds, stream,
                         src, ldv, work, bstride, st.nst, p.n3, st.d_off, st.d_gpos,
                         st.group, p.line[2], twiddles<T>(p, 2));
    if (path == kPathSplit)
      launch_split<T>(false, p, st, work, bstride, tbuf, static_cast<const T*>(potential),
                      nullptr, nbc, stream);
    else if (path != kPathTeam ||
             !launch_planes_team<T>(p, st, work, bstride, static_cast<const T*>(potential),
                                    twiddles<T>(p, 0), nbc, stream))
      hipLaunchKernelGGL((k_planes<T, kPlaneApply>), gp, dim3(kPlaneThreads), plane_lds,
                         stream, work, bstride, st.nst, st.d_plane, st.d_xcols, st.nxcols,
                         static_cast<const T*>(potential), static_cast<T*>(nullptr), p.n1,
                         p.n2, p.line[0], p.line[1], twiddles<T>(p, 0), twiddles<T>(p, 1));
    if (!launch_sticks_team<T>(false, p, st, src, ldv, work, bstride,
                               static_cast<const T*>(kinetic), dst, ldr, twiddles<T>(p, 2),
                               inv_n, nbc, stream))
      hipLaunchKernelGGL(k_sticks_forward<T>, gs, dim3(kStickThreads), stick_lds, stream,
                         work, bstride, src, ldv, static_cast<const T*>(kinetic), dst, ldr,
                         st.nst, p.n3, st.d_off, st.d_gpos, st.group, p.line[2],
                         twiddles<T>(p, 2), inv_n);
  }
  return hipGetLastError() == hipSuccess ? QEPW_OK : QEPW_ERR_HIP;
}

int apply_local(const Plan& p, int precision, const void* vectors, int64_t ldv, int nb,
                const void* potential, const void* kinetic, void* result, int64_t ldr,
                void* workspace, size_t ws_bytes, hipStream_t stream) {
  if (precision == QEPW_C128)
    return apply_local_t<double>(p, vectors, ldv, nb, potential, kinetic, result, ldr,
                                 workspace, ws_bytes, stream);
  if (precision == QEPW_C64)
    return apply_local_t<float>(p, vectors, ldv, nb, potential, kinetic, result, ldr,
                                workspace, ws_bytes, stream);
  return QEPW_ERR_ARG;
}

static bool group_size(int n) {
#define QEPW_GROUP_IS(NN) if (n == NN) return true;
  QEPW_GROUP_SIZES(QEPW_GROUP_IS)
#undef QEPW_GROUP_IS
  return false;
}

// Whether a problem's FFT phase can join a grouped launch: square planes and
// a z length among QEPW_GROUP_SIZES, the team plane path, its whole band
// block in one workspace chunk.
template <typename T>
static bool group_eligible(const Plan& p, int nb, size_t ws_bytes) {
  const StickTables& st = p.sphere;
  if (!(p.n1 == p.n2 && group_size(p.n1) && group_size(p.n3))) return false;
  if (plane_path<T>(p, st) != kPathTeam || !team_available<T>(p, st)) return false;
  return (size_t)nb * band_bytes_t<T>(p, st) <= ws_bytes && nb <= 65535;
}

template <typename T>
int apply_local_group_t(int count, const Plan* const* plans, const void* const* vectors,
                        const int64_t* ldv, const int* nb, const void* const* potentials,
                        const void* const* kinetics, void* const* results, const int64_t* ldr,
                        void* const* workspaces, const size_t* ws_bytes, hipStream_t stream) {
  // problems that cannot join run on their own; the rest in groups of at most
  // kMaxFftGroup, three launches per group
  std::vector<int> members;
  for (int i = 0; i < count; ++i) {
    if (nb[i] <= 0) continue;
    if (group_eligible<T>(*plans[i], nb[i], ws_bytes[i])) {
      members.push_back(i);
    } else {
      const int rc = apply_local_t<T>(*plans[i], vectors[i], ldv[i], nb[i], potentials[i],
                                      kinetics[i], results[i], ldr[i], workspaces[i],
                                      ws_bytes[i], stream);
      if (rc != QEPW_OK) return rc;
    }
  }
  for (size_t g0 = 0; g0 < members.size(); g0 += kMaxFftGroup) {
    const int m = (int)std::min<size_t>(kMaxFftGroup, members.size() - g0);
    StickGroupArgs<T> sa{};
    PlaneGroupArgs<T> pa{};
    sa.count = pa.count = m;
    size_t stick_lds = 0, plane_lds = 0;
    int64_t stotal = 0, ptotal = 0;
    for (int j = 0; j < m; ++j) {
      const int i = members[g0 + j];
      const Plan& p = *plans[i];
      const StickTables& st = p.sphere;
      const int teams_z = team_teams(p.n3);
      sa.start[j] = (int)stotal;
      sa.groups[j] = (st.nst + teams_z - 1) / teams_z;
      sa.n[j] = p.n3;
      sa.nst[j] = st.nst;
      sa.psi[j] = static_cast<const cplx<T>*>(vectors[i]);
      sa.ldpsi[j] = ldv[i];
      sa.work[j] = static_cast<cplx<T>*>(workspaces[i]);
      sa.bstride[j] = (int64_t)p.n3 * st.nst;
      sa.off[j] = st.d_off;
      sa.gpos[j] = st.d_gpos;
      sa.tw[j] = twiddles<T>(p, 2);
      sa.kin[j] = static_cast<const T*>(kinetics[i]);
      sa.out[j] = static_cast<cplx<T>*>(results[i]);
      sa.ldout[j] = ldr[i];
      sa.inv_n[j] = T(1.0 / ((double)p.n1 * p.n2 * p.n3));
      stotal += (int64_t)sa.groups[j] * nb[i];
      stick_lds = std::max(stick_lds, (size_t)teams_z * (p.n3 + 1) * sizeof(cplx<T>));
      pa.start[j] = (int)ptotal;
      pa.n[j] = p.n1;
      pa.n3[j] = p.n3;
      pa.prune[j] = team_prunable(p.n1, st.slot_mask) ? 1 : 0;
      pa.nst[j] = st.nst;
      pa.nxc[j] = st.nxcols;
      pa.work[j] = sa.work[j];
      pa.bstride[j] = sa.bstride[j];
      pa.colpos[j] = st.d_colpos;
      pa.colmap[j] = st.d_colmap;
      pa.V[j] = static_cast<const T*>(potentials[i]);
      pa.tw[j] = twiddles<T>(p, 0);
      ptotal += (int64_t)p.n3 * nb[i];
      plane_lds = std::max(plane_lds, plane_team_lds_bytes<T>(p.n1, st.nxcols));
    }
    sa.start[m] = (int)stotal;
    pa.start[m] = (int)ptotal;
    if (stotal > 0x7fffffff || ptotal > 0x7fffffff) return QEPW_ERR_UNSUPPORTED;
    hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sticks_group<T, false>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)stick_lds);
    hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sticks_group<T, true>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)stick_lds);
    hipFuncSetAttribute(reinterpret_cast<const void*>(&k_planes_group<T>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)plane_lds);
    hipLaunchKernelGGL((k_sticks_group<T, false>), dim3((unsigned)stotal), dim3(kTeamThreads),
                       stick_lds, stream, sa);
    hipLaunchKernelGGL(k_planes_group<T>, dim3((unsigned)ptotal), dim3(kPlaneTeamThreads),
                       plane_lds, stream, pa);
    hipLaunchKernelGGL((k_sticks_group<T, true>), dim3((unsigned)stotal), dim3(kTeamThreads),
                       stick_lds, stream, sa);
  }
  return hipGetLastError() == hipSuccess ? QEPW_OK : QEPW_ERR_HIP;
}

int apply_local_group(int count, const Plan* const* plans, int precision,
                      const void* const* vectors, const int64_t* ldv, const int* nb,
                      const void* const* potentials, const void* const* kinetics,
                      void* const* results, const int64_t* ldr, void* const* workspaces,
                      const size_t* ws_bytes, hipStream_t stream) {
  if (precision == QEPW_C128)
    return apply_local_group_t<double>(count, plans, vectors, ldv, nb, potentials, kinetics,
                                       results, ldr, workspaces, ws_bytes, stream);
  if (precision == QEPW_C64)
    return apply_local_group_t<float>(count, plans, vectors, ldv, nb, potentials, kinetics,
                                      results, ldr, workspaces, ws_bytes, stream);
  return QEPW_ERR_ARG;
}

// ---- density accumulation -------------------------------------------------
// rho(r) = scale * sum_b w_b |psi_b(r)|^2 with psi_b(r) the unnormalised
// backward FFT (reference: accumulate_density_serial, _native_fft.pyx:
// 2592-2648).  The backward z stage is apply_local's (the team kernel where
// n3 has one, k_sticks_inverse otherwise); the plane stage
// loops over a fixed slice of bands inside each workgroup and accumulates in
// LDS / registers, writing one partial plane per (slice, z); a last kernel
// sums the slices in a fixed order.  No atomics: results are reproducible.

template <typename T, int N, bool PRUNE>
__global__ __launch_bounds__(kTeamThreads) void k_density_team(
    const cplx<T>* __restrict__ work, int64_t band_stride, int nst,
    const int* __restrict__ colpos, const int* __restrict__ colmap, int nxc, int nb,
    int bands_per_slice, const double* __restrict__ weights, double* __restrict__ partial,
    int accumulate, const cplx<T>* __restrict__ tw) {
  using G = TeamGeom<N>;
  constexpr int L = G::L, M = G::M;
  constexpr unsigned KEEP = PRUNE ? G::pruned : G::all;
  // rows of a team: y = team + r * teams, r < R; each lane keeps the density
  // of its outputs x = t + L k of those rows in registers across the bands
  constexpr int R = (N + G::teams - 1) / G::teams;
  extern __shared__ unsigned char smem_raw[];
  constexpr int NP = N + 1;  // padded column stride and zero column, as in k_planes_team
  cplx<T>* Q = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* scratch_all = Q + (nxc + 1) * NP;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int team_in_wave = lane / M;
  const int t = lane - team_in_wave * M;
  const bool in_team = team_in_wave < G::per_wave;
  const int team = wave * G::per_wave + team_in_wave;
  cplx<T>* S = scratch_all + team * NP;
  const int z = blockIdx.x;
  const int slice = blockIdx.y;
  const int b0 = slice * bands_per_slice;
  const int b1 = min(nb, b0 + bands_per_slice);
  const cplx<T> zero = mk<T>(T(0), T(0));
  TeamTwiddles<T, N> twr;
  twr.load(t, tw);
  int cq[L];
#pragma unroll
  for (int s = 0; s < L; ++s) {
    const int c = (t < M && ((KEEP >> s) & 1u)) ? colmap[t + M * s] : -1;
    cq[s] = (c < 0 ? nxc : c) * NP;
  }
  double acc[R][M];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int k = 0; k < M; ++k) acc[r][k] = 0.0;
  // the zero column, read by the row pass at the inactive x positions (the
  // column pass below never writes it; its closing barrier orders this)
  for (int i = tid; i < NP; i += kTeamThreads) Q[nxc * NP + i] = zero;
  for (int b = b0; b < b1; ++b) {
    const cplx<T>* w = work + (int64_t)b * band_stride + (int64_t)z * nst;
    const double wb = weights[b];
    // backward y-FFT of every active column, its inputs read straight from
    // the stick buffer (colpos: the stick at (c, y), or -1; every position of
    // a column first, then every stick load), its outputs written to the
    // column
    if (in_team) {
      for (int c = team; c < nxc; c += G::teams) {
        if (c != team) wave_lds_sync();
        cplx<T>* col = Q + c * NP;
        const int* pos = colpos + c * N;
        cplx<T> bb[L], a[M];
        if (t < M) {
          int st[L];
#pragma unroll
          for (int s = 0; s < L; ++s) st[s] = ((KEEP >> s) & 1u) ? pos[t + M * s] : -1;
#pragma unroll
          for (int s = 0; s < L; ++s) {
            bb[s] = zero;
            if ((KEEP >> s) & 1u) {
              const cplx<T> v = w[st[s] >= 0 ? st[s] : 0];
              bb[s] = st[s] >= 0 ? v : zero;
            }
          }
        }
        dif_front<T, N, +1>(t, S, twr, bb);
        wave_lds_sync();
        dif_back<T, N, +1>(t, S, a);
        if (t < L) {
#pragma unroll
          for (int k = 0; k < M; ++k) col[t + L * k] = a[k];
        }
      }
    }
    __syncthreads();
    // backward x-FFT of this team's rows, |psi|^2 accumulated in registers
    if (in_team) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int y = team + r * G::teams;
        if (y < N) {
          if (r > 0) wave_lds_sync();
          cplx<T> bb[L], a[M];
          if (t < M) {
#pragma unroll
            for (int s = 0; s < L; ++s)
              bb[s] = ((KEEP >> s) & 1u) ? Q[cq[s] + y] : zero;
          }
          dif_front<T, N, +1>(t, S, twr, bb);
          wave_lds_sync();
          dif_back<T, N, +1>(t, S, a);
          if (t < L) {
#pragma unroll
            for (int k = 0; k < M; ++k)
              acc[r][k] += wb * ((double)a[k].x * (double)a[k].x +
                                 (double)a[k].y * (double)a[k].y);
          }
        }
      }
    }
    __syncthreads();
  }
  // (y, x) of the plane, (n3, n2, n1) layout: one writer per element
  double* out = partial + ((int64_t)slice * gridDim.x + z) * N * N;
  if (in_team && t < L) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int y = team + r * G::teams;
      if (y < N) {
#pragma unroll
        for (int k = 0; k < M; ++k) {
          const int i = y * N + t + L * k;
          out[i] = (accumulate ? out[i] : 0.0) + acc[r][k];
        }
      }
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kPlaneThreads) void k_density_plane(
    const cplx<T>* __restrict__ work, int64_t band_stride, int nst,
    const int* __restrict__ plane_of, const int* __restrict__ xcols, int nxcols, int nb,
    int bands_per_slice, const double* __restrict__ weights, double* __restrict__ partial,
    int accumulate, int n1, int n2, LinePlan p1, LinePlan p2,
    const cplx<T>* __restrict__ tw1, const cplx<T>* __restrict__ tw2) {
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* P = reinterpret_cast<cplx<T>*>(smem_raw);
  constexpr int kPer = kPlanePoints;  // plane points owned per thread (np <= NT * 15)
  const int tid = threadIdx.x;
  const int z = blockIdx.x;
  const int slice = blockIdx.y;
  const int b0 = slice * bands_per_slice;
  const int b1 = min(nb, b0 + bands_per_slice);
  const int np = n1 * n2;
  const cplx<T> zero = mk<T>(T(0), T(0));
  double acc[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) acc[k] = 0.0;
  for (int b = b0; b < b1; ++b) {
    const cplx<T>* w = work + (int64_t)b * band_stride + (int64_t)z * nst;
    const double wb = weights[b];
    for (int i = tid; i < np; i += kPlaneThreads) P[i] = zero;
    __syncthreads();
    for (int s = tid; s < nst; s += kPlaneThreads) P[plane_of[s]] = w[s];
    __syncthreads();
    lds_fft<T, kPlaneThreads, kPlanePoints>(P, nxcols, n2, 1, xcols, p2, tw2, +1, tid);
    lds_fft<T, kPlaneThreads, kPlanePoints>(P, n2, 1, n2, nullptr, p1, tw1, +1, tid);
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int i = tid + k * kPlaneThreads;
      if (i < np) {
        const cplx<T> v = P[i];
        acc[k] += wb * ((double)v.x * (double)v.x + (double)v.y * (double)v.y);
      }
    }
    __syncthreads();
  }
  double* out = partial + ((int64_t)slice * gridDim.x + z) * np;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int i = tid + k * kPlaneThreads;  // LDS plane index x * n2 + y
    if (i < np) {
      const int x = i / n2, y = i - x * n2;
      const int o = y * n1 + x;
      out[o] = (accumulate ? out[o] : 0.0) + acc[k];
    }
  }
}

__global__ void k_sum_slices(const double* __restrict__ partial, int slices, int64_t n,
                             double scale, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int k = 0; k < slices; ++k) s += partial[(int64_t)k * n + i];
  out[i] = scale * s;
}

// Density rows of the split stage: a workgroup owns rows y0..y0+nr of plane
// z and, for every band of its slice, gathers them from the column lines of
// k_split_cols_inverse, runs the backward x-FFT and adds w_b |psi(r)|^2 to
// its register accumulators; the sum over its bands is written (or added)
// once, so the result does not depend on how bands are scheduled.
template <typename T>
__global__ __launch_bounds__(kSplitThreads) void k_split_rows_density(
    const cplx<T>* __restrict__ tbuf, int64_t tband, const int* __restrict__ xcols, int nxc,
    int n1, int n2, int n3, int nb, int bands_per_slice, const double* __restrict__ weights,
    double* __restrict__ partial, int accumulate, int group, LinePlan p1,
    const cplx<T>* __restrict__ tw1) {
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* Ls = reinterpret_cast<cplx<T>*>(smem_raw);
  constexpr int kPer = kPlanePoints;  // row points owned per thread (nr n1 <= NT * kPer)
  const int tid = threadIdx.x;
  const int y0 = blockIdx.x * group;
  const int nr = min(group, n2 - y0);
  const int z = blockIdx.y;
  const int slice = blockIdx.z;
  const int b0 = slice * bands_per_slice;
  const int b1 = min(nb, b0 + bands_per_slice);
  const int npts = nr * n1;
  const cplx<T> zero = mk<T>(T(0), T(0));
  double acc[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) acc[k] = 0.0;
  for (int b = b0; b < b1; ++b) {
    const double wb = weights[b];
    for (int i = tid; i < npts; i += kSplitThreads) Ls[i] = zero;
    __syncthreads();
    const cplx<T>* base = tbuf + (int64_t)b * tband + (int64_t)z * nxc * n2 + y0;
    for (int i = tid; i < nr * nxc; i += kSplitThreads) {
      const int r = i % nr, c = i / nr;
      Ls[r * n1 + xcols[c]] = base[(int64_t)c * n2 + r];
    }
    __syncthreads();
    lds_fft<T, kSplitThreads, kPlanePoints>(Ls, nr, n1, 1, nullptr, p1, tw1, +1, tid);
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int i = tid + k * kSplitThreads;
      if (i < npts) {
        const cplx<T> v = Ls[i];
        acc[k] += wb * ((double)v.x * (double)v.x + (double)v.y * (double)v.y);
      }
    }
    __syncthreads();
  }
  // partial[slice] is an (n3, n2, n1) grid; Ls index r * n1 + x is row y0 + r
  double* out = partial + ((int64_t)slice * n3 + z) * ((int64_t)n1 * n2) + (int64_t)y0 * n1;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int i = tid + k * kSplitThreads;
    if (i < npts) out[i] = (accumulate ? out[i] : 0.0) + acc[k];
  }
}

constexpr int kDensitySlices = 16;

size_t density_workspace_bytes(const Plan& p, int precision, int nb) {
  const size_t ngrid = (size_t)p.n1 * p.n2 * p.n3;
  return apply_local_band_bytes(p, precision) * (size_t)(nb > 0 ? nb : 1) +
         (size_t)kDensitySlices * ngrid * sizeof(double) + 256;
}

template <typename T>
static size_t density_team_lds(int n, int nxc) {
  // the density accumulates in registers: the plane kernel's LDS
  return team_lds_bytes<T>(n, nxc);
}

template <typename T>
int accumulate_density_t(const Plan& p, const void* vectors, int64_t ldv, int nb,
                         const double* weights, double scale, double* out, void* workspace,
                         size_t ws_bytes, hipStream_t stream) {
  const StickTables& st = p.sphere;
  const size_t ngrid = (size_t)p.n1 * p.n2 * p.n3;
  const size_t np = (size_t)p.n1 * p.n2;
  if (st.group * p.n3 > lds_fft_capacity(p.line[2], kStickThreads, kStickPts))
    return QEPW_ERR_UNSUPPORTED;
  const size_t team_lds =
      (p.n1 == p.n2 && st.d_colpos && st.d_colmap) ? density_team_lds<T>(p.n1, st.nxcols) : 0;
  const bool team = team_lds > 0 && team_lds <= (size_t)kMaxLdsBytes;
  // planes that neither the team kernel nor one workgroup's LDS can hold go
  // through the split stage's column lines, as in apply_local
  const bool split = !team && !generic_plane_fits<T>(p) && split_available<T>(p, st);
  if (!team && !split) {
    int rc = check_fft_limits<T>(p, st);
    if (rc) return rc;
  }
  const size_t partial_bytes = (size_t)kDensitySlices * ngrid * sizeof(double);
  const size_t stick_band = (size_t)p.n3 * st.nst * sizeof(cplx<T>);
  const int64_t tband = split ? split_band_elems(p, st) : 0;
  const size_t band_bytes = stick_band + (size_t)tband * sizeof(cplx<T>);
  if (ws_bytes < partial_bytes + band_bytes) return QEPW_ERR_WORKSPACE;
  double* partial = static_cast<double*>(workspace);
  cplx<T>* work = reinterpret_cast<cplx<T>*>(static_cast<char*>(workspace) + partial_bytes);
  int chunk = (int)((ws_bytes - partial_bytes) / band_bytes);
  if (chunk > 65535) chunk = 65535;
  // the split stage's column lines follow the chunk's stick buffers
  cplx<T>* tbuf = work + (int64_t)chunk * p.n3 * st.nst;
  const size_t stick_lds = (size_t)st.group * (p.n3 + 1) * sizeof(cplx<T>);
  const size_t plane_lds = np * sizeof(cplx<T>);
  set_lds_attrs<T>(stick_lds, 0);
  if (!team && !split)
    hipFuncSetAttribute(reinterpret_cast<const void*>(&k_density_plane<T>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)plane_lds);
  const int gc = split ? split_group<T>(p.line[1]) : 0;
  const int gr = split ? split_group<T>(p.line[0]) : 0;
  const size_t lc = (size_t)gc * p.n2 * sizeof(cplx<T>);
  const size_t lr = (size_t)gr * p.n1 * sizeof(cplx<T>);
  if (split) {
    hipFuncSetAttribute(reinterpret_cast<const void*>(&k_split_cols_inverse<T>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lc);
    hipFuncSetAttribute(reinterpret_cast<const void*>(&k_split_rows_density<T>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lr);
  }
  const int64_t bstride = (int64_t)p.n3 * st.nst;
  if (nb <= 0) {
    hipLaunchKernelGGL(k_sum_slices, dim3((unsigned)((ngrid + 255) / 256)), dim3(256), 0,
                       stream, partial, 0, (int64_t)ngrid, scale, out);
    return hipGetLastError() == hipSuccess ? QEPW_OK : QEPW_ERR_HIP;
  }
  for (int b0 = 0; b0 < nb; b0 += chunk) {
    const int nbc = (nb - b0) < chunk ? (nb - b0) : chunk;
    const cplx<T>* src = static_cast<const cplx<T>*>(vectors) + (int64_t)b0 * ldv;
    // the z stage of apply_local: the team kernel where n3 has one
    if (!launch_sticks_team<T>(true, p, st, src, ldv, work, bstride, nullptr, nullptr, 0,
                               twiddles<T>(p, 2), T(0), nbc, stream))
      hipLaunchKernelGGL(k_sticks_inverse<T>, dim3(st.ngroups, nbc), dim3(kStickThreads),
                         stick_lds, stream, src, ldv, work, bstride, st.nst, p.n3, st.d_off,
                         st.d_gpos, st.group, p.line[2], twiddles<T>(p, 2));
    const int per = (nbc + kDensitySlices - 1) / kDensitySlices;
    const dim3 grid(p.n3, kDensitySlices);
    const int acc = b0 > 0 ? 1 : 0;
    bool launched = false;
    if (team) {
#define QEPW_LAUNCH_DENSITY_K(NN, PR)                                                        \
  do {                                                                                       \
    hipFuncSetAttribute(reinterpret_cast<const void*>(&k_density_team<T, NN, PR>),          \
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)team_lds);          \
    hipLaunchKernelGGL((k_density_team<T, NN, PR>), grid, dim3(kTeamThreads), team_lds,     \
                       stream, work, bstride, st.nst, st.d_colpos, st.d_colmap, st.nxcols,  \
                       nbc, per, weights + b0, partial, acc, twiddles<T>(p, 0));             \
  } while (0)
#define QEPW_LAUNCH_DENSITY(NN)                                                              \
  if (!launched && p.n1 == NN) {                                                             \
    if ((st.slot_mask & ~TeamGeom<NN>::pruned) == 0u)                                        \
      QEPW_LAUNCH_DENSITY_K(NN, true);                                                       \
    else                                                                                     \
      QEPW_LAUNCH_DENSITY_K(NN, false);                                                      \
    launched = true;                                                                         \
  }
      QEPW_TEAM_SIZES(QEPW_LAUNCH_DENSITY)
#undef QEPW_LAUNCH_DENSITY
#undef QEPW_LAUNCH_DENSITY_K
    }
    if (!launched && split) {
      const dim3 gcols((unsigned)((st.nxcols + gc - 1) / gc), (unsigned)p.n3, (unsigned)nbc);
      hipLaunchKernelGGL(k_split_cols_inverse<T>, gcols, dim3(kSplitThreads), lc, stream, work,
                         bstride, st.nst, st.d_colstart, st.d_stickq, st.nxcols, p.n2, tbuf,
                         tband, gc, p.line[1], twiddles<T>(p, 1));
      const dim3 grows((unsigned)((p.n2 + gr - 1) / gr), (unsigned)p.n3, kDensitySlices);
      hipLaunchKernelGGL(k_split_rows_density<T>, grows, dim3(kSplitThreads), lr, stream, tbuf,
                         tband, st.d_xcols, st.nxcols, p.n1, p.n2, p.n3, nbc, per, weights + b0,
                         partial, acc, gr, p.line[0], twiddles<T>(p, 0));
      launched = true;
    }
    if (!launched)
      hipLaunchKernelGGL(k_density_plane<T>, grid, dim3(kPlaneThreads), plane_lds, stream, work,
                         bstride, st.nst, st.d_plane, st.d_xcols, st.nxcols, nbc, per,
                         weights + b0, partial, acc, p.n1, p.n2, p.line[0], p.line[1],
                         twiddles<T>(p, 0), twiddles<T>(p, 1));
  }
  hipLaunchKernelGGL(k_sum_slices, dim3((unsigned)((ngrid + 255) / 256)), dim3(256), 0, stream,
                     partial, kDensitySlices, (int64_t)ngrid, scale, out);
  return hipGetLastError() == hipSuccess ? QEPW_OK : QEPW_ERR_HIP;
}

int accumulate_density(const Plan& p, int precision, const void* vectors, int64_t ldv, int nb,
                       const double* weights, double scale, double* out, void* workspace,
                       size_t ws_bytes, hipStream_t stream) {
  if (precision == QEPW_C128)
    return accumulate_density_t<double>(p, vectors, ldv, nb, weights, scale, out, workspace,
                                        ws_bytes, stream);
  if (precision == QEPW_C64)
    return accumulate_density_t<float>(p, vectors, ldv, nb, weights, scale, out, workspace,
                                       ws_bytes, stream);
  return QEPW_ERR_ARG;
}

size_t inverse_real_bytes(const Plan& p) {
  const size_t grid = (size_t)p.n1 * p.n2 * p.n3 * sizeof(double2);
  // the split stage adds one line buffer of the whole grid
  return plane_path<double>(p, p.full) == kPathSplit ? 2 * grid : grid;
}

// V(r) = Re backward-FFT(V(G)) of a C-order (n1, n2, n3) complex grid, written
// in the (n3, n2, n1) layout the plane kernels read
// (reference: _native_fft.pyx:2188-2202 inverse_real; basis.py:777-792).
int inverse_real(const Plan& p, const void* grid_g, void* out_real, void* workspace,
                 size_t ws_bytes, hipStream_t stream) {
  const StickTables& st = p.full;
  const int path = plane_path<double>(p, st);
  if (path == kPathNone) return QEPW_ERR_UNSUPPORTED;
  if (ws_bytes < inverse_real_bytes(p)) return QEPW_ERR_WORKSPACE;
  const size_t stick_lds = (size_t)st.group * (p.n3 + 1) * sizeof(double2);
  const size_t plane_lds = (size_t)p.n1 * p.n2 * sizeof(double2);
  set_lds_attrs<double>(stick_lds, path == kPathSplit ? 0 : plane_lds);
  const int64_t bstride = (int64_t)p.n3 * st.nst;
  double2* work = static_cast<double2*>(workspace);
  hipLaunchKernelGGL(k_sticks_inverse<double>, dim3(st.ngroups, 1), dim3(kStickThreads),
                     stick_lds, stream, static_cast<const double2*>(grid_g), (int64_t)0, work,
                     bstride, st.nst, p.n3, st.d_off, st.d_gpos, st.group, p.line[2],
                     twiddles<double>(p, 2));
  if (path == kPathSplit)
    launch_split<double>(true, p, st, work, bstride, work + bstride, nullptr,
                         static_cast<double*>(out_real), 1, stream);
  else
    hipLaunchKernelGGL((k_planes<double, kPlaneReal>), dim3(p.n3, 1), dim3(kPlaneThreads),
                       plane_lds, stream, work, bstride, st.nst, st.d_plane, st.d_xcols,
                       st.nxcols, static_cast<const double*>(nullptr),
                       static_cast<double*>(out_real), p.n1, p.n2, p.line[0], p.line[1],
                       twiddles<double>(p, 0), twiddles<double>(p, 1));
  return hipGetLastError() == hipSuccess ? QEPW_OK : QEPW_ERR_HIP;
}

}  // namespace qepw
