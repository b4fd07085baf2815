// Register-resident DFTs and the "team" line FFT used by the xy-plane kernel.
//
// RegDft<T, M, SIGN>: an M-point DFT (M 2,3,5-smooth, <= 16) held in one
// lane's registers, X[k] = sum_j v[j] exp(SIGN 2 pi i jk/M), computed by
// recursive decimation in time M = R*S:
//   A_j1[k1] = DFT_S over j2 of v[j1 + R j2];
//   X[k1 + S k2] = DFT_R over j1 of (w_M^(j1 k1) A_j1[k1]).
// Twiddles are compile-time constants (a constexpr sine/cosine evaluated by
// the compiler from their Taylor series after exact octant reduction).
//
// team_fft<T, N, SIGN>: one line of length N = L*M (M >= L, both <= 16) by a
// team of M lanes of one wave: lanes j1 < L take the M-point DFTs of the
// decimated sequences x[j1 + L j2], multiply by w_N^(j1 k1) and exchange
// through the team's LDS scratch; lanes k1 < M then take the L-point DFTs
// that give X[k1 + M k2].  The exchange needs no workgroup barrier: the team
// lives in one wave, whose LDS operations complete in program order.
#pragma once

#include "qepw_fft_lds.h"

namespace qepw {

// ---- compile-time unit roots -------------------------------------------
constexpr double kPiD = 3.14159265358979323846264338327950288;

constexpr double cx_sin_small(double x) {  // |x| <= pi/4
  double term = x, sum = x;
  for (int i = 1; i < 14; ++i) {
    term *= -x * x / ((2.0 * i) * (2.0 * i + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double cx_cos_small(double x) {  // |x| <= pi/4
  double term = 1.0, sum = 1.0;
  for (int i = 1; i < 14; ++i) {
    term *= -x * x / ((2.0 * i - 1.0) * (2.0 * i));
    sum += term;
  }
  return sum;
}
// cos / sin of 2 pi k / n with k reduced exactly into an octant.
constexpr double cx_cos2pi(int k, int n) {
  k %= n;
  if (k < 0) k += n;
  // theta = 2 pi k / n = (pi/4) * (8k/n); octant o = floor(8k/n)
  const int o = (8 * k) / n;
  const double r = 2.0 * kPiD * (double)(8 * k - o * n) / (8.0 * n);  // in [0, pi/4)
  switch (o) {
    case 0: return cx_cos_small(r);
    case 1: return cx_sin_small(kPiD / 4 - r);
    case 2: return -cx_sin_small(r);
    case 3: return -cx_cos_small(kPiD / 4 - r);
    case 4: return -cx_cos_small(r);
    case 5: return -cx_sin_small(kPiD / 4 - r);
    case 6: return cx_sin_small(r);
    default: return cx_cos_small(kPiD / 4 - r);
  }
}
constexpr double cx_sin2pi(int k, int n) {
  k %= n;
  if (k < 0) k += n;
  const int o = (8 * k) / n;
  const double r = 2.0 * kPiD * (double)(8 * k - o * n) / (8.0 * n);
  switch (o) {
    case 0: return cx_sin_small(r);
    case 1: return cx_cos_small(kPiD / 4 - r);
    case 2: return cx_cos_small(r);
    case 3: return cx_sin_small(kPiD / 4 - r);
    case 4: return -cx_sin_small(r);
    case 5: return -cx_cos_small(kPiD / 4 - r);
    case 6: return -cx_cos_small(r);
    default: return -cx_sin_small(kPiD / 4 - r);
  }
}

template <int M> struct UnitRoots {
  // w[k] = exp(+2 pi i k / M)
  double c[M > 0 ? M : 1];
  double s[M > 0 ? M : 1];
  constexpr UnitRoots() : c(), s() {
    for (int k = 0; k < M; ++k) {
      c[k] = cx_cos2pi(k, M);
      s[k] = cx_sin2pi(k, M);
    }
  }
};

constexpr int smallest_radix(int m) {
  return (m % 4 == 0) ? 4 : (m % 2 == 0) ? 2 : (m % 3 == 0) ? 3 : 5;
}

template <typename T, int M, int SIGN> struct RegDft {
  static __device__ __forceinline__ void run(cplx<T>* v) {
    constexpr int R = smallest_radix(M);
    constexpr int S = M / R;
    constexpr UnitRoots<M> w{};
    cplx<T> a[R][S];
#pragma unroll
    for (int j1 = 0; j1 < R; ++j1)
#pragma unroll
      for (int j2 = 0; j2 < S; ++j2) a[j1][j2] = v[j1 + R * j2];
#pragma unroll
    for (int j1 = 0; j1 < R; ++j1) RegDft<T, S, SIGN>::run(a[j1]);
#pragma unroll
    for (int k1 = 0; k1 < S; ++k1) {
      cplx<T> b[R];
#pragma unroll
      for (int j1 = 0; j1 < R; ++j1) {
        const int e = (j1 * k1) % M;
        if (e == 0) {
          b[j1] = a[j1][k1];
        } else {
          cplx<T> tw;
          tw.x = (T)w.c[e];
          tw.y = (T)(SIGN > 0 ? w.s[e] : -w.s[e]);
          b[j1] = cmul(a[j1][k1], tw);
        }
      }
      SmallDft<T, R>::run(b, SIGN);
#pragma unroll
      for (int k2 = 0; k2 < R; ++k2) v[k1 + S * k2] = b[k2];
    }
  }
};

template <typename T, int SIGN> struct RegDft<T, 1, SIGN> {
  static __device__ __forceinline__ void run(cplx<T>*) {}
};

constexpr int team_L(int n) {
  int best = 0;
  for (int l = 1; l <= 16; ++l)
    if (n % l == 0 && n / l <= 16 && n / l >= l) best = l;
  return best;
}

// Orders this wave's LDS accesses around a team exchange (no workgroup sync).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-lane twiddles of a team FFT, w[k] = w_N^(t k) for the lane's index t
// in its team: k < M for t < L (the decimation-in-time split below) and
// k < L for t < M (the decimation-in-frequency split); the largest index,
// (L-1)(M-1), stays inside the N-entry table.  They are the same for every
// line of a plane, so a kernel loads them into registers once.
template <typename T, int N> struct TeamTwiddles {
  static constexpr int L = team_L(N);
  static constexpr int M = N / L;
  cplx<T> w[M];
  __device__ __forceinline__ void load(int t, const cplx<T>* __restrict__ tw) {
#pragma unroll
    for (int k = 0; k < M; ++k)
      w[k] = (t < M && (t < L || k < L)) ? tw[t * k] : mk<T>(T(1), T(0));
  }
};

// ---- team FFT halves ------------------------------------------------------
// One N = L*M point line by the M lanes of a team, split at its exchange so
// that callers can chain transforms in registers.  Two factorisations:
//
//   DIT: lanes t < L hold inputs x[t + L s] (s < M);  dit_front takes the
//        M-point DFTs and twiddles and writes S[t*M + k];  after a wave sync
//        dit_back (lanes t < M) reads S[j*M + t] (j < L) and leaves
//        X[t + M k] (k < L) in the lane's registers.
//   DIF: lanes t < M hold inputs x[t + M s] (s < L) -- the DIT's output
//        order;  dif_front takes the L-point DFTs and twiddles and writes
//        S[k*M + t];  after a wave sync dif_back (lanes t < L) reads
//        S[t*M + j] (j < M) and leaves X[t + L k] (k < M) -- the DIT's
//        input order.
//
// Every lane's exchange writes land on exactly the slots that the same lane
// read in the other factorisation's back half, so a DIT and a DIF may follow
// each other (in either order) on one scratch line with no wave sync between
// them; two DITs (or two DIFs) in a row need one.  Twiddles by w^0 = 1 are
// exact, so they are applied to every lane without a branch.
template <typename T, int N, int SIGN>
__device__ __forceinline__ void dit_front(int t, cplx<T>* S, const TeamTwiddles<T, N>& tw,
                                          cplx<T> (&a)[N / team_L(N)]) {
  constexpr int L = team_L(N);
  constexpr int M = N / L;
  if (t < L) {
    RegDft<T, M, SIGN>::run(a);
#pragma unroll
    for (int k = 1; k < M; ++k) {
      a[k] = SIGN > 0 ? cmulc(a[k], tw.w[k]) : cmul(a[k], tw.w[k]);
    }
#pragma unroll
    for (int k = 0; k < M; ++k) S[t * M + k] = a[k];
  }
}

template <typename T, int N, int SIGN>
__device__ __forceinline__ void dit_back(int t, const cplx<T>* S, cplx<T> (&b)[team_L(N)]) {
  constexpr int L = team_L(N);
  constexpr int M = N / L;
  if (t < M) {
#pragma unroll
    for (int j = 0; j < L; ++j) b[j] = S[j * M + t];
    RegDft<T, L, SIGN>::run(b);
  }
}

template <typename T, int N, int SIGN>
__device__ __forceinline__ void dif_front(int t, cplx<T>* S, const TeamTwiddles<T, N>& tw,
                                          cplx<T> (&b)[team_L(N)]) {
  constexpr int L = team_L(N);
  constexpr int M = N / L;
  if (t < M) {
    RegDft<T, L, SIGN>::run(b);
#pragma unroll
    for (int k = 1; k < L; ++k) {
      b[k] = SIGN > 0 ? cmulc(b[k], tw.w[k]) : cmul(b[k], tw.w[k]);
    }
#pragma unroll
    for (int k = 0; k < L; ++k) S[k * M + t] = b[k];
  }
}

template <typename T, int N, int SIGN>
__device__ __forceinline__ void dif_back(int t, const cplx<T>* S,
                                         cplx<T> (&a)[N / team_L(N)]) {
  constexpr int L = team_L(N);
  constexpr int M = N / L;
  if (t < L) {
#pragma unroll
    for (int j = 0; j < M; ++j) a[j] = S[t * M + j];
    RegDft<T, M, SIGN>::run(a);
  }
}

// One N-point line by the M lanes of a team; ``t`` is the lane's index in
// the team; ``scratch`` holds N elements private to the team; ``tw`` holds
// the lane's twiddles.  load(j, s) gives input j = t + L s (s < M, lanes
// t < L); store(k, s, v) receives output k = t + M s (s < L, lanes t < M),
// each exactly once; the slot index s lets callers keep per-lane tables in
// registers.
template <typename T, int N, int SIGN, class Load, class Store>
__device__ __forceinline__ void team_fft(int t, cplx<T>* scratch, const TeamTwiddles<T, N>& tw,
                                         Load load, Store store) {
  constexpr int L = team_L(N);
  constexpr int M = N / L;
  static_assert(L >= 1 && M >= L && M <= 16, "unsupported team FFT length");
  if (t < L) {
    cplx<T> a[M];
#pragma unroll
    for (int j2 = 0; j2 < M; ++j2) a[j2] = load(t + L * j2, j2);
    RegDft<T, M, SIGN>::run(a);
#pragma unroll
    for (int k1 = 0; k1 < M; ++k1) {
      cplx<T> v = a[k1];
      if (k1 > 0 && t > 0) {
        cplx<T> w = tw.w[k1];
        if (SIGN > 0) w.y = -w.y;
        v = cmul(v, w);
      }
      scratch[t * M + k1] = v;
    }
  }
  wave_lds_sync();
  if (t < M) {
    cplx<T> b[L];
#pragma unroll
    for (int j1 = 0; j1 < L; ++j1) b[j1] = scratch[j1 * M + t];
    RegDft<T, L, SIGN>::run(b);
#pragma unroll
    for (int k2 = 0; k2 < L; ++k2) store(t + M * k2, k2, b[k2]);
  }
  wave_lds_sync();
}

// team_fft on a line held in LDS, with the line itself as the exchange
// buffer (every input is in registers before the first write): line[k] is
// replaced by X[k], k < N.  The team lives in one wave, so wave-level syncs
// order its LDS accesses.
template <typename T, int N, int SIGN>
__device__ __forceinline__ void team_fft_inplace(int t, cplx<T>* line,
                                                 const TeamTwiddles<T, N>& tw) {
  constexpr int L = team_L(N);
  constexpr int M = N / L;
  cplx<T> a[M];
  if (t < L) {
#pragma unroll
    for (int j2 = 0; j2 < M; ++j2) a[j2] = line[t + L * j2];
    RegDft<T, M, SIGN>::run(a);
#pragma unroll
    for (int k1 = 1; k1 < M; ++k1) {
      if (t > 0) {
        cplx<T> w = tw.w[k1];
        if (SIGN > 0) w.y = -w.y;
        a[k1] = cmul(a[k1], w);
      }
    }
  }
  wave_lds_sync();
  if (t < L) {
#pragma unroll
    for (int k1 = 0; k1 < M; ++k1) line[t * M + k1] = a[k1];
  }
  wave_lds_sync();
  cplx<T> b[L];
  if (t < M) {
#pragma unroll
    for (int j1 = 0; j1 < L; ++j1) b[j1] = line[j1 * M + t];
    RegDft<T, L, SIGN>::run(b);
  }
  wave_lds_sync();
  if (t < M) {
#pragma unroll
    for (int k2 = 0; k2 < L; ++k2) line[t + M * k2] = b[k2];
  }
  wave_lds_sync();
}

}  // namespace qepw
