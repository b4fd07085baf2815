// Workgroup-cooperative mixed-radix FFT of many lines held in LDS.
//
// Method: the Stockham autosort formulation of Govindaraju, Lloyd, Dotsenko,
// Smith and Manferdelli, "High Performance Discrete Fourier Transforms on
// Graphics Processors" (SC'08), generalised to mixed radices 2/3/4/5:
//   for each stage of radix R (Ns = product of earlier radices):
//     butterfly j in [0, n/R): read v_r = x[j + r n/R], twiddle
//     v_r *= w_n^(r (j mod Ns) n/(Ns R)), R-point DFT, write
//     y[(j/Ns) Ns R + (j mod Ns) + r Ns] = v_r.
// Output is in natural order.  The transform runs in place in LDS: every
// thread first reads all of its butterflies into registers, the workgroup
// synchronises, then every thread writes.  ``PTS`` bounds the complex points
// a thread holds per stage; the host checks that lines*n <= NT*R*(PTS/R) for
// every stage before a launch.
//
// sign = -1 is the forward transform (exp(-2 pi i jk/n)), +1 the backward
// one; neither is normalised.  ``tw`` holds w_n^k = exp(-2 pi i k/n),
// k in [0, n), computed on the host in long double.
#pragma once

#include "qepw_common.h"

namespace qepw {

template <typename T, int R> struct SmallDft;

template <typename T> struct SmallDft<T, 2> {
  static __device__ __forceinline__ void run(cplx<T> (&v)[2], int) {
    cplx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b); v[1] = csub(a, b);
  }
};

template <typename T> struct SmallDft<T, 3> {
  static __device__ __forceinline__ void run(cplx<T> (&v)[3], int sign) {
    const T h = T(0.866025403784438646763723170752936183);  // sqrt(3)/2
    cplx<T> t1 = cadd(v[1], v[2]);
    cplx<T> t2 = csub(v[0], cscale(t1, T(0.5)));
    cplx<T> t3 = cmuli(cscale(csub(v[1], v[2]), h), sign);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, t3);
    v[2] = csub(t2, t3);
  }
};

template <typename T> struct SmallDft<T, 4> {
  static __device__ __forceinline__ void run(cplx<T> (&v)[4], int sign) {
    cplx<T> s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    cplx<T> s13 = cadd(v[1], v[3]), d13 = cmuli(csub(v[1], v[3]), sign);
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
  }
};

template <typename T> struct SmallDft<T, 5> {
  static __device__ __forceinline__ void run(cplx<T> (&v)[5], int sign) {
    const T c1 = T(0.309016994374947424102293417182819059);   // cos(2pi/5)
    const T c2 = T(-0.809016994374947424102293417182819059);  // cos(4pi/5)
    const T s1 = T(0.951056516295153572116439333379382143);   // sin(2pi/5)
    const T s2 = T(0.587785252292473129168705954639072769);   // sin(4pi/5)
    cplx<T> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    cplx<T> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    cplx<T> r1 = cadd(v[0], cadd(cscale(a1, c1), cscale(a2, c2)));
    cplx<T> r2 = cadd(v[0], cadd(cscale(a1, c2), cscale(a2, c1)));
    cplx<T> i1 = cmuli(cadd(cscale(b1, s1), cscale(b2, s2)), sign);
    cplx<T> i2 = cmuli(csub(cscale(b1, s2), cscale(b2, s1)), sign);
    v[0] = cadd(v[0], cadd(a1, a2));
    v[1] = cadd(r1, i1);
    v[4] = csub(r1, i1);
    v[2] = cadd(r2, i2);
    v[3] = csub(r2, i2);
  }
};

template <typename T, int R, int NT, int PTS>
__device__ __forceinline__ void lds_stage(cplx<T>* buf, int nlines, int line_stride,
                                          int elem_stride, const int* __restrict__ line_map,
                                          int n, int ns,
                                          const cplx<T>* __restrict__ tw, int sign,
                                          int tid) {
  constexpr int MAXB = (PTS / R) > 0 ? (PTS / R) : 1;
  const int nbf = n / R;
  const int total = nlines * nbf;
  const int twstep = n / (ns * R);
  cplx<T> v[MAXB][R];
  int out_base[MAXB];
#pragma unroll
  for (int k = 0; k < MAXB; ++k) {
    const int t = tid + k * NT;
    if (t < total) {
      const int line = t / nbf;
      const int j = t - line * nbf;
      const int jm = j % ns;
      const int base = (line_map ? line_map[line] : line) * line_stride;
#pragma unroll
      for (int r = 0; r < R; ++r) v[k][r] = buf[base + (j + r * nbf) * elem_stride];
      if (ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cplx<T> w = tw[jm * r * twstep];
          if (sign > 0) w.y = -w.y;
          v[k][r] = cmul(v[k][r], w);
        }
      }
      SmallDft<T, R>::run(v[k], sign);
      out_base[k] = base + ((j / ns) * ns * R + jm) * elem_stride;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < MAXB; ++k) {
    const int t = tid + k * NT;
    if (t < total) {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[out_base[k] + r * ns * elem_stride] = v[k][r];
    }
  }
  __syncthreads();
}

// FFT of ``nlines`` lines of length p.n; element e of line l sits at
// buf[m(l)*line_stride + e*elem_stride] with m = line_map (identity when
// null).  Must be called by all NT threads of the workgroup.
template <typename T, int NT, int PTS>
__device__ __forceinline__ void lds_fft(cplx<T>* buf, int nlines, int line_stride,
                                        int elem_stride, const int* __restrict__ line_map,
                                        const LinePlan& p,
                                        const cplx<T>* __restrict__ tw, int sign,
                                        int tid) {
  int ns = 1;
  for (int s = 0; s < p.nstages; ++s) {
    const int R = p.radix[s];
    if (R == 4)
      lds_stage<T, 4, NT, PTS>(buf, nlines, line_stride, elem_stride, line_map, p.n, ns, tw, sign, tid);
    else if (R == 2)
      lds_stage<T, 2, NT, PTS>(buf, nlines, line_stride, elem_stride, line_map, p.n, ns, tw, sign, tid);
    else if (R == 3)
      lds_stage<T, 3, NT, PTS>(buf, nlines, line_stride, elem_stride, line_map, p.n, ns, tw, sign, tid);
    else
      lds_stage<T, 5, NT, PTS>(buf, nlines, line_stride, elem_stride, line_map, p.n, ns, tw, sign, tid);
    ns *= R;
  }
}

// Largest lines*n that lds_fft<.., NT, PTS> accepts for radix set of ``p``.
inline int lds_fft_capacity(const LinePlan& p, int nt, int pts) {
  int cap = 1 << 30;
  for (int s = 0; s < p.nstages; ++s) {
    const int r = p.radix[s];
    const int maxb = (pts / r) > 0 ? (pts / r) : 1;
    const int c = nt * maxb * r;
    if (c < cap) cap = c;
  }
  return cap;
}

}  // namespace qepw
