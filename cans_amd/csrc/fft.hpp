// LDS-resident batched complex FFT building block (radix 8/4/2/3 Stockham).
//
// One "team" of T threads transforms one line of N = 2^m (or 3 * 2^m) complex values that
// sits in LDS, addressed through an element map.  The LDS element type L may be
// float or double, and the butterflies compute in L: FP32 lines give FP32
// transforms (twiddles from a table in L: the double table rounded once), FP64
// lines FP64 ones.  Twiddle products use fused multiply-adds.  T = max(1, N/8): every stage
// leaves each thread exactly 8 (or N, when N < 8) values in registers.  All teams of a
// work-group run the same N and call this together (it contains barriers).
//
// Formulation: the Stockham autosort iteration of Govindaraju et al., "High
// performance discrete Fourier transforms on graphics processors" (SC'08):
// stage with radix R and span Ns maps butterfly j to inputs j + r*N/R and
// outputs (j/Ns)*Ns*R + j%Ns + r*Ns, after twiddling input r by W_{Ns*R}^{r*(j%Ns)}.
#pragma once
#include "common.hpp"

namespace cans {

template <typename T>
struct Rsqrt2;
template <>
struct Rsqrt2<double> {
  static constexpr double v = 0.70710678118654752440;
};
template <>
struct Rsqrt2<float> {
  static constexpr float v = 0.70710678f;
};

// multiply by s*i (s = -1 forward, +1 inverse)
template <typename T, int s>
__device__ __forceinline__ cplx<T> mul_si(cplx<T> x) {
  return {-(T)s * x.im, (T)s * x.re};
}

template <typename T, int s>
__device__ __forceinline__ void bfly2(cplx<T>* v) {
  cplx<T> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <typename T, int s>
__device__ __forceinline__ void bfly4(cplx<T>* v) {
  cplx<T> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
  cplx<T> b0 = cadd(v[1], v[3]), b1 = mul_si<T, s>(csub(v[1], v[3]));
  v[0] = cadd(a0, b0);
  v[2] = csub(a0, b0);
  v[1] = cadd(a1, b1);
  v[3] = csub(a1, b1);
}

template <typename T, int s>
__device__ __forceinline__ void bfly8(cplx<T>* v) {
  cplx<T> e[4] = {v[0], v[2], v[4], v[6]};
  cplx<T> o[4] = {v[1], v[3], v[5], v[7]};
  bfly4<T, s>(e);
  bfly4<T, s>(o);
  const T r = Rsqrt2<T>::v;
  // W8^1 = (1 + s i)/sqrt2, W8^2 = s i, W8^3 = (-1 + s i)/sqrt2
  cplx<T> o1 = {(o[1].re - (T)s * o[1].im) * r, (o[1].im + (T)s * o[1].re) * r};
  cplx<T> o2 = mul_si<T, s>(o[2]);
  cplx<T> o3 = {(-o[3].re - (T)s * o[3].im) * r, (-o[3].im + (T)s * o[3].re) * r};
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

template <typename T, int s, int R>
__device__ __forceinline__ void bfly(cplx<T>* v) {
  if constexpr (R == 2) bfly2<T, s>(v);
  else if constexpr (R == 4) bfly4<T, s>(v);
  else bfly8<T, s>(v);
}

// LDS element maps: contiguous line with one pad slot every 16 elements (keeps the
// strided Stockham writes free of bank conflicts), or a line interleaved with S others.
struct IxPad {
  int base;
  __device__ __forceinline__ int operator()(int e) const { return base + e + (e >> 4); }
};
struct IxStride {
  int base, S;
  __device__ __forceinline__ int operator()(int e) const { return base + e * S; }
};
__host__ __device__ constexpr int padded_len(int N) { return N + (N >> 4) + 1; }

template <typename L>
__device__ __forceinline__ cplx<double> widen(cplx<L> x) {
  return {(double)x.re, (double)x.im};
}
template <typename L>
__device__ __forceinline__ cplx<L> narrow(cplx<double> x) {
  return {(L)x.re, (L)x.im};
}
template <typename L>
__device__ __forceinline__ cplx<L> narrow(double re, double im) {
  return {(L)re, (L)im};
}

// twiddle products with fused multiply-adds
template <typename T>
__device__ __forceinline__ cplx<T> tmul(cplx<T> a, cplx<T> b) {
#pragma clang fp contract(fast)
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ cplx<T> tmulc(cplx<T> a, cplx<T> b) {
#pragma clang fp contract(fast)
  return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}

// one Stockham stage: R radix, NBF butterflies per thread, arithmetic in L
// act = false: a lane outside every team (work-groups whose size is not a multiple
// of the team size); it takes part in the barriers only
template <typename L, int s, int R, int NBF, class IX>
__device__ __forceinline__ void fft_stage(cplx<L>* buf, IX ix, int N, int lNs, int lt, int Tn,
                                          const cplx<L>* __restrict__ tw, bool act = true) {
  typedef L T;
  constexpr int LR = R == 2 ? 1 : R == 4 ? 2 : 3;
  const int NR = N >> LR;
  const int Ns = 1 << lNs;
  cplx<T> v[NBF][R];
  if (act) {
#pragma unroll
    for (int bb = 0; bb < NBF; bb++) {
      const int j = lt + bb * Tn;
#pragma unroll
      for (int r = 0; r < R; r++) v[bb][r] = buf[ix(j + r * NR)];
    }
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (int bb = 0; bb < NBF; bb++) {
      const int j = lt + bb * Tn;
      const int k = j & (Ns - 1);
      if (k) {
        const int step = k * (N >> (lNs + LR));  // k * N/(Ns*R)
        // W^r for r = 1..R-1 from the table entries W^1, W^2, W^4 (at most one product
        // deep), so only a few twiddles are live at a time
        const cplx<T> w1 = tw[step];
        if constexpr (R == 2) {
          v[bb][1] = (s < 0) ? tmul(v[bb][1], w1) : tmulc(v[bb][1], w1);
        } else {
          const cplx<T> w2 = tw[2 * step];
          const cplx<T> w3 = tmul(w1, w2);
          v[bb][1] = (s < 0) ? tmul(v[bb][1], w1) : tmulc(v[bb][1], w1);
          v[bb][2] = (s < 0) ? tmul(v[bb][2], w2) : tmulc(v[bb][2], w2);
          v[bb][3] = (s < 0) ? tmul(v[bb][3], w3) : tmulc(v[bb][3], w3);
          if constexpr (R == 8) {
            const cplx<T> w4 = tw[4 * step];
            const cplx<T> w5 = tmul(w4, w1), w6 = tmul(w4, w2), w7 = tmul(w4, w3);
            v[bb][4] = (s < 0) ? tmul(v[bb][4], w4) : tmulc(v[bb][4], w4);
            v[bb][5] = (s < 0) ? tmul(v[bb][5], w5) : tmulc(v[bb][5], w5);
            v[bb][6] = (s < 0) ? tmul(v[bb][6], w6) : tmulc(v[bb][6], w6);
            v[bb][7] = (s < 0) ? tmul(v[bb][7], w7) : tmulc(v[bb][7], w7);
          }
        }
      }
      bfly<T, s, R>(v[bb]);
      const int idxD = ((j >> lNs) << (lNs + LR)) + k;
#pragma unroll
      for (int r = 0; r < R; r++) buf[ix(idxD + r * Ns)] = v[bb][r];
    }
  }
  __syncthreads();
}

// radix 3: y_r = sum_q x_q W3^{r q}, W3 = exp(s 2 pi i/3) = -1/2 + s i sqrt(3)/2
template <typename T, int s>
__device__ __forceinline__ void bfly3(cplx<T>* v) {
  const T h = (T)0.86602540378443864676;  // sqrt(3)/2
  const cplx<T> t1 = cadd(v[1], v[2]);
  const cplx<T> t2 = {v[0].re - (T)0.5 * t1.re, v[0].im - (T)0.5 * t1.im};
  const cplx<T> d = csub(v[1], v[2]);
  const cplx<T> t3 = mul_si<T, s>(cplx<T>{h * d.re, h * d.im});
  v[0] = cadd(v[0], t1);
  v[1] = cadd(t2, t3);
  v[2] = csub(t2, t3);
}

// Last stage of a line of N = 3 * 2^m (m >= 3): radix 3 with span Ns = N/3.  The
// N/3 butterflies do not split evenly over the team's N/8 threads; each thread takes
// up to three of them (j = lt, lt + Tn, lt + 2 Tn below N/3).
template <typename L, int s, class IX>
__device__ __forceinline__ void fft_stage3(cplx<L>* buf, IX ix, int N, int lt, int Tn,
                                           const cplx<L>* __restrict__ tw, bool act = true) {
  typedef L T;
  const int NR = N / 3;
  cplx<T> v[3][3];
#pragma unroll
  for (int bb = 0; bb < 3; bb++) {
    const int j = lt + bb * Tn;
    if (act && j < NR) {
#pragma unroll
      for (int r = 0; r < 3; r++) v[bb][r] = buf[ix(j + r * NR)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int bb = 0; bb < 3; bb++) {
    const int j = lt + bb * Tn;
    if (act && j < NR) {
      if (j) {  // W_N^{r j}, r = 1, 2 (span Ns = N/3: k = j, step = j)
        const cplx<T> w1 = tw[j];
        const cplx<T> w2 = tw[2 * j];
        v[bb][1] = (s < 0) ? tmul(v[bb][1], w1) : tmulc(v[bb][1], w1);
        v[bb][2] = (s < 0) ? tmul(v[bb][2], w2) : tmulc(v[bb][2], w2);
      }
      bfly3<T, s>(v[bb]);
#pragma unroll
      for (int r = 0; r < 3; r++) buf[ix(j + r * NR)] = v[bb][r];
    }
  }
  __syncthreads();
}

// full transform of one line (unnormalised), s = -1 forward, +1 inverse.
// N = 2^logN, or N = 3 * 2^logN with logN >= 3 when r3 is set (the power-of-two
// stages run first with spans 1, 8, ..., then one radix-3 stage).
template <typename L, int s, class IX>
__device__ __forceinline__ void fft_line(cplx<L>* buf, IX ix, int N, int logN, int lt, int Tn,
                         const cplx<L>* __restrict__ tw, int r3 = 0, bool act = true) {
  if (logN == 0 && !r3) return;
  if (logN == 1 && !r3) {
    fft_stage<L, s, 2, 1>(buf, ix, N, 0, lt, Tn, tw, act);
    return;
  }
  if (logN == 2 && !r3) {
    fft_stage<L, s, 4, 1>(buf, ix, N, 0, lt, Tn, tw, act);
    return;
  }
  int lNs = 0;
  int rem = logN;
  while (rem >= 3) {
    fft_stage<L, s, 8, 1>(buf, ix, N, lNs, lt, Tn, tw, act);
    lNs += 3;
    rem -= 3;
  }
  if (rem == 2) fft_stage<L, s, 4, 2>(buf, ix, N, lNs, lt, Tn, tw, act);
  else if (rem == 1) fft_stage<L, s, 2, 4>(buf, ix, N, lNs, lt, Tn, tw, act);
  if (r3) fft_stage3<L, s>(buf, ix, N, lt, Tn, tw, act);
}

}  // namespace cans
