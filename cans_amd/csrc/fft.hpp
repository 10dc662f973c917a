// LDS-resident batched complex FFT building block (radix 8/4/2, then 3/5, Stockham).
//
// One "team" of T threads transforms one line of N = 2^m (or 2^m 3^a 5^b) complex values that
// sits in LDS, addressed through an element map.  The LDS element type L may be
// float or double, and the butterflies compute in L: FP32 lines give FP32
// transforms (twiddles from a table in L: the double table rounded once), FP64
// lines FP64 ones.  Twiddle products use fused multiply-adds.  T = max(1, N/8): every
// power-of-two stage leaves each thread exactly 8 (or N, when N < 8) values in registers, an
// odd stage of radix P up to P * ceil(8/P).  All teams of a
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

// radix 5: y_r = sum_q x_q W5^{r q}, W5 = exp(s 2 pi i/5).  With t1 = x1 + x4, t2 = x2 + x3,
// t3 = x1 - x4, t4 = x2 - x3, c_n = cos(2 pi n/5), s_n = sin(2 pi n/5):
// y_{1,4} = x0 + c1 t1 + c2 t2 +- s i (s1 t3 + s2 t4), y_{2,3} = x0 + c2 t1 + c1 t2 +- s i (s2 t3 - s1 t4)
template <typename T, int s>
__device__ __forceinline__ void bfly5(cplx<T>* v) {
  const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
  const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
  const cplx<T> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  const cplx<T> t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  const cplx<T> a1 = {v[0].re + c1 * t1.re + c2 * t2.re, v[0].im + c1 * t1.im + c2 * t2.im};
  const cplx<T> a2 = {v[0].re + c2 * t1.re + c1 * t2.re, v[0].im + c2 * t1.im + c1 * t2.im};
  const cplx<T> b1 = mul_si<T, s>(cplx<T>{s1 * t3.re + s2 * t4.re, s1 * t3.im + s2 * t4.im});
  const cplx<T> b2 = mul_si<T, s>(cplx<T>{s2 * t3.re - s1 * t4.re, s2 * t3.im - s1 * t4.im});
  v[0] = cadd(v[0], cadd(t1, t2));
  v[1] = cadd(a1, b1);
  v[4] = csub(a1, b1);
  v[2] = cadd(a2, b2);
  v[3] = csub(a2, b2);
}

// One Stockham stage of odd radix P (3 or 5) with span Ns (Ns * P divides N), after the
// power-of-two stages of a line of N = 2^m * (odd part), m >= 3.  The N/P butterflies do
// not split evenly over the team's N/8 threads; each thread takes up to ceil(8/P) of them
// (j = lt, lt + Tn, ... below N/P).
template <typename L, int s, int P, class IX>
__device__ __forceinline__ void fft_stage_odd(cplx<L>* buf, IX ix, int N, int Ns, int lt, int Tn,
                                              const cplx<L>* __restrict__ tw, bool act = true) {
  typedef L T;
  constexpr int NBF = (8 + P - 1) / P;
  const int NR = N / P;
  const int step = N / (Ns * P);  // W_{Ns P}^{r k} = W_N^{r k step}
  cplx<T> v[NBF][P];
#pragma unroll
  for (int bb = 0; bb < NBF; bb++) {
    const int j = lt + bb * Tn;
    if (act && j < NR) {
#pragma unroll
      for (int r = 0; r < P; r++) v[bb][r] = buf[ix(j + r * NR)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int bb = 0; bb < NBF; bb++) {
    const int j = lt + bb * Tn;
    if (act && j < NR) {
      const int q = j / Ns, k = j - q * Ns;
      if (k) {
        const cplx<T> w1 = tw[k * step];
        const cplx<T> w2 = tw[2 * k * step];
        v[bb][1] = (s < 0) ? tmul(v[bb][1], w1) : tmulc(v[bb][1], w1);
        v[bb][2] = (s < 0) ? tmul(v[bb][2], w2) : tmulc(v[bb][2], w2);
        if constexpr (P == 5) {
          const cplx<T> w3 = tw[3 * k * step];
          const cplx<T> w4 = tw[4 * k * step];
          v[bb][3] = (s < 0) ? tmul(v[bb][3], w3) : tmulc(v[bb][3], w3);
          v[bb][4] = (s < 0) ? tmul(v[bb][4], w4) : tmulc(v[bb][4], w4);
        }
      }
      if constexpr (P == 3) bfly3<T, s>(v[bb]);
      else bfly5<T, s>(v[bb]);
      const int idxD = q * Ns * P + k;
#pragma unroll
      for (int r = 0; r < P; r++) buf[ix(idxD + r * Ns)] = v[bb][r];
    }
  }
  __syncthreads();
}

// full transform of one line (unnormalised), s = -1 forward, +1 inverse.
// N = 2^logN * odd, where odd (the odd part of N) is 1 or a product of at most three
// factors 3 and 5, and logN >= 3 when odd > 1.  The power-of-two stages run first with
// spans 1, 8, ..., then one radix-3 stage per factor 3 and one radix-5 stage per factor 5.
// r4: radix-4 power-of-two stages (two butterflies a thread, spans 1, 4, 16, ...) instead
// of radix 8: twice the LDS round trips, half the values a butterfly holds (the launch-shape
// tuner chooses).
template <typename L, int s, class IX>
__device__ __forceinline__ void fft_line(cplx<L>* buf, IX ix, int N, int logN, int lt, int Tn,
                         const cplx<L>* __restrict__ tw, int odd = 1, bool act = true, bool r4 = false) {
  if (logN == 0 && odd == 1) return;
  if (logN == 1 && odd == 1) {
    fft_stage<L, s, 2, 1>(buf, ix, N, 0, lt, Tn, tw, act);
    return;
  }
  if (logN == 2 && odd == 1) {
    fft_stage<L, s, 4, 1>(buf, ix, N, 0, lt, Tn, tw, act);
    return;
  }
  int lNs = 0;
  int rem = logN;
  if (r4) {
    for (; rem >= 2; lNs += 2, rem -= 2) fft_stage<L, s, 4, 2>(buf, ix, N, lNs, lt, Tn, tw, act);
  } else {
    for (; rem >= 3; lNs += 3, rem -= 3) fft_stage<L, s, 8, 1>(buf, ix, N, lNs, lt, Tn, tw, act);
    if (rem == 2) fft_stage<L, s, 4, 2>(buf, ix, N, lNs, lt, Tn, tw, act);
  }
  if (rem == 1) fft_stage<L, s, 2, 4>(buf, ix, N, lNs, lt, Tn, tw, act);
  int Ns = 1 << logN;
  for (; odd % 3 == 0; odd /= 3, Ns *= 3) fft_stage_odd<L, s, 3>(buf, ix, N, Ns, lt, Tn, tw, act);
  for (; odd % 5 == 0; odd /= 5, Ns *= 5) fft_stage_odd<L, s, 5>(buf, ix, N, Ns, lt, Tn, tw, act);
}

}  // namespace cans
