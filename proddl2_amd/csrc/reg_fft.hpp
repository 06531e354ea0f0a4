// reg_fft.hpp -- compile-time register FFTs and the "four-step" line FFT used
// by the specialised scan kernels.
//
// rdft<R, DIR>(v): in-place DFT of R values held in registers, natural order in
// and out.  Sizes 2, 3, 4, 5, 8, 16 are hand-written (fft_lds.hpp Dft<>); every
// other size R = P*Q is built by one Cooley-Tukey split with compile-time
// twiddles W_R^{j1 k2}.
//
// A line FFT of N = A*B points is done by B threads (lane g of the line):
//   phase 1: lane g holds x[g + B m], m < A; DFT_A over m; twiddle W_N^{g k1};
//   exchange through LDS;
//   phase 2: lane g holds Y_{g'}[k1] for k1 = g + B j (j < J = ceil(A/B)),
//   all g'; DFT_B over g' gives X[k1 + A k2].
// X[k1 + A k2] = sum_g W_B^{g k2} W_N^{g k1} sum_m x[g + B m] W_A^{m k1}.
#pragma once
#include "fft_lds.hpp"
#include "twiddle_consts.hpp"

constexpr bool rdft_is_base(int R) { return R == 1 || R == 2 || R == 3 || R == 4 || R == 5 || R == 8 || R == 16; }

constexpr int rdft_factor(int R) {
  return (R % 4 == 0) ? 4 : (R % 2 == 0) ? 2 : (R % 3 == 0) ? 3 : (R % 5 == 0) ? 5 : R;
}

template <int R, int DIR, bool BASE = rdft_is_base(R)>
struct RDft;

template <int DIR>
struct RDft<1, DIR, true> {
  __device__ __forceinline__ static void run(float2*) {}
};

template <int R, int DIR>
struct RDft<R, DIR, true> {
  __device__ __forceinline__ static void run(float2* v) { Dft<R, DIR>::run(v, nullptr, 0); }
};

template <int R, int DIR>
struct RDft<R, DIR, false> {
  static constexpr int P = rdft_factor(R);
  static constexpr int Q = R / P;
  static_assert(P > 1 && Q > 1 && R <= 32, "unsupported register DFT size");
  __device__ __forceinline__ static void run(float2* v) {
    float2 y[P][Q];
#pragma unroll
    for (int j1 = 0; j1 < P; ++j1) {
      float2 t[Q];
#pragma unroll
      for (int j2 = 0; j2 < Q; ++j2) t[j2] = v[j1 + P * j2];
      RDft<Q, DIR>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < Q; ++k2) y[j1][k2] = t[k2];
    }
#pragma unroll
    for (int j1 = 1; j1 < P; ++j1)
#pragma unroll
      for (int k2 = 1; k2 < Q; ++k2) {
        const int m = (j1 * k2) % R;
        const float c = kCos[R][m], s = DIR * kSin[R][m];
        const float2 a = y[j1][k2];
        y[j1][k2] = make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
      }
#pragma unroll
    for (int k2 = 0; k2 < Q; ++k2) {
      float2 t[P];
#pragma unroll
      for (int j1 = 0; j1 < P; ++j1) t[j1] = y[j1][k2];
      RDft<P, DIR>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) v[k2 + Q * k1] = t[k1];
    }
  }
};

template <int R, int DIR>
__device__ __forceinline__ void rdft(float2* v) {
  RDft<R, DIR>::run(v);
}

// Phase 1 of the four-step FFT: v[m] = x[g + B m] -> DFT_A, then twiddle by
// W_N^{g k1} (DIR = +1 uses the conjugate).  twN[m] = exp(-2 pi i m / N).
template <int A, int B, int DIR>
__device__ __forceinline__ void four_step_phase1(float2 (&v)[A], int g, const float2* __restrict__ twN) {
  rdft<A, DIR>(v);
#pragma unroll
  for (int k1 = 1; k1 < A; ++k1) {
    float2 w = twN[g * k1];
    if (DIR > 0) w.y = -w.y;
    v[k1] = cmul(v[k1], w);
  }
}

// DFT of A values of which only the first MZ may be nonzero (natural order in
// and out): with R = A / MZ and k = r + R q,
//   X[r + R q] = sum_{m < MZ} (x[m] W_A^{m r}) W_MZ^{m q},
// i.e. R twiddled DFTs of length MZ.  MZ not dividing A falls back to rdft<A>.
template <int A, int MZ, int DIR>
__device__ __forceinline__ void rdft_zt(float2 (&v)[A]) {
  if constexpr (MZ >= A || A % MZ != 0) {
    rdft<A, DIR>(v);
  } else {
    constexpr int R = A / MZ;
    float2 y[R][MZ];
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
      for (int m = 0; m < MZ; ++m) {
        const int e = (m * r) % A;
        const float c = kCos[A][e], s = DIR * kSin[A][e];
        const float2 a = v[m];
        y[r][m] = e == 0 ? a : make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
      }
      rdft<MZ, DIR>(y[r]);
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int q = 0; q < MZ; ++q) v[r + R * q] = y[r][q];
  }
}

// Phase 1 with a zero tail: x[g + B m] = 0 for m >= MZ.
template <int A, int B, int MZ, int DIR>
__device__ __forceinline__ void four_step_phase1_zt(float2 (&v)[A], int g, const float2* __restrict__ twN) {
  rdft_zt<A, MZ, DIR>(v);
#pragma unroll
  for (int k1 = 1; k1 < A; ++k1) {
    float2 w = twN[g * k1];
    if (DIR > 0) w.y = -w.y;
    v[k1] = cmul(v[k1], w);
  }
}

// Inverse-order variant used after accumulation: lane holds X[k1 + A k2] for
// its k1 (all k2) in u[k2]; DFT_B over k2 (direction DIR) then twiddle by
// W_N^{DIR-sign k1 g}, giving P[k1][g] in u[g].
template <int A, int B, int DIR>
__device__ __forceinline__ void four_step_dif_phase1(float2 (&u)[B], int k1, const float2* __restrict__ twN) {
  rdft<B, DIR>(u);
#pragma unroll
  for (int g = 1; g < B; ++g) {
    float2 w = twN[g * k1];
    if (DIR > 0) w.y = -w.y;
    u[g] = cmul(u[g], w);
  }
}
