// fft_lds.hpp -- mixed-radix FP32 FFT stages over LDS for gfx950.
//
// One workgroup transforms many lines that live in LDS.  Each stage is a
// Stockham autosort pass (Govindaraju, Lloyd, Dotsenko, Smith, Manferdelli,
// "High Performance Discrete Fourier Transforms on Graphics Processors",
// SC'08): butterfly j of a radix-R stage reads R points at stride n/R,
// twiddles them by W_{Ns*R}^{r*(j mod Ns)}, runs a size-R DFT in registers and
// writes them at stride Ns.  The pass works in place: every thread first loads
// all of its butterflies into registers, the workgroup synchronises, then
// every thread writes.  A thread holds at most PD_MAXV complex values per
// stage; the host checks each plan against that bound before launch.
//
// Sign convention is FFTW's: DIR = -1 forward e^{-2 pi i jk/n}, DIR = +1
// backward; no normalisation.
#pragma once
// FP contraction only within one source expression (a * b + c written as
// such), never chosen by the optimiser across statements: the column pass
// inlines the same transform twice (one copy per rotation of a pair), and
// optimiser-chosen fusion could round the two copies differently, making a
// rotation's result depend on its place in the batch.
#pragma clang fp contract(on)
#include <hip/hip_runtime.h>

#ifndef PD_MAXV
#define PD_MAXV 32
#endif
#define PD_MAX_STAGES 8

// A stage is {radix, ns, tws, 0}: ns is the product of the radices of the
// previous stages, tws = n / (ns * radix) the twiddle table stride.  Stages
// live in device memory so that no kernel indexes a private array.
struct FftPlan {
  int n;
  int nstages;
  const int4* st;    // nstages entries
  const float2* tw;  // tw[m] = exp(-2 pi i m / n), m < n (FP64-computed, rounded)
};
struct FftStage {
  int radix, ns, tws;
};

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by DIR*i: forward -i, backward +i
template <int DIR>
__device__ __forceinline__ float2 mul_dir_i(float2 a) {
  return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

template <int DIR>
__device__ __forceinline__ void dft2(float2& a, float2& b) {
  float2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <int DIR>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = mul_dir_i<DIR>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

template <int R, int DIR>
struct Dft;

template <int DIR>
struct Dft<2, DIR> {
  __device__ __forceinline__ static void run(float2* v, const float2*, int) { dft2<DIR>(v[0], v[1]); }
};

template <int DIR>
struct Dft<4, DIR> {
  __device__ __forceinline__ static void run(float2* v, const float2*, int) { dft4<DIR>(v[0], v[1], v[2], v[3]); }
};

template <int DIR>
struct Dft<3, DIR> {
  __device__ __forceinline__ static void run(float2* v, const float2*, int) {
    const float s = DIR * 0.866025403784438646763723170752936183f;
    float2 t = cadd(v[1], v[2]), u = csub(v[1], v[2]);
    float2 m = make_float2(v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y);
    float2 isu = make_float2(-s * u.y, s * u.x);
    v[0] = cadd(v[0], t);
    v[1] = cadd(m, isu);
    v[2] = csub(m, isu);
  }
};

template <int DIR>
struct Dft<5, DIR> {
  __device__ __forceinline__ static void run(float2* v, const float2*, int) {
    const float c1 = 0.309016994374947424102293417182819059f;   // cos(2pi/5)
    const float c2 = -0.809016994374947424102293417182819059f;  // cos(4pi/5)
    const float s1 = DIR * 0.951056516295153572116439333379382143f;
    const float s2 = DIR * 0.587785252292473129168705954639072769f;
    float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    float2 u1 = csub(v[1], v[4]), u2 = csub(v[2], v[3]);
    float2 a = v[0];
    float2 p1 = make_float2(a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y);
    float2 p2 = make_float2(a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y);
    // q1 = i*(s1 u1 + s2 u2), q2 = i*(s2 u1 - s1 u2)
    float2 r1 = make_float2(s1 * u1.x + s2 * u2.x, s1 * u1.y + s2 * u2.y);
    float2 r2 = make_float2(s2 * u1.x - s1 * u2.x, s2 * u1.y - s1 * u2.y);
    float2 q1 = make_float2(-r1.y, r1.x), q2 = make_float2(-r2.y, r2.x);
    v[0] = make_float2(a.x + t1.x + t2.x, a.y + t1.y + t2.y);
    v[1] = cadd(p1, q1);
    v[4] = csub(p1, q1);
    v[2] = cadd(p2, q2);
    v[3] = csub(p2, q2);
  }
};

template <int DIR>
struct Dft<8, DIR> {
  // j = j1 + 4 j2, k = k1 + 2 k2: radix-2 over j2, twiddle W8^{j1 k1}, radix-4 over j1.
  __device__ __forceinline__ static void run(float2* v, const float2*, int) {
    const float r = 0.707106781186547524400844362104849039f;
    float2 e[4], o[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      e[j] = cadd(v[j], v[j + 4]);
      o[j] = csub(v[j], v[j + 4]);
    }
    // o[j] *= W8^j (W8 = e^{DIR i pi/4})
    o[1] = make_float2(r * (o[1].x - DIR * o[1].y), r * (o[1].y + DIR * o[1].x));
    o[2] = mul_dir_i<DIR>(o[2]);
    o[3] = make_float2(r * (-o[3].x - DIR * o[3].y), r * (-o[3].y + DIR * o[3].x));
    dft4<DIR>(e[0], e[1], e[2], e[3]);
    dft4<DIR>(o[0], o[1], o[2], o[3]);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
      v[2 * k2] = e[k2];
      v[2 * k2 + 1] = o[k2];
    }
  }
};

template <int DIR>
struct Dft<16, DIR> {
  // j = j1 + 4 j2, k = k1 + 4 k2: radix-4 over j2, twiddle W16^{j1 k1}, radix-4 over j1.
  __device__ __forceinline__ static void run(float2* v, const float2*, int) {
    const float c1 = 0.923879532511286756128183189396788933f;  // cos(pi/8)
    const float s1 = 0.382683432365089771728459984030398866f;  // sin(pi/8)
    const float r = 0.707106781186547524400844362104849039f;
    float2 y[4][4];
#pragma unroll
    for (int j1 = 0; j1 < 4; ++j1) {
      float2 a0 = v[j1], a1 = v[j1 + 4], a2 = v[j1 + 8], a3 = v[j1 + 12];
      dft4<DIR>(a0, a1, a2, a3);
      y[j1][0] = a0; y[j1][1] = a1; y[j1][2] = a2; y[j1][3] = a3;
    }
    // twiddles W16^{j1 k1}, W16 = e^{DIR i pi/8}: cos(m pi/8) + DIR i sin(m pi/8)
    const float cw[10] = {1.f, c1, r, s1, 0.f, -s1, -r, -c1, -1.f, -c1};
    const float sw[10] = {0.f, s1, r, c1, 1.f, c1, r, s1, 0.f, -s1};
#pragma unroll
    for (int j1 = 1; j1 < 4; ++j1)
#pragma unroll
      for (int k1 = 1; k1 < 4; ++k1) {
        const int m = j1 * k1;  // <= 9
        const float c = cw[m], s = DIR * sw[m];
        float2 a = y[j1][k1];
        y[j1][k1] = make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
      }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      float2 a0 = y[0][k1], a1 = y[1][k1], a2 = y[2][k1], a3 = y[3][k1];
      dft4<DIR>(a0, a1, a2, a3);
      v[k1] = a0; v[k1 + 4] = a1; v[k1 + 8] = a2; v[k1 + 12] = a3;
    }
  }
};

// Generic O(R^2) DFT for the remaining primes (7, 11, 13) using W_R^m = tw[m*n/R].
template <int R, int DIR>
struct Dft {
  __device__ __forceinline__ static void run(float2* v, const float2* tw, int n) {
    float2 u[R];
    const int step = n / R;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      float2 acc = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w = tw[((r * q) % R) * step];
        if (DIR > 0) w.y = -w.y;
        acc = cadd(acc, cmul(v[r], w));
      }
      u[q] = acc;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = u[q];
  }
};

// One Stockham stage over `lines` lines in LDS.  Element i of line l is at
// buf[l*lstride + i*estride].
template <int R, int DIR>
__device__ __noinline__ void fft_stage(float2* buf, const FftPlan P, const FftStage S, int lines,
                                          int lstride, int estride) {
  constexpr int MAXT = (PD_MAXV / R) > 0 ? (PD_MAXV / R) : 1;
  const int nb = P.n / R;
  const int tasks = lines * nb;
  const int nth = blockDim.x, tid = threadIdx.x;
  float2 v[MAXT][R];
#pragma unroll
  for (int q = 0; q < MAXT; ++q) {
    const int task = tid + q * nth;
    if (task < tasks) {
      const int line = task % lines, j = task / lines;
      const float2* base = buf + line * lstride;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = base[(j + r * nb) * estride];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXT; ++q) {
    const int task = tid + q * nth;
    if (task < tasks) {
      const int line = task % lines, j = task / lines;
      const int k = j % S.ns;
      if (S.ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = P.tw[r * k * S.tws];
          if (DIR > 0) w.y = -w.y;
          v[q][r] = cmul(v[q][r], w);
        }
      }
      Dft<R, DIR>::run(v[q], P.tw, P.n);
      const int idxD = (j / S.ns) * S.ns * R + k;
      float2* base = buf + line * lstride;
#pragma unroll
      for (int r = 0; r < R; ++r) base[(idxD + r * S.ns) * estride] = v[q][r];
    }
  }
  __syncthreads();
}

template <int DIR>
__device__ void fft_lds(float2* buf, const FftPlan P, int lines, int lstride, int estride) {
  for (int s = 0; s < P.nstages; ++s) {
    const int4 s4 = P.st[s];
    const FftStage S = {s4.x, s4.y, s4.z};
    switch (S.radix) {
      case 2: fft_stage<2, DIR>(buf, P, S, lines, lstride, estride); break;
      case 3: fft_stage<3, DIR>(buf, P, S, lines, lstride, estride); break;
      case 4: fft_stage<4, DIR>(buf, P, S, lines, lstride, estride); break;
      case 5: fft_stage<5, DIR>(buf, P, S, lines, lstride, estride); break;
      case 7: fft_stage<7, DIR>(buf, P, S, lines, lstride, estride); break;
      case 8: fft_stage<8, DIR>(buf, P, S, lines, lstride, estride); break;
      case 11: fft_stage<11, DIR>(buf, P, S, lines, lstride, estride); break;
      case 13: fft_stage<13, DIR>(buf, P, S, lines, lstride, estride); break;
      case 16: fft_stage<16, DIR>(buf, P, S, lines, lstride, estride); break;
      default: break;
    }
  }
}

// Real-to-complex post-processing of rows that hold a length-M complex FFT Z
// of the packed real sequence z[m] = x[2m] + i x[2m+1] (N = 2M real points):
// X[k] = E + W_N^k O, X[M-k] = conj(E - W_N^k O), E = (Z[k] + conj Z[M-k])/2,
// O = (Z[k] - conj Z[M-k])/(2i).  Writes X[0..M] in place (row needs M+1 slots).
// tw2[k] = exp(-2 pi i k / N), k < M.  Each pair (k, M-k) is owned by one task.
__device__ __forceinline__ void r2c_post(float2* buf, int rows, int lstride, int M, const float2* tw2) {
  const int half = M / 2 + 1;
  const int tasks = rows * half;
  for (int task = threadIdx.x; task < tasks; task += blockDim.x) {
    const int row = task % rows, k = task / rows;
    float2* z = buf + row * lstride;
    const int mk = (M - k) % M;
    const float2 zk = z[k], zm = z[mk];
    const float2 E = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
    // O = (zk - conj zm) / (2i) = (-i/2) * (zk - conj zm)
    const float2 d = make_float2(zk.x - zm.x, zk.y + zm.y);
    const float2 O = make_float2(0.5f * d.y, -0.5f * d.x);
    const float2 WO = cmul(tw2[k], O);
    if (k == 0) {
      z[0] = make_float2(E.x + WO.x, 0.f);
      z[M] = make_float2(E.x - WO.x, 0.f);
    } else {
      z[k] = cadd(E, WO);
      if (mk != k) z[mk] = cconj(csub(E, WO));
    }
  }
  __syncthreads();
}

// Inverse of r2c_post (FFTW c2r semantics: the imaginary parts of X[0] and
// X[M] are ignored).  Produces Z[k] = (X[k] + conj X[M-k]) + i W_N^{-k} (X[k] - conj X[M-k])
// for k < M, so that a backward length-M FFT of Z yields N * x packed.
__device__ __forceinline__ void c2r_pre(float2* buf, int rows, int lstride, int M, const float2* tw2) {
  const int half = M / 2 + 1;
  const int tasks = rows * half;
  for (int task = threadIdx.x; task < tasks; task += blockDim.x) {
    const int row = task % rows, k = task / rows;
    float2* z = buf + row * lstride;
    float2 xk = z[k], xm = z[M - k];
    if (k == 0) { xk.y = 0.f; xm.y = 0.f; }
    const float2 E2 = make_float2(xk.x + xm.x, xk.y - xm.y);
    const float2 D = make_float2(xk.x - xm.x, xk.y + xm.y);
    const float2 O2 = cmulc(D, tw2[k]);  // D * W^{-k}
    const float2 iO2 = make_float2(-O2.y, O2.x);
    z[k] = cadd(E2, iO2);
    if (k != 0 && (M - k) != k) {
      // Z[M-k] = conj(E2) + i conj(O2)
      z[M - k] = make_float2(E2.x + O2.y, -E2.y + O2.x);
    }
  }
  __syncthreads();
}
