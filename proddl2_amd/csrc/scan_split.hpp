// scan_split.hpp -- plane passes split into row and column kernels, for grids
// whose y-z half plane does not fit in LDS (e.g. 256 x 129 complex = 264 KiB
// at 256^3) or in the register plan of the whole-plane kernels.
//
// Forward ligand:  k_lig_rows (deposit + r2c along z, 32 rows of the ligand box
//                  per block) -> L1 rows; k_cols_fft<-1> (FFT along y in
//                  place, 16 columns per block; rows outside the box are 0).
// Inverse:         k_cols_fft<+1> on S1, then k_inv_rows_topk_fast<N2> (register
//                  FFTs, wave-level selection under the rotation's threshold) or,
//                  for other N2, k_inv_rows_topk (c2r along z in chunks of 32 rows,
//                  1/N, and the plane's K best cells kept in a block-wide list
//                  across the chunks).
// Receptor:        k_rec_rows + k_cols_fft<-1>, then the usual k_col_fwd.
// N1 = N2 = 256:   each plane pass in one 1,024-thread kernel that holds a
//                  plane on chip in registers and LDS: k_lig_plane_split
//                  (ligand boxes of up to 128 rows; each row's z spectrum by
//                  direct summation over its atoms) and k_inv_plane_topk_split.
//
// The same transforms (r2c along z, FFT along y, and their inverses; the FFT
// factorisations differ, so results agree to rounding) and the same selection
// rule ((value, offset) lexicographic, NaN as +inf) as the whole-plane kernels
// of scan_kernels.hpp; the row and column kernels take one extra HBM round
// trip of the plane per pass.
#pragma once

#include "scan_fast.hpp"
#include "scan_kernels.hpp"

constexpr int kSplitRows = 32;     // rows per block in the row kernels (at most; ScanGeom::zrows)
constexpr int kSplitCols = 16;     // columns per block in the column kernels
constexpr int kSplitColStride = kSplitCols + 1;
constexpr int kSplitThreads = 256;

// grid (p0, nrot, ceil(p1 / zrows)), LDS zrows * ls complex.
__global__ __launch_bounds__(kSplitThreads) void k_lig_rows(ScanGeom g, FftPlan pz,
                                                            const float2* __restrict__ tw2,
                                                            const int* __restrict__ cells, int n_atoms,
                                                            const float* __restrict__ w,
                                                            float2* __restrict__ L1) {
  extern __shared__ float2 buf[];
  const int p = blockIdx.x, b = blockIdx.y;
  const int r0 = g.lo[1] + blockIdx.z * g.zrows;
  const int nr = min(g.zrows, g.lo[1] + g.p[1] - r0);
  const int i0 = g.lo[0] + p;
  const int plane = g.n1 * g.n2;
  const int* cb = cells + (size_t)b * n_atoms;
  float* fb = reinterpret_cast<float*>(buf);
  const int tot = g.zrows * g.ls;
  for (int t = 0; t < g.T; ++t) {
    for (int e = threadIdx.x; e < tot; e += blockDim.x) buf[e] = make_float2(0.f, 0.f);
    __syncthreads();
    const float* wt = w + (size_t)t * n_atoms;
    for (int j = threadIdx.x; j < n_atoms; j += blockDim.x) {
      const int off = cb[j];
      if (off >= 0 && off / plane == i0) {
        const int r = off - i0 * plane;
        const int i1 = r / g.n2, i2 = r - i1 * g.n2;
        if (i1 >= r0 && i1 < r0 + nr) atomicAdd(&fb[(size_t)(i1 - r0) * 2 * g.ls + zreal(g, i2)], wt[j]);
      }
    }
    __syncthreads();
    fft_lds<-1>(buf, pz, nr, g.ls, 1);
    r2c_finish(g, buf, nr, tw2);
    float2* dst = L1 + (((size_t)b * g.T + t) * g.pp + (g.lo[0] - g.pl) + p) * (size_t)g.M + (size_t)r0 * g.h2;
    for (int e = threadIdx.x; e < nr * g.h2; e += blockDim.x) {
      const int rr = e / g.h2, k2 = e - rr * g.h2;
      dst[e] = buf[rr * g.ls + k2];
    }
    __syncthreads();
  }
}

// FFT along y of planes laid out [y][k2] (row stride h2), in place.  Plane q
// (q / grp groups of grp planes) starts at base + ((q / grp) * pitch + off +
// q % grp) * M; rows outside [ylo, yhi) are read as 0.
// grid (n_planes, ceil(h2 / kSplitCols)), LDS n1 * kSplitColStride complex.
template <int DIR>
__global__ __launch_bounds__(kSplitThreads) void k_cols_fft(ScanGeom g, FftPlan py, float2* __restrict__ base,
                                                            int grp, int pitch, int off, int ylo, int yhi) {
  extern __shared__ float2 buf[];
  const int c0 = blockIdx.y * kSplitCols;
  const int cw = min(kSplitCols, g.h2 - c0);
  const int q = blockIdx.x;
  float2* pl = base + ((size_t)(q / grp) * pitch + off + q % grp) * g.M + c0;
  const int nel = g.n1 * kSplitCols;
  for (int e = threadIdx.x; e < nel; e += blockDim.x) {
    const int y = e / kSplitCols, c = e - y * kSplitCols;
    float2 v = make_float2(0.f, 0.f);
    if (c < cw && y >= ylo && y < yhi) v = pl[(size_t)y * g.h2 + c];
    buf[y * kSplitColStride + c] = v;
  }
  __syncthreads();
  fft_lds<DIR>(buf, py, kSplitCols, 1, kSplitColStride);
  for (int e = threadIdx.x; e < nel; e += blockDim.x) {
    const int y = e / kSplitCols, c = e - y * kSplitCols;
    if (c < cw) pl[(size_t)y * g.h2 + c] = buf[y * kSplitColStride + c];
  }
}

// The same y FFT with N1 known at compile time: a tile of N1 rows x CB
// columns (CB = 256 / B, 128-byte row pieces) is read into LDS with coalesced
// loads, each column is transformed by B lanes with the register four-step
// FFT (one exchange through the tile), and the tile is written back
// coalesced.  grid (n_planes, ceil(h2 / CB)), block 256, LDS N1 * (CB + 1)
// complex.  Same plane addressing and row window as k_cols_fft.
template <int N1, int DIR>
__global__ __launch_bounds__(256) void k_cols_fast(ScanGeom g, const float2* __restrict__ twN1,
                                                   float2* __restrict__ base, int grp, int pitch, int off, int ylo,
                                                   int yhi) {
  using S = Split<N1>;
  constexpr int A = S::A, B = S::B, J = S::J, CB = 256 / B, CBS = CB + 1;
  static_assert(A * B == N1 && 256 % B == 0, "column split");
  extern __shared__ float2 tile[];
  const int c0 = blockIdx.y * CB;
  const int cw = min(CB, g.h2 - c0);
  const int q = blockIdx.x;
  float2* pl = base + ((size_t)(q / grp) * pitch + off + q % grp) * g.M + c0;
  const int cl = threadIdx.x % CB, lg = threadIdx.x / CB;
  // the lane's twiddles W_N1^{lg k1}, loaded with the tile rather than after it
  float2 tw[A];
#pragma unroll
  for (int k1 = 1; k1 < A; ++k1) {
    tw[k1] = twN1[lg * k1];
    if (DIR > 0) tw[k1].y = -tw[k1].y;
  }
  // N1 * CB / 256 elements per thread, all loads in flight before the first
  // LDS write (a rolled loop would wait for each load in turn)
  static_assert(N1 * CB % 256 == 0, "whole tile rows per pass");
  constexpr int NE = N1 * CB / 256;
  float2 ld[NE];
#pragma unroll
  for (int i = 0; i < NE; ++i) {
    const int e = threadIdx.x + 256 * i;
    const int y = e / CB, c = e % CB;
    ld[i] = (c < cw && y >= ylo && y < yhi) ? pl[(size_t)y * g.h2 + c] : make_float2(0.f, 0.f);
  }
#pragma unroll
  for (int i = 0; i < NE; ++i) {
    const int e = threadIdx.x + 256 * i;
    tile[(e / CB) * CBS + e % CB] = ld[i];
  }
  __syncthreads();
  float2 x[A];
#pragma unroll
  for (int m = 0; m < A; ++m) x[m] = tile[(lg + B * m) * CBS + cl];
  rdft<A, DIR>(x);
#pragma unroll
  for (int k1 = 1; k1 < A; ++k1) x[k1] = cmul(x[k1], tw[k1]);
  __syncthreads();
#pragma unroll
  for (int k1 = 0; k1 < A; ++k1) tile[(k1 * B + lg) * CBS + cl] = x[k1];
  __syncthreads();
  float2 w[J][B];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int k1 = lg + B * j;
    if (k1 < A)
#pragma unroll
      for (int qq = 0; qq < B; ++qq) w[j][qq] = tile[(k1 * B + qq) * CBS + cl];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int k1 = lg + B * j;
    if (k1 < A) {
      rdft<B, DIR>(w[j]);
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) tile[(k1 + A * k2) * CBS + cl] = w[j][k2];
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NE; ++i) {
    const int e = threadIdx.x + 256 * i;
    const int y = e / CB, c = e % CB;
    if (c < cw) pl[(size_t)y * g.h2 + c] = tile[y * CBS + c];
  }
}

// Receptor rows: dense real rows -> r2c(z).  grid (n0, ceil(n1 / zrows)).
__global__ __launch_bounds__(kSplitThreads) void k_rec_rows(ScanGeom g, FftPlan pz, const float2* __restrict__ tw2,
                                                            const float* __restrict__ R,
                                                            float2* __restrict__ out) {
  extern __shared__ float2 buf[];
  const int i0 = blockIdx.x;
  const int r0 = blockIdx.y * g.zrows;
  const int nr = min(g.zrows, g.n1 - r0);
  float* fb = reinterpret_cast<float*>(buf);
  for (int e = threadIdx.x; e < g.zrows * g.ls; e += blockDim.x) buf[e] = make_float2(0.f, 0.f);
  __syncthreads();
  const float* src = R + ((size_t)i0 * g.n1 + r0) * g.n2;
  for (int e = threadIdx.x; e < nr * g.n2; e += blockDim.x) {
    const int rr = e / g.n2, i2 = e - rr * g.n2;
    fb[(size_t)rr * 2 * g.ls + zreal(g, i2)] = src[e];
  }
  __syncthreads();
  fft_lds<-1>(buf, pz, nr, g.ls, 1);
  r2c_finish(g, buf, nr, tw2);
  float2* dst = out + (size_t)i0 * g.M + (size_t)r0 * g.h2;
  for (int e = threadIdx.x; e < nr * g.h2; e += blockDim.x) {
    const int rr = e / g.h2, k2 = e - rr * g.h2;
    dst[e] = buf[rr * g.ls + k2];
  }
}

// Inverse rows + per-plane top-K.  grid (n0, nrot), LDS zrows * ls complex.
// The plane's K best (value, offset) pairs are kept sorted in LDS while the
// chunks of rows stream through.
__global__ __launch_bounds__(kSplitThreads) void k_inv_rows_topk(ScanGeom g, FftPlan pz,
                                                                 const float2* __restrict__ tw2,
                                                                 const float2* __restrict__ S1, float scale,
                                                                 int K, float* __restrict__ cval,
                                                                 int* __restrict__ coff,
                                                                 float* __restrict__ grid_out) {
  extern __shared__ float2 buf[];
  __shared__ float red_v[16];
  __shared__ int red_o[16];
  __shared__ float lv[32];
  __shared__ int lo[32];
  const float INF = __int_as_float(0x7f800000);
  const int i0 = blockIdx.x, b = blockIdx.y;
  const float2* src = S1 + ((size_t)b * g.n0 + i0) * g.M;
  const int npl = g.n1 * g.n2;
  const int rs = 2 * g.ls;
  float* fb = reinterpret_cast<float*>(buf);
  if (threadIdx.x < 32) {
    lv[threadIdx.x] = INF;
    lo[threadIdx.x] = INT_MAX;
  }
  for (int r0 = 0; r0 < g.n1; r0 += g.zrows) {
    const int nr = min(g.zrows, g.n1 - r0);
    __syncthreads();
    for (int e = threadIdx.x; e < nr * g.h2; e += blockDim.x) {
      const int rr = e / g.h2, k2 = e - rr * g.h2;
      buf[rr * g.ls + k2] = src[(size_t)r0 * g.h2 + e];
    }
    __syncthreads();
    c2r_start(g, buf, nr, tw2);
    fft_lds<+1>(buf, pz, nr, g.ls, 1);
    float bv = INF;
    int be = INT_MAX;
    for (int e = threadIdx.x; e < nr * g.n2; e += blockDim.x) {
      const int rr = e / g.n2, i2 = e - rr * g.n2;
      const int o = r0 * g.n2 + e;
      float v = fb[rr * rs + zreal(g, i2)] * scale;
      if (grid_out) grid_out[((size_t)b * g.n0 + i0) * npl + o] = v;
      if (v != v) v = INF;
      fb[rr * rs + zreal(g, i2)] = v;
      if (pd_better(v, o, bv, be)) { bv = v; be = o; }
    }
    __syncthreads();
    // Move this chunk's cells that beat the list's K-th entry into the list.
    for (;;) {
      float v = bv;
      int o = be;
      block_argmin(v, o, red_v, red_o);
      if (o == INT_MAX || !pd_better(v, o, lv[K - 1], lo[K - 1])) break;
      if (threadIdx.x == 0) {
        int k = K - 1;
        while (k > 0 && pd_better(v, o, lv[k - 1], lo[k - 1])) {
          lv[k] = lv[k - 1];
          lo[k] = lo[k - 1];
          --k;
        }
        lv[k] = v;
        lo[k] = o;
      }
      if (o == be) {
        const int rr = o / g.n2 - r0, i2 = o % g.n2;
        fb[rr * rs + zreal(g, i2)] = INF;
        bv = INF;
        be = INT_MAX;
        for (int e = threadIdx.x; e < nr * g.n2; e += blockDim.x) {
          const int r2 = e / g.n2, j2 = e - r2 * g.n2;
          const float x = fb[r2 * rs + zreal(g, j2)];
          const int oe = r0 * g.n2 + e;
          if (x != INF && pd_better(x, oe, bv, be)) { bv = x; be = oe; }
        }
      }
      __syncthreads();
    }
  }
  __syncthreads();
  if (threadIdx.x < K) {
    const size_t at = ((size_t)b * g.n0 + i0) * K + threadIdx.x;
    cval[at] = lv[threadIdx.x];
    coff[at] = lo[threadIdx.x] == INT_MAX ? -1 : i0 * npl + lo[threadIdx.x];
  }
}

// ---------------------------------------------------------------------------
// Inverse rows + top-K with register FFTs (N2 known at compile time).  grid
// (nrot, n0): the rotation varies fastest, as in the whole-plane kernel, so
// that the rotation-wide threshold thr[b] tightens while the planes run.
// Chunks of LZ = 512 / B2 rows: c2r pre-step on loading from S1, half-length
// backward four-step FFT (one LDS exchange), scale, then per wave the K best
// under the threshold; a block-wide running list of the K best carries over
// the chunks.  LDS LZ * LS complex.
// ---------------------------------------------------------------------------

template <int N2>
__global__ __launch_bounds__(512) void k_inv_rows_topk_fast(ScanGeom g, const float2* __restrict__ twM2,
                                                            const float2* __restrict__ tw2,
                                                            const float2* __restrict__ S1, float scale, int K,
                                                            float* __restrict__ cval, int* __restrict__ coff,
                                                            float* __restrict__ grid_out,
                                                            unsigned* __restrict__ thr) {
  constexpr int M2 = N2 / 2, H2 = M2 + 1, LS = H2 | 1, NT = 512, NW = NT / 64, L = 4;
  using SZ = Split<M2>;
  constexpr int A2 = SZ::A, B2 = SZ::B, J2 = SZ::J;
  constexpr int LZ = NT / B2;
  extern __shared__ float2 buf[];
  __shared__ uint64_t wtop[NW][32];
  __shared__ uint64_t best[32];
  __shared__ uint32_t s_thr;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int b = blockIdx.x, i0 = blockIdx.y;
  const int N1 = g.n1;
  const float2* src = S1 + ((size_t)b * g.n0 + i0) * (size_t)(N1 * H2);
  const float INF = __int_as_float(0x7f800000);
  const int plane = N1 * N2, base = i0 * plane;
  const int gz = tid / LZ, rz = tid % LZ;
  static_assert(LZ % 64 == 0, "one lane group per wave");
  const int gzu = __builtin_amdgcn_readfirstlane(gz);  // wave-uniform twiddle reads
  // best and s_thr are first read after the barriers of the first chunk; the
  // threshold read goes out after that chunk's loads so its latency overlaps
  // theirs.
  if (tid < 32) best[tid] = kNoKey;
  for (int r0 = 0; r0 < N1; r0 += LZ) {
    const int row = r0 + rz;
    const bool rok = row < N1;
    float2 z[A2];
#pragma unroll
    for (int m = 0; m < A2; ++m) {
      const int k = gz + B2 * m;
      float2 xk = make_float2(0.f, 0.f), xm = xk;
      if (rok) {
        xk = src[row * H2 + k];
        xm = src[row * H2 + M2 - k];
      }
      if (k == 0) { xk.y = 0.f; xm.y = 0.f; }
      z[m] = c2r_point(xk, xm, tw2[gzu + B2 * m]);
    }
    if (r0 == 0 && tid == 0)
      s_thr = thr ? __hip_atomic_load(thr + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0xFFFFFFFFu;
    four_step_phase1<A2, B2, +1>(z, gzu, twM2);
    if (rok)
#pragma unroll
      for (int k1 = 0; k1 < A2; ++k1) buf[rz * LS + k1 * B2 + gz] = z[k1];
    __syncthreads();
    float2 u[J2][B2];
#pragma unroll
    for (int j = 0; j < J2; ++j) {
      const int k1 = gz + B2 * j;
      if (rok && k1 < A2) {
#pragma unroll
        for (int q = 0; q < B2; ++q) u[j][q] = buf[rz * LS + k1 * B2 + q];
        rdft<B2, +1>(u[j]);
      }
    }
    __syncthreads();
    float lmin = INF;
#pragma unroll
    for (int j = 0; j < J2; ++j) {
      const int k1 = gz + B2 * j;
      if (rok && k1 < A2) {
#pragma unroll
        for (int k2 = 0; k2 < B2; ++k2) {
          const int e = row * N2 + 2 * (k1 + A2 * k2);
          float vx = u[j][k2].x * scale, vy = u[j][k2].y * scale;
          if (grid_out) {
            float* go = grid_out + ((size_t)b * g.n0 + i0) * (size_t)plane;
            go[e] = vx;
            go[e + 1] = vy;
          }
          if (vx != vx) vx = INF;
          if (vy != vy) vy = INF;
          u[j][k2] = make_float2(vx, vy);
          lmin = fminf(lmin, fminf(vx, vy));
        }
      }
    }
    // threshold: the rotation's (read once) and this plane's running K-th best
    const uint64_t bk = best[K - 1];
    float tval = s_thr == 0xFFFFFFFFu ? INF : pd_key_value((uint64_t)s_thr << 32);
    if (bk != kNoKey) tval = fminf(tval, pd_key_value(bk));
    if (__ballot(lmin <= tval) == 0ull) {
      if (lane < K) wtop[wave][lane] = kNoKey;
    } else {
      const float tau = fminf(wave_kth_smallest(lmin, K), tval);
      uint64_t h[L];
#pragma unroll
      for (int i = 0; i < L; ++i) h[i] = kNoKey;
#pragma unroll
      for (int j = 0; j < J2; ++j) {
        const int k1 = gz + B2 * j;
        if (rok && k1 < A2) {
#pragma unroll
          for (int k2 = 0; k2 < B2; ++k2) {
            const int e = base + row * N2 + 2 * (k1 + A2 * k2);
            const float2 x = u[j][k2];
            if (x.x <= tau) list_insert(h, pd_key(x.x, e));
            if (x.y <= tau) list_insert(h, pd_key(x.y, e + 1));
          }
        }
      }
      uint64_t last = 0;
      int popped = 0;
      for (int rk = 0; rk < K; ++rk) {
        const uint64_t m = wave_min_u64(h[0]);
        if (lane == 0) wtop[wave][rk] = m;
        if (m == kNoKey) {
          if (lane > rk && lane < K) wtop[wave][lane] = kNoKey;
          break;
        }
        if (h[0] == m) {
          last = m;
          list_pop(h);
          ++popped;
        }
        if (h[0] == kNoKey && popped > 0 && popped % L == 0) {
#pragma unroll
          for (int j = 0; j < J2; ++j) {
            const int k1 = gz + B2 * j;
            if (rok && k1 < A2) {
#pragma unroll
              for (int k2 = 0; k2 < B2; ++k2) {
                const int e = base + row * N2 + 2 * (k1 + A2 * k2);
                const uint64_t kx = pd_key(u[j][k2].x, e), ky = pd_key(u[j][k2].y, e + 1);
                if (kx > last) list_insert(h, kx);
                if (ky > last) list_insert(h, ky);
              }
            }
          }
        }
      }
    }
    __syncthreads();
    // wave 0 merges the waves' candidates with the running list
    if (wave == 0) {
      uint64_t hh[5];
#pragma unroll
      for (int i = 0; i < 5; ++i) hh[i] = kNoKey;
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        const int idx = lane + 64 * q;
        if (idx < NW * K)
          list_insert(hh, wtop[idx / K][idx % K]);
        else if (idx < NW * K + K)
          list_insert(hh, best[idx - NW * K]);
      }
      for (int rk = 0; rk < K; ++rk) {
        const uint64_t m = wave_min_u64(hh[0]);
        if (lane == 0) best[rk] = m;
        if (m != kNoKey && hh[0] == m) list_pop(hh);
      }
    }
    __syncthreads();
  }
  if (tid < K) {
    const uint64_t m = best[tid];
    const size_t at = ((size_t)b * g.n0 + i0) * K + tid;
    cval[at] = m == kNoKey ? INF : pd_key_value(m);
    coff[at] = m == kNoKey ? -1 : (int)(uint32_t)m;
  }
  if (thr && tid == 0 && best[K - 1] != kNoKey) atomicMin(thr + b, (unsigned)(best[K - 1] >> 32));
}

// ---------------------------------------------------------------------------
// Inverse y pass, inverse rows and top-K in one kernel, for N1 = 256 and
// N2 = 256: the x plane of S1 is read from HBM once and stays on chip.
// grid (nrot, n0), block 1024 = 16 waves; lane c and wave r.
//   1. Loads: for the column sets s (kz = c + 64 s, s < M2 / 64), the lane
//      holds rows r + 16 m (m < 16); each wave load reads 512 contiguous
//      bytes of one row.  Column kz = M2 (one more than the lanes hold) is
//      transformed by lanes 0..15 of wave 0 alone, with the same four-step
//      (exchange through a 2 KiB LDS array); its real part goes to LDS.
//   2. Backward y FFT, four-step 16 x 16: DFT over m, twiddle W_256^{r k1}
//      (wave-uniform), one exchange of the set through LDS (128 KiB), DFT
//      over the 16 waves.  The lane then holds rows r + 16 k2.
//   3. Per half of the rows (128 rows x M2 + 1 complex, 129 KiB of LDS): the
//      lanes write their rows of the half, then the rows take the c2r
//      pre-step, the half-length backward four-step FFT (LZ = 128 rows of 8
//      lanes, exchange in place), the scale, and the K-best selection under
//      the rotation's threshold with a block-wide running list, as in
//      k_inv_rows_topk_fast.
// Same arithmetic, in the same order, as k_cols_fast<N1, +1> followed by
// k_inv_rows_topk_fast<N2>, so the correlation values are the same bits.
// LDS: max(16 * 16 * 64, 128 * (M2 + 1)) complex + lists.
// ---------------------------------------------------------------------------
template <int N1, int N2>
__global__ __launch_bounds__(1024) void k_inv_plane_topk_split(ScanGeom g, const float2* __restrict__ twN1,
                                                               const float2* __restrict__ twM2,
                                                               const float2* __restrict__ tw2,
                                                               const float2* __restrict__ S1, float scale, int K,
                                                               float* __restrict__ cval, int* __restrict__ coff,
                                                               float* __restrict__ grid_out,
                                                               unsigned* __restrict__ thr) {
  constexpr int NT = 1024, NW = NT / 64, L = 4;
  constexpr int A1 = 16, B1 = NW;  // y four-step: 16 values per lane x 16 waves
  static_assert(A1 * B1 == N1, "y split");
  constexpr int M2 = N2 / 2, H2 = M2 + 1, LS = H2 | 1, NS = M2 / 64;
  static_assert(M2 % 64 == 0, "whole column sets");
  using SZ = Split<M2>;
  constexpr int A2 = SZ::A, B2 = SZ::B, J2 = SZ::J;
  constexpr int LZ = NT / B2, NH = N1 / LZ;
  static_assert(N1 % LZ == 0 && LZ % 64 == 0 && A1 % NH == 0, "halves of whole rows");
  constexpr int KH = A1 / NH;  // k2 values per half
  extern __shared__ float2 buf[];
  __shared__ uint64_t wtop[NW][32];
  __shared__ uint64_t best[32];
  __shared__ uint32_t s_thr;
  __shared__ int s_any[NH];
  __shared__ float2 s_nyq[N1];  // exchange of the kz = M2 column's FFT
  __shared__ float s_g[N1];     // that column after the y FFT (real part)
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int b = blockIdx.x, i0 = blockIdx.y;
  const float2* src = S1 + ((size_t)b * g.n0 + i0) * (size_t)(N1 * H2);
  const float INF = __int_as_float(0x7f800000);
  const int plane = N1 * N2, base = i0 * plane;
  float2 x[NS][A1];
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int m = 0; m < A1; ++m) x[s][m] = src[(wave + B1 * m) * H2 + lane + 64 * s];
  if (wave == 0 && lane < B1) {
    // column kz = M2: the same four-step on lanes 0..15 of one wave, lane r
    // holding rows r + 16 m; the exchange stays within the wave
    float2 v[A1];
#pragma unroll
    for (int m = 0; m < A1; ++m) v[m] = src[(lane + B1 * m) * H2 + M2];
    rdft<A1, +1>(v);
#pragma unroll
    for (int k1 = 1; k1 < A1; ++k1) {
      float2 w = twN1[lane * k1];
      w.y = -w.y;
      v[k1] = cmul(v[k1], w);
    }
#pragma unroll
    for (int k1 = 0; k1 < A1; ++k1) s_nyq[k1 * B1 + lane] = v[k1];
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < B1; ++q) v[q] = s_nyq[lane * B1 + q];
    rdft<B1, +1>(v);
#pragma unroll
    for (int k2 = 0; k2 < B1; ++k2) s_g[lane + A1 * k2] = v[k2].x;
  }
  if (tid < 32) best[tid] = kNoKey;
  if (tid < NH) s_any[tid] = 0;
  if (tid == 0)
    s_thr = thr ? __hip_atomic_load(thr + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0xFFFFFFFFu;
  // backward y FFT of each set
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    rdft<A1, +1>(x[s]);
#pragma unroll
    for (int k1 = 1; k1 < A1; ++k1) {
      float2 w = twN1[wave * k1];
      w.y = -w.y;
      x[s][k1] = cmul(x[s][k1], w);
    }
    if (s > 0) __syncthreads();  // the previous set's reads are done
#pragma unroll
    for (int k1 = 0; k1 < A1; ++k1) buf[(k1 * B1 + wave) * 64 + lane] = x[s][k1];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < B1; ++q) x[s][q] = buf[(wave * B1 + q) * 64 + lane];
    rdft<B1, +1>(x[s]);  // x[s][k2] is row wave + 16 k2
    // finish the DFT here: held as final values the rows of the second half
    // take 2 registers each, as shared butterfly operands they take more
#pragma unroll
    for (int k = 0; k < A1; ++k) asm volatile("" : "+v"(x[s][k].x), "+v"(x[s][k].y));
  }
  const int gz = tid / LZ, rz = tid % LZ;
  const int gzu = __builtin_amdgcn_readfirstlane(gz);
#pragma unroll
  for (int hh = 0; hh < NH; ++hh) {
    __syncthreads();  // LDS reads of the previous phase are done
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
      for (int k = 0; k < KH; ++k) {
        const int yl = wave + B1 * k;
        const float2 v = x[s][KH * hh + k];
        buf[yl * LS + lane + 64 * s] = v;
        if (s == 0 && lane == 0) buf[yl * LS + M2] = make_float2(s_g[LZ * hh + yl], 0.f);
      }
    __syncthreads();
    const int row = LZ * hh + rz;
    float2 z[A2];
#pragma unroll
    for (int m = 0; m < A2; ++m) {
      const int k = gz + B2 * m;
      float2 xk = buf[rz * LS + k], xm = buf[rz * LS + M2 - k];
      if (k == 0) { xk.y = 0.f; xm.y = 0.f; }
      z[m] = c2r_point(xk, xm, tw2[gzu + B2 * m]);
    }
    four_step_phase1<A2, B2, +1>(z, gzu, twM2);
    __syncthreads();  // every row of the half is read before the exchange
#pragma unroll
    for (int k1 = 0; k1 < A2; ++k1) buf[rz * LS + k1 * B2 + gz] = z[k1];
    __syncthreads();
    // second DFT of each k1; the results go back to the slots the lane read
    // (no other lane reads them), and the selection reads them from there,
    // which keeps the registers for the rows of the next half
    float lmin = INF;
#pragma unroll
    for (int j = 0; j < J2; ++j) {
      const int k1 = gz + B2 * j;
      if (k1 < A2) {
        float2 u[B2];
#pragma unroll
        for (int q = 0; q < B2; ++q) u[q] = buf[rz * LS + k1 * B2 + q];
        rdft<B2, +1>(u);
#pragma unroll
        for (int k2 = 0; k2 < B2; ++k2) {
          const int e = row * N2 + 2 * (k1 + A2 * k2);
          float vx = u[k2].x * scale, vy = u[k2].y * scale;
          if (grid_out) {
            float* go = grid_out + ((size_t)b * g.n0 + i0) * (size_t)plane;
            go[e] = vx;
            go[e + 1] = vy;
          }
          if (vx != vx) vx = INF;
          if (vy != vy) vy = INF;
          buf[rz * LS + k1 * B2 + k2] = make_float2(vx, vy);
          lmin = fminf(lmin, fminf(vx, vy));
        }
      }
    }
    asm volatile("" ::: "memory");  // read the values back from LDS below
    const uint64_t bk = best[K - 1];
    float tval = s_thr == 0xFFFFFFFFu ? INF : pd_key_value((uint64_t)s_thr << 32);
    if (bk != kNoKey) tval = fminf(tval, pd_key_value(bk));
    if (__ballot(lmin <= tval) == 0ull) {
      if (lane < K) wtop[wave][lane] = kNoKey;
    } else {
      if (lane == 0) s_any[hh] = 1;
      const float tau = fminf(wave_kth_smallest(lmin, K), tval);
      uint64_t h[L];
#pragma unroll
      for (int i = 0; i < L; ++i) h[i] = kNoKey;
#pragma unroll
      for (int j = 0; j < J2; ++j) {
        const int k1 = gz + B2 * j;
        if (k1 < A2) {
#pragma unroll
          for (int k2 = 0; k2 < B2; ++k2) {
            const int e = base + row * N2 + 2 * (k1 + A2 * k2);
            const float2 v = buf[rz * LS + k1 * B2 + k2];
            if (v.x <= tau) list_insert(h, pd_key(v.x, e));
            if (v.y <= tau) list_insert(h, pd_key(v.y, e + 1));
          }
        }
      }
      uint64_t last = 0;
      int popped = 0;
      for (int rk = 0; rk < K; ++rk) {
        const uint64_t m = wave_min_u64(h[0]);
        if (lane == 0) wtop[wave][rk] = m;
        if (m == kNoKey) {
          if (lane > rk && lane < K) wtop[wave][lane] = kNoKey;
          break;
        }
        if (h[0] == m) {
          last = m;
          list_pop(h);
          ++popped;
        }
        if (h[0] == kNoKey && popped > 0 && popped % L == 0) {
          // rescan (rare): a rolled loop, so that the values are not held in
          // registers next to the rows of the next half
#pragma unroll
          for (int j = 0; j < J2; ++j) {
            const int k1 = gz + B2 * j;
            if (k1 < A2) {
#pragma unroll 1
              for (int k2 = 0; k2 < B2; ++k2) {
                const int e = base + row * N2 + 2 * (k1 + A2 * k2);
                const float2 v = buf[rz * LS + k1 * B2 + k2];
                const uint64_t kx = pd_key(v.x, e), ky = pd_key(v.y, e + 1);
                if (kx > last) list_insert(h, kx);
                if (ky > last) list_insert(h, ky);
              }
            }
          }
        }
      }
    }
    __syncthreads();
    // wave 0 merges the waves' candidates with the running list, when any
    // wave had one under the threshold
    if (wave == 0 && s_any[hh]) {
      constexpr int NQ = (NW * 32 + 32 + 63) / 64;
      uint64_t hq[NQ];
#pragma unroll
      for (int i = 0; i < NQ; ++i) hq[i] = kNoKey;
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int idx = lane + 64 * q;
        if (idx < NW * K)
          list_insert(hq, wtop[idx / K][idx % K]);
        else if (idx < NW * K + K)
          list_insert(hq, best[idx - NW * K]);
      }
      for (int rk = 0; rk < K; ++rk) {
        const uint64_t m = wave_min_u64(hq[0]);
        if (m == kNoKey) break;  // best[rk..] hold no key already
        if (lane == 0) best[rk] = m;
        if (hq[0] == m) list_pop(hq);
      }
    }
  }
  __syncthreads();
  if (tid < K) {
    const uint64_t m = best[tid];
    const size_t at = ((size_t)b * g.n0 + i0) * K + tid;
    cval[at] = m == kNoKey ? INF : pd_key_value(m);
    coff[at] = m == kNoKey ? -1 : (int)(uint32_t)m;
  }
  if (thr && tid == 0 && best[K - 1] != kNoKey) atomicMin(thr + b, (unsigned)(best[K - 1] >> 32));
}

// ---------------------------------------------------------------------------
// Ligand plane pass in one kernel for N1 = 256 and N2 = 256 with at most 128
// box rows: the plane's spectrum is written once, as whole rows.
// grid (p0 padded to 16, nrot, T), block 1024 = 16 waves; one term per block:
//   1. the z spectrum of each box row by direct summation over the row's
//      atoms, X[row][kz] = sum_j w_j exp(-2 pi i kz z_j / N2) for kz = 0..M2,
//      in ascending atom order (the plane's atoms are listed by row, then
//      each row is put in atom order), one row per wave, lane c taking
//      columns c + 64 q; X[0] and X[M2] are real.  A plane with more atoms
//      than the list holds is summed from the atom cells directly (same
//      order, slower);
//   2. the lanes load their column sets (lane c, wave r: kz = c + 64 s, rows
//      r + 16 m, zero outside the box), forward y FFT four-step 16 x 16 with
//      one LDS exchange per set; a wave's stores cover 512 contiguous bytes
//      of a row (plane-major) or 32-byte pieces of the column pass's
//      interleaved order (g.il), whose 16-plane groups run on one XCD so that
//      its L2 merges the pieces into whole lines.  Column kz = M2 is
//      transformed by lanes 0..15 of wave 0 alone.
// LDS: 128 * LS complex (the box rows, then the y exchange), the twiddles and
// the atom list.
// ---------------------------------------------------------------------------
constexpr int kPlaneSplitRows = 128;  // box rows the one-kernel ligand pass holds
constexpr int kPlaneSplitList = 1536;  // atoms of one plane the one-kernel ligand pass lists

template <int N1, int N2>
__global__ __launch_bounds__(1024) void k_lig_plane_split(ScanGeom g, const float2* __restrict__ twN1,
                                                          const float2* __restrict__ tw2,
                                                          const int* __restrict__ cells, int n_atoms,
                                                          const float* __restrict__ w, float2* __restrict__ L1) {
  constexpr int NT = 1024, NW = NT / 64;
  constexpr int A1 = 16, B1 = NW;
  static_assert(A1 * B1 == N1, "y split");
  constexpr int M2 = N2 / 2, H2 = M2 + 1, LS = H2 | 1, NS = M2 / 64;
  static_assert(M2 % 64 == 0, "whole column sets");
  extern __shared__ float2 buf[];
  __shared__ float2 s_nyq[N1];
  __shared__ float2 s_tw[N2];
  __shared__ int s_row[kPlaneSplitRows + 1];  // row r's atoms: [s_row[r], s_row[r + 1])
  __shared__ int s_cur[kPlaneSplitRows];
  __shared__ int s_z[kPlaneSplitList];
  __shared__ int s_ra[kPlaneSplitList];
  __shared__ float s_w[kPlaneSplitList];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  // blocks go to the 8 XCDs round-robin in launch order: block L (on XCD
  // L % 8) takes item ((L / 128) * 8 + L % 8) * 16 + (L / 8) % 16 of the
  // plane-fastest items, so each group of 16 planes stays on one XCD (a last
  // partial group of 128 blocks keeps item L)
  const int PX = gridDim.x, Lb = blockIdx.x + PX * (blockIdx.y + gridDim.y * blockIdx.z);
  const int nblk = PX * gridDim.y * gridDim.z;
  const int item = Lb < nblk / 128 * 128 ? ((Lb >> 7) * 8 + (Lb & 7)) * 16 + ((Lb >> 3) & 15) : Lb;
  const int p = item % PX, b = (item / PX) % gridDim.y, t = item / (PX * gridDim.y);
  if (p >= g.p[0]) return;  // padding plane: the whole block leaves before any barrier
  const int i0 = g.lo[0] + p, lo1 = g.lo[1], p1 = g.p[1];
  const int plane = N1 * N2;
  const int* cb = cells + (size_t)b * n_atoms;
  const float* wt = w + (size_t)t * n_atoms;
  for (int i = tid; i < N2; i += NT) s_tw[i] = tw2[i];
  for (int i = tid; i <= p1; i += NT) s_row[i] = 0;
  __syncthreads();
  // box row * N2 + z of atom j if it lies in this plane, else -1
  auto key_of = [&](int j) -> int {
    const int off = cb[j];
    if (off < 0 || off / plane != i0) return -1;
    const int r = off - i0 * plane;
    const int y = r / N2;
    return (y - lo1) * N2 + (r - y * N2);
  };
  for (int j = tid; j < n_atoms; j += NT) {
    const int k = key_of(j);
    if (k >= 0) atomicAdd(&s_row[k / N2 + 1], 1);
  }
  __syncthreads();
  if (tid == 0)
    for (int i = 1; i <= p1; ++i) s_row[i] += s_row[i - 1];
  __syncthreads();
  const bool listed = s_row[p1] <= kPlaneSplitList;
  if (listed) {
    for (int i = tid; i < p1; i += NT) s_cur[i] = s_row[i];
    __syncthreads();
    for (int j = tid; j < n_atoms; j += NT) {
      const int k = key_of(j);
      if (k >= 0) {
        const int r = k / N2;
        const int pos = atomicAdd(&s_cur[r], 1);
        s_z[pos] = k - r * N2;
        s_ra[pos] = j;
        s_w[pos] = wt[j];
      }
    }
    __syncthreads();
    // each row in atom order (insertion sort of its few entries)
    for (int r = tid; r < p1; r += NT) {
      const int a0 = s_row[r], a1 = s_row[r + 1];
      for (int i = a0 + 1; i < a1; ++i) {
        const int ja = s_ra[i], za = s_z[i];
        const float wa = s_w[i];
        int q = i - 1;
        while (q >= a0 && s_ra[q] > ja) {
          s_ra[q + 1] = s_ra[q];
          s_z[q + 1] = s_z[q];
          s_w[q + 1] = s_w[q];
          --q;
        }
        s_ra[q + 1] = ja;
        s_z[q + 1] = za;
        s_w[q + 1] = wa;
      }
    }
    __syncthreads();
  }
  {
    constexpr int NQ = (H2 + 63) / 64;  // columns per lane
    for (int r = wave; r < p1; r += NW) {
      float2 a[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) a[q] = make_float2(0.f, 0.f);
      auto add = [&](int z, float wv) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          const int c = lane + 64 * q;
          const float2 ph = s_tw[(c * z) % N2];
          a[q].x = fmaf(wv, ph.x, a[q].x);
          a[q].y = fmaf(wv, c == 0 || c == M2 ? 0.f : ph.y, a[q].y);
        }
      };
      if (listed) {
        const int a0 = s_row[r], a1 = s_row[r + 1];
        for (int i = a0; i < a1; ++i) add(s_z[i], s_w[i]);
      } else {
        for (int j = 0; j < n_atoms; ++j) {
          const int k = key_of(j);
          if (k >= 0 && k / N2 == r) add(k - r * N2, wt[j]);
        }
      }
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int c = lane + 64 * q;
        if (c <= M2) buf[r * LS + c] = a[q];
      }
    }
  }
  __syncthreads();
  // stores through a descriptor over the (rotation, term)'s planes, 32-bit
  // offsets; column m of plane q sits at l1_off(g, q, m)
  constexpr int PM = N1 * H2;
  const __amdgpu_buffer_rsrc_t rd = pd_rsrc(L1 + ((size_t)b * g.T + t) * g.pp * (size_t)PM, (uint32_t)g.pp * PM * 8u);
  const int q = g.lo[0] - g.pl + p;
  const uint32_t qb = g.il ? (uint32_t)((q / g.il) * (PM >> g.ilsh) * 64 + (q % g.il) * (64 / g.il)) : (uint32_t)q * PM;
  const uint32_t cmask = g.il ? (64u / g.il) - 1u : 0u;
  auto off8 = [&](int m) -> uint32_t {  // byte offset of column m of this plane
    return g.il ? (qb + ((uint32_t)(m >> g.ilsh) << 6) + ((uint32_t)m & cmask)) * 8u : (qb + (uint32_t)m) * 8u;
  };
  if (wave == 0 && lane < B1) {
    // column kz = M2: the same four-step on 16 lanes, the exchange inside the
    // wave
    float2 v[A1];
#pragma unroll
    for (int m = 0; m < A1; ++m) {
      const int rr = lane + B1 * m - lo1;
      v[m] = (rr >= 0 && rr < p1) ? buf[rr * LS + M2] : make_float2(0.f, 0.f);
    }
    four_step_phase1<A1, B1, -1>(v, lane, twN1);
#pragma unroll
    for (int k1 = 0; k1 < A1; ++k1) s_nyq[k1 * B1 + lane] = v[k1];
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < B1; ++q) v[q] = s_nyq[lane * B1 + q];
    rdft<B1, -1>(v);
#pragma unroll
    for (int k2 = 0; k2 < B1; ++k2) pd_st2(rd, off8((lane + A1 * k2) * H2 + M2), 0u, v[k2]);
  }
  float2 x[NS][A1];
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int m = 0; m < A1; ++m) {
      const int rr = wave + B1 * m - lo1;  // wave-uniform
      x[s][m] = (rr >= 0 && rr < p1) ? buf[rr * LS + lane + 64 * s] : make_float2(0.f, 0.f);
    }
  __syncthreads();  // the rows are read: the buffer takes the y exchange
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    rdft<A1, -1>(x[s]);
#pragma unroll
    for (int k1 = 1; k1 < A1; ++k1) x[s][k1] = cmul(x[s][k1], twN1[wave * k1]);
    if (s > 0) __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < A1; ++k1) buf[(k1 * B1 + wave) * 64 + lane] = x[s][k1];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < B1; ++q) x[s][q] = buf[(wave * B1 + q) * 64 + lane];
    rdft<B1, -1>(x[s]);
#pragma unroll
    for (int k2 = 0; k2 < B1; ++k2) pd_st2(rd, off8((wave + A1 * k2) * H2 + lane + 64 * s), 0u, x[s][k2]);
  }
}
