// scan_fast.hpp -- the three hot scan kernels specialised on grid edge.
//
// Same dataflow as the generic kernels in scan_kernels.hpp (ligand plane pass,
// column pass, inverse plane pass + top-K), but every 1D transform is a
// compile-time four-step FFT held in registers (reg_fft.hpp): B lanes per
// line, one LDS exchange per transform, inputs loaded straight from HBM into
// registers and outputs stored straight from registers.
#pragma once
#include <type_traits>

#include "reg_fft.hpp"
#include "scan_kernels.hpp"

// Minimum waves per SIMD requested from the register allocator.
#ifndef PD_COL_WAVES
#define PD_COL_WAVES 1
#endif
#ifndef PD_PLANE_WAVES
#define PD_PLANE_WAVES 1
#endif
#ifndef PD_LIG_WAVES
#define PD_LIG_WAVES PD_PLANE_WAVES
#endif

// N = A * B with a given B.
template <int N, int BB>
struct SplitB {
  static constexpr int B = BB, A = N / BB, J = (A + BB - 1) / BB;
};

// N = A * B with B lanes per line; B = 0 marks an unsupported edge.
template <int N>
struct Split {
  static constexpr int B = (N % 8 == 0 && N / 8 <= 20) ? 8 : (N % 16 == 0 && N / 16 <= 16) ? 16
                          : (N % 4 == 0 && N / 4 <= 20) ? 4 : 0;
  static constexpr int A = B ? N / B : 0;
  static constexpr int J = B ? (A + B - 1) / B : 0;
};

// Lanes per column of the column pass: 16 lanes (N = 16 A, 4 <= A <= 16) where
// the split of the plane passes would give each lane three or more rounds of
// B values (144, 160: 8 lanes x 18 or 20 values), or everywhere with
// PD_COL_B16.  Measured: 16 lanes take the 160^3 column pass from 20.2 to
// 14.2 us per rotation and 144^3 from 15.9 to 10.9, but slow 64^3-128^3
// (128^3: 161 -> 242 ms per 30,000 rotations).
#ifndef PD_COL_B16
#define PD_COL_B16 0
#endif
template <int N>
struct ColSplit {
  static constexpr bool W =
      (PD_COL_B16 || Split<N>::J >= 3) && N % 16 == 0 && N / 16 >= 4 && N / 16 <= 16;
  static constexpr int B = W ? 16 : Split<N>::B;
  static constexpr int A = W ? N / 16 : Split<N>::A;
  static constexpr int J = B ? (A + B - 1) / B : 0;
};

// Raw buffer access: a per-wave descriptor (base in SGPRs), a 32-bit per-lane
// byte offset and a uniform byte offset, so no 64-bit address lives in VGPRs.
// Offsets must stay below the descriptor's byte count (hardware range check).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t pd_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float2 pd_ld2(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ float4 pd_ld4(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ void pd_st2(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, float2 v) {
  using I2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(I2, v), r, (int)voff, (int)soff, 0);
}

// ---------------------------------------------------------------------------
// Top-K helpers.  A candidate is a 64-bit key (order-preserving bits of the
// FP32 value, offset): unsigned key order is the (value, offset) order, i.e.
// lower energy first and, among equal energies, the earlier row-major cell
// (the cell the reference's bounded heap keeps, Common/queue.hpp:233-239).
// ---------------------------------------------------------------------------
constexpr uint64_t kNoKey = ~0ull;

__device__ __forceinline__ uint64_t pd_key(float v, int off) {
  uint32_t b = __float_as_uint(v);
  // -0.0 and +0.0 are equal to the reference's operator<: one key, so that
  // among them the earlier cell wins as among any equal values
  if (b == 0x80000000u) b = 0u;
  b = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  return ((uint64_t)b << 32) | (uint32_t)off;
}
__device__ __forceinline__ float pd_key_value(uint64_t k) {
  uint32_t b = (uint32_t)(k >> 32);
  b = (b & 0x80000000u) ? (b & 0x7fffffffu) : ~b;
  return __uint_as_float(b);
}

// One DPP step of a wave minimum: lanes whose source lane is out of range
// keep their own value (old operand = own value).
template <int CTRL, int ROWMASK>
__device__ __forceinline__ uint64_t dpp_min_step(uint64_t k) {
  const uint32_t lo = (uint32_t)k, hi = (uint32_t)(k >> 32);
  const uint32_t l2 = (uint32_t)__builtin_amdgcn_update_dpp((int)lo, (int)lo, CTRL, ROWMASK, 0xf, false);
  const uint32_t h2 = (uint32_t)__builtin_amdgcn_update_dpp((int)hi, (int)hi, CTRL, ROWMASK, 0xf, false);
  const uint64_t o = ((uint64_t)h2 << 32) | l2;
  return o < k ? o : k;
}

// Minimum of a 64-bit key over the 64 lanes of the wave, returned to every lane:
// prefix minima within rows (row_shr 1,2,4,8), then row_bcast 15 and 31
// carry them across rows so lane 63 holds the wave minimum.
__device__ __forceinline__ uint64_t wave_min_u64(uint64_t k) {
  k = dpp_min_step<0x111, 0xf>(k);
  k = dpp_min_step<0x112, 0xf>(k);
  k = dpp_min_step<0x114, 0xf>(k);
  k = dpp_min_step<0x118, 0xf>(k);
  k = dpp_min_step<0x142, 0xa>(k);
  k = dpp_min_step<0x143, 0xc>(k);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)k, 63);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(k >> 32), 63);
  return ((uint64_t)hi << 32) | lo;
}

// Minimum of an FP32 value over the wave (same DPP pattern), to every lane.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_fmin_step(float x) {
  const float y = __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(x), __float_as_int(x), CTRL, ROWMASK, 0xf, false));
  return fminf(x, y);
}
__device__ __forceinline__ float wave_min_f32(float x) {
  x = dpp_fmin_step<0x111, 0xf>(x);
  x = dpp_fmin_step<0x112, 0xf>(x);
  x = dpp_fmin_step<0x114, 0xf>(x);
  x = dpp_fmin_step<0x118, 0xf>(x);
  x = dpp_fmin_step<0x142, 0xa>(x);
  x = dpp_fmin_step<0x143, 0xc>(x);
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63));
}

// K-th smallest of the 64 lane values x (each lane one value): K rounds of
// wave minimum, each round retiring the lowest lane holding the minimum.
// At least K of the wave's values are <= the result.
__device__ __forceinline__ float wave_kth_smallest(float x, int K) {
  const int lane = threadIdx.x & 63;
  float m = x;
  for (int r = 0; r < K; ++r) {
    m = wave_min_f32(x);
    const unsigned long long hit = __ballot(x == m);
    if (hit == 0ull) break;  // only NaN lanes left
    if (lane == __ffsll((long long)hit) - 1) x = __int_as_float(0x7f800000);
    if (m == __int_as_float(0x7f800000)) break;
  }
  return m;
}

// Insert key into an ascending list of L keys (drops the largest).
template <int L>
__device__ __forceinline__ void list_insert(uint64_t (&h)[L], uint64_t k) {
#pragma unroll
  for (int i = 0; i < L; ++i) {
    const uint64_t a = h[i];
    const bool lt = k < a;
    h[i] = lt ? k : a;
    k = lt ? a : k;
  }
}

template <int L>
__device__ __forceinline__ void list_pop(uint64_t (&h)[L]) {
#pragma unroll
  for (int i = 0; i + 1 < L; ++i) h[i] = h[i + 1];
  h[L - 1] = kNoKey;
}

// Orders this wave's LDS accesses: every lane's earlier LDS writes are visible
// to every lane's later LDS reads (and the reverse), without a block barrier.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// Merge per rotation, one wave per rotation: the K best of the n0 per-plane
// lists (each ascending in key order, at most K entries, offset -1 = empty).
// The rotation's keys are staged in LDS; lane L owns planes L + 64 q and holds
// each one's current head; K rounds of a wave minimum take the next best and
// advance the list(s) it came from.  Values >= maxValCorr end the list, the
// reference's queue bound (CorrelationProcessor::init, docking_fft.hpp:283-299).  grid (nrot), block 256:
// all four waves stage the keys (many loads in flight), wave 0 merges.  LDS
// n0 * K * 8 bytes; n0 <= 64 * kMergePlanes.
// ---------------------------------------------------------------------------
constexpr int kMergePlanes = 8;

__global__ __launch_bounds__(256) void k_merge_wave(ScanGeom g, int K, const float* __restrict__ cval,
                                                    const int* __restrict__ coff, double max_val, int rot0,
                                                    PeakOut* __restrict__ peaks, int* __restrict__ n_out) {
  extern __shared__ uint64_t keys[];
  const int b = blockIdx.x, lane = threadIdx.x & 63;
  const int nc = g.n0 * K;
#pragma unroll 4
  for (int e = threadIdx.x; e < nc; e += 256) {
    float v = cval[(size_t)b * nc + e];
    const int o = coff[(size_t)b * nc + e];
    if (v != v) v = __int_as_float(0x7f800000);
    keys[e] = o < 0 ? kNoKey : pd_key(v, o);
  }
  __syncthreads();
  if (threadIdx.x >= 64) return;
  int cur[kMergePlanes];
  uint64_t hk[kMergePlanes];
  uint64_t head = kNoKey;
#pragma unroll
  for (int q = 0; q < kMergePlanes; ++q) {
    const int p = lane + 64 * q;
    cur[q] = 0;
    hk[q] = p < g.n0 ? keys[p * K] : kNoKey;
    head = hk[q] < head ? hk[q] : head;
  }
  PeakOut* out = peaks + (size_t)b * K;
  int count = 0;
  for (int r = 0; r < K; ++r) {
    const uint64_t m = wave_min_u64(head);  // the same in every lane
    if (m == kNoKey) break;
    const float v = pd_key_value(m);
    if (!((double)v < max_val)) break;
    if (lane == 0) {
      PeakOut pk;
      const int o = (int)(uint32_t)m;
      pk.rot = rot0 + b;
      pk.rank = r;
      pk.offset = o;
      const int x0 = o / (g.n1 * g.n2), rem = o - x0 * g.n1 * g.n2;
      const int x1 = rem / g.n2, x2 = rem - x1 * g.n2;
      const int x[3] = {x0, x1, x2}, n[3] = {g.n0, g.n1, g.n2};
      for (int d = 0; d < 3; ++d) {
        int y = -x[d];
        if (y < -(n[d] / 2)) y += n[d];
        pk.s[d] = y;
      }
      pk.value = v;
      pk.reserved = 0.f;
      pk.value64 = __longlong_as_double(0x7ff8000000000000LL);
      out[r] = pk;
    }
    count = r + 1;
    // advance every list whose head is this key (one entry per list)
    head = kNoKey;
#pragma unroll
    for (int q = 0; q < kMergePlanes; ++q) {
      if (hk[q] == m) {
        const int p = lane + 64 * q;
        ++cur[q];
        hk[q] = cur[q] < K ? keys[p * K + cur[q]] : kNoKey;
      }
      head = hk[q] < head ? hk[q] : head;
    }
  }
  if (lane == 0) n_out[b] = count;
}

// X[k] of the real row from the packed half-length spectrum Z (r2c post-step).
__device__ __forceinline__ float2 r2c_point(float2 zk, float2 zm, float2 w) {
  const float2 E = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
  const float2 d = make_float2(zk.x - zm.x, zk.y + zm.y);
  const float2 O = make_float2(0.5f * d.y, -0.5f * d.x);
  return cadd(E, cmul(w, O));
}

// Z[k] for the backward half-length FFT from X[k], X[M-k] (c2r pre-step).
__device__ __forceinline__ float2 c2r_point(float2 xk, float2 xm, float2 w) {
  const float2 E2 = make_float2(xk.x + xm.x, xk.y - xm.y);
  const float2 D = make_float2(xk.x - xm.x, xk.y + xm.y);
  const float2 O2 = cmulc(D, w);
  return make_float2(E2.x - O2.y, E2.y + O2.x);
}

// ---------------------------------------------------------------------------
// Receptor spectra in the column pass's lane order.  Lane L = (lg, cl) of wave
// w in column tile `tile` needs, per term t, R_t[k1 + A k2][col] for
// k1 = lg + B j (j < J; k1 clamped to A - 1) and k2 < B, col = tile C + w CW + cl
// (clamped to the tile's first column past the last one, as the column pass
// clamps its lanes).  Those J B values, in order f = j B + k2, are stored in
// pairs: pair i of lane L at float4 index (((t ntile + tile) 4 + w) NI + i) 64 + L,
// NI = J B / 2, so that one load per pair reads 1 KiB contiguous per wave.
// grid ceil(T * ntile * 4 * NI * 64 / 256), block 256.
// ---------------------------------------------------------------------------
template <int N0>
struct ColLanes {
  using S = ColSplit<N0>;
  static constexpr int A = S::A, B = S::B, J = S::J, C = 256 / B, CW = 64 / B, NI = J * B / 2;
  static_assert(B % 2 == 0, "pairs of receptor values per lane");
};

template <int N0>
__global__ __launch_bounds__(256) void k_rhat_lanes(ScanGeom g, const float2* __restrict__ Rhat,
                                                    float4* __restrict__ RL) {
  using P = ColLanes<N0>;
  const int M = g.M, ntile = (M + P::C - 1) / P::C;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)g.T * ntile * 4 * P::NI * 64) return;
  const int L = (int)(e % 64);
  long long r = e / 64;
  const int i = (int)(r % P::NI);
  r /= P::NI;
  const int w = (int)(r % 4);
  r /= 4;
  const int tile = (int)(r % ntile), t = (int)(r / ntile);
  const int lg = L / P::CW, cl = L % P::CW;
  const int m0 = tile * P::C;
  const int c = m0 + w * P::CW + cl < M ? m0 + w * P::CW + cl : m0;
  float2 v[2];
  for (int q = 0; q < 2; ++q) {
    const int f = 2 * i + q, j = f / P::B, k2 = f % P::B;
    const int k1 = min(lg + P::B * j, P::A - 1);
    v[q] = Rhat[((size_t)t * g.n0 + k1 + P::A * k2) * M + c];
  }
  RL[e] = make_float4(v[0].x, v[0].y, v[1].x, v[1].y);
}

// ---------------------------------------------------------------------------
// Column pass.  1D grid of 8 * ceil(ntile/8) * nyb blocks, block 256,
// C = 256/B lines per tile.  Blocks are dealt to the 8 XCDs round-robin, so
// block id b runs on XCD b % 8; the mapping below gives every rotation group
// of a column tile the same XCD, and those groups run back to back, so the
// tile's receptor spectra (T * N0 * C complex) are read once into that XCD's
// L2 and reused nyb times.
// ---------------------------------------------------------------------------
// Waves per SIMD the column pass is compiled for: three at 144 and 160 while
// MZ <= 3 (at most 20 bytes per lane of spills: column pass 144^3 10.4 -> 10.2
// us per rotation, 160^3 12.6 -> 12.4), else what the registers give (two).
constexpr int col_waves(int n, int mz) { return (n == 144 || n == 160) && mz <= 3 ? 3 : PD_COL_WAVES; }

template <int N0, int MZ>
__global__ __launch_bounds__(256, col_waves(N0, MZ)) void k_col_fast(ScanGeom g, const float2* __restrict__ twN0,
                                                  const float2* __restrict__ L1,
                                                  const float4* __restrict__ RL,
                                                  float2* __restrict__ S1, int bpb, int nrot, int nyb) {
  using S = ColSplit<N0>;
  constexpr int NI = ColLanes<N0>::NI;
  constexpr int A = S::A, B = S::B, J = S::J, C = 256 / B;
  // Each wave owns CW = 64/B columns; lane (lg, cl) is lane lg of column cl.
  // The exchange of a column stays inside its wave: slot (k1, q, cl) at
  // k1*XS + q*CW + cl, XS padded so each half-wave of a 64-bit access hits
  // 64 distinct banks.  The twiddle table shares the array (read per use
  // rather than held in registers).
  constexpr int CW = 64 / B, XS = B * CW + 8, WX = A * XS;
  __shared__ float2 xbuf[4 * WX + N0];
  float2* twl = xbuf + 4 * WX;
  for (int i = threadIdx.x; i < N0; i += blockDim.x) twl[i] = twN0[i];
  __syncthreads();
  // wave index as a uniform value: the receptor loads' scalar offsets depend
  // on it, and taken from threadIdx each would need a waterfall loop
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int lg = lane / CW, cl = lane % CW;
  __builtin_assume(lg >= 0 && lg < B);
  float2* xb = xbuf + wave * WX;
  const int col = wave * CW + cl;
  const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
  const int tile = (local / nyb) * 8 + xcd, rg = local % nyb;
  const int M = g.M, T = g.T;
  const int m0 = tile * C;
  if (m0 >= M) return;
  const bool colok = m0 + col < M;
  const bool sync_st = g.col_sync && m0 + C <= M;  // whole tile: no wave leaves early
  const int bbeg = rg * bpb, bend = min(nrot, bbeg + bpb);
  // L1 holds pp planes (a multiple of B) per (rotation, term): x position
  // lg + B m is a stored plane for every lane (m < pp / B) or for none, and
  // positions m >= MZ >= pp / B are zero (not loaded).
  const int pp = g.pp;
  // In the plane-major order, lanes past the last column (last tile only) redo
  // column 0 of the tile: the same loads and arithmetic give the same bits, so
  // their stores are harmless duplicates.
  const uint32_t vcol = colok ? (uint32_t)col * 8u : 0u;
  const uint32_t pM8 = (uint32_t)pp * M * 8u;
  // receptor spectra of all terms in lane order (k_rhat_lanes) and the
  // batch's ligand planes
  const int ntile = (M + C - 1) / C;
  const __amdgpu_buffer_rsrc_t rr = pd_rsrc(RL, (uint32_t)T * ntile * 4 * NI * 1024u);
  const __amdgpu_buffer_rsrc_t rl = pd_rsrc(L1 + (size_t)bbeg * T * pp * M, (uint32_t)(bend - bbeg) * T * pM8);
  const __amdgpu_buffer_rsrc_t rs = pd_rsrc(S1 + (size_t)bbeg * g.n0 * M, (uint32_t)(bend - bbeg) * g.n0 * M * 8u);
  // zero-size buffer: every load through it is out of range and returns 0
  const __amdgpu_buffer_rsrc_t rz = pd_rsrc(L1, 0u);
  // Ligand order (g.il): plane-major, or blocks of B planes x CW columns in
  // lane order.  The interleaved order is used only when M is a multiple of
  // CW, so a wave's columns all exist or none does; a wave past the last
  // column has nothing to do (no block barrier follows).
  const bool il = g.il != 0;
  if (il && !colok) return;
  const int colg = (m0 + wave * CW) / CW;  // the wave's column block
  // Rotations go in pairs, so that one load of a term's receptor values
  // serves both.  Iteration j: pair j / 2T, term (j / 2) % T, member j % 2;
  // an odd last rotation is paired with a phantom whose planes read 0 and
  // whose stores fall outside rs (dropped by the range check).
  const int nrel = bend - bbeg;
  const int nj = ((nrel + 1) / 2) * 2 * T;
  // Every memory instruction is issued unconditionally, so that the count of
  // outstanding loads is the same on every path and each wait covers only
  // what it consumes: a position m in [pp / B, MZ), a phantom rotation or an
  // iteration past the last reads 0 through rz.  The planes of (rotation
  // bbeg + rel, term t) start at (rel T + t) pp M complex in rl.
  auto load_planes = [&](int j, float2 (&x)[A]) {
    const int rel = 2 * (j / (2 * T)) + (j & 1), t = (j >> 1) % T;
    const uint32_t st = (uint32_t)(rel * T + t) * pM8;
#pragma unroll
    for (int m = 0; m < A; ++m) {
      if (m < MZ) {
        const bool in = j < nj && rel < nrel && B * m < pp;
        // plane-major: plane lg + B m, column m0 + col; interleaved: one
        // 512-byte block per (m, wave), read in lane order
        const uint32_t o = il ? ((uint32_t)(m * (M / CW) + colg) * 64u + (uint32_t)lane) * 8u
                              : (uint32_t)(lg + B * m) * M * 8u + vcol + (uint32_t)m0 * 8u;
        x[m] = pd_ld2(in ? rl : rz, in ? o : 0u, in ? st : 0u);
      } else {
        x[m] = make_float2(0.f, 0.f);
      }
    }
  };
  float2 acc0[J][B], acc1[J][B], rv[J][B];
#pragma unroll
  for (int j = 0; j < J; ++j)
#pragma unroll
    for (int k = 0; k < B; ++k) acc0[j][k] = acc1[j][k] = make_float2(0.f, 0.f);
  // Forward transform of v (iteration j) into acc, with the next iteration's
  // planes loaded into vn; member 0 first loads the term's receptor values.
  auto fwd = [&](int j, float2 (&v)[A], float2 (&vn)[A], float2 (&acc)[J][B], bool load_rv) {
    if (load_rv) {
      const int t = (j >> 1) % T;
      const uint32_t rbase = ((uint32_t)(t * ntile + tile) * 4u + (uint32_t)wave) * NI * 1024u;
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        const float4 q = pd_ld4(rr, (uint32_t)lane * 16u, rbase + (uint32_t)i * 1024u);
        rv[(2 * i) / B][(2 * i) % B] = make_float2(q.x, q.y);
        rv[(2 * i + 1) / B][(2 * i + 1) % B] = make_float2(q.z, q.w);
      }
    }
    // Loads complete in issue order: the receptor values (L2 hits) go out
    // before the prefetch (HBM), so that the accumulation below does not
    // wait for the prefetch, and v's planes, older than both, are waited for
    // alone.  The barriers keep the scheduler from reordering the groups.
    __builtin_amdgcn_sched_barrier(0);
    load_planes(j + 1, vn);
    __builtin_amdgcn_sched_barrier(0);
    four_step_phase1_zt<A, B, MZ, -1>(v, lg, twl);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) xb[k1 * XS + lg * CW + cl] = v[k1];
    wave_lds_sync();
    float2 w[J][B];
#pragma unroll
    for (int jj = 0; jj < J; ++jj) {
      const int k1 = lg + B * jj;
      if (k1 < A)
#pragma unroll
        for (int q = 0; q < B; ++q) w[jj][q] = xb[k1 * XS + q * CW + cl];
    }
    wave_lds_sync();
#pragma unroll
    for (int jj = 0; jj < J; ++jj) {
      const int k1 = lg + B * jj;
      if (k1 < A) {
        rdft<B, -1>(w[jj]);
        // acc += w conj(rv) as four FMAs (explicit, so both members of a
        // rotation pair round alike): measured 6.30 -> 6.20 us per rotation at 128^3
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) {
          const float2 a = w[jj][k2], c = rv[jj][k2];
          float2 s = acc[jj][k2];
          s.x = fmaf(a.x, c.x, s.x);
          s.x = fmaf(a.y, c.y, s.x);
          s.y = fmaf(a.y, c.x, s.y);
          s.y = fmaf(-a.x, c.y, s.y);
          acc[jj][k2] = s;
        }
      }
    }
  };
  // Inverse along x of a finished rotation (DFT_B over k2, twiddle, exchange,
  // DFT_A over k1), stored to S1, and the accumulator cleared.
  const uint32_t sout = (uint32_t)nrel * g.n0 * M * 8u;  // first byte past rs
  auto inv_store = [&](float2 (&acc)[J][B], int rel) {
#pragma unroll
    for (int jj = 0; jj < J; ++jj) {
      const int k1 = lg + B * jj;
      if (k1 < A) {
        four_step_dif_phase1<A, B, +1>(acc[jj], k1, twl);
#pragma unroll
        for (int q = 0; q < B; ++q) xb[k1 * XS + q * CW + cl] = acc[jj][q];
      }
    }
    wave_lds_sync();
    float2 x[A];
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) x[k1] = xb[k1 * XS + lg * CW + cl];
    wave_lds_sync();
    rdft<A, +1>(x);
    const bool ok = rel < nrel;
    if (sync_st) __syncthreads();
    const uint32_t sb = ok ? ((uint32_t)rel * g.n0 * M + m0) * 8u : 0u;
#pragma unroll
    for (int m = 0; m < A; ++m) pd_st2(rs, ok ? (uint32_t)(lg + B * m) * M * 8u + vcol : sout, sb, x[m]);
#pragma unroll
    for (int jj = 0; jj < J; ++jj)
#pragma unroll
      for (int k = 0; k < B; ++k) acc[jj][k] = make_float2(0.f, 0.f);
  };
  // two plane buffers in turn, so no copy waits on the prefetch
  float2 va[A], vb[A];
  load_planes(0, va);
#pragma unroll 1
  for (int j = 0; j < nj; j += 2) {
    fwd(j, va, vb, acc0, true);
    fwd(j + 1, vb, va, acc1, false);
    if ((j >> 1) % T == T - 1) {
      const int rel0 = 2 * (j / (2 * T));
      inv_store(acc0, rel0);
      inv_store(acc1, rel0 + 1);
    }
  }
}

// ---------------------------------------------------------------------------
// Ligand plane pass.  grid (p0, nrot), block 512, LDS N1 * LS complex.
constexpr int kLigTerms = 4;  // terms whose weights the plane's atom list keeps in LDS
// Atoms of one plane the list holds: as many as keep the block within 80 KB
// of LDS (two blocks per CU: 128^3 ligand pass 198 -> 156 us per launch of 64
// rotations), at least 384, at most 1024 (where the plane alone takes more
// than 80 KB).  A plane with more atoms is summed from the atom cells.
template <int N1, int N2>
struct LigList {
  static constexpr int kBuf = N1 * ((N2 / 2 + 1) | 1) * 8;        // dynamic: the plane
  static constexpr int kFixed = N2 * 8 + (2 * N1 + 1) * 4 + 256;   // twiddles, row starts, cursors
  static constexpr int kPer = 8 + 4 * kLigTerms;                   // bytes per listed atom
  static constexpr int kFit = (80 * 1024 - kBuf - kFixed) / kPer / 32 * 32;
  static constexpr int value = kFit >= 384 ? (kFit < 1024 ? kFit : 1024) : 1024;
};
// ---------------------------------------------------------------------------
// Threads per block of the two whole-plane kernels: 1,024 where the plane
// takes more than 80 KB of LDS (144, 160: one block per CU either way), so
// that each thread holds half as many values and the CU runs four waves per
// SIMD instead of two (ligand pass 144^3: 10.2 -> 9.3 us per rotation, 160^3:
// 8.2 -> 7.3, same bits).  The launch must use the same count.
constexpr int plane_threads(int n) { return n >= 144 ? 1024 : 512; }

template <int N1, int N2, int NT = plane_threads(N1)>
__global__ __launch_bounds__(NT, PD_LIG_WAVES) void k_lig_plane_fwd_fast(ScanGeom g, const float2* __restrict__ twN1,
                                                            const float2* __restrict__ twM2,
                                                            const float2* __restrict__ tw2,
                                                            const int* __restrict__ cells, int n_atoms,
                                                            const float* __restrict__ w,
                                                            float2* __restrict__ L1) {
  constexpr int M2 = N2 / 2, H2 = M2 + 1, LS = H2 | 1;
  using SY = Split<N1>;
  constexpr int A1 = SY::A, B1 = SY::B, J1 = SY::J;
  constexpr int LY = NT / B1;
  // Columns 0 and M2 are real along y.  With fewer lane groups than H2
  // columns (128 at 512 threads) they travel packed as one complex column
  // X[.][0] + i X[.][M2] in slot 0 and are split by Hermitian symmetry after
  // the y pass; with LY >= H2 (up to 120 at 512 threads, 160 at 1,024) each
  // has lanes of its own that would otherwise idle, and no split follows
  // (ligand pass 96^3: 2.20 -> 2.06 us per rotation, 120^3: 2.71 -> 2.53,
  // 160^3: 7.21 -> 6.90).  At 144 the unpacked pass spills more (100 against
  // 72 bytes per lane) and measured 3% slower, so 144 stays packed.
  constexpr bool packed = LY < H2 || N1 == 144;
  constexpr int NC = packed ? M2 : H2;  // columns the y pass transforms
  constexpr int RY = (NC + LY - 1) / LY;
  extern __shared__ float2 buf[];
  const int tid = threadIdx.x;
  // Blocks go to the 8 XCDs round-robin in launch order.  The 16 planes of one
  // interleaved group share each 512-byte block of L1 in 32-byte pieces, so
  // they run on one XCD, whose L2 then merges the pieces into whole lines:
  // block L (on XCD L % 8) takes item ((L / 128) * 8 + L % 8) * 16 + (L / 8) % 16
  // of the plane-fastest items over a plane count padded to 16 (gridDim.x);
  // a last partial group of 128 blocks keeps item L.
  const int PX = gridDim.x, Lb = blockIdx.x + PX * blockIdx.y, nblk = PX * gridDim.y;
  const int item = Lb < nblk / 128 * 128 ? ((Lb >> 7) * 8 + (Lb & 7)) * 16 + ((Lb >> 3) & 15) : Lb;
  const int p = item % PX, b = item / PX;
  if (p >= g.p[0]) return;  // padding plane: the whole block leaves before any barrier
  const int i0 = g.lo[0] + p;
  const int lo1 = g.lo[1], p1 = g.p[1];
  const int plane = N1 * N2;
  const int* cb = cells + (size_t)b * n_atoms;
  const int gy = tid / LY, cy = tid % LY;
  const int lane = tid & 63, wave = tid >> 6;
  // The z spectra of the plane's box rows by direct summation over each row's
  // atoms (a handful per row, against a length-N2 transform of every row):
  //   X_t[row][kz] = sum_j w_tj exp(-2 pi i kz z_j / N2),
  // summed in ascending atom order, so that every run rounds alike.  The
  // plane's atoms are listed by box row (counted, placed with atomics, then
  // each row put in atom order), with the weights of up to kLigTerms terms.
  // A plane with more than kLigList atoms is summed from the atom cells
  // directly (same order, slower).
  __shared__ float2 s_tw[N2];
  __shared__ int s_row[N1 + 1];  // row r's atoms: [s_row[r], s_row[r + 1])
  __shared__ int s_cur[N1];
  constexpr int kLigList = LigList<N1, N2>::value;
  __shared__ int s_z[kLigList];
  __shared__ int s_ra[kLigList];
  __shared__ float s_w[kLigTerms][kLigList];
  const bool wl = g.T <= kLigTerms;
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
#pragma unroll 4
  for (int j = tid; j < n_atoms; j += NT) {
    const int k = key_of(j);
    if (k >= 0) atomicAdd(&s_row[k / N2 + 1], 1);
  }
  __syncthreads();
  if (tid == 0)
    for (int i = 1; i <= p1; ++i) s_row[i] += s_row[i - 1];
  __syncthreads();
  const bool listed = s_row[p1] <= kLigList;
  if (listed) {
    for (int i = tid; i < p1; i += NT) s_cur[i] = s_row[i];
    __syncthreads();
#pragma unroll 4
    for (int j = tid; j < n_atoms; j += NT) {
      const int k = key_of(j);
      if (k >= 0) {
        const int r = k / N2;
        const int pos = atomicAdd(&s_cur[r], 1);
        s_z[pos] = k - r * N2;
        s_ra[pos] = j;
        if (wl)
          for (int t = 0; t < g.T; ++t) s_w[t][pos] = w[(size_t)t * n_atoms + j];
      }
    }
    __syncthreads();
    // each row in atom order (insertion sort of its few entries)
    for (int r = tid; r < p1; r += NT) {
      const int a0 = s_row[r], a1 = s_row[r + 1];
      for (int i = a0 + 1; i < a1; ++i) {
        const int ja = s_ra[i], za = s_z[i];
        float wa[kLigTerms];
        for (int t = 0; t < kLigTerms; ++t) wa[t] = wl && t < g.T ? s_w[t][i] : 0.f;
        int q = i - 1;
        while (q >= a0 && s_ra[q] > ja) {
          s_ra[q + 1] = s_ra[q];
          s_z[q + 1] = s_z[q];
          if (wl)
            for (int t = 0; t < g.T; ++t) s_w[t][q + 1] = s_w[t][q];
          --q;
        }
        s_ra[q + 1] = ja;
        s_z[q + 1] = za;
        if (wl)
          for (int t = 0; t < g.T; ++t) s_w[t][q + 1] = wa[t];
      }
    }
    __syncthreads();
  }
  constexpr int NQ = (NC + 63) / 64;  // columns per lane in the summation
  for (int t = 0; t < g.T; ++t) {
    // Box rows of term t into buf, one row per wave at a time, lane c taking
    // columns c + 64 q: slot c in [1, M2) holds X[c]; slot 0 holds
    // X[0] + i X[M2] (both real) when packed, else X[0] and slot M2 X[M2].
    const float* wt = w + (size_t)t * n_atoms;
    for (int r = wave; r < p1; r += NT / 64) {
      float2 a[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) a[q] = make_float2(0.f, 0.f);
      auto add = [&](int z, float wv) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          const int c = lane + 64 * q;
          const float2 ph = s_tw[(c * z) % N2];
          const float py = packed ? (c == 0 ? s_tw[(M2 * z) % N2].x : ph.y) : (c == 0 || c == M2 ? 0.f : ph.y);
          a[q].x = fmaf(wv, ph.x, a[q].x);
          a[q].y = fmaf(wv, py, a[q].y);
        }
      };
      if (listed) {
        const int a0 = s_row[r], a1 = s_row[r + 1];
        for (int i = a0; i < a1; ++i) add(s_z[i], wl ? s_w[t][i] : wt[s_ra[i]]);
      } else {
        for (int j = 0; j < n_atoms; ++j) {
          const int k = key_of(j);
          if (k >= 0 && k / N2 == r) add(k - r * N2, wt[j]);
        }
      }
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int c = lane + 64 * q;
        if (c < NC) buf[(lo1 + r) * LS + c] = a[q];
      }
    }
    __syncthreads();
    // y pass: forward FFT along y of NC columns, store to HBM; packed, the
    // spectra of columns 0 and M2 are split afterwards by Hermitian symmetry.
    {
      float2 v[RY][A1];
#pragma unroll
      for (int r = 0; r < RY; ++r) {
        const int col = r * LY + cy;
        if (col < NC) {
#pragma unroll
          for (int m = 0; m < A1; ++m) {
            const int i1 = gy + B1 * m;
            const int rr = i1 - lo1;
            v[r][m] = rr >= 0 && rr < p1 ? buf[i1 * LS + col] : make_float2(0.f, 0.f);
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < RY; ++r) {
        const int col = r * LY + cy;
        if (col < NC) {
          four_step_phase1<A1, B1, -1>(v[r], gy, twN1);
#pragma unroll
          for (int k1 = 0; k1 < A1; ++k1) buf[(k1 * B1 + gy) * LS + col] = v[r][k1];
        }
      }
      __syncthreads();
      // stores through a buffer descriptor over the (rotation, term)'s planes:
      // 32-bit offsets instead of one 64-bit address per stored value; in the
      // interleaved order (g.il) column m of plane q sits at l1_off(g, q, m)
      float2* dst = L1 + ((size_t)b * g.T + t) * g.pp * (size_t)(N1 * H2);
      const int q = g.lo[0] - g.pl + p;
      const __amdgpu_buffer_rsrc_t rd = pd_rsrc(dst, (uint32_t)g.pp * (N1 * H2 * 8));
      const uint32_t qb = g.il ? (uint32_t)(((q / g.il) * ((N1 * H2) >> g.ilsh)) * 64 + (q % g.il) * (64 / g.il))
                               : (uint32_t)q * (N1 * H2);
      const uint32_t cmask = g.il ? (64u / g.il) - 1u : 0u;
      auto off8 = [&](int m) -> uint32_t {  // byte offset of column m of this plane
        return g.il ? (qb + ((uint32_t)(m >> g.ilsh) << 6) + ((uint32_t)m & cmask)) * 8u : (qb + (uint32_t)m) * 8u;
      };
      // Column m + k2 A1 H2 lies k2 kstep bytes past column m (the same for
      // every lane) when A1 H2 is a whole number of the interleave's column
      // blocks (the interleave is the column pass's B = ColSplit<N1>::B lanes
      // on these cubic grids, or none): then one offset per lane and a scalar
      // offset per store replace the per-store address arithmetic.
      constexpr bool lin = (A1 * H2) % (64 / ColSplit<N1>::B) == 0;
      const uint32_t kstep = g.il ? ((uint32_t)(A1 * H2) >> g.ilsh) * 512u : (uint32_t)(A1 * H2) * 8u;
#pragma unroll
      for (int r = 0; r < RY; ++r) {
        const int col = r * LY + cy;
#pragma unroll
        for (int j = 0; j < J1; ++j) {
          const int k1 = gy + B1 * j;
          if (col < NC && k1 < A1) {
            float2 x[B1];
#pragma unroll
            for (int q = 0; q < B1; ++q) x[q] = buf[(k1 * B1 + q) * LS + col];
            rdft<B1, -1>(x);
            if (!packed || col != 0) {
              if constexpr (lin) {
                const uint32_t o = off8(k1 * H2 + col);
#pragma unroll
                for (int k2 = 0; k2 < B1; ++k2) pd_st2(rd, o, (uint32_t)k2 * kstep, x[k2]);
              } else {
#pragma unroll
                for (int k2 = 0; k2 < B1; ++k2) pd_st2(rd, off8((k1 + A1 * k2) * H2 + col), 0u, x[k2]);
              }
            } else {
              // packed spectrum C[k] parked in the unused slot column M2
#pragma unroll
              for (int k2 = 0; k2 < B1; ++k2) buf[(k1 + A1 * k2) * LS + M2] = x[k2];
            }
          }
        }
      }
      __syncthreads();
      if constexpr (packed) {
        // Y0[k] = (C[k] + conj C[-k]) / 2, YM[k] = (C[k] - conj C[-k]) / (2i)
        for (int k = tid; k < N1; k += NT) {
          const float2 c = buf[k * LS + M2], cm = buf[((N1 - k) % N1) * LS + M2];
          pd_st2(rd, off8(k * H2), 0u, make_float2(0.5f * (c.x + cm.x), 0.5f * (c.y - cm.y)));
          pd_st2(rd, off8(k * H2 + M2), 0u, make_float2(0.5f * (c.y + cm.y), -0.5f * (c.x - cm.x)));
        }
        __syncthreads();
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Inverse plane pass + per-plane top-K.  grid (nrot, n0), block NT.
// ---------------------------------------------------------------------------
template <int N1, int N2, int NT = plane_threads(N1)>
__global__ __launch_bounds__(NT, PD_PLANE_WAVES) void k_plane_inv_topk_fast(ScanGeom g, const float2* __restrict__ twN1,
                                                             const float2* __restrict__ twM2,
                                                             const float2* __restrict__ tw2,
                                                             const float2* __restrict__ S1, float scale,
                                                             int K, float* __restrict__ cval,
                                                             int* __restrict__ coff,
                                                             float* __restrict__ grid_out,
                                                             unsigned* __restrict__ thr) {
  constexpr int M2 = N2 / 2, H2 = M2 + 1, LS = H2 | 1;
  using SY = Split<N1>;
  // z pass: 4 lanes per row at 1,024 threads (144, 160: one round of N1 rows
  // instead of two with 8 lanes, so fewer values per lane)
  using SZ = std::conditional_t<(NT >= 1024 && M2 % 4 == 0 && N1 * 4 <= NT), SplitB<M2, 4>, Split<M2>>;
  constexpr int A1 = SY::A, B1 = SY::B, J1 = SY::J;
  constexpr int A2 = SZ::A, B2 = SZ::B, J2 = SZ::J;
  constexpr int LY = NT / B1;
  // Columns 0 and M2 are Hermitian along y (their transforms are real).  With
  // fewer lane groups than H2 columns (128 at 512 threads) they travel packed
  // as one complex column X[.][0] + i X[.][M2] in slot 0, and the z pass
  // unpacks real and imaginary parts.  With LY >= H2 (up to 120 at 512
  // threads, 144 and 160 at 1,024) column M2 has lanes of its own that would
  // otherwise idle, and both are transformed alone: no packed loads, which
  // hold 2 A1 more registers per lane and wait in every lane (inverse pass
  // 96^3: 2.12 -> 1.94 us per rotation, 120^3: 3.27 -> 2.91).
  constexpr bool packed = LY < H2;
  constexpr int NC = packed ? M2 : H2;  // columns the y pass transforms
  constexpr int RY = (NC + LY - 1) / LY;
  constexpr int LZ = NT / B2, RZ = (N1 + LZ - 1) / LZ;
  extern __shared__ float2 buf[];
  const int tid = threadIdx.x;
  // grid (nrot, n0): the rotation varies fastest, so each rotation's planes run
  // spread over time and later planes find a tight threshold
  const int b = blockIdx.x, i0 = blockIdx.y;
  const float2* src = S1 + ((size_t)b * g.n0 + i0) * (size_t)(N1 * H2);
  // The rotation's selection threshold, read once per block before the
  // transforms (an older value is larger, hence still a valid bound).  The
  // read is device-coherent: each XCD has its own L2, which the other XCDs'
  // atomics do not update.  It is issued after the plane's loads, so that its
  // latency overlaps theirs; the barriers of the y pass publish it.
  __shared__ uint32_t s_thr;
  // y pass: backward FFT along y of every column (NC of them).  Packed, every
  // lane loads column M2 of its rows with an ordinary vector load issued with
  // the others: a divergent, scalar or LDS-staged load costs an extra memory
  // round trip before the transform.  The empty asm consumes those values in
  // every lane, so the compiler cannot sink the loads into the column-0 branch.
  {
    const int gy = tid / LY, cy = tid % LY;
    // LY is a multiple of 64: gy is wave-uniform, and so are the twiddle reads
    static_assert(LY % 64 == 0, "one lane group per wave");
    const int gyu = __builtin_amdgcn_readfirstlane(gy);
    float2 v[RY][A1];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const int col = r * LY + cy;
      if (col < NC) {
        if constexpr (packed) {
          float2 ym[A1];
#pragma unroll
          for (int m = 0; m < A1; ++m) {
            const int row = gy + B1 * m;
            v[r][m] = src[row * H2 + col];
            ym[m] = src[row * H2 + M2];
          }
#pragma unroll
          for (int m = 0; m < A1; ++m) {
            asm volatile("" ::"v"(ym[m].x), "v"(ym[m].y));
            if (col == 0) v[r][m] = make_float2(v[r][m].x - ym[m].y, v[r][m].y + ym[m].x);
          }
        } else {
#pragma unroll
          for (int m = 0; m < A1; ++m) v[r][m] = src[(gy + B1 * m) * H2 + col];
        }
      }
    }
    if (tid == 0) s_thr = thr ? __hip_atomic_load(thr + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0xFFFFFFFFu;
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const int col = r * LY + cy;
      if (col < NC) {
        four_step_phase1<A1, B1, +1>(v[r], gyu, twN1);
#pragma unroll
        for (int k1 = 0; k1 < A1; ++k1) buf[(k1 * B1 + gy) * LS + col] = v[r][k1];
      }
    }
    __syncthreads();
    float2 w[RY][J1][B1];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const int col = r * LY + cy;
#pragma unroll
      for (int j = 0; j < J1; ++j) {
        const int k1 = gy + B1 * j;
        if (col < NC && k1 < A1)
#pragma unroll
          for (int q = 0; q < B1; ++q) w[r][j][q] = buf[(k1 * B1 + q) * LS + col];
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const int col = r * LY + cy;
#pragma unroll
      for (int j = 0; j < J1; ++j) {
        const int k1 = gy + B1 * j;
        if (col < NC && k1 < A1) {
          rdft<B1, +1>(w[r][j]);
#pragma unroll
          for (int k2 = 0; k2 < B1; ++k2) buf[(k1 + A1 * k2) * LS + col] = w[r][j][k2];
        }
      }
    }
    __syncthreads();
  }
  // z pass: c2r pre-step on the fly, backward half-length FFT of every row
  const int gz = tid / LZ, rz = tid % LZ;
  static_assert(LZ % 64 == 0, "one lane group per wave");
  const int gzu = __builtin_amdgcn_readfirstlane(gz);  // wave-uniform twiddle reads
  float2 u[RZ][J2][B2];
  {
    float2 z[RZ][A2];
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
      const int row = r * LZ + rz;
      if (row < N1)
#pragma unroll
        for (int m = 0; m < A2; ++m) {
          const int k = gz + B2 * m, ku = gzu + B2 * m;
          float2 xk = buf[row * LS + k], xm;
          if (k == 0) {  // X[0] and X[M2] are real: packed as Re and Im, or in slots 0 and M2
            if constexpr (packed)
              xm = make_float2(xk.y, 0.f);
            else
              xm = make_float2(buf[row * LS + M2].x, 0.f);
            xk.y = 0.f;
          } else {
            xm = buf[row * LS + M2 - k];
          }
          z[r][m] = c2r_point(xk, xm, tw2[ku]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
      const int row = r * LZ + rz;
      if (row < N1) {
        four_step_phase1<A2, B2, +1>(z[r], gzu, twM2);
#pragma unroll
        for (int k1 = 0; k1 < A2; ++k1) buf[row * LS + k1 * B2 + gz] = z[r][k1];
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
      const int row = r * LZ + rz;
#pragma unroll
      for (int j = 0; j < J2; ++j) {
        const int k1 = gz + B2 * j;
        if (row < N1 && k1 < A2) {
#pragma unroll
          for (int q = 0; q < B2; ++q) u[r][j][q] = buf[row * LS + k1 * B2 + q];
          rdft<B2, +1>(u[r][j]);
        }
      }
    }
  }
  // scale, optional dump, lane minimum
  constexpr int L = 4;
  const float INF = __int_as_float(0x7f800000);
  const int plane = N1 * N2;
  const int base = i0 * plane;
  float lmin = INF;
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    const int row = r * LZ + rz;
#pragma unroll
    for (int j = 0; j < J2; ++j) {
      const int k1 = gz + B2 * j;
      if (row < N1 && k1 < A2) {
#pragma unroll
        for (int k2 = 0; k2 < B2; ++k2) {
          const int e = row * N2 + 2 * (k1 + A2 * k2);
          float vx = u[r][j][k2].x * scale, vy = u[r][j][k2].y * scale;
          if (grid_out) {
            float* go = grid_out + ((size_t)b * g.n0 + i0) * (size_t)plane;
            go[e] = vx;
            go[e + 1] = vy;
          }
          if (vx != vx) vx = INF;
          if (vy != vy) vy = INF;
          u[r][j][k2] = make_float2(vx, vy);
          lmin = fminf(lmin, fminf(vx, vy));
        }
      }
    }
  }
  // Rotation-wide threshold thr[b] (order-preserving bits of an FP32 value):
  // every finished plane lowers it to its own K-th best, which bounds the
  // rotation's K-th best from above, so no cell above it can be selected.  A
  // wave whose cells all lie above it skips the selection.
  __shared__ uint64_t wtop[NT / 64][32];
  const int lane = tid & 63, wave = tid >> 6;
  const uint32_t tk = s_thr;
  const float tval = tk == 0xFFFFFFFFu ? INF : pd_key_value((uint64_t)tk << 32);
  if (__ballot(lmin <= tval) == 0ull) {
    if (lane < K) wtop[wave][lane] = kNoKey;
  } else {
  // The wave's K best values are all <= tau, the K-th smallest lane minimum;
  // only those enter the lane's sorted list of L keys.
  const float tau = fminf(wave_kth_smallest(lmin, K), tval);
  uint64_t h[L];
#pragma unroll
  for (int i = 0; i < L; ++i) h[i] = kNoKey;
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    const int row = r * LZ + rz;
#pragma unroll
    for (int j = 0; j < J2; ++j) {
      const int k1 = gz + B2 * j;
      if (row < N1 && k1 < A2) {
#pragma unroll
        for (int k2 = 0; k2 < B2; ++k2) {
          const int e = base + row * N2 + 2 * (k1 + A2 * k2);
          const float2 x = u[r][j][k2];
          if (x.x <= tau) list_insert(h, pd_key(x.x, e));
          if (x.y <= tau) list_insert(h, pd_key(x.y, e + 1));
        }
      }
    }
  }
  // per-wave K best: K rounds of a DPP wave minimum over the list heads; the
  // lane that supplied the minimum pops it and, once its list is used up,
  // refills it with its next L keys.
  uint64_t last = 0;
  int popped = 0;
  for (int rk = 0; rk < K; ++rk) {
    const uint64_t m = wave_min_u64(h[0]);
    if (lane == 0) wtop[wave][rk] = m;
    if (m == kNoKey) {  // nothing left in the wave (uniform)
      if (lane > rk && lane < K) wtop[wave][lane] = kNoKey;
      break;
    }
    if (m != kNoKey && h[0] == m) {
      last = m;
      list_pop(h);
      ++popped;
    }
    if (h[0] == kNoKey && popped > 0 && popped % L == 0) {
#pragma unroll
      for (int r = 0; r < RZ; ++r) {
        const int row = r * LZ + rz;
#pragma unroll
        for (int j = 0; j < J2; ++j) {
          const int k1 = gz + B2 * j;
          if (row < N1 && k1 < A2) {
#pragma unroll
            for (int k2 = 0; k2 < B2; ++k2) {
              const int e = base + row * N2 + 2 * (k1 + A2 * k2);
              const uint64_t kx = pd_key(u[r][j][k2].x, e), ky = pd_key(u[r][j][k2].y, e + 1);
              if (kx > last) list_insert(h, kx);
              if (ky > last) list_insert(h, ky);
            }
          }
        }
      }
    }
  }
  }
  __syncthreads();
  // block merge by wave 0: NT/64 * K keys (K <= 32), at most HM per lane
  if (wave == 0) {
    constexpr int NW = NT / 64, HM = NW / 2;
    uint64_t hh[HM];
#pragma unroll
    for (int i = 0; i < HM; ++i) hh[i] = kNoKey;
#pragma unroll
    for (int q = 0; q < HM; ++q) {
      const int idx = lane + 64 * q;
      if (idx < NW * K) list_insert(hh, wtop[idx / K][idx % K]);
    }
    float* ov = cval + ((size_t)b * g.n0 + i0) * K;
    int* oo = coff + ((size_t)b * g.n0 + i0) * K;
    for (int rk = 0; rk < K; ++rk) {
      const uint64_t m = wave_min_u64(hh[0]);
      if (lane == 0) {
        ov[rk] = m == kNoKey ? INF : pd_key_value(m);
        oo[rk] = m == kNoKey ? -1 : (int)(uint32_t)m;
        if (thr && rk == K - 1 && m != kNoKey) atomicMin(thr + b, (unsigned)(m >> 32));
      }
      if (m != kNoKey && hh[0] == m) list_pop(hh);
    }
  }
}
