// prop_kernels.hpp -- HIP kernels of the learned rotation proposal (gfx950).
//
// Per complex: interatomic-vector maps of receptor and ligand on a 16^3 grid,
// a 3-layer 3D convolution per molecule, then per rotation 24 rotation
// functions (receptor map against the rotated ligand map), standardised, and
// an MLP 30 -> 384 -> 384 -> 1 whose two wide layers run on MFMA (BF16
// operands, FP32 accumulation).  Training kernels give the exact gradient of
// the mean squared error through all of it, and an Adam step.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace prop {

constexpr int G = 16;             // map edge
constexpr int NG = G * G * G;     // map cells
constexpr int CIN = 2;            // map channels
constexpr int C = 8;              // conv channels
constexpr int NSH = 3;            // shells
constexpr int NF = C * NSH;       // rotation functions per rotation
constexpr int NIN = NF + 6;       // MLP inputs
constexpr int NINP = 32;          // padded row of the MLP input
constexpr int HID = 384;          // hidden width
constexpr double DX = 0.9;        // map spacing (A)
constexpr double SIGMA = 0.5;     // Gaussian width of one pair vector (A)
constexpr double TRUNC = 2.0;     // per-axis window of one pair vector (A)
constexpr double RMAX = 7.2;      // longest pair vector (A)
constexpr double RSURF = 3.3;     // coordination radius (A)
constexpr int NSURF = 12;         // fewer neighbours than this: surface atom
constexpr float SHELL_R[NSH + 1] = {2.0f, 3.6f, 4.8f, 7.2f};

// Parameter layout (floats), mirrored by pd_prop_get_params.
constexpr int CW1 = C * CIN * 27, CB1 = C, CW2 = C * C * 27, CB2 = C, CW3 = C * C * 27, CB3 = C;
constexpr int CONV_STACK = CW1 + CB1 + CW2 + CB2 + CW3 + CB3;  // 3912
constexpr int OFF_W1 = 2 * CONV_STACK;
constexpr int OFF_B1 = OFF_W1 + HID * NIN;
constexpr int OFF_W2 = OFF_B1 + HID;
constexpr int OFF_B2 = OFF_W2 + HID * HID;
constexpr int OFF_W3 = OFF_B2 + HID;
constexpr int OFF_B3 = OFF_W3 + HID;
constexpr int NPARAM = OFF_B3 + 1;  // 167953

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Round to the nearest BF16 (ties to even), returned as FP32.
__device__ __host__ inline float bf16r(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  float r;
  __builtin_memcpy(&r, &u, 4);
  return r;
}
__device__ inline unsigned short bf16bits(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__device__ inline float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ---------------------------------------------------------------------------
// Interatomic-vector maps
// ---------------------------------------------------------------------------
#pragma clang fp contract(off)

// surf[i] = 1 if atom i has fewer than NSURF other atoms within RSURF.
__global__ __launch_bounds__(256) void k_patt_surface(int n, const double* __restrict__ pos,
                                                      float* __restrict__ surf) {
  __shared__ double sp[3 * 256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double xi = 0, yi = 0, zi = 0;
  if (i < n) {
    xi = pos[3 * i];
    yi = pos[3 * i + 1];
    zi = pos[3 * i + 2];
  }
  int cnt = 0;
  const double r2 = RSURF * RSURF;
  for (int j0 = 0; j0 < n; j0 += 256) {
    __syncthreads();
    const int j = j0 + threadIdx.x;
    if (j < n) {
      sp[3 * threadIdx.x] = pos[3 * j];
      sp[3 * threadIdx.x + 1] = pos[3 * j + 1];
      sp[3 * threadIdx.x + 2] = pos[3 * j + 2];
    }
    __syncthreads();
    const int m = min(256, n - j0);
    if (i < n)
      for (int jj = 0; jj < m; ++jj) {
        const double dx = sp[3 * jj] - xi, dy = sp[3 * jj + 1] - yi, dz = sp[3 * jj + 2] - zi;
        const double d2 = dx * dx + dy * dy + dz * dz;
        cnt += (j0 + jj != i && d2 < r2) ? 1 : 0;
      }
  }
  if (i < n) surf[i] = cnt < NSURF ? 1.0f : 0.0f;
}

// Per-axis window of one pair-vector component: cells k with |g_k - v| <= TRUNC.
__device__ inline int axis_window(double v, double* e) {
  int lo = (int)ceil((v - TRUNC) / DX + 7.5);
  int hi = (int)floor((v + TRUNC) / DX + 7.5);
  lo = max(lo, 0);
  hi = min(hi, G - 1);
  int cnt = 0;
  for (int k = lo; k <= hi && cnt < 5; ++k, ++cnt) {
    const double d = (k - 7.5) * DX - v;
    e[cnt] = exp(-(d * d) / (2.0 * SIGMA * SIGMA));
  }
  return (cnt << 8) | (lo & 255);
}

// Block b sums the pairs (i, j), i = b, b + gridDim.x, ..., into its own
// partial map (2 x NG floats, LDS atomics); k_patt_reduce adds the partials.
__global__ __launch_bounds__(256) void k_patt_splat(int n, const double* __restrict__ pos,
                                                    const float* __restrict__ surf,
                                                    float* __restrict__ partial) {
  __shared__ float acc[CIN * NG];
  for (int p = threadIdx.x; p < CIN * NG; p += 256) acc[p] = 0.0f;
  __syncthreads();
  const double rmax2 = RMAX * RMAX;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    const float si = surf[i];
    for (int j = threadIdx.x; j < n; j += 256) {
      if (j == i) continue;
      const double vx = pos[3 * j] - xi, vy = pos[3 * j + 1] - yi, vz = pos[3 * j + 2] - zi;
      const double d2 = vx * vx + vy * vy + vz * vz;
      if (d2 >= rmax2) continue;
      const bool both = si * surf[j] != 0.0f;
      double ex[5], ey[5], ez[5];
      const int wx = axis_window(vx, ex), wy = axis_window(vy, ey), wz = axis_window(vz, ez);
      const int nx = wx >> 8, ny = wy >> 8, nz = wz >> 8;
      const int lx = wx & 255, ly = wy & 255, lz = wz & 255;
      for (int a = 0; a < nx; ++a)
        for (int b = 0; b < ny; ++b) {
          const double exy = ex[a] * ey[b];
          const int row = ((lx + a) * G + (ly + b)) * G + lz;
          for (int c = 0; c < nz; ++c) {
            const float w = (float)(exy * ez[c]);
            atomicAdd(&acc[row + c], w);
            if (both) atomicAdd(&acc[NG + row + c], w);
          }
        }
    }
  }
  __syncthreads();
  float* dst = partial + (size_t)blockIdx.x * CIN * NG;
  for (int p = threadIdx.x; p < CIN * NG; p += 256) dst[p] = acc[p];
}

__global__ __launch_bounds__(256) void k_patt_reduce(int nblk, const float* __restrict__ partial, float scale,
                                                     float* __restrict__ out) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= CIN * NG) return;
  float s = 0.0f;
  for (int b = 0; b < nblk; ++b) s += partial[(size_t)b * CIN * NG + e];
  out[e] = s * scale;
}

#pragma clang fp contract(fast)

// ---------------------------------------------------------------------------
// 3x3x3 convolutions on the 16^3 maps (zero padding); z = molecule.
// ---------------------------------------------------------------------------
struct ConvFwd {
  const float* x;    // [cin][NG]
  const float* w;    // [cout][cin][27]
  const float* b;    // [cout]
  const float* res;  // [CIN][NG] added to channel co % CIN, or null
  float* y;          // [cout][NG]
};

__global__ __launch_bounds__(256) void k_conv_fwd(ConvFwd a0, ConvFwd a1, int cin, int relu) {
  const ConvFwd a = blockIdx.z ? a1 : a0;
  const int p = blockIdx.x * 256 + threadIdx.x;
  const int co = blockIdx.y;
  const int i = p >> 8, j = (p >> 4) & 15, k = p & 15;
  float s = a.b[co];
  for (int ci = 0; ci < cin; ++ci) {
    const float* x = a.x + ci * NG;
    const float* w = a.w + (co * cin + ci) * 27;
    for (int t = 0; t < 27; ++t) {
      const int ii = i + t / 9 - 1, jj = j + (t / 3) % 3 - 1, kk = k + t % 3 - 1;
      if ((unsigned)ii < G && (unsigned)jj < G && (unsigned)kk < G) s += bf16r(w[t]) * x[(ii * G + jj) * G + kk];
    }
  }
  if (relu) s = fmaxf(s, 0.0f);
  if (a.res) s += a.res[(co % CIN) * NG + p];
  a.y[co * NG + p] = s;
}

struct ConvBwdData {
  const float* dy;    // [cout][NG]
  const float* w;     // [cout][cin][27]
  const float* mask;  // [cin][NG]: gradient kept where mask > 0
  float* dx;          // [cin][NG]
};

__global__ __launch_bounds__(256) void k_conv_bwd_data(ConvBwdData a0, ConvBwdData a1, int cin, int cout) {
  const ConvBwdData a = blockIdx.z ? a1 : a0;
  const int p = blockIdx.x * 256 + threadIdx.x;
  const int ci = blockIdx.y;
  const int i = p >> 8, j = (p >> 4) & 15, k = p & 15;
  float s = 0.0f;
  for (int co = 0; co < cout; ++co) {
    const float* dy = a.dy + co * NG;
    const float* w = a.w + (co * cin + ci) * 27;
    for (int t = 0; t < 27; ++t) {
      // output cell q reads p through tap t when q + off(t) = p
      const int ii = i - (t / 9 - 1), jj = j - ((t / 3) % 3 - 1), kk = k - (t % 3 - 1);
      if ((unsigned)ii < G && (unsigned)jj < G && (unsigned)kk < G) s += bf16r(w[t]) * dy[(ii * G + jj) * G + kk];
    }
  }
  if (a.mask[ci * NG + p] <= 0.0f) s = 0.0f;
  a.dx[ci * NG + p] = s;
}

struct ConvBwdW {
  const float* dy;  // [cout][NG]
  const float* x;   // [cin][NG]
  float* dw;        // [cout][cin][27]
  float* db;        // [cout]
};

// Block (t, ci, co + cout * molecule) reduces over the NG cells.
__global__ __launch_bounds__(256) void k_conv_bwd_w(ConvBwdW a0, ConvBwdW a1, int cin, int cout) {
  const int mol = blockIdx.z / cout, co = blockIdx.z % cout;
  const ConvBwdW a = mol ? a1 : a0;
  const int t = blockIdx.x, ci = blockIdx.y;
  const int di = t / 9 - 1, dj = (t / 3) % 3 - 1, dk = t % 3 - 1;
  const float* dy = a.dy + co * NG;
  const float* x = a.x + ci * NG;
  float s = 0.0f, sb = 0.0f;
  for (int p = threadIdx.x; p < NG; p += 256) {
    const int i = p >> 8, j = (p >> 4) & 15, k = p & 15;
    const int ii = i + di, jj = j + dj, kk = k + dk;
    const float g = dy[p];
    if ((unsigned)ii < G && (unsigned)jj < G && (unsigned)kk < G) s += g * x[(ii * G + jj) * G + kk];
    sb += g;
  }
  __shared__ float red[2][4];
  s = wave_sum(s);
  sb = wave_sum(sb);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = s;
    red[1][threadIdx.x >> 6] = sb;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a.dw[(co * cin + ci) * 27 + t] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    if (t == 0 && ci == 0) a.db[co] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
  }
}

// ---------------------------------------------------------------------------
// Rotation functions
// ---------------------------------------------------------------------------
// Listed grid vectors (shell order): vlist[v] = cell index of vector v.
__global__ void k_rf_pack(int nv, const int* __restrict__ vlist, const float* __restrict__ F, float* __restrict__ Fv) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= nv * C) return;
  const int v = e / C, c = e % C;
  Fv[e] = F[c * NG + vlist[v]];
}

struct Shells {
  int off[NSH + 1];
};

__device__ inline void load_rot(const double* r, float R[9]) {
#pragma unroll
  for (int q = 0; q < 9; ++q) R[q] = (float)r[q];
}

// Trilinear corners of q = R^T v in map index space (align corners: cell k at
// (k - 7.5) * DX).  Returns corner cell indices (or -1) and weights.
__device__ inline void corners(const float R[9], int p, int cell[8], float wt[8]) {
  const float vx = ((p >> 8) - 7.5f) * (float)DX, vy = (((p >> 4) & 15) - 7.5f) * (float)DX,
              vz = ((p & 15) - 7.5f) * (float)DX;
  const float inv = (float)(1.0 / DX);
  const float ux = (R[0] * vx + R[3] * vy + R[6] * vz) * inv + 7.5f;
  const float uy = (R[1] * vx + R[4] * vy + R[7] * vz) * inv + 7.5f;
  const float uz = (R[2] * vx + R[5] * vy + R[8] * vz) * inv + 7.5f;
  const float fx0 = floorf(ux), fy0 = floorf(uy), fz0 = floorf(uz);
  const int ix = (int)fx0, iy = (int)fy0, iz = (int)fz0;
  const float fx = ux - fx0, fy = uy - fy0, fz = uz - fz0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int a = q >> 2, b = (q >> 1) & 1, c = q & 1;
    const int x = ix + a, y = iy + b, z = iz + c;
    const bool ok = (unsigned)x < G && (unsigned)y < G && (unsigned)z < G;
    cell[q] = ok ? (x * G + y) * G + z : -1;
    wt[q] = (a ? fx : 1.0f - fx) * (b ? fy : 1.0f - fy) * (c ? fz : 1.0f - fz);
  }
}

constexpr int RF_THREADS = 1024;
constexpr int RF_WAVES = RF_THREADS / 64;

// One wave per rotation; the ligand map (C x NG) sits in LDS as [cell][channel].
__global__ __launch_bounds__(RF_THREADS) void k_rf_fwd(int n_rot, const double* __restrict__ rots,
                                                       const float* __restrict__ FL, const float* __restrict__ FRv,
                                                       const int* __restrict__ vlist, Shells sh,
                                                       float* __restrict__ rf) {
  extern __shared__ float4 lds4[];
  float* lds = reinterpret_cast<float*>(lds4);
  for (int e = threadIdx.x; e < C * NG; e += RF_THREADS) lds[(e % NG) * C + e / NG] = FL[e];
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int r = blockIdx.x * RF_WAVES + wave; r < n_rot; r += gridDim.x * RF_WAVES) {
    float R[9];
    load_rot(rots + 9 * (size_t)r, R);
    float out[NF];
#pragma unroll
    for (int s = 0; s < NSH; ++s) {
      float acc[C];
#pragma unroll
      for (int c = 0; c < C; ++c) acc[c] = 0.0f;
      for (int v = sh.off[s] + lane; v < sh.off[s + 1]; v += 64) {
        int cell[8];
        float wt[8];
        corners(R, vlist[v], cell, wt);
        float smp[C];
#pragma unroll
        for (int c = 0; c < C; ++c) smp[c] = 0.0f;
#pragma unroll
        for (int q = 0; q < 8; ++q)
          if (cell[q] >= 0) {
            const float4 a = lds4[cell[q] * 2], b = lds4[cell[q] * 2 + 1];
            smp[0] += wt[q] * a.x; smp[1] += wt[q] * a.y; smp[2] += wt[q] * a.z; smp[3] += wt[q] * a.w;
            smp[4] += wt[q] * b.x; smp[5] += wt[q] * b.y; smp[6] += wt[q] * b.z; smp[7] += wt[q] * b.w;
          }
        const float4 fa = reinterpret_cast<const float4*>(FRv)[v * 2];
        const float4 fb = reinterpret_cast<const float4*>(FRv)[v * 2 + 1];
        acc[0] += fa.x * smp[0]; acc[1] += fa.y * smp[1]; acc[2] += fa.z * smp[2]; acc[3] += fa.w * smp[3];
        acc[4] += fb.x * smp[4]; acc[5] += fb.y * smp[5]; acc[6] += fb.z * smp[6]; acc[7] += fb.w * smp[7];
      }
#pragma unroll
      for (int c = 0; c < C; ++c) out[c * NSH + s] = wave_sum(acc[c]);
    }
    if (lane < NF) {
      float o = 0.0f;
#pragma unroll
      for (int q = 0; q < NF; ++q) o = lane == q ? out[q] : o;
      rf[(size_t)r * NF + lane] = o;
    }
  }
}

// dFRv[v][c] = sum_r drf[r][c][s(v)] * sample_r(v)[c]; block b handles a chunk of
// rotations and writes its partial sums (nv x C) to part[b].
__global__ __launch_bounds__(RF_THREADS) void k_rf_bwd_rec(int n_rot, const double* __restrict__ rots,
                                                           const float* __restrict__ FL,
                                                           const int* __restrict__ vlist, int nv, Shells sh,
                                                           const float* __restrict__ drf, int chunk,
                                                           float* __restrict__ part) {
  extern __shared__ float4 lds4[];
  float* lds = reinterpret_cast<float*>(lds4);
  for (int e = threadIdx.x; e < C * NG; e += RF_THREADS) lds[(e % NG) * C + e / NG] = FL[e];
  __syncthreads();
  const int r0 = blockIdx.x * chunk, r1 = min(n_rot, r0 + chunk);
  for (int v = threadIdx.x; v < nv; v += RF_THREADS) {
    const int s = v < sh.off[1] ? 0 : (v < sh.off[2] ? 1 : 2);
    const int p = vlist[v];
    float acc[C];
#pragma unroll
    for (int c = 0; c < C; ++c) acc[c] = 0.0f;
    for (int r = r0; r < r1; ++r) {
      float R[9];
      load_rot(rots + 9 * (size_t)r, R);
      int cell[8];
      float wt[8];
      corners(R, p, cell, wt);
      float smp[C];
#pragma unroll
      for (int c = 0; c < C; ++c) smp[c] = 0.0f;
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (cell[q] >= 0) {
          const float4 a = lds4[cell[q] * 2], b = lds4[cell[q] * 2 + 1];
          smp[0] += wt[q] * a.x; smp[1] += wt[q] * a.y; smp[2] += wt[q] * a.z; smp[3] += wt[q] * a.w;
          smp[4] += wt[q] * b.x; smp[5] += wt[q] * b.y; smp[6] += wt[q] * b.z; smp[7] += wt[q] * b.w;
        }
      const float* g = drf + (size_t)r * NF;
#pragma unroll
      for (int c = 0; c < C; ++c) acc[c] += g[c * NSH + s] * smp[c];
    }
    float* dst = part + ((size_t)blockIdx.x * nv + v) * C;
#pragma unroll
    for (int c = 0; c < C; ++c) dst[c] = acc[c];
  }
}

// dFL[cell][c] += drf[r][c][s] * FRv[v][c] * weight(corner); one wave per
// rotation, LDS accumulator, partial map per block.
__global__ __launch_bounds__(RF_THREADS) void k_rf_bwd_lig(int n_rot, const double* __restrict__ rots,
                                                           const float* __restrict__ FRv,
                                                           const int* __restrict__ vlist, Shells sh,
                                                           const float* __restrict__ drf, float* __restrict__ part) {
  extern __shared__ float4 lds4[];
  float* acc = reinterpret_cast<float*>(lds4);
  for (int e = threadIdx.x; e < C * NG; e += RF_THREADS) acc[e] = 0.0f;
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int r = blockIdx.x * RF_WAVES + wave; r < n_rot; r += gridDim.x * RF_WAVES) {
    float R[9];
    load_rot(rots + 9 * (size_t)r, R);
    const float* g = drf + (size_t)r * NF;
    for (int s = 0; s < NSH; ++s) {
      float gs[C];
#pragma unroll
      for (int c = 0; c < C; ++c) gs[c] = g[c * NSH + s];
      for (int v = sh.off[s] + lane; v < sh.off[s + 1]; v += 64) {
        int cell[8];
        float wt[8];
        corners(R, vlist[v], cell, wt);
        const float4 fa = reinterpret_cast<const float4*>(FRv)[v * 2];
        const float4 fb = reinterpret_cast<const float4*>(FRv)[v * 2 + 1];
        const float f[C] = {fa.x, fa.y, fa.z, fa.w, fb.x, fb.y, fb.z, fb.w};
#pragma unroll
        for (int q = 0; q < 8; ++q)
          if (cell[q] >= 0) {
#pragma unroll
            for (int c = 0; c < C; ++c) atomicAdd(&acc[cell[q] * C + c], wt[q] * gs[c] * f[c]);
          }
      }
    }
  }
  __syncthreads();
  float* dst = part + (size_t)blockIdx.x * C * NG;
  for (int e = threadIdx.x; e < C * NG; e += RF_THREADS) dst[e] = acc[e];
}

// out[e] = sum over nblk partials (e < n), in block order.
__global__ void k_sum_partials(int nblk, int n, const float* __restrict__ part, float* __restrict__ out) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= n) return;
  float s = 0.0f;
  for (int b = 0; b < nblk; ++b) s += part[(size_t)b * n + e];
  out[e] = s;
}

// dF_R[c][vlist[v]] = dFRv[v][c] (other cells stay zero).
__global__ void k_rf_unpack(int nv, const int* __restrict__ vlist, const float* __restrict__ dFv, float* __restrict__ dF) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= nv * C) return;
  const int v = e / C, c = e % C;
  dF[c * NG + vlist[v]] = dFv[e];
}

// [cell][c] -> [c][cell]
__global__ void k_cell_to_channel(const float* __restrict__ a, float* __restrict__ b) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= C * NG) return;
  const int c = e / NG, p = e % NG;
  b[e] = a[p * C + c];
}

__global__ void k_fill(float* p, long long n, float v) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e < n) p[e] = v;
}

// ---------------------------------------------------------------------------
// Standardisation over rotations (per column), and the MLP input
// ---------------------------------------------------------------------------
// stats[2c] = mean, stats[2c+1] = 1 / (std + eps), std with n - 1; one block per column.
__global__ __launch_bounds__(256) void k_colstats(int n, const float* __restrict__ x, int ld, float eps,
                                                  float* __restrict__ stats) {
  const int c = blockIdx.x;
  double s = 0.0;
  for (int r = threadIdx.x; r < n; r += 256) s += x[(size_t)r * ld + c];
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  const double mean = red[0] / n;
  __syncthreads();
  double q = 0.0;
  for (int r = threadIdx.x; r < n; r += 256) {
    const double d = x[(size_t)r * ld + c] - mean;
    q += d * d;
  }
  red[threadIdx.x] = q;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double sd = n > 1 ? sqrt(red[0] / (n - 1)) : 0.0;
    stats[2 * c] = (float)mean;
    stats[2 * c + 1] = (float)(1.0 / (sd + eps));
  }
}

// X0[r] = [(rf - mean) * rstd (24), R00, R10, R20, R01, R11, R21, 0, 0]
__global__ void k_build_input(int n, const float* __restrict__ rf, const float* __restrict__ stats,
                              const double* __restrict__ rots, float* __restrict__ X0) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= n * NINP) return;
  const int r = e / NINP, c = e % NINP;
  float v = 0.0f;
  if (c < NF) {
    v = (rf[(size_t)r * NF + c] - stats[2 * c]) * stats[2 * c + 1];
  } else if (c < NIN) {
    const int q = c - NF;  // column (q / 3) of R, row q % 3
    v = (float)rots[9 * (size_t)r + (q % 3) * 3 + q / 3];
  }
  X0[e] = v;
}

// y = (e - mean) * rstd (one column of stats)
__global__ void k_standardize(int n, const float* __restrict__ e, const float* __restrict__ stats, float* __restrict__ y) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r < n) y[r] = (e[r] - stats[0]) * stats[1];
}

// Backward of the standardisation: per column c (one block each)
// drf = a (dz - mean(dz)) - b (x - m), a = 1/(sd + eps), b = sum(dz (x - m)) / ((sd + eps)^2 (n - 1) sd).
__global__ __launch_bounds__(256) void k_zscore_bwd(int n, const float* __restrict__ rf, const float* __restrict__ stats,
                                                    float eps, const float* __restrict__ dX0,
                                                    float* __restrict__ drf) {
  const int c = blockIdx.x;
  const double m = stats[2 * c];
  double s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int r = threadIdx.x; r < n; r += 256) {
    const double dz = dX0[(size_t)r * NINP + c];
    const double d = rf[(size_t)r * NF + c] - m;
    s1 += dz;
    s2 += dz * d;
    s3 += d * d;
  }
  __shared__ double red[3][256];
  red[0][threadIdx.x] = s1;
  red[1][threadIdx.x] = s2;
  red[2][threadIdx.x] = s3;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      for (int q = 0; q < 3; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + o];
    __syncthreads();
  }
  const double sd = n > 1 ? sqrt(red[2][0] / (n - 1)) : 0.0;
  const double a = 1.0 / (sd + eps);
  const double mdz = red[0][0] / n;
  const double b = sd > 0.0 ? red[1][0] / ((sd + eps) * (sd + eps) * (n - 1) * sd) : 0.0;
  for (int r = threadIdx.x; r < n; r += 256) {
    const double dz = dX0[(size_t)r * NINP + c];
    const double d = rf[(size_t)r * NF + c] - m;
    drf[(size_t)r * NF + c] = (float)(a * (dz - mdz) - b * d);
  }
}

// ---------------------------------------------------------------------------
// GEMM on MFMA: C[m][n] = sum_k A(m, k) B(k, n), operands rounded to BF16,
// FP32 accumulation.  A(m, k) = A[m * sam + k * sak], B(k, n) = B[k * sbk + n * sbn].
// 64 x 64 tile per block (4 waves, 32 x 32 each = 2 x 2 MFMA 16x16x32 tiles).
// Epilogue 0: store; 1: + bias[n], ReLU; 2: keep where mask[m * ldm + n] > 0.
// ---------------------------------------------------------------------------
constexpr int GT = 64, GK = 32, GKP = GK + 8;

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const float* __restrict__ A, long long sam,
                                                   long long sak, const float* __restrict__ B, long long sbk,
                                                   long long sbn, float* __restrict__ Cm, long long ldc,
                                                   const float* __restrict__ bias, const float* __restrict__ mask,
                                                   long long ldm) {
  __shared__ __attribute__((aligned(16))) unsigned short As[GT][GKP];
  __shared__ __attribute__((aligned(16))) unsigned short Bs[GT][GKP];
  const int tid = threadIdx.x;
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  const int wave = tid >> 6, lane = tid & 63;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const bool a_krow = sak == 1, b_kcol = sbk == 1;
  for (int k0 = 0; k0 < K; k0 += GK) {
#pragma unroll
    for (int q = 0; q < (GT * GK) / 256; ++q) {
      const int e = tid + 256 * q;
      int m, k;
      if (a_krow) {
        m = e >> 5;
        k = e & 31;
      } else {
        m = e & 63;
        k = e >> 6;
      }
      const int gm = m0 + m, gk = k0 + k;
      const float v = (gm < M && gk < K) ? A[gm * sam + gk * sak] : 0.0f;
      As[m][k] = bf16bits(v);
      int n;
      if (b_kcol) {
        n = e >> 5;
        k = e & 31;
      } else {
        n = e & 63;
        k = e >> 6;
      }
      const int gn = n0 + n, gk2 = k0 + k;
      const float w = (gn < N && gk2 < K) ? B[gk2 * sbk + gn * sbn] : 0.0f;
      Bs[n][k] = bf16bits(w);
    }
    __syncthreads();
    // one gfx950 16x16x32 BF16 MFMA per output tile and k tile: lane l holds
    // row (or column) l % 16 and the 8 k values 8 (l / 16) .. 8 (l / 16) + 7
    static_assert(GK == 32, "one 16x16x32 MFMA step per k tile");
    {
      bf16x8 af[2], bfr[2];
#pragma unroll
      for (int a = 0; a < 2; ++a)
        af[a] = *reinterpret_cast<const bf16x8*>(&As[wm + a * 16 + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
      for (int b = 0; b < 2; ++b)
        bfr[b] = *reinterpret_cast<const bf16x8*>(&Bs[wn + b * 16 + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + a * 16 + 4 * (lane >> 4) + r;
        const int n = n0 + wn + b * 16 + (lane & 15);
        if (m < M && n < N) {
          float v = acc[a][b][r];
          if (EPI == 1) v = fmaxf(v + bias[n], 0.0f);
          if (EPI == 2) v = mask[m * ldm + n] > 0.0f ? v : 0.0f;
          Cm[m * ldc + n] = v;
        }
      }
}

// out[r] = b3 + sum_i H2[r][i] bf16(W3[i]); one wave per row.
__global__ __launch_bounds__(256) void k_out_fwd(int n, const float* __restrict__ H2, const float* __restrict__ W3,
                                                 const float* __restrict__ b3, float* __restrict__ out) {
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (r >= n) return;
  float s = 0.0f;
  for (int i = lane; i < HID; i += 64) s += H2[(size_t)r * HID + i] * bf16r(W3[i]);
  s = wave_sum(s);
  if (lane == 0) out[r] = s + b3[0];
}

// loss = mean((out - y)^2), dout = 2 (out - y) / n; one block.
__global__ __launch_bounds__(1024) void k_mse(int n, const float* __restrict__ out, const float* __restrict__ y,
                                              float* __restrict__ dout, float* __restrict__ loss) {
  double s = 0.0;
  for (int r = threadIdx.x; r < n; r += 1024) {
    const float d = out[r] - y[r];
    dout[r] = 2.0f * d / (float)n;
    s += (double)d * d;
  }
  __shared__ double red[1024];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = (float)(red[0] / n);
}

// dZ2[r][i] = dout[r] bf16(W3[i]) where H2 > 0.
__global__ void k_out_bwd(int n, const float* __restrict__ dout, const float* __restrict__ W3,
                          const float* __restrict__ H2, float* __restrict__ dZ2) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long long)n * HID) return;
  const int r = (int)(e / HID), i = (int)(e % HID);
  dZ2[e] = H2[e] > 0.0f ? dout[r] * bf16r(W3[i]) : 0.0f;
}

// out[c] = sum_r X[r][c] * (scale ? scale[r] : 1), columns in tiles of 64.
__global__ __launch_bounds__(256) void k_colsum(int n, int ncol, const float* __restrict__ X, int ld,
                                                const float* __restrict__ scale, float* __restrict__ out) {
  const int c = blockIdx.x * 64 + (threadIdx.x & 63), w = threadIdx.x >> 6;
  float s = 0.0f;
  if (c < ncol)
    for (int r = w; r < n; r += 4) s += X[(size_t)r * ld + c] * (scale ? scale[r] : 1.0f);
  __shared__ float red[4][64];
  red[w][threadIdx.x & 63] = s;
  __syncthreads();
  if (w == 0 && c < ncol) out[c] = (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

// sum of dout (one block) -> out[0]
__global__ __launch_bounds__(1024) void k_sum(int n, const float* __restrict__ x, float* __restrict__ out) {
  float s = 0.0f;
  for (int r = threadIdx.x; r < n; r += 1024) s += x[r];
  __shared__ float red[16];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.0f;
    for (int q = 0; q < 16; ++q) t += red[q];
    out[0] = t;
  }
}

// Adam, the same arithmetic as torch.optim.Adam (no weight decay).
__global__ void k_adam(int n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float step_size, float b1, float b2, float eps, float sqrt_bc2) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= n) return;
  const float gg = g[e];
  const float mm = b1 * m[e] + (1.0f - b1) * gg;
  const float vv = b2 * v[e] + (1.0f - b2) * gg * gg;
  m[e] = mm;
  v[e] = vv;
  p[e] -= step_size * mm / (sqrtf(vv) / sqrt_bc2 + eps);
}

// ---------------------------------------------------------------------------
// Selection of the n_keep lowest scores (one block), and gathers
// ---------------------------------------------------------------------------
__device__ inline uint32_t order_key(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

constexpr int SEL_THREADS = 1024;

__global__ __launch_bounds__(SEL_THREADS) void k_select_lowest(int n, const float* __restrict__ s, int keep,
                                                               int* __restrict__ idx) {
  __shared__ unsigned hist[256];
  __shared__ unsigned sel_bin, sel_want;
  __shared__ unsigned wsum[2][SEL_THREADS / 64];
  unsigned prefix = 0, mask = 0, want = (unsigned)keep;
  for (int pass = 0; pass < 4; ++pass) {
    const int shift = 24 - 8 * pass;
    for (int b = threadIdx.x; b < 256; b += SEL_THREADS) hist[b] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += SEL_THREADS) {
      const uint32_t u = order_key(s[i]);
      if ((u & mask) == prefix) atomicAdd(&hist[(u >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned cum = 0, b = 0;
      for (; b < 255; ++b) {
        if (cum + hist[b] >= want) break;
        cum += hist[b];
      }
      sel_bin = b;
      sel_want = want - cum;
    }
    __syncthreads();
    prefix |= sel_bin << shift;
    mask |= 255u << shift;
    want = sel_want;
    __syncthreads();
  }
  // keys < prefix are kept, plus the first `want` keys == prefix in index order
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const unsigned long long lt_mask = (1ull << lane) - 1ull;
  unsigned out_base = 0, ties_before = 0;
  for (int base = 0; base < n; base += SEL_THREADS) {
    const int i = base + threadIdx.x;
    uint32_t u = 0;
    if (i < n) u = order_key(s[i]);
    const bool lt = i < n && u < prefix, eq = i < n && u == prefix;
    const unsigned long long beq = __ballot(eq);
    const unsigned eq_rank_w = __popcll(beq & lt_mask);
    if (lane == 0) wsum[0][wave] = __popcll(beq);
    __syncthreads();
    unsigned eq_before = 0, eq_total = 0;
    for (int q = 0; q < SEL_THREADS / 64; ++q) {
      eq_before += q < wave ? wsum[0][q] : 0u;
      eq_total += wsum[0][q];
    }
    const bool take = lt || (eq && ties_before + eq_before + eq_rank_w < want);
    const unsigned long long bt = __ballot(take);
    if (lane == 0) wsum[1][wave] = __popcll(bt);
    __syncthreads();
    unsigned t_before = 0, t_total = 0;
    for (int q = 0; q < SEL_THREADS / 64; ++q) {
      t_before += q < wave ? wsum[1][q] : 0u;
      t_total += wsum[1][q];
    }
    if (take) idx[out_base + t_before + __popcll(bt & lt_mask)] = i;
    out_base += t_total;
    ties_before += eq_total;
    __syncthreads();
  }
}

__global__ void k_gather_rots(int n, const int* __restrict__ idx, const double* __restrict__ in, double* __restrict__ out) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= n * 9) return;
  out[e] = in[9 * (size_t)idx[e / 9] + e % 9];
}

}  // namespace prop
