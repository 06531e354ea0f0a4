// scan_kernels.hpp -- HIP kernels of the FFT rigid-body correlation scan.
//
// Per batch of rotations the scan runs five kernels:
//   k_bin            FP64 rotation + cell binning of every ligand atom
//                    (Rotation::operator(), transformation.hpp:206-223;
//                    Grid::operator()(SpatialCoord), grid.hpp:442-451)
//   k_lig_plane_fwd  per (rotation, x-plane inside the ligand box): deposit the
//                    atoms of that plane for each term into LDS
//                    (MolForceRadial::projectLigand, docking_molforce.hpp:351-368),
//                    r2c along z and a complex FFT along y, written once to HBM
//   k_col            per tile of (y,z) frequency columns: FFT along x of every
//                    term, multiply by conj(receptor spectrum), sum over terms
//                    (FFTCorrelator::correlate + collectTotal, docking_fft.hpp:622-633,
//                    docking_molforce.hpp:248-269, fused by linearity), inverse FFT
//                    along x, written once to HBM
//   k_plane_inv_topk per (rotation, x-plane): inverse FFT along y, c2r along z,
//                    scale 1/N and keep the K best cells of the plane
//   k_merge          per rotation: K best of all planes, maxValCorr filter,
//                    offset -> displacement (CorrelationProcessor::selectFromFFT,
//                    docking_fft.hpp:316-389; Math/fft.hpp:199-206)
// plus the once-per-complex receptor kernels (FP64 projection, forward FFT) and
// FP64 rescoring of selected poses.
#pragma once
#include <climits>
#include <cstdint>

#include "fft_lds.hpp"

struct ScanGeom {
  int n0, n1, n2;
  int h2;       // n2/2 + 1
  // Length of the complex FFT along z: n2/2 for an even n2 (two reals packed
  // per complex value, then the r2c post-step), n2 for an odd n2 (the real row
  // as complex values with imaginary parts 0; the first h2 outputs are the r2c
  // result).
  int m2;
  int odd2;     // n2 is odd
  int ls;       // LDS row stride (complex) in the plane kernels, odd, >= h2 (>= n2 when odd2)
  int zrows;    // rows per block of the split path's row kernels (<= 32, so that the z plan fits)
  int M;        // n1 * h2: size of one (k1,k2) frequency plane
  int T;        // terms
  int lo[3];    // ligand box: cells [lo, lo+p) along each axis
  int p[3];
  // Ligand spectra L1 hold pp planes per (rotation, term); plane q is grid
  // plane pl + q (pl = lo[0]; planes q >= p[0] stay 0).  The column pass takes
  // plane q as x position q and correlates it with the spectra of the receptor
  // grid shifted by pl along x: shifting both grids leaves the correlation
  // unchanged, and the ligand's nonzero positions start at 0.
  int pl, pp;
  // L1 order within a (rotation, term): il = 0 plane after plane; il = B (the
  // fast column pass's lanes per column, a power of two) in 512-byte blocks of
  // B planes x 64/B columns, the order in which one wave of that pass reads a
  // position of its transform.  ilsh = log2(64 / B).
  int il, ilsh;
  // Fast column pass: the block's four waves meet at a barrier before a
  // rotation's stores, so that their pieces of each 128-byte line of S1 (and
  // their next loads) go out together (whole tiles only).  Set for N0 >= 96:
  // at 256 the pass moves the same bytes 16% faster, at 128 4% faster, and the
  // 96^3 and 120^3 scans run 0.4% and 1.1% faster.
  int col_sync;
  double h;
  double zero_h[3];  // centred-geometry zero_ - h per axis
  double zsc[3];     // centred-geometry cell-0 centre per axis
};

// Float offset of real element i2 within an LDS row of the z transform:
// packed two per complex value for an even n2, the real part of slot i2 for an
// odd n2.
__device__ __forceinline__ int zreal(const ScanGeom& g, int i2) { return g.odd2 ? 2 * i2 : i2; }

// After the forward z FFT of `rows` rows: the r2c post-step for an even n2;
// nothing for an odd n2 (the full complex FFT already holds the half spectrum).
__device__ __forceinline__ void r2c_finish(const ScanGeom& g, float2* buf, int rows, const float2* tw2) {
  if (!g.odd2) r2c_post(buf, rows, g.ls, g.m2, tw2);
}

// Before the backward z FFT of `rows` rows holding X[0..n2/2]: the c2r
// pre-step for an even n2; for an odd n2 the Hermitian completion
// X[n2-k] = conj X[k] (k = 1..n2/2) with Im X[0] ignored, as FFTW's c2r does.
__device__ __forceinline__ void c2r_start(const ScanGeom& g, float2* buf, int rows, const float2* tw2) {
  if (!g.odd2) {
    c2r_pre(buf, rows, g.ls, g.m2, tw2);
    return;
  }
  const int hz = g.n2 / 2;
  for (int task = threadIdx.x; task < rows * (hz + 1); task += blockDim.x) {
    const int row = task % rows, k = task / rows;
    float2* z = buf + row * g.ls;
    if (k == 0)
      z[0].y = 0.f;
    else
      z[g.n2 - k] = cconj(z[k]);  // writes slots hz+1..n2-1, reads 1..hz
  }
  __syncthreads();
}

// Offset (complex) of plane q, column m within one (rotation, term) of L1.
__host__ __device__ __forceinline__ size_t l1_off(const ScanGeom& g, int q, int m) {
  if (!g.il) return (size_t)q * g.M + m;
  const int cw = 64 / g.il;
  return ((size_t)(q / g.il) * (g.M >> g.ilsh) + (m >> g.ilsh)) * 64 + (q % g.il) * cw + (m & (cw - 1));
}

struct PeakOut {  // layout identical to pd_peak
  int32_t rot, rank, s[3], offset;
  float value, reserved;
  double value64;
};

// ---------------------------------------------------------------------------
// FP64 helpers: no contraction, so results equal the CPU's IEEE per-op rounding.
// ---------------------------------------------------------------------------

__device__ __forceinline__ int to_logical_d(double x, double zero_h, double h) {
#pragma clang fp contract(off)
  const double u = (x - zero_h) / h;
  if (!(u > -1.0e9 && u < 1.0e9)) return INT_MIN;  // NaN / overflow: reported as out of grid
  return static_cast<int>(u) + (0 - 1);
}

__device__ __forceinline__ void rotate_d(const double* m, double x0, double x1, double x2, double* y) {
#pragma clang fp contract(off)
  for (int i = 0; i < 3; ++i) {
    double acc = 0;
    acc += m[3 * i] * x0;
    acc += m[3 * i + 1] * x1;
    acc += m[3 * i + 2] * x2;
    y[i] = acc;
  }
}

__device__ __forceinline__ double field_f2_d(int kind, const double* p, double rP2, double cutP2) {
#pragma clang fp contract(off)
  switch (kind) {
    case 0: {  // LJ repulsion, potentials_inc.hpp:311-314
      const double A = rP2 * rP2 * rP2 + p[1];
      return p[0] / A / A;
    }
    case 1: {  // LJ attraction, potentials_inc.hpp:349-352
      const double A = rP2 * rP2 * rP2 + p[1];
      return -p[0] / A;
    }
    case 2: {  // full soft-core LJ with shifted force, potentials_inc.hpp:214-219
      const double A = rP2 * rP2 * rP2 + p[2];
      return (p[0] / A - p[1]) / A + p[3] + p[4] * sqrt(rP2);
    }
    case 3:
      return rP2 <= p[2] ? p[0] : p[1];
    case 4:
      return p[0] / rP2;
    case 5: {
      const double u = 1.0 - rP2 / cutP2;
      return p[0] * (u * u);
    }
  }
  return 0.0;
}

// One atom's contribution to one cell (Grid/project.hpp:150-157).
__device__ __forceinline__ double project_add_d(double acc, double x0, double x1, double x2,
                                                const double* c, int kind, const double* p,
                                                double cutP2, double minDist2) {
#pragma clang fp contract(off)
  const double d0 = x0 - c[0], d1 = x1 - c[1], d2 = x2 - c[2];
  double rP2 = d0 * d0 + d1 * d1 + d2 * d2;
  if (rP2 <= cutP2) {
    if (rP2 < minDist2) rP2 = minDist2;
    acc += field_f2_d(kind, p, rP2, cutP2);
  }
  return acc;
}

// ---------------------------------------------------------------------------
// Small kernels
// ---------------------------------------------------------------------------

__global__ void k_fill_int(int* p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// tw[m] = exp(-2 pi i m / n) evaluated in FP64, rounded to FP32.
__global__ void k_twiddles(int n, float2* tw) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < n) {
    double s, c;
    sincospi(-2.0 * (double)m / (double)n, &s, &c);
    tw[m] = make_float2((float)c, (float)s);
  }
}

__global__ void k_to_float(const double* __restrict__ in, float* __restrict__ out, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)in[i];
}

// Cell of every rotated ligand atom; -1 (and err bit 1) when outside the box.
__global__ void k_bin(ScanGeom g, const double* __restrict__ rots, int nrot,
                      const double* __restrict__ q, int n_atoms, int* __restrict__ cells,
                      int* __restrict__ err, unsigned* __restrict__ thr) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  // reset the batch's per-rotation selection thresholds (no threshold yet)
  if (thr && idx < nrot) thr[idx] = 0xFFFFFFFFu;
  if (idx >= (long long)nrot * n_atoms) return;
  const int b = (int)(idx / n_atoms), j = (int)(idx - (long long)b * n_atoms);
  double y[3];
  rotate_d(rots + 9 * (size_t)b, q[3 * j], q[3 * j + 1], q[3 * j + 2], y);
  int c[3];
  bool ok = true;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    c[d] = to_logical_d(y[d], g.zero_h[d], g.h);
    ok = ok && c[d] >= g.lo[d] && c[d] < g.lo[d] + g.p[d];
  }
  cells[idx] = ok ? (c[0] * g.n1 + c[1]) * g.n2 + c[2] : -1;
  if (!ok) atomicOr(err, 1);
}

// ---------------------------------------------------------------------------
// Ligand plane pass: deposit + r2c(z) + FFT(y) for one x-plane, all terms.
// grid (p0, nrot), block 512, LDS n1*ls complex.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(512) void k_lig_plane_fwd(ScanGeom g, FftPlan pz, FftPlan py,
                                                       const float2* __restrict__ tw2,
                                                       const int* __restrict__ cells, int n_atoms,
                                                       const float* __restrict__ w,
                                                       float2* __restrict__ L1) {
  extern __shared__ float2 buf[];
  const int p = blockIdx.x, b = blockIdx.y;
  const int i0 = g.lo[0] + p;
  const int plane = g.n1 * g.n2;
  const int* cb = cells + (size_t)b * n_atoms;
  float* fb = reinterpret_cast<float*>(buf);
  const int tot = g.n1 * g.ls;
  for (int t = 0; t < g.T; ++t) {
    for (int e = threadIdx.x; e < tot; e += blockDim.x) buf[e] = make_float2(0.f, 0.f);
    __syncthreads();
    const float* wt = w + (size_t)t * n_atoms;
    for (int j = threadIdx.x; j < n_atoms; j += blockDim.x) {
      const int off = cb[j];
      if (off >= 0 && off / plane == i0) {
        const int r = off - i0 * plane;
        const int i1 = r / g.n2, i2 = r - i1 * g.n2;
        atomicAdd(&fb[(size_t)i1 * 2 * g.ls + zreal(g, i2)], wt[j]);
      }
    }
    __syncthreads();
    fft_lds<-1>(buf + g.lo[1] * g.ls, pz, g.p[1], g.ls, 1);
    r2c_finish(g, buf + g.lo[1] * g.ls, g.p[1], tw2);
    fft_lds<-1>(buf, py, g.h2, 1, g.ls);
    float2* dst = L1 + ((size_t)b * g.T + t) * g.pp * (size_t)g.M;
    const int q = g.lo[0] - g.pl + p;
    for (int e = threadIdx.x; e < g.M; e += blockDim.x) {
      const int k1 = e / g.h2, k2 = e - k1 * g.h2;
      dst[l1_off(g, q, e)] = buf[k1 * g.ls + k2];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Column pass: FFT(x) per term, sum_t L_t * conj(R_t), inverse FFT(x).
// grid (ceil(M/C), ceil(nrot/bpb)), block 256, LDS n0*C complex.
// Each thread holds at most PD_COL_PER accumulator values (host-checked).
// ---------------------------------------------------------------------------
#define PD_COL_PER 16
__global__ __launch_bounds__(256) void k_col(ScanGeom g, FftPlan px, const float2* __restrict__ L1,
                                             const float2* __restrict__ Rhat,
                                             float2* __restrict__ S1, int C, int bpb, int nrot) {
  extern __shared__ float2 buf[];
  const int m0 = blockIdx.x * C;
  const int cw = min(C, g.M - m0);
  const int bbeg = blockIdx.y * bpb, bend = min(nrot, bbeg + bpb);
  const int nel = g.n0 * C;
  for (int b = bbeg; b < bend; ++b) {
    float2 acc[PD_COL_PER];
#pragma unroll
    for (int q = 0; q < PD_COL_PER; ++q) acc[q] = make_float2(0.f, 0.f);
    for (int t = 0; t < g.T; ++t) {
      const float2* src = L1 + ((size_t)b * g.T + t) * g.pp * (size_t)g.M;
      for (int e = threadIdx.x; e < nel; e += blockDim.x) {
        const int i0 = e / C, c = e - i0 * C;
        float2 v = make_float2(0.f, 0.f);
        if (i0 < g.pp && c < cw) v = src[l1_off(g, i0, m0 + c)];
        buf[e] = v;
      }
      __syncthreads();
      fft_lds<-1>(buf, px, C, 1, C);
      const float2* R = Rhat + (size_t)t * g.n0 * g.M + m0;
#pragma unroll
      for (int q = 0; q < PD_COL_PER; ++q) {
        const int e = threadIdx.x + q * blockDim.x;
        if (e < nel) {
          const int k0 = e / C, c = e - k0 * C;
          if (c < cw) acc[q] = cadd(acc[q], cmulc(buf[e], R[(size_t)k0 * g.M + c]));
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < PD_COL_PER; ++q) {
      const int e = threadIdx.x + q * blockDim.x;
      if (e < nel) buf[e] = acc[q];
    }
    __syncthreads();
    fft_lds<+1>(buf, px, C, 1, C);
    float2* dst = S1 + (size_t)b * g.n0 * g.M + m0;
    for (int e = threadIdx.x; e < nel; e += blockDim.x) {
      const int i0 = e / C, c = e - i0 * C;
      if (c < cw) dst[(size_t)i0 * g.M + c] = buf[e];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Block-wide lexicographic (value, offset) argmin; NaN counts as +inf.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool pd_better(float v1, int o1, float v2, int o2) {
  return v1 < v2 || (v1 == v2 && o1 < o2);
}

__device__ __forceinline__ void block_argmin(float& v, int& o, float* red_v, int* red_o) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const float ov = __shfl_xor(v, off, 64);
    const int oo = __shfl_xor(o, off, 64);
    if (pd_better(ov, oo, v, o)) { v = ov; o = oo; }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if (lane == 0) { red_v[wave] = v; red_o[wave] = o; }
  __syncthreads();
  v = red_v[0];
  o = red_o[0];
  for (int k = 1; k < nw; ++k)
    if (pd_better(red_v[k], red_o[k], v, o)) { v = red_v[k]; o = red_o[k]; }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Inverse plane pass + per-plane top-K.  grid (n0, nrot), block 512,
// LDS n1*ls complex.  Candidates: cval/coff[(b*n0 + i0)*K + r]; offset -1 and
// +inf when the plane has fewer than K cells.  grid_out (optional) receives
// the scaled real plane.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(512) void k_plane_inv_topk(ScanGeom g, FftPlan pz, FftPlan py,
                                                        const float2* __restrict__ tw2,
                                                        const float2* __restrict__ S1, float scale,
                                                        int K, float* __restrict__ cval,
                                                        int* __restrict__ coff,
                                                        float* __restrict__ grid_out) {
  extern __shared__ float2 buf[];
  __shared__ float red_v[16];
  __shared__ int red_o[16];
  const int i0 = blockIdx.x, b = blockIdx.y;
  const float2* src = S1 + ((size_t)b * g.n0 + i0) * g.M;
  for (int e = threadIdx.x; e < g.M; e += blockDim.x) {
    const int k1 = e / g.h2, k2 = e - k1 * g.h2;
    buf[k1 * g.ls + k2] = src[e];
  }
  __syncthreads();
  fft_lds<+1>(buf, py, g.h2, 1, g.ls);
  c2r_start(g, buf, g.n1, tw2);
  fft_lds<+1>(buf, pz, g.n1, g.ls, 1);
  float* fb = reinterpret_cast<float*>(buf);
  const int npl = g.n1 * g.n2;
  const int rs = 2 * g.ls;  // float stride of one row
  float bv = __int_as_float(0x7f800000);
  int be = INT_MAX;
  for (int e = threadIdx.x; e < npl; e += blockDim.x) {
    const int i1 = e / g.n2, i2 = e - i1 * g.n2;
    float v = fb[i1 * rs + zreal(g, i2)] * scale;
    if (grid_out) grid_out[((size_t)b * g.n0 + i0) * npl + e] = v;
    if (v != v) v = __int_as_float(0x7f800000);
    fb[i1 * rs + zreal(g, i2)] = v;
    if (pd_better(v, e, bv, be)) { bv = v; be = e; }
  }
  __syncthreads();
  float* ov = cval + ((size_t)b * g.n0 + i0) * K;
  int* oo = coff + ((size_t)b * g.n0 + i0) * K;
  for (int r = 0; r < K; ++r) {
    float v = bv;
    int o = be;
    block_argmin(v, o, red_v, red_o);
    if (threadIdx.x == 0) {
      ov[r] = v;
      oo[r] = (o == INT_MAX) ? -1 : i0 * npl + o;
    }
    if (o != INT_MAX && o == be) {
      const int i1 = o / g.n2, i2 = o - i1 * g.n2;
      fb[i1 * rs + zreal(g, i2)] = __int_as_float(0x7f800000);
      bv = __int_as_float(0x7f800000);
      be = INT_MAX;
      for (int e = threadIdx.x; e < npl; e += blockDim.x) {
        const int j1 = e / g.n2, j2 = e - j1 * g.n2;
        const float x = fb[j1 * rs + zreal(g, j2)];
        if (x != __int_as_float(0x7f800000) && pd_better(x, e, bv, be)) { bv = x; be = e; }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Merge per rotation: K best of n0*K candidates, ascending, value < maxValCorr.
// grid (nrot), block 256, LDS n0*K*(4+4) bytes.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_merge(ScanGeom g, int K, const float* __restrict__ cval,
                                               const int* __restrict__ coff, double max_val,
                                               int rot0, PeakOut* __restrict__ peaks,
                                               int* __restrict__ n_out) {
  extern __shared__ float smem[];
  __shared__ float red_v[8];
  __shared__ int red_o[8];
  __shared__ int s_count;
  const int b = blockIdx.x;
  const int nc = g.n0 * K;
  float* sv = smem;
  int* so = reinterpret_cast<int*>(smem + nc);
  const float INF = __int_as_float(0x7f800000);
  float bv = INF;
  int be = INT_MAX;
  for (int e = threadIdx.x; e < nc; e += blockDim.x) {
    float v = cval[(size_t)b * nc + e];
    int o = coff[(size_t)b * nc + e];
    if (o < 0) { v = INF; o = INT_MAX; }
    sv[e] = v;
    so[e] = o;
    if (o != INT_MAX && pd_better(v, o, bv, be)) { bv = v; be = o; }
  }
  if (threadIdx.x == 0) s_count = 0;
  __syncthreads();
  PeakOut* out = peaks + (size_t)b * K;
  bool done = false;
  for (int r = 0; r < K; ++r) {
    float v = bv;
    int o = be;
    block_argmin(v, o, red_v, red_o);
    if (o == INT_MAX || !((double)v < max_val)) done = true;  // uniform across the block
    if (done) break;
    if (threadIdx.x == 0) {
      PeakOut pk;
      pk.rot = rot0 + b;
      pk.rank = r;
      pk.offset = o;
      const int x0 = o / (g.n1 * g.n2), rem = o - x0 * g.n1 * g.n2;
      const int x1 = rem / g.n2, x2 = rem - x1 * g.n2;
      const int x[3] = {x0, x1, x2}, n[3] = {g.n0, g.n1, g.n2};
      for (int d = 0; d < 3; ++d) {
        int y = -1 * x[d];
        if (y < -1 * (n[d] / 2)) y += n[d];
        pk.s[d] = y;
      }
      pk.value = v;
      pk.reserved = 0.f;
      pk.value64 = __longlong_as_double(0x7ff8000000000000LL);
      out[r] = pk;
      s_count = r + 1;
    }
    if (o == be) {
      for (int e = threadIdx.x; e < nc; e += blockDim.x)
        if (so[e] == o) { sv[e] = INF; so[e] = INT_MAX; }
      bv = INF;
      be = INT_MAX;
      for (int e = threadIdx.x; e < nc; e += blockDim.x)
        if (so[e] != INT_MAX && pd_better(sv[e], so[e], bv, be)) { bv = sv[e]; be = so[e]; }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) n_out[b] = s_count;
}

// ---------------------------------------------------------------------------
// Receptor: FP64 radial-field projection (Grid/project.hpp:143-176).
// Bricks of 4(x) x 8(y) x 8(z) cells, 256 threads.  Atoms are visited in
// ascending order, so every cell accumulates exactly as the CPU loop does.
// region: n_atoms x 6 ints {lo0,hi0,lo1,hi1,lo2,hi2}, already cropped.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_project(ScanGeom g, int n_atoms, const double* __restrict__ pos,
                                                 const int* __restrict__ region, int kind,
                                                 const double* __restrict__ params, int np,
                                                 double cutP2, double minDist2,
                                                 double* __restrict__ grid) {
  __shared__ int list[256];
  __shared__ int wcnt[4];
  const int nbx = (g.n2 + 7) / 8, nby = (g.n1 + 7) / 8;
  const int bid = blockIdx.x;
  const int bx = bid % nbx, by = (bid / nbx) % nby, bz = bid / (nbx * nby);
  const int tid = threadIdx.x;
  const int i0 = bz * 4 + (tid >> 6), i1 = by * 8 + ((tid >> 3) & 7), i2 = bx * 8 + (tid & 7);
  const bool valid = i0 < g.n0 && i1 < g.n1 && i2 < g.n2;
  const int blo0 = bz * 4, bhi0 = min(g.n0 - 1, bz * 4 + 3);
  const int blo1 = by * 8, bhi1 = min(g.n1 - 1, by * 8 + 7);
  const int blo2 = bx * 8, bhi2 = min(g.n2 - 1, bx * 8 + 7);
  double x0, x1, x2;
  {
#pragma clang fp contract(off)
    x0 = g.zsc[0] + g.h * i0;
    x1 = g.zsc[1] + g.h * i1;
    x2 = g.zsc[2] + g.h * i2;
  }
  double acc = 0.0;
  const int lane = tid & 63, wave = tid >> 6;
  for (int chunk = 0; chunk < n_atoms; chunk += 256) {
    const int a = chunk + tid;
    bool hit = false;
    if (a < n_atoms) {
      const int* rg = region + 6 * (size_t)a;
      hit = rg[0] <= bhi0 && rg[1] >= blo0 && rg[2] <= bhi1 && rg[3] >= blo1 && rg[4] <= bhi2 &&
            rg[5] >= blo2;
    }
    const unsigned long long mask = __ballot(hit);
    const int before = __popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) wcnt[wave] = __popcll(mask);
    __syncthreads();
    int base = 0;
    for (int w = 0; w < wave; ++w) base += wcnt[w];
    const int total = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
    if (hit) list[base + before] = a;
    __syncthreads();
    if (valid) {
      for (int k = 0; k < total; ++k) {
        const int at = list[k];
        const int* rg = region + 6 * (size_t)at;
        if (i0 < rg[0] || i0 > rg[1] || i1 < rg[2] || i1 > rg[3] || i2 < rg[4] || i2 > rg[5]) continue;
acc = project_add_d(acc, x0, x1, x2, pos + 3 * (size_t)at, kind, params + (size_t)np * at,
                            cutP2, minDist2);
      }
    }
    __syncthreads();
  }
  if (valid) grid[((size_t)i0 * g.n1 + i1) * g.n2 + i2] = acc;
}

// Receptor plane pass: dense real plane -> r2c(z) -> FFT(y).  grid (n0), block 512.
__global__ __launch_bounds__(512) void k_rec_plane_fwd(ScanGeom g, FftPlan pz, FftPlan py,
                                                       const float2* __restrict__ tw2,
                                                       const float* __restrict__ R,
                                                       float2* __restrict__ out) {
  extern __shared__ float2 buf[];
  const int i0 = blockIdx.x;
  float* fb = reinterpret_cast<float*>(buf);
  const int npl = g.n1 * g.n2;
  for (int e = threadIdx.x; e < g.n1 * g.ls; e += blockDim.x) buf[e] = make_float2(0.f, 0.f);
  __syncthreads();
  for (int e = threadIdx.x; e < npl; e += blockDim.x) {
    const int i1 = e / g.n2, i2 = e - i1 * g.n2;
    fb[(size_t)i1 * 2 * g.ls + zreal(g, i2)] = R[(size_t)i0 * npl + e];
  }
  __syncthreads();
  fft_lds<-1>(buf, pz, g.n1, g.ls, 1);
  r2c_finish(g, buf, g.n1, tw2);
  fft_lds<-1>(buf, py, g.h2, 1, g.ls);
  float2* dst = out + (size_t)i0 * g.M;
  for (int e = threadIdx.x; e < g.M; e += blockDim.x) {
    const int k1 = e / g.h2, k2 = e - k1 * g.h2;
    dst[e] = buf[k1 * g.ls + k2];
  }
}

// Receptor column pass: FFT(x) only.  grid (ceil(M/C)), block 256, LDS n0*C.
__global__ __launch_bounds__(256) void k_col_fwd(ScanGeom g, FftPlan px, const float2* __restrict__ in,
                                                 float2* __restrict__ out, int C, int shift) {
  extern __shared__ float2 buf[];
  const int m0 = blockIdx.x * C;
  const int cw = min(C, g.M - m0);
  const int nel = g.n0 * C;
  for (int e = threadIdx.x; e < nel; e += blockDim.x) {
    const int i0 = e / C, c = e - i0 * C;
    const int src = (i0 + shift) % g.n0;  // position i0 of the grid shifted by -shift
    buf[e] = c < cw ? in[(size_t)src * g.M + m0 + c] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_lds<-1>(buf, px, C, 1, C);
  for (int e = threadIdx.x; e < nel; e += blockDim.x) {
    const int k0 = e / C, c = e - k0 * C;
    if (c < cw) out[(size_t)k0 * g.M + m0 + c] = buf[e];
  }
}

// ---------------------------------------------------------------------------
// FP64 rescoring: E = sum_t sum_j w_tj R_t((l_j + s) mod N).  grid (n), block 256.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_rescore(ScanGeom g, const double* __restrict__ rots,
                                                 const int* __restrict__ s,
                                                 const double* __restrict__ q, int n_atoms,
                                                 const double* __restrict__ w64,
                                                 const double* __restrict__ R64,
                                                 double* __restrict__ out) {
  __shared__ double red[4];
  const int pose = blockIdx.x;
  const size_t N = (size_t)g.n0 * g.n1 * g.n2;
  const int n[3] = {g.n0, g.n1, g.n2};
  double acc = 0.0;
  for (int j = threadIdx.x; j < n_atoms; j += blockDim.x) {
    double y[3];
    rotate_d(rots + 9 * (size_t)pose, q[3 * j], q[3 * j + 1], q[3 * j + 2], y);
    int c[3];
    for (int d = 0; d < 3; ++d) {
      int l = to_logical_d(y[d], g.zero_h[d], g.h);
      if (l == INT_MIN) l = 0;
      c[d] = (((l + s[3 * pose + d]) % n[d]) + n[d]) % n[d];
    }
    const size_t off = ((size_t)c[0] * g.n1 + c[1]) * g.n2 + c[2];
    for (int t = 0; t < g.T; ++t) acc += w64[(size_t)t * n_atoms + j] * R64[t * N + off];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[pose] = red[0] + red[1] + red[2] + red[3];
}
