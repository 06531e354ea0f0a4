// scan_post.hpp -- many translations per rotation, and the reference's
// post-filters, on the GPU.
//
// CorrelationProcessor keeps the maxNTransInp best translations of a rotation
// (docking_fft.hpp:316-345, 451-468) and hands them to TranSymm and then
// TranClust (Worker::findTranslationalMinima, docking_app.hpp:417-447).  The
// scan's per-plane lists hold at most 32 cells, so with post-filters (or when
// the caller asks for the candidate list itself) the inverse pass also writes
// the rotation's real correlation grid, and k_select_post, one block per
// rotation:
//   1. finds the maxNTransInp-th smallest (value, offset) key by radix select
//      over the grid (four 8-bit digits of the order-preserving value bits,
//      then, only when equal values straddle the cut, of the offset), so the
//      kept set is exactly the bounded heap's (Common/queue.hpp:221-251: below
//      maxValCorr, and among equal values the earlier row-major cell);
//   2. gathers those keys into LDS and sorts them (bitonic), ascending;
//   3. TranSymm: keeps, in order, the translations whose pose passes the
//      C_n symmetry test (docking_fft.hpp:146-230, Geom/symmetry.hpp:17-62);
//   4. TranClust: greedy leaders with radius sqrt(3) r, visited in ascending
//      -value, i.e. from the highest energy down: ClusterMatrix::cluster
//      overwrites its density with the weights -value (Geom/cluster.hpp:291-292)
//      and sorts ascending (:308-309); only the first maxNTrans leaders are
//      reported (docking_fft.hpp:89-112, 429-447), and a row's leadership
//      depends only on the leaders before it, so the walk stops there;
//   5. writes the first maxNTrans translations (or the whole candidate list).
// Among exactly equal values the sort order here is by offset; the reference's
// follows its heap (sort_heap) and std::sort.
#pragma once

#include "scan_fast.hpp"

constexpr int kCandMax = 4096;   // candidates per rotation (maxNTransInp) on this path
constexpr int kSelThreads = 1024;

// Post-filter parameters (pd_postfilter, with the derived neighbour radius).
struct PostGeom {
  int K;            // translations reported per rotation (maxNTrans)
  int k_inp;        // candidates per rotation (maxNTransInp), <= kCandMax
  int n_symm;       // TranSymm when > 1
  int cluster;      // TranClust when != 0
  int candidates;   // write the candidate list (k_inp per rotation) instead
  double max_val;   // maxValCorr
  double mol_radius, max_rmsd;
  double clust_rad2;  // neighbours: dist2 <= clust_rad2
  double fa[12], fb[12];  // pose = fa * (Translation(h s) * rot) * fb; 9 rotation (row-major) + 3
};

// Displacement of a correlation offset (WrappedIndex::unwrap, Math/fft.hpp:199-206).
__device__ __forceinline__ void pd_unwrap(const ScanGeom& g, int o, int* s) {
  const int x0 = o / (g.n1 * g.n2), rem = o - x0 * g.n1 * g.n2;
  const int x1 = rem / g.n2, x2 = rem - x1 * g.n2;
  const int x[3] = {x0, x1, x2}, n[3] = {g.n0, g.n1, g.n2};
  for (int d = 0; d < 3; ++d) {
    int y = -x[d];
    if (y < -(n[d] / 2)) y += n[d];
    s[d] = y;
  }
}

// Lanes holding the same bin add together: one LDS atomic per distinct bin of
// the wave (bin < 0: nothing).  Every lane of the wave must call it.
__device__ __forceinline__ void wave_hist_add(uint32_t* hist, int bin) {
  uint64_t active = __ballot(bin >= 0);
  while (active) {
    const int first = __ffsll((long long)active) - 1;
    const int lb = __shfl(bin, first, 64);
    const uint64_t same = __ballot(bin == lb) & active;
    if ((int)(threadIdx.x & 63) == first) atomicAdd(&hist[lb], (uint32_t)__popcll(same));
    active &= ~same;
  }
}

// Rigid transformation x -> m x + v (Geom/transformation.hpp:244-360), FP64.
struct Rigid64 {
  double m[9], v[3];
};

__device__ __forceinline__ void rigid_compose(const Rigid64& a, const Rigid64& b, Rigid64& r) {
#pragma clang fp contract(off)
  // m = a.m b.m (bztinyvec.hpp:70-85), v = a.m b.v + a.v (transformation.hpp:323-328)
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double z = 0;
      for (int k = 0; k < 3; ++k) z += a.m[3 * i + k] * b.m[3 * k + j];
      r.m[3 * i + j] = z;
    }
  for (int i = 0; i < 3; ++i) {
    double acc = 0;
    for (int j = 0; j < 3; ++j) acc += a.m[3 * i + j] * b.v[j];
    r.v[i] = acc + a.v[i];
  }
}

// RMSD of the n-fold pose at radius*e1, radius*e2, radius*e3 and 0
// (SymmetryCheckerCn::operator(), Geom/symmetry.hpp:35-51).
__device__ double symm_rmsd(const Rigid64& t, int n_symm, double radius) {
#pragma clang fp contract(off)
  Rigid64 tn = t, tmp;
  for (int i = 1; i < n_symm; ++i) {
    rigid_compose(t, tn, tmp);
    tn = tmp;
  }
  double sum = 0;
  for (int k = 0; k < 4; ++k) {
    double u[3] = {0, 0, 0};
    if (k < 3) u[k] = radius;
    double d2 = 0;
    for (int i = 0; i < 3; ++i) {
      double acc = 0;
      for (int j = 0; j < 3; ++j) acc += tn.m[3 * i + j] * u[j];
      const double d = (acc + tn.v[i]) - u[i];
      d2 += d * d;
    }
    sum += d2;
  }
  return sqrt(sum / 4);
}

__device__ __forceinline__ void write_peak(const ScanGeom& g, PeakOut* out, int rot, int rank, uint64_t key) {
  PeakOut pk;
  const int o = (int)(uint32_t)key;
  pk.rot = rot;
  pk.rank = rank;
  pk.offset = o;
  pd_unwrap(g, o, pk.s);
  pk.value = pd_key_value(key);
  pk.reserved = 0.f;
  pk.value64 = __longlong_as_double(0x7ff8000000000000LL);
  out[rank] = pk;
}

// grid (nrot), block kSelThreads.  grid: nrot x N real correlation grids;
// rots: the batch's rotations; out: nrot x (candidates ? k_inp : K).
__global__ __launch_bounds__(kSelThreads) void k_select_post(ScanGeom g, PostGeom pg, const float* __restrict__ grid,
                                                            const double* __restrict__ rots, int rot0,
                                                            PeakOut* __restrict__ out, int* __restrict__ n_out) {
#pragma clang fp contract(off)
  __shared__ uint32_t hist[256];
  __shared__ uint64_t list[kCandMax];
  __shared__ uint8_t keep[kCandMax];
  __shared__ uint32_t s_digit, s_rem, s_cnt, s_all, s_m;
  __shared__ double lead_t[32][3];
  const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
  const int N = g.n0 * g.n1 * g.n2;
  const float* G = grid + (size_t)b * N;

  // 1. radix select of the k_inp-th smallest (value, offset) key
  uint32_t vprefix = 0, vmask = 0, oprefix = 0, omask = 0, rem = (uint32_t)pg.k_inp;
  bool all = false, by_offset = false;
  for (int pass = 0; pass < 8; ++pass) {
    const bool off_pass = pass >= 4;
    const int shift = 24 - 8 * (pass & 3);
    if (off_pass && !by_offset) break;
    for (int i = tid; i < 256; i += kSelThreads) hist[i] = 0;
    __syncthreads();
    for (int base = 0; base < N; base += kSelThreads) {
      const int e = base + tid;
      int bin = -1;
      if (e < N) {
        const float v = G[e];
        if ((double)v < pg.max_val) {
          const uint32_t k = (uint32_t)(pd_key(v, 0) >> 32);
          if (!off_pass) {
            if ((k & vmask) == vprefix) bin = (int)((k >> shift) & 255u);
          } else if (k == vprefix && ((uint32_t)e & omask) == oprefix) {
            bin = (int)(((uint32_t)e >> shift) & 255u);
          }
        }
      }
      wave_hist_add(hist, bin);
    }
    __syncthreads();
    if (tid == 0) {
      uint32_t tot = 0;
      for (int x = 0; x < 256; ++x) tot += hist[x];
      if (pass == 0 && tot <= rem) {
        s_all = 1;  // no more candidates than wanted: all of them
      } else {
        uint32_t cum = 0;
        int d = 0;
        while (d < 255 && cum + hist[d] < rem) cum += hist[d++];
        s_all = 0;
        s_digit = (uint32_t)d;
        s_rem = rem - cum;
        s_cnt = hist[d];
      }
    }
    __syncthreads();
    if (s_all) {
      all = true;
      break;
    }
    rem = s_rem;
    if (!off_pass) {
      vprefix |= s_digit << shift;
      vmask |= 0xFFu << shift;
      // after the last value digit: only when some equal values fall beyond the cut
      if (pass == 3) by_offset = rem < s_cnt;
    } else {
      oprefix |= s_digit << shift;
      omask |= 0xFFu << shift;
    }
  }
  const uint32_t olim = by_offset ? oprefix : 0xFFFFFFFFu;

  // 2. gather the kept keys and sort them
  __syncthreads();  // every thread has read s_cnt of the last pass
  if (tid == 0) s_cnt = 0;
  __syncthreads();
  for (int e = tid; e < N; e += kSelThreads) {
    const float v = G[e];
    if (!((double)v < pg.max_val)) continue;
    const uint64_t key = pd_key(v, e);
    const uint32_t k = (uint32_t)(key >> 32);
    if (all || k < vprefix || (k == vprefix && (uint32_t)e <= olim)) {
      const uint32_t slot = atomicAdd(&s_cnt, 1u);
      if (slot < (uint32_t)kCandMax) list[slot] = key;
    }
  }
  __syncthreads();
  int m = (int)min(s_cnt, (uint32_t)pg.k_inp);
  int P = 1;
  while (P < m) P <<= 1;
  for (int i = m + tid; i < P; i += kSelThreads) list[i] = kNoKey;
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < P; i += kSelThreads) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const uint64_t a = list[i], c = list[ixj];
          if ((a > c) == ((i & k) == 0)) {
            list[i] = c;
            list[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  const int rot = rot0 + b;
  if (pg.candidates) {
    PeakOut* o = out + (size_t)b * pg.k_inp;
    for (int i = tid; i < m; i += kSelThreads) write_peak(g, o, rot, i, list[i]);
    if (tid == 0) n_out[b] = m;
    return;
  }

  // 3. TranSymm: keep, in order, the symmetric poses
  if (pg.n_symm > 1) {
    Rigid64 A, Bf, R;
    for (int i = 0; i < 9; ++i) {
      A.m[i] = pg.fa[i];
      Bf.m[i] = pg.fb[i];
      R.m[i] = rots[9 * (size_t)b + i];
    }
    for (int i = 0; i < 3; ++i) {
      A.v[i] = pg.fa[9 + i];
      Bf.v[i] = pg.fb[9 + i];
    }
    for (int i = tid; i < m; i += kSelThreads) {
      int s[3];
      pd_unwrap(g, (int)(uint32_t)list[i], s);
      Rigid64 X = R, AX, T;
      for (int d = 0; d < 3; ++d) X.v[d] = g.h * (double)s[d];  // Translation(h s) * Rotation
      rigid_compose(A, X, AX);
      rigid_compose(AX, Bf, T);
      keep[i] = symm_rmsd(T, pg.n_symm, pg.mol_radius) < pg.max_rmsd;
    }
    __syncthreads();
    if (tid < 64) {  // one wave compacts in order; each write goes to a slot already read
      int w = 0;
      for (int base = 0; base < m; base += 64) {
        const int i = base + lane;
        const uint64_t key = i < m ? list[i] : kNoKey;
        const bool f = i < m && keep[i];
        const uint64_t mask = __ballot(f);
        if (f) list[w + __popcll(mask & ((1ull << lane) - 1ull))] = key;
        w += __popcll(mask);
      }
      if (lane == 0) s_m = (uint32_t)w;
    }
    __syncthreads();
    m = (int)s_m;
  }

  PeakOut* o = out + (size_t)b * pg.K;
  if (!pg.cluster) {
    const int nout = min(m, pg.K);
    for (int i = tid; i < nout; i += kSelThreads) write_peak(g, o, rot, i, list[i]);
    if (tid == 0) n_out[b] = nout;
    return;
  }

  // 4. TranClust: one wave walks from the highest value down; a translation
  // leads unless an earlier leader lies within the radius; stop at K leaders
  if (tid >= 64) return;
  int nl = 0;
  for (int idx = m - 1; idx >= 0 && nl < pg.K; --idx) {
    const uint64_t key = list[idx];
    int s[3];
    pd_unwrap(g, (int)(uint32_t)key, s);
    double t[3];
    for (int d = 0; d < 3; ++d) t[d] = g.h * (double)s[d];
    bool near = false;
    if (lane < nl) {
      double d2 = 0;
      for (int d = 0; d < 3; ++d) {
        const double x = t[d] - lead_t[lane][d];
        d2 += x * x;
      }
      near = d2 <= pg.clust_rad2;
    }
    if (__ballot(near) == 0ull) {
      if (lane == 0) {
        for (int d = 0; d < 3; ++d) lead_t[nl][d] = t[d];
        write_peak(g, o, rot, nl, key);
      }
      ++nl;
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (lane == 0) n_out[b] = nl;
}
