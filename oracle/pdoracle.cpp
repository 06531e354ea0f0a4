// pdoracle.cpp -- FP64 CPU oracle (TEST INFRASTRUCTURE ONLY; see pdoracle.h).
//
// A restatement of the PRODDL2 worker's rotation scan.  Each function cites
// the reference file:line it follows.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg load this library.

#include "pdoracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Geometry (Grid/gridcoord.hpp)
// ---------------------------------------------------------------------------

struct Axis {
  double h;
  double zero_h;             // zero_ - h       (gridcoord.hpp:172)
  double zero_start_center;  // zero_ + h/2.0   (gridcoord.hpp:173)
  int n;
};

// Centred geometry ctor, gridcoord.hpp:127-142 followed by setup() 166-175.
Axis centered_axis(double h, int n) {
  // geom(h, 0): zero_start_ = 0 - h*0 = 0
  double zero_start0 = 0.0 - h * 0;
  double s_lower = zero_start0 + h * 0;              // toSpatialLeft(0)      :218-221
  double s_upper = zero_start0 + h + h * (n - 1);    // toSpatialRight(n-1)   :226-229
  double s_size = s_upper - s_lower;
  double zero = -1. * s_size / 2;                     // :139
  Axis a;
  a.h = h;
  a.n = n;
  double zero_start = zero - h * 0;                   // :171 (offset 0)
  a.zero_h = zero - h;                                // :172
  a.zero_start_center = zero_start + h / 2.0;         // :173
  return a;
}

// toLogical, gridcoord.hpp:185-192: int((x - zero_h)/h) + (offset - 1).
inline int to_logical(const Axis& a, double x) {
  return static_cast<int>((x - a.zero_h) / a.h) + (0 - 1);
}

// toSpatial, gridcoord.hpp:145-148.
inline double to_spatial(const Axis& a, int i) { return a.zero_start_center + a.h * i; }

// ---------------------------------------------------------------------------
// FFT: own mixed-radix Stockham transform in FP64.  Sign -1 is FFTW's forward
// (e^{-2 pi i jk/n}), +1 its backward; neither normalises.
// ---------------------------------------------------------------------------

struct Plan1 {
  int n = 0;
  std::vector<int> radix;
  std::vector<double> tw;  // tw[2m], tw[2m+1] = cos, sin of -2 pi m / n
};

std::vector<int> factorize(int n) {
  std::vector<int> r;
  while (n % 4 == 0) { r.push_back(4); n /= 4; }
  while (n % 2 == 0) { r.push_back(2); n /= 2; }
  for (int p = 3; p * p <= n; p += 2)
    while (n % p == 0) { r.push_back(p); n /= p; }
  if (n > 1) r.push_back(n);
  return r;
}

const Plan1& get_plan(int n) {
  thread_local std::map<int, std::unique_ptr<Plan1>> cache;
  auto it = cache.find(n);
  if (it != cache.end()) return *it->second;
  auto p = std::make_unique<Plan1>();
  p->n = n;
  p->radix = factorize(n);
  p->tw.resize(2 * (size_t)n);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int m = 0; m < n; ++m) {
    long double ang = -two_pi * (long double)m / (long double)n;
    p->tw[2 * m] = (double)std::cos(ang);
    p->tw[2 * m + 1] = (double)std::sin(ang);
  }
  const Plan1& ref = *p;
  cache.emplace(n, std::move(p));
  return ref;
}

// Complex FFT of one line; Stockham autosort (Govindaraju et al., SC'08,
// "High Performance Discrete Fourier Transforms on Graphics Processors").
void fft_line(const Plan1& P, int sign, double* x, double* scratch) {
  const int n = P.n;
  if (n == 1) return;
  double* in = x;
  double* out = scratch;
  int Ns = 1;
  int maxR = 1;
  for (int R : P.radix) maxR = std::max(maxR, R);
  std::vector<double> vbuf(2 * (size_t)maxR), ubuf(2 * (size_t)maxR);
  double* v = vbuf.data();
  double* u = ubuf.data();
  for (int R : P.radix) {
    const int nb = n / R;
    const int tstride = n / (Ns * R);
    for (int j = 0; j < nb; ++j) {
      const int k = j % Ns;
      for (int r = 0; r < R; ++r) {
        double a = in[2 * (j + r * nb)], b = in[2 * (j + r * nb) + 1];
        int m = (r * k * tstride) % n;
        double c = P.tw[2 * m], s = -sign * P.tw[2 * m + 1];  // forward: tw, backward: conj(tw)
        v[2 * r] = a * c - b * s;
        v[2 * r + 1] = a * s + b * c;
      }
      if (R == 2) {
        u[0] = v[0] + v[2]; u[1] = v[1] + v[3];
        u[2] = v[0] - v[2]; u[3] = v[1] - v[3];
      } else if (R == 4) {
        double t0r = v[0] + v[4], t0i = v[1] + v[5];
        double t1r = v[0] - v[4], t1i = v[1] - v[5];
        double t2r = v[2] + v[6], t2i = v[3] + v[7];
        double t3r = v[2] - v[6], t3i = v[3] - v[7];
        // multiply t3 by sign*i : forward (-i), backward (+i)
        double mr = -sign * t3i, mi = sign * t3r;
        u[0] = t0r + t2r; u[1] = t0i + t2i;
        u[4] = t0r - t2r; u[5] = t0i - t2i;
        u[2] = t1r + mr;  u[3] = t1i + mi;
        u[6] = t1r - mr;  u[7] = t1i - mi;
      } else {
        // generic O(R^2) DFT with W_R^{q} = tw[q * n/R]
        for (int q = 0; q < R; ++q) {
          double sr = 0, si = 0;
          for (int r = 0; r < R; ++r) {
            int m = ((r * q) % R) * nb;
            double c = P.tw[2 * m], s = -sign * P.tw[2 * m + 1];  // forward: tw, backward: conj(tw)
            sr += v[2 * r] * c - v[2 * r + 1] * s;
            si += v[2 * r] * s + v[2 * r + 1] * c;
          }
          u[2 * q] = sr;
          u[2 * q + 1] = si;
        }
      }
      const int idxD = (j / Ns) * Ns * R + k;
      for (int r = 0; r < R; ++r) {
        out[2 * (idxD + r * Ns)] = u[2 * r];
        out[2 * (idxD + r * Ns) + 1] = u[2 * r + 1];
      }
    }
    std::swap(in, out);
    Ns *= R;
  }
  if (in != x) std::memcpy(x, in, sizeof(double) * 2 * n);
}

// In-place complex FFTs along one axis of a 3D complex array of shape s.
void fft_axis(const int* s, int axis, int sign, double* data) {
  const int n = s[axis];
  const Plan1& P = get_plan(n);
  std::vector<double> line(2 * (size_t)n), scratch(2 * (size_t)n);
  const size_t st2 = 1, st1 = (size_t)s[2], st0 = (size_t)s[1] * s[2];
  const size_t stride = axis == 0 ? st0 : (axis == 1 ? st1 : st2);
  const int a = axis == 0 ? 1 : 0, b = axis == 2 ? 1 : 2;
  const size_t sta = a == 0 ? st0 : st1, stb = b == 1 ? st1 : st2;
  for (int ia = 0; ia < s[a]; ++ia)
    for (int ib = 0; ib < s[b]; ++ib) {
      size_t base = ia * sta + ib * stb;
      for (int i = 0; i < n; ++i) {
        line[2 * i] = data[2 * (base + i * stride)];
        line[2 * i + 1] = data[2 * (base + i * stride) + 1];
      }
      fft_line(P, sign, line.data(), scratch.data());
      for (int i = 0; i < n; ++i) {
        data[2 * (base + i * stride)] = line[2 * i];
        data[2 * (base + i * stride) + 1] = line[2 * i + 1];
      }
    }
}

void rfft3(const int* n, const double* real, double* cplx) {
  const int H = n[2] / 2 + 1;
  const Plan1& P = get_plan(n[2]);
  std::vector<double> line(2 * (size_t)n[2]), scratch(2 * (size_t)n[2]);
  for (int i0 = 0; i0 < n[0]; ++i0)
    for (int i1 = 0; i1 < n[1]; ++i1) {
      const double* row = real + ((size_t)i0 * n[1] + i1) * n[2];
      for (int i = 0; i < n[2]; ++i) { line[2 * i] = row[i]; line[2 * i + 1] = 0.0; }
      fft_line(P, -1, line.data(), scratch.data());
      double* o = cplx + 2 * (((size_t)i0 * n[1] + i1) * H);
      std::memcpy(o, line.data(), sizeof(double) * 2 * H);
    }
  int s[3] = {n[0], n[1], H};
  fft_axis(s, 1, -1, cplx);
  fft_axis(s, 0, -1, cplx);
}

// Unnormalised c2r with FFTW semantics: the imaginary parts FFTW's c2r ignores
// (those of self-conjugate bins along the last axis) are dropped.
void irfft3(const int* n, const double* cplx_in, double* real) {
  const int H = n[2] / 2 + 1;
  std::vector<double> c(cplx_in, cplx_in + 2 * (size_t)n[0] * n[1] * H);
  int s[3] = {n[0], n[1], H};
  fft_axis(s, 0, +1, c.data());
  fft_axis(s, 1, +1, c.data());
  const Plan1& P = get_plan(n[2]);
  std::vector<double> line(2 * (size_t)n[2]), scratch(2 * (size_t)n[2]);
  for (int i0 = 0; i0 < n[0]; ++i0)
    for (int i1 = 0; i1 < n[1]; ++i1) {
      const double* row = c.data() + 2 * (((size_t)i0 * n[1] + i1) * H);
      for (int k = 0; k < H; ++k) { line[2 * k] = row[2 * k]; line[2 * k + 1] = row[2 * k + 1]; }
      line[1] = 0.0;
      if (n[2] % 2 == 0) line[2 * (n[2] / 2) + 1] = 0.0;
      for (int k = H; k < n[2]; ++k) {  // Hermitian completion
        line[2 * k] = row[2 * (n[2] - k)];
        line[2 * k + 1] = -row[2 * (n[2] - k) + 1];
      }
      fft_line(P, +1, line.data(), scratch.data());
      double* o = real + ((size_t)i0 * n[1] + i1) * n[2];
      for (int i = 0; i < n[2]; ++i) o[i] = line[2 * i];
    }
}

// ---------------------------------------------------------------------------
// Radial fields (potentials_inc.hpp) and the new config terms.
// ---------------------------------------------------------------------------

inline double pow3(double x) { return x * x * x; }          // Common/math.hpp:20
inline double pow6(double x) { double t1 = x * x * x; return t1 * t1; }  // :23

inline double field_f2(int kind, const double* p, double rP2, double cutP2) {
  switch (kind) {
    case PDO_FIELD_LJ_REP: {  // potentials_inc.hpp:311-314
      double A = pow3(rP2) + p[1];
      return p[0] / A / A;
    }
    case PDO_FIELD_LJ_ATT: {  // potentials_inc.hpp:349-352
      double A = pow3(rP2) + p[1];
      return -p[0] / A;
    }
    case PDO_FIELD_LJ_FULL: {  // potentials_inc.hpp:214-219
      double A = pow3(rP2) + p[2];
      return (p[0] / A - p[1]) / A + p[3] + p[4] * std::sqrt(rP2);
    }
    case PDO_FIELD_TWOSTEP:
      return rP2 <= p[2] ? p[0] : p[1];
    case PDO_FIELD_COULOMB:
      return p[0] / rP2;
    case PDO_FIELD_DESOLV: {
      double u = 1.0 - rP2 / cutP2;
      return p[0] * (u * u);
    }
  }
  return 0.0;
}

inline void rotate(const double* m, const double* x, double* y) {
  // blitz_ext::product (bztinyvec.hpp:52-65): y(i) = 0; y(i) += m(i,j)*x(j)
  for (int i = 0; i < 3; ++i) {
    double acc = 0;
    for (int j = 0; j < 3; ++j) acc += m[3 * i + j] * x[j];
    y[i] = acc;
  }
}

struct Grid3 {
  int n[3];
  Axis ax[3];
  size_t N() const { return (size_t)n[0] * n[1] * n[2]; }
};

Grid3 make_grid(const int* n, double h) {
  Grid3 g;
  for (int d = 0; d < 3; ++d) { g.n[d] = n[d]; g.ax[d] = centered_axis(h, n[d]); }
  return g;
}

// projectField, Grid/project.hpp:143-159, region from :89-98 and crop grid.hpp:613-620.
void project_field(const Grid3& g, int kind, const double* p, double cut, double minDist2,
                   const double* c, double* grid) {
  const double cutP2 = cut * cut;
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = std::max(0, to_logical(g.ax[d], c[d] - cut));
    hi[d] = std::min(g.n[d] - 1, to_logical(g.ax[d], c[d] + cut));
  }
  for (int i0 = lo[0]; i0 <= hi[0]; ++i0)
    for (int i1 = lo[1]; i1 <= hi[1]; ++i1)
      for (int i2 = lo[2]; i2 <= hi[2]; ++i2) {
        double d0 = to_spatial(g.ax[0], i0) - c[0];
        double d1 = to_spatial(g.ax[1], i1) - c[1];
        double d2 = to_spatial(g.ax[2], i2) - c[2];
        double rP2 = d0 * d0 + d1 * d1 + d2 * d2;  // blitz::sum(v*v), left to right
        if (rP2 <= cutP2) {
          if (rP2 < minDist2) rP2 = minDist2;
          grid[((size_t)i0 * g.n[1] + i1) * g.n[2] + i2] += field_f2(kind, p, rP2, cutP2);
        }
      }
}

int project_ligand(const Grid3& g, int n_atoms, const double* pos, const double* w, double* grid) {
  std::fill(grid, grid + g.N(), 0.0);
  for (int j = 0; j < n_atoms; ++j) {
    int c[3];
    for (int d = 0; d < 3; ++d) {
      c[d] = to_logical(g.ax[d], pos[3 * j + d]);
      if (c[d] < 0 || c[d] >= g.n[d]) return j + 1;
    }
    grid[((size_t)c[0] * g.n[1] + c[1]) * g.n[2] + c[2]] += w[j];
  }
  return 0;
}

void correlate(const int* n, int T, const double* rec_spec, const double* lig, double* out,
               std::vector<double>& spec, std::vector<double>& tmp) {
  const size_t N = (size_t)n[0] * n[1] * n[2];
  const size_t H = (size_t)n[0] * n[1] * (n[2] / 2 + 1);
  spec.resize(2 * H);
  tmp.resize(N);
  const int Nint = n[0] * n[1] * n[2];
  for (int t = 0; t < T; ++t) {
    rfft3(n, lig + t * N, spec.data());
    const double* R = rec_spec + 2 * H * t;
    for (size_t i = 0; i < H; ++i) {  // arraysC(lig) *= conj(arraysC(rec))
      double a = spec[2 * i], b = spec[2 * i + 1], c = R[2 * i], d = -R[2 * i + 1];
      spec[2 * i] = a * c - b * d;
      spec[2 * i + 1] = a * d + b * c;
    }
    double* dst = t == 0 ? out : tmp.data();
    irfft3(n, spec.data(), dst);
    for (size_t i = 0; i < N; ++i) dst[i] /= Nint;  // /= N (docking_fft.hpp:630-631)
    if (t > 0)
      for (size_t i = 0; i < N; ++i) out[i] += dst[i];  // collectTotal, molforce.hpp:259-265
  }
}

// BoundPriorityQueue<int, vector<int>, cmp, double, cmp> (Common/queue.hpp:49-280)
// fed by selectIntoQueue (docking_fft.hpp:451-468): same std heap calls.
struct ValCmp {
  const double* data;
  bool operator()(int x, int y) const { return data[x] < data[y]; }
};

int select(const int* n, double h, const double* grid, int max_inp, double max_val, int n_max,
           int* offsets, int* s, double* tran, double* values) {
  std::vector<int> c;
  c.reserve(max_inp);
  ValCmp comp{grid};
  const int Ntot = n[0] * n[1] * n[2];
  for (int off = 0; off < Ntot; ++off) {  // row-major, last index fastest
    if (!(grid[off] < max_val)) continue;                  // queue.hpp:225-226
    if ((int)c.size() >= max_inp) {                        // queue.hpp:233-239
      if (max_inp > 0 && comp(off, c.front())) {
        std::pop_heap(c.begin(), c.end(), comp);
        c.pop_back();
      } else {
        continue;
      }
    }
    c.push_back(off);
    std::push_heap(c.begin(), c.end(), comp);
  }
  std::sort_heap(c.begin(), c.end(), comp);               // queue.hpp:184-193
  int nout = std::min((int)c.size(), n_max);              // getTranValuesFilled :429-447
  const int stride[3] = {n[1] * n[2], n[2], 1};
  for (int i = 0; i < nout; ++i) {
    int ind = c[i];
    if (offsets) offsets[i] = ind;
    if (values) values[i] = grid[ind];
    int x[3];
    for (int d = 0; d < 3; ++d) { x[d] = ind / stride[d]; ind %= stride[d]; }
    for (int d = 0; d < 3; ++d) {  // WrappedIndex::unwrap, Math/fft.hpp:199-206
      int y = -1 * x[d];
      if (y < -1 * (n[d] / 2)) y += n[d];
      if (s) s[3 * i + d] = y;
      if (tran) tran[3 * i + d] = h * y;  // toSpatialDiff, gridcoord.hpp:150-153
    }
  }
  return nout;
}

}  // namespace

extern "C" {

int pdo_field_nparams(int kind) {
  switch (kind) {
    case PDO_FIELD_LJ_REP: return 2;
    case PDO_FIELD_LJ_ATT: return 2;
    case PDO_FIELD_LJ_FULL: return 5;
    case PDO_FIELD_TWOSTEP: return 3;
    case PDO_FIELD_COULOMB: return 1;
    case PDO_FIELD_DESOLV: return 1;
  }
  return -1;
}

int pdo_find_first_prime_factoring(int startN, int endN, const int* primes, int nPrimes, int* exps,
                                   int* resultN) {
  for (int N = (startN <= 1 ? 2 : startN); N < endN; ++N) {
    int n = N;
    for (int ip = 0; ip < nPrimes; ++ip) {
      int prime = primes[ip], e = 0;
      while (n % prime == 0) { n /= prime; ++e; }
      exps[ip] = e;
      if (n == 1) {
        for (int jp = ip + 1; jp < nPrimes; ++jp) exps[jp] = 0;
        *resultN = N;
        return 1;
      }
    }
  }
  return 0;
}

int pdo_find_best_size(int startN, int* resultN) {
  static const int primes[6] = {2, 3, 5, 7, 11, 13};  // Math/fft.cc:8
  int exps[6];
  const int endN = 2147483647;
  while (true) {
    if (pdo_find_first_prime_factoring(startN, endN, primes, 6, exps, resultN)) {
      if (exps[4] + exps[5] <= 1) return 1;
      startN = *resultN + 1;
    } else {
      return 0;
    }
  }
}

void pdo_geometry_centered(double h, int n, double* zero_h, double* zsc) {
  Axis a = centered_axis(h, n);
  *zero_h = a.zero_h;
  *zsc = a.zero_start_center;
}

int pdo_to_logical(double h, double zero_h, double x) {
  Axis a;
  a.h = h;
  a.zero_h = zero_h;
  return to_logical(a, x);
}

int pdo_grid_size_from_box(double lo, double hi, double h) {
  // Geom(grStep, boxDiag(0)): zero_ = lo, zero_h_ = lo - h; findBestSize(toLogical(hi))
  Axis a;
  a.h = h;
  a.zero_h = lo - h;
  int start = to_logical(a, hi), res = 0;
  if (!pdo_find_best_size(start, &res)) return -1;
  return res;
}

void pdo_euler_to_matrix(const double* ang, double* m) {
  double a = ang[0], b = ang[1], g = ang[2];
  double sa = std::sin(a), ca = std::cos(a), sb = std::sin(b), cb = std::cos(b),
         sg = std::sin(g), cg = std::cos(g);
  m[0] = cg * ca;                m[1] = cg * sa;                m[2] = -sg;
  m[3] = sb * sg * ca - cb * sa; m[4] = sb * sg * sa + cb * ca; m[5] = cg * sb;
  m[6] = cb * sg * ca + sb * sa; m[7] = cb * sg * sa - sb * ca; m[8] = cg * cb;
}

void pdo_matrix_to_euler(const double* m, double* ang, int* unique) {
  const double HALF_PI = 4.0 * std::atan(1.0) / 2.;
  double a0, a1, a2;
  int u;
  if (m[2] < 1.0) {
    if (m[2] > -1.0) {
      a0 = std::atan2(-m[5], m[8]);
      a1 = std::asin(m[2]);
      a2 = std::atan2(-m[1], m[0]);
      u = 1;
    } else {
      a0 = -std::atan2(m[3], m[4]);
      a1 = -HALF_PI;
      a2 = 0.0;
      u = 0;
    }
  } else {
    a0 = std::atan2(m[3], m[4]);
    a1 = HALF_PI;
    a2 = 0.0;
    u = 0;
  }
  ang[0] = -a2;
  ang[1] = -a0;
  ang[2] = -a1;
  if (unique) *unique = u;
}

int pdo_load_rotational_grid(const char* path, double* mats, int cap) {
  std::ifstream in(path);
  if (!in.good()) return -1;
  int n = -1;
  in >> n;
  if (n <= 0) return -1;
  in.ignore(1000, '\n');
  const double Pi = 4.0 * std::atan(1.0);
  const double degToRad = Pi / 180.;
  for (int i = 0; i < n; ++i) {
    double a[3];
    in >> a[0] >> a[1] >> a[2];
    if (!in.good() && i != n - 1) return -2;
    for (int d = 0; d < 3; ++d) a[d] *= degToRad;
    if (i < cap) pdo_euler_to_matrix(a, mats + 9 * i);
  }
  return n;
}

void pdo_rotate_points(const double* m, int n, const double* in, double* out) {
  for (int j = 0; j < n; ++j) rotate(m, in + 3 * j, out + 3 * j);
}

void pdo_matmul3(const double* a, const double* b, double* c) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double z = 0;
      for (int k = 0; k < 3; ++k) z += a[3 * i + k] * b[3 * k + j];
      c[3 * i + j] = z;
    }
}

void pdo_project_receptor(const int* n, double h, int n_atoms, const double* pos, int kind,
                          const double* params, double cutoff, double min_dist, double* grid) {
  Grid3 g = make_grid(n, h);
  std::fill(grid, grid + g.N(), 0.0);
  // MolForceRadialR::projectReceptor (docking_molforce.hpp:307-316): minDist
  // defaults to grid.minSpatialStep()/4.
  if (min_dist < 0) min_dist = h / 4.;
  const double minDist2 = min_dist * min_dist;
  const int np = pdo_field_nparams(kind);
  for (int a = 0; a < n_atoms; ++a)
    project_field(g, kind, params + (size_t)np * a, cutoff, minDist2, pos + 3 * a, grid);
}

int pdo_project_ligand(const int* n, double h, int n_atoms, const double* pos, const double* w,
                       double* grid) {
  Grid3 g = make_grid(n, h);
  return project_ligand(g, n_atoms, pos, w, grid);
}

void pdo_cells(const int* n, double h, int n_atoms, const double* pos, int* cells) {
  Grid3 g = make_grid(n, h);
  for (int j = 0; j < n_atoms; ++j)
    for (int d = 0; d < 3; ++d) cells[3 * j + d] = to_logical(g.ax[d], pos[3 * j + d]);
}

void pdo_ljcomp_params(int n_rec, const double* eps_r, const double* sigma_r,
                       const double* alpha_r, int n_lig, const double* eps_l,
                       const double* sigma_l, double* rep, double* att, double* w_rep,
                       double* w_att) {
  double sum = 0;  // blitz::mean(params.sigma(iLig)) (docking_molforce.hpp:475)
  for (int j = 0; j < n_lig; ++j) sum += sigma_l[j];
  double sigma_avg_l = sum / n_lig;
  for (int a = 0; a < n_rec; ++a) {  // LJRep/LJAttr ctors :485-497, :551-563
    double sigma = std::sqrt(sigma_r[a]);
    double sigmaAlpha = std::sqrt(sigma_r[a] * sigma_avg_l);
    double eps = std::sqrt(eps_r[a]);
    double s6 = pow6(sigma);                       // potentials_inc.hpp:297-301
    double as6 = alpha_r[a] * pow6(sigmaAlpha);
    double e4s6 = 4. * eps * s6;
    double e4s12 = e4s6 * s6;
    rep[2 * a] = e4s12;
    rep[2 * a + 1] = as6;
    att[2 * a] = e4s6;
    att[2 * a + 1] = as6;
  }
  for (int j = 0; j < n_lig; ++j) {  // :505-510, :571-576
    w_rep[j] = std::sqrt(eps_l[j]) * pow6(sigma_l[j]);
    w_att[j] = std::sqrt(eps_l[j]) * pow3(sigma_l[j]);
  }
}

void pdo_ljavg_params(int n_rec, const double* eps_r, const double* sigma_r,
                      const double* alpha_r, int n_lig, const double* eps_l,
                      const double* sigma_l, int mix, double cutoff, double* params,
                      double* w) {
  double se = 0, ss = 0;  // docking_molforce.hpp:395-396
  for (int j = 0; j < n_lig; ++j) { se += eps_l[j]; ss += sigma_l[j]; }
  double eps_avg_l = se / n_lig, sigma_avg_l = ss / n_lig;
  for (int a = 0; a < n_rec; ++a) {  // :406-432
    double sigma = mix == 0 ? (sigma_r[a] + sigma_avg_l) / 2. : std::sqrt(sigma_r[a] * sigma_avg_l);
    double eps = std::sqrt(eps_r[a] * eps_avg_l);
    double alpha = alpha_r[a];
    // SoftCoreLJ ctor, potentials_inc.hpp:185-208, CUT_POLICY_SHIFT_FORCE
    double s6 = pow6(sigma), as6 = alpha * s6, e4s6 = 4. * eps * s6, e4s12 = e4s6 * s6;
    double cut = cutoff, cutP2 = cut * cut;
    // g2(cutP2, (cut,0,0)) :221-228
    double r6 = pow3(cutP2), A = r6 + as6, B = s6 / (r6 + as6) - 1.;
    double C = -6. * (cutP2 * cutP2) * (e4s6 * A * B + e4s12) / pow3(A);
    double gcut0 = cut * C + cut * (0.0 / std::sqrt(cutP2));
    double p0[5] = {e4s12, e4s6, as6, 0.0, 0.0};
    double fcut = field_f2(PDO_FIELD_LJ_FULL, p0, cutP2, cutP2);
    double fshift = -fcut;
    double gshift = -gcut0;
    fshift -= gshift * cut;
    double* p = params + 5 * a;
    p[0] = e4s12; p[1] = e4s6; p[2] = as6; p[3] = fshift; p[4] = gshift;
  }
  for (int j = 0; j < n_lig; ++j) w[j] = 1;  // :439-446
}

void pdo_fft1(int n, int sign, double* data) {
  const Plan1& P = get_plan(n);
  std::vector<double> scratch(2 * (size_t)n);
  fft_line(P, sign, data, scratch.data());
}

void pdo_rfft3(const int* n, const double* real, double* cplx) { rfft3(n, real, cplx); }
void pdo_irfft3(const int* n, const double* cplx, double* real) { irfft3(n, cplx, real); }

void pdo_correlate(const int* n, int n_terms, const double* rec_spectra, const double* lig_grids,
                   double* out) {
  std::vector<double> spec, tmp;
  correlate(n, n_terms, rec_spectra, lig_grids, out, spec, tmp);
}

int pdo_select(const int* n, double h, const double* grid, int max_inp, double max_val, int n_max,
               int* offsets, int* s, double* tran, double* values) {
  return select(n, h, grid, max_inp, max_val, n_max, offsets, s, tran, values);
}

double pdo_direct_energy(const int* n, int n_terms, const double* rec_grids, int n_lig,
                         const int* cells, const double* w, const int* s) {
  const size_t N = (size_t)n[0] * n[1] * n[2];
  double e = 0;
  for (int t = 0; t < n_terms; ++t)
    for (int j = 0; j < n_lig; ++j) {
      int c[3];
      for (int d = 0; d < 3; ++d) c[d] = (((cells[3 * j + d] + s[d]) % n[d]) + n[d]) % n[d];
      e += w[(size_t)t * n_lig + j] * rec_grids[t * N + ((size_t)c[0] * n[1] + c[1]) * n[2] + c[2]];
    }
  return e;
}

int pdo_scan(const int* n, double h, int n_terms, const double* rec_spectra, int n_lig,
             const double* lig_pos, const double* w, int n_rot, const double* rots, int max_inp,
             double max_val, int K, int n_threads, int* n_out, int* offsets, int* s,
             double* values, double* corr_out) {
  if (n_threads < 1) n_threads = 1;
  const size_t N = (size_t)n[0] * n[1] * n[2];
  // Build the FFT plans once per thread lazily (thread_local cache).
  std::vector<int> status(n_threads, 0);
  auto work = [&](int tid) {
    Grid3 g = make_grid(n, h);
    std::vector<double> pos((size_t)3 * n_lig), lig((size_t)n_terms * N), corr(N), spec, tmp;
    for (int r = tid; r < n_rot; r += n_threads) {
      // Worker::scanRotation: resetLigPosRot + currentRot(ligPosRot)
      pdo_rotate_points(rots + 9 * r, n_lig, lig_pos, pos.data());
      for (int t = 0; t < n_terms; ++t) {
        int bad = project_ligand(g, n_lig, pos.data(), w + (size_t)t * n_lig, lig.data() + t * N);
        if (bad) { status[tid] = r + 1; return; }
      }
      double* c = corr_out ? corr_out + (size_t)r * N : corr.data();
      correlate(n, n_terms, rec_spectra, lig.data(), c, spec, tmp);
      n_out[r] = select(n, h, c, max_inp, max_val, K, offsets + (size_t)r * K, s + (size_t)3 * r * K,
                        nullptr, values + (size_t)r * K);
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < n_threads; ++t) th.emplace_back(work, t);
  work(0);
  for (auto& t : th) t.join();
  for (int t = 0; t < n_threads; ++t)
    if (status[t]) return status[t];
  return 0;
}

// ---------------------------------------------------------------------------
// Post-filters (TranSymm, TranClust) on one rotation's translations.
// ---------------------------------------------------------------------------

namespace {

// RotationTranslation: x -> m x + v (Geom/transformation.hpp:244-360).
struct Rigid {
  double m[9];
  double v[3];
};

Rigid rigid_from(const double* t12) {
  Rigid r;
  for (int i = 0; i < 9; ++i) r.m[i] = t12[i];
  for (int i = 0; i < 3; ++i) r.v[i] = t12[9 + i];
  return r;
}

// blitz_ext::product(tensor, vector) (bztinyvec.hpp:52-65)
void product_mv(const double* m, const double* x, double* y) {
  for (int i = 0; i < 3; ++i) {
    double acc = 0;
    for (int j = 0; j < 3; ++j) acc += m[3 * i + j] * x[j];
    y[i] = acc;
  }
}

// RotationTranslation::operator*(RotationTranslation) (transformation.hpp:323-328):
// m = a.m b.m, v = a.m b.v + a.v
Rigid compose(const Rigid& a, const Rigid& b) {
  Rigid r;
  pdo_matmul3(a.m, b.m, r.m);
  double p[3];
  product_mv(a.m, b.v, p);
  for (int i = 0; i < 3; ++i) r.v[i] = p[i] + a.v[i];
  return r;
}

// SymmetryCheckerCn::operator() (Geom/symmetry.hpp:35-51)
double symm_rmsd(const Rigid& t, int n_symm, double radius) {
  Rigid tn = t;
  for (int i = 1; i < n_symm; ++i) tn = compose(t, tn);
  const double u[4][3] = {{radius, 0, 0}, {0, radius, 0}, {0, 0, radius}, {0, 0, 0}};
  double sum = 0;
  for (int k = 0; k < 4; ++k) {
    double y[3];
    product_mv(tn.m, u[k], y);  // RotationTranslation::operator()(Point), :330-332
    double d2 = 0;
    for (int i = 0; i < 3; ++i) {
      const double d = (y[i] + tn.v[i]) - u[k][i];
      d2 += d * d;
    }
    sum += d2;
  }
  return std::sqrt(sum / 4);
}

}  // namespace

int pdo_tran_symm(int n, double* tran, double* values, int* tag, const double* rot, const double* frame_a,
                  const double* frame_b, int n_symm, double mol_radius, double max_rmsd, double* rmsd) {
  const Rigid A = rigid_from(frame_a), B = rigid_from(frame_b);
  Rigid R;
  for (int i = 0; i < 9; ++i) R.m[i] = rot[i];
  int out = 0;
  for (int i = 0; i < n; ++i) {
    // Translation(tran) * Rotation -> RotationTranslation(rot, tran) (transformation.hpp:84-86),
    // then m_tR * t * m_tL, left to right (docking_molstruct.hpp:41-43)
    for (int d = 0; d < 3; ++d) R.v[d] = tran[3 * i + d];
    const Rigid t = compose(compose(A, R), B);
    const double r = symm_rmsd(t, n_symm, mol_radius);
    if (rmsd) rmsd[i] = r;
    if (r < max_rmsd) {
      for (int d = 0; d < 3; ++d) tran[3 * out + d] = tran[3 * i + d];
      values[out] = values[i];
      if (tag) tag[out] = tag[i];
      ++out;
    }
  }
  return out;
}

int pdo_tran_clust(int n, double* tran, double* values, int* tag, double cluster_radius, int max_out,
                   int cap) {
  if (n <= 0) return 0;
  // weights = -value (docking_fft.hpp:71); the density is then overwritten by
  // the weights (cluster.hpp:291-292), so rows are visited in ascending -value
  const double r2c = cluster_radius * cluster_radius;
  const double cutoff2 = r2c * 3;                   // cluster.hpp:160-162 (M_col = 3)
  const double cutoff = std::sqrt(cutoff2);
  const double rad2 = cutoff * cutoff;              // SphereIter: radius = cutoff (partpoints.hpp:551-562)
  std::vector<double> dens(n);
  for (int i = 0; i < n; ++i) dens[i] = -values[i];
  std::vector<int> ind(n), place(n), leader(n);
  for (int i = 0; i < n; ++i) ind[i] = leader[i] = i;
  std::sort(ind.begin(), ind.end(), [&](int a, int b) { return dens[a] < dens[b]; });  // :308-309
  for (int i = 0; i < n; ++i) place[ind[i]] = i;
  std::vector<int> leaders;
  for (int id = 0; id < n; ++id) {
    const int i = ind[id];
    int best = place[i], lead = i;
    // every leader inserted so far within the radius (dist2 <= rad2,
    // partpoints.hpp:584) is a candidate; the earliest in density order wins
    for (int k : leaders) {
      double d2 = 0;
      for (int d = 0; d < 3; ++d) {
        const double t = tran[3 * i + d] - tran[3 * k + d];
        d2 += t * t;
      }
      if (d2 <= rad2 && place[k] < best) {
        best = place[k];
        lead = k;
      }
    }
    if (lead == i)
      leaders.push_back(i);
    else
      leader[i] = lead;
  }
  // getClusters (cluster.hpp:395-414): leaders in density order; then
  // TranClust::process copies them to the front (docking_fft.hpp:95-112)
  int n_out = (int)leaders.size();
  n_out = std::min(n_out, cap);
  n_out = std::min(n_out, max_out);
  const std::vector<double> t3(tran, tran + 3 * (size_t)n), v(values, values + n);
  std::vector<int> tg;
  if (tag) tg.assign(tag, tag + n);
  for (int c = 0; c < n_out; ++c) {
    const int r = leaders[c];
    for (int d = 0; d < 3; ++d) tran[3 * c + d] = t3[3 * r + d];
    values[c] = v[r];
    if (tag) tag[c] = tg[r];
  }
  return n_out;
}

int pdo_select_post(const int* n, double h, const double* grid, int max_inp, double max_val, int K,
                    const double* rot, const double* frame_a, const double* frame_b, int n_symm,
                    double mol_radius, double max_rmsd, double cluster_radius, int* offsets, int* s,
                    double* values) {
  std::vector<int> off(max_inp), ss(3 * (size_t)max_inp);
  std::vector<double> tran(3 * (size_t)max_inp), val(max_inp);
  int m = select(n, h, grid, max_inp, max_val, max_inp, off.data(), ss.data(), tran.data(), val.data());
  std::vector<int> tag(m);
  for (int i = 0; i < m; ++i) tag[i] = i;
  if (n_symm > 1)
    m = pdo_tran_symm(m, tran.data(), val.data(), tag.data(), rot, frame_a, frame_b, n_symm, mol_radius, max_rmsd,
                      nullptr);
  if (cluster_radius > 0) m = pdo_tran_clust(m, tran.data(), val.data(), tag.data(), cluster_radius, K, max_inp);
  m = std::min(m, K);
  for (int i = 0; i < m; ++i) {
    const int t = tag[i];
    offsets[i] = off[t];
    for (int d = 0; d < 3; ++d) s[3 * i + d] = ss[3 * t + d];
    values[i] = val[i];
  }
  return m;
}

int pdo_write_scan_record(const char* path, int append, const double* rot, int n,
                          const double* tran, const double* values) {
  FILE* f = std::fopen(path, append ? "ab" : "wb");
  if (!f) return -1;
  int32_t sig = 2014011401;  // RotFftScanIO_Bin::SIGNATURE (docking_io_bin.hpp:23)
  std::fwrite(&sig, 4, 1, f);
  std::fwrite(rot, 8, 9, f);
  uint64_t nn = (uint64_t)n;
  std::fwrite(&nn, 8, 1, f);
  for (int i = 0; i < n; ++i) {
    std::fwrite(tran + 3 * i, 8, 3, f);
    std::fwrite(values + i, 8, 1, f);
  }
  std::fclose(f);
  return 0;
}

int pdo_read_scan_records(const char* path, int cap, int maxn, double* rots, int* counts,
                          double* tran, double* values) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return -1;
  int nrec = 0;
  while (nrec < cap) {
    int32_t sig;
    if (std::fread(&sig, 4, 1, f) != 1) break;
    if (sig != 2014011401) { std::fclose(f); return -2; }
    double rot[9];
    uint64_t nn;
    if (std::fread(rot, 8, 9, f) != 9 || std::fread(&nn, 8, 1, f) != 1) break;
    std::vector<double> rec(4 * nn);
    if (nn && std::fread(rec.data(), 8, 4 * nn, f) != 4 * nn) break;
    std::memcpy(rots + 9 * nrec, rot, sizeof rot);
    counts[nrec] = (int)nn;
    for (uint64_t i = 0; i < nn && (int)i < maxn; ++i) {
      for (int d = 0; d < 3; ++d) tran[((size_t)nrec * maxn + i) * 3 + d] = rec[4 * i + d];
      values[(size_t)nrec * maxn + i] = rec[4 * i + 3];
    }
    ++nrec;
  }
  std::fclose(f);
  return nrec;
}

}  // extern "C"
