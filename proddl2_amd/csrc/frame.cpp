// frame.cpp -- the docking frame's geometry on the host: an approximate
// minimum-volume bounding box of the receptor and the diameter of a point set.
//
// The reference orients the receptor along its minimum-volume box
// (MolStruct::moveIntoMinBox, docking_molstruct.hpp:199-245, through
// Geom::Bounding::Box with fixedDirections = false, Geom/bounding.hpp:135-183)
// and centres the ligand on its diameter (Geom::Bounding::Diameter with
// eps = 0, bounding.hpp:33-97).  Both call the vendored gdiam library.  This
// file implements the published method instead (G. Barequet and S. Har-Peled,
// "Efficiently approximating the minimum-volume bounding box of a point set in
// three dimensions", J. Algorithms 38 (2001) 91-109):
//
//   1. a constant-factor box: one axis along the diameter, the second along the
//      diameter of the points projected onto the plane orthogonal to it, the
//      third orthogonal to both (or the axis-parallel box when smaller);
//   2. the box with one axis fixed along a direction v is found exactly in the
//      plane orthogonal to v: the smallest-area rectangle enclosing the
//      projected points has a side collinear with an edge of their convex hull
//      (H. Freeman and R. Shapira, CACM 18 (1975) 409-413), so every hull edge
//      is tried;
//   3. local improvement: each of the current box's axes in turn is kept and
//      the other two re-optimised (step 2), a candidate replacing the box only
//      when its volume is smaller; 10 rounds;
//   4. a grid of directions: every primitive integer vector (a, b, c) with
//      |a|, |b| <= 5 and 0 <= c <= 5, taken in the improved box's axes, seeds a
//      box (step 2) that is improved (step 3); the smallest volume wins (the
//      first in enumeration order among equal volumes), and is improved once
//      more.
//
// The reference's grid size 5 and sample of all points (bounding.hpp:145,
// docking_molstruct.hpp:203) are the parameters used here.  The box axes
// come out as a proper rotation: ordered by decreasing extent, the largest
// component of the first two made positive, the third their cross product.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <new>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "../../include/pdfft.h"

namespace pd_detail {
std::string& last_error();
}

namespace {

pd_status invalid(const char* msg) {
  pd_detail::last_error() = msg;
  return PD_ERR_INVALID_ARGUMENT;
}

struct V3 {
  double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 unit(V3 a) {
  const double l = norm(a);
  return l > 0 ? (1.0 / l) * a : a;
}

struct P2 {
  double x, y;
};
inline double cross2(P2 o, P2 a, P2 b) { return (a.x - o.x) * (b.y - o.y) - (a.y - o.y) * (b.x - o.x); }

struct OBox {
  V3 ax[3];
  double lo[3], hi[3];
  double volume() const { return (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]); }
};

using Pts = std::vector<V3>;

void bound(OBox& b, const Pts& p) {
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = HUGE_VAL;
    b.hi[k] = -HUGE_VAL;
  }
  for (const V3& q : p)
    for (int k = 0; k < 3; ++k) {
      const double d = dot(b.ax[k], q);
      b.lo[k] = std::min(b.lo[k], d);
      b.hi[k] = std::max(b.hi[k], d);
    }
}

OBox make_box(V3 a0, V3 a1, V3 a2, const Pts& p) {
  OBox b;
  b.ax[0] = a0;
  b.ax[1] = a1;
  b.ax[2] = a2;
  bound(b, p);
  return b;
}

// An orthonormal pair spanning the plane orthogonal to the unit vector v: the
// coordinate axis least aligned with v, made orthogonal to it.
void plane_basis(V3 v, V3& e1, V3& e2) {
  const double ax = std::fabs(v.x), ay = std::fabs(v.y), az = std::fabs(v.z);
  const V3 a = (ax <= ay && ax <= az) ? V3{1, 0, 0} : (ay <= az ? V3{0, 1, 0} : V3{0, 0, 1});
  e1 = unit(cross(v, a));
  e2 = cross(v, e1);
}

// Convex hull, counter-clockwise, without collinear points (Andrew's monotone
// chain).  `q` is reordered.
std::vector<P2> hull2(std::vector<P2>& q) {
  std::sort(q.begin(), q.end(), [](P2 a, P2 b) { return a.x < b.x || (a.x == b.x && a.y < b.y); });
  q.erase(std::unique(q.begin(), q.end(), [](P2 a, P2 b) { return a.x == b.x && a.y == b.y; }), q.end());
  const int n = (int)q.size();
  if (n < 3) return q;
  std::vector<P2> h(2 * n);
  int k = 0;
  for (int i = 0; i < n; ++i) {
    while (k >= 2 && cross2(h[k - 2], h[k - 1], q[i]) <= 0) --k;
    h[k++] = q[i];
  }
  for (int i = n - 2, t = k + 1; i >= 0; --i) {
    while (k >= t && cross2(h[k - 2], h[k - 1], q[i]) <= 0) --k;
    h[k++] = q[i];
  }
  h.resize(k - 1);
  return h;
}

// The smallest box with one axis along v (step 2).
OBox best_box_along(V3 v, const Pts& p) {
  v = unit(v);
  V3 e1, e2;
  plane_basis(v, e1, e2);
  std::vector<P2> q(p.size());
  for (size_t i = 0; i < p.size(); ++i) q[i] = {dot(p[i], e1), dot(p[i], e2)};
  const std::vector<P2> h = hull2(q);
  P2 best_u{1, 0};
  if (h.size() == 2) {
    best_u = {h[1].x - h[0].x, h[1].y - h[0].y};
  } else if (h.size() >= 3) {
    double best = HUGE_VAL;
    const int m = (int)h.size();
    for (int i = 0; i < m; ++i) {
      const P2 a = h[i], b = h[(i + 1) % m];
      const double l = std::hypot(b.x - a.x, b.y - a.y);
      if (!(l > 0)) continue;
      const P2 u{(b.x - a.x) / l, (b.y - a.y) / l};
      double s0 = HUGE_VAL, s1 = -HUGE_VAL, t0 = HUGE_VAL, t1 = -HUGE_VAL;
      for (const P2& c : h) {
        const double s = c.x * u.x + c.y * u.y, t = -c.x * u.y + c.y * u.x;
        s0 = std::min(s0, s);
        s1 = std::max(s1, s);
        t0 = std::min(t0, t);
        t1 = std::max(t1, t);
      }
      const double area = (s1 - s0) * (t1 - t0);
      if (area < best) {
        best = area;
        best_u = u;
      }
    }
  }
  const double l = std::hypot(best_u.x, best_u.y);
  const P2 u{best_u.x / l, best_u.y / l};
  const V3 a1 = unit(u.x * e1 + u.y * e2);
  const V3 a2 = unit(cross(v, a1));
  return make_box(v, a1, a2, p);
}

// Step 3.
OBox improve(OBox b, const Pts& p, int rounds) {
  for (int r = 0; r < rounds; ++r) {
    const OBox c = best_box_along(b.ax[r % 3], p);
    if (c.volume() < b.volume()) b = c;
  }
  return b;
}

// Indices of the two points farthest apart (the first such pair in (i, j)
// order, i < j).
void diameter(const Pts& p, int& bi, int& bj) {
  bi = 0;
  bj = p.size() > 1 ? 1 : 0;
  double best = -1;
  const int n = (int)p.size();
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      const V3 d = p[i] - p[j];
      const double s = dot(d, d);
      if (s > best) {
        best = s;
        bi = i;
        bj = j;
      }
    }
}

// Step 1.
OBox constant_box(const Pts& p) {
  OBox ap = make_box({1, 0, 0}, {0, 1, 0}, {0, 0, 1}, p);
  int i, j;
  diameter(p, i, j);
  const V3 d1 = unit(p[i] - p[j]);
  if (!(norm(d1) > 0)) return ap;
  V3 e1, e2;
  plane_basis(d1, e1, e2);
  std::vector<P2> q(p.size());
  for (size_t k = 0; k < p.size(); ++k) q[k] = {dot(p[k], e1), dot(p[k], e2)};
  const std::vector<P2> h = hull2(q);
  P2 w{1, 0};
  double best = -1;
  for (size_t a = 0; a < h.size(); ++a)
    for (size_t b = a + 1; b < h.size(); ++b) {
      const double s = (h[a].x - h[b].x) * (h[a].x - h[b].x) + (h[a].y - h[b].y) * (h[a].y - h[b].y);
      if (s > best) {
        best = s;
        w = {h[a].x - h[b].x, h[a].y - h[b].y};
      }
    }
  const V3 d2 = unit(w.x * e1 + w.y * e2);
  if (!(norm(d2) > 0)) return ap;
  const OBox b = make_box(d1, d2, unit(cross(d1, d2)), p);
  return ap.volume() < b.volume() ? ap : b;
}

int gcd(int a, int b) {
  a = std::abs(a);
  b = std::abs(b);
  while (b) {
    const int t = a % b;
    a = b;
    b = t;
  }
  return a;
}

OBox min_volume_box(const Pts& p) {
  const OBox base = improve(constant_box(p), p, 10);
  constexpr int G = 5;
  std::vector<V3> dirs;
  for (int a = -G; a <= G; ++a)
    for (int b = -G; b <= G; ++b)
      for (int c = 0; c <= G; ++c) {
        if (gcd(gcd(a, b), c) != 1) continue;  // zero vector and non-primitive multiples
        dirs.push_back((double)a * base.ax[0] + (double)b * base.ax[1] + (double)c * base.ax[2]);
      }
  // Each direction's box is independent of the others: evaluate them on a few
  // threads, then reduce in enumeration order.
  std::vector<OBox> cand(dirs.size());
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const size_t nt = std::min<size_t>(hw, dirs.size());
  auto work = [&](size_t t) {
    for (size_t k = t; k < dirs.size(); k += nt) cand[k] = improve(best_box_along(dirs[k], p), p, 10);
  };
  std::vector<std::thread> pool;
  size_t started = 0;
  try {
    for (; started < nt; ++started) pool.emplace_back(work, started);
  } catch (const std::system_error&) {
    // no more threads: this thread takes the slices that did not start
  }
  for (size_t t = started; t < nt; ++t) work(t);
  for (auto& th : pool) th.join();
  OBox best = base;
  for (const OBox& c : cand)
    if (c.volume() < best.volume()) best = c;
  return improve(best, p, 10);
}

// Axes ordered by decreasing extent, the first two with their largest
// component positive, the third their cross product; extents recomputed.
OBox canonical(const OBox& b, const Pts& p) {
  int ord[3] = {0, 1, 2};
  std::stable_sort(ord, ord + 3, [&](int x, int y) { return b.hi[x] - b.lo[x] > b.hi[y] - b.lo[y]; });
  V3 a[2];
  for (int k = 0; k < 2; ++k) {
    V3 v = b.ax[ord[k]];
    const double c[3] = {v.x, v.y, v.z};
    int m = 0;
    for (int i = 1; i < 3; ++i)
      if (std::fabs(c[i]) > std::fabs(c[m])) m = i;
    if (c[m] < 0) v = -1.0 * v;
    a[k] = v;
  }
  return make_box(a[0], a[1], unit(cross(a[0], a[1])), p);
}

Pts points(int n, const double* pos) {
  Pts p(n);
  for (int i = 0; i < n; ++i) p[i] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
  return p;
}

bool finite(int n, const double* pos) {
  for (int i = 0; i < 3 * n; ++i)
    if (!std::isfinite(pos[i])) return false;
  return true;
}

}  // namespace

extern "C" pd_status pd_min_volume_box(int n, const double* pos, double* axes, double* lo, double* hi) {
  if (n < 1 || !pos || !axes || !lo || !hi) return invalid("pd_min_volume_box: null pointer or no points");
  if (!finite(n, pos)) return invalid("pd_min_volume_box: non-finite coordinate");
  OBox b;
  try {
    const Pts p = points(n, pos);
    b = canonical(min_volume_box(p), p);
  } catch (const std::bad_alloc&) {
    return invalid("pd_min_volume_box: out of host memory");
  }
  for (int k = 0; k < 3; ++k) {
    axes[3 * k] = b.ax[k].x;
    axes[3 * k + 1] = b.ax[k].y;
    axes[3 * k + 2] = b.ax[k].z;
    lo[k] = b.lo[k];
    hi[k] = b.hi[k];
  }
  return PD_OK;
}

extern "C" pd_status pd_diameter_pair(int n, const double* pos, int* i, int* j) {
  if (n < 1 || !pos || !i || !j) return invalid("pd_diameter_pair: null pointer or no points");
  if (!finite(n, pos)) return invalid("pd_diameter_pair: non-finite coordinate");
  try {
    diameter(points(n, pos), *i, *j);
  } catch (const std::bad_alloc&) {
    return invalid("pd_diameter_pair: out of host memory");
  }
  return PD_OK;
}
