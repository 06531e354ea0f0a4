// gdiam_shim.cpp -- C entry points over the reference's vendored gdiam
// library (src/PRODDL/External/Gdiam/gdiam.cpp through
// src/PRODDL/Geom/gdiam_simple.cpp), compiled from the reference's own source
// files by `make ref` into oracle/_ref.  Test infrastructure only: it checks
// the library's own minimum-volume box (pd_min_volume_box) and diameter
// against what the reference's MolStruct::moveIntoMinBox computes
// (docking_molstruct.hpp:199-245, Geom/bounding.hpp:33-183).
#include "PRODDL/Geom/gdiam_simple.hpp"

extern "C" {

// Geom::Bounding::Box(points, fixedDirections = false, grid_n, sample_n):
// left/right: the box corners (0,0,0) and (1,1,1) in the input frame,
// dirs: 3 x 3, row k the box direction k.
void gdref_mvbb(const double* pts, int n, int grid_n, int sample_n, double* left, double* right, double* dirs) {
  PRODDL::Geom::Gdiam::simp_approx_mvbb_grid_sample(pts, n, left, right, dirs, grid_n, sample_n < 0 ? n : sample_n);
}

// Geom::Bounding::Diameter(points, eps): the two end points.
void gdref_diam(const double* pts, int n, double eps, double* p, double* q) {
  PRODDL::Geom::Gdiam::simp_approx_diam_pair(pts, n, p, q, eps);
}
}
