// proddl_dock_fft.cpp -- drop-in for the reference's proddl-dock-fft worker
// (src/PRODDL/docking_main.cpp:50-157, include/PRODDL/docking_app.hpp:316-513).
//
//   proddl-dock-fft --task rot-scan --options OPT.json --molforce-params MF.pdmf
//                   --fft-rot-grid-start S --fft-rot-grid-end E --fft-rot-scan-res OUT.bin
//   proddl-dock-fft --task gather --options OPT.json --molforce-params MF.pdmf
//                   --fft-rot-scan-list LIST.txt --fft-res RES.bin
//
// Same flags, same option keys (gridStep, cutOffFft, anglesFile, maxNTrans,
// maxNTransInp, maxValCorr, potentialName, startingSpin, maxNRigidMatches) and
// the same RotFftScanIO_Bin output records (docking_io_bin.hpp:21-76).  The
// scan itself runs in libpdfft.so on the GPU through the C ABI.
//
// Differences from the reference, all forced by this image: molforce
// parameters come in a small binary container (PDMF, written by
// proddl2_amd.mfp) because HDF5 is absent; the receptor frame is its
// approximate minimum-volume box comes from the library's own implementation
// of the published method (pd_min_volume_box) instead of the vendored gdiam:
// the same box, its axes possibly in another order and sign
// (docking_molstruct.hpp:199-245).
#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "pdfft.h"

namespace {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// ---------------------------------------------------------------------------
// Options: a JSON object of numbers, strings and booleans (nested objects are
// parsed and kept under "outer.inner" keys), like the reference's Options
// (Common/options_io_json.cpp:9-74).  Command-line values override, with
// dashes turned into underscores (Common/argparse.hpp:25-27).
// ---------------------------------------------------------------------------
class Options {
 public:
  std::map<std::string, std::string> kv;

  bool has(const std::string& k) const { return kv.count(k) != 0; }
  std::string str(const std::string& k) const {
    auto it = kv.find(k);
    if (it == kv.end()) throw Error("options_key_error: missing option '" + k + "'");
    return it->second;
  }
  double num(const std::string& k) const { return std::stod(str(k)); }
  double num(const std::string& k, double d) const { return has(k) ? num(k) : d; }
  int integer(const std::string& k) const { return (int)std::lround(num(k)); }
  int integer(const std::string& k, int d) const { return has(k) ? integer(k) : d; }
  std::string str(const std::string& k, const std::string& d) const { return has(k) ? str(k) : d; }
  bool flag(const std::string& k, bool d) const {
    if (!has(k)) return d;
    const std::string v = str(k);
    return v == "true" || v == "1";
  }
};

class JsonReader {
 public:
  explicit JsonReader(const std::string& s) : s_(s) {}
  void parse(Options& o) {
    ws();
    object(o, "");
  }

 private:
  const std::string& s_;
  size_t i_ = 0;
  void ws() {
    while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) ++i_;
  }
  char peek() {
    ws();
    if (i_ >= s_.size()) throw Error("JSON: unexpected end");
    return s_[i_];
  }
  void expect(char c) {
    if (peek() != c) throw Error(std::string("JSON: expected '") + c + "'");
    ++i_;
  }
  std::string string() {
    expect('"');
    std::string out;
    while (i_ < s_.size() && s_[i_] != '"') {
      if (s_[i_] == '\\' && i_ + 1 < s_.size()) ++i_;
      out += s_[i_++];
    }
    expect('"');
    return out;
  }
  void object(Options& o, const std::string& prefix) {
    expect('{');
    if (peek() == '}') { ++i_; return; }
    while (true) {
      std::string key = string();
      expect(':');
      const std::string full = prefix.empty() ? key : prefix + "." + key;
      char c = peek();
      if (c == '{') {
        object(o, full);
      } else if (c == '"') {
        o.kv[full] = string();
      } else if (c == '[') {
        throw Error("JSON: arrays are not supported in options (key " + full + ")");
      } else {
        size_t b = i_;
        while (i_ < s_.size() && s_[i_] != ',' && s_[i_] != '}' && !std::isspace((unsigned char)s_[i_])) ++i_;
        std::string tok = s_.substr(b, i_ - b);
        if (tok == "null") throw Error("JSON: null is not supported in options (key " + full + ")");
        o.kv[full] = tok;
      }
      if (peek() == ',') { ++i_; continue; }
      expect('}');
      return;
    }
  }
};

std::string read_file(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in.good()) throw Error("Could not open input file: " + path);
  std::stringstream ss;
  ss << in.rdbuf();
  return ss.str();
}

// ---------------------------------------------------------------------------
// Molforce parameters (docking_io_hdf5.hpp:19-71): pos(n,3), mass, eps, sigma,
// alpha (n), mol_offsets(>=2, 2), attribute mix; PDMF container.
// ---------------------------------------------------------------------------
struct Array {
  char dtype = 'd';
  std::vector<uint64_t> shape;
  std::vector<double> d;
  std::vector<int32_t> i;
};

std::map<std::string, Array> read_pdmf(const std::string& path) {
  std::string raw = read_file(path);
  size_t at = 0;
  auto take = [&](void* dst, size_t n) {
    if (at + n > raw.size()) throw Error("truncated molforce file: " + path);
    std::memcpy(dst, raw.data() + at, n);
    at += n;
  };
  char magic[4];
  take(magic, 4);
  if (std::memcmp(magic, "PDMF", 4) != 0) throw Error("not a PDMF molforce file: " + path);
  uint32_t version = 0, count = 0;
  take(&version, 4);
  take(&count, 4);
  if (version != 1) throw Error("unsupported PDMF version");
  std::map<std::string, Array> out;
  for (uint32_t e = 0; e < count; ++e) {
    uint32_t nl = 0;
    take(&nl, 4);
    std::string name(nl, '\0');
    take(&name[0], nl);
    Array a;
    uint8_t dt = 0;
    uint32_t nd = 0;
    take(&dt, 1);
    take(&nd, 4);
    a.dtype = (char)dt;
    a.shape.resize(nd);
    uint64_t total = 1;
    for (uint32_t k = 0; k < nd; ++k) {
      take(&a.shape[k], 8);
      total *= a.shape[k];
    }
    if (a.dtype == 'd') {
      a.d.resize(total);
      take(a.d.data(), 8 * total);
    } else if (a.dtype == 'i') {
      a.i.resize(total);
      take(a.i.data(), 4 * total);
    } else {
      throw Error("PDMF: unknown dtype in " + name);
    }
    out[name] = std::move(a);
  }
  return out;
}

struct Molecule {
  std::vector<double> pos;  // n x 3
  std::vector<double> mass, eps, sigma, alpha;
  int n() const { return (int)mass.size(); }
};

struct MolForceParams {
  Molecule mol[2];  // iRec = 0, iLig = 1
  int mix = 1;
};

MolForceParams load_molforce(const std::string& path) {
  auto ds = read_pdmf(path);
  for (const char* k : {"pos", "mass", "eps", "sigma", "alpha", "mol_offsets", "mix"})
    if (!ds.count(k)) throw Error(std::string("molforce file lacks dataset '") + k + "'");
  MolForceParams p;
  p.mix = ds["mix"].i.empty() ? (int)ds["mix"].d.at(0) : ds["mix"].i.at(0);
  const Array& off = ds["mol_offsets"];
  if (off.shape.size() != 2 || off.shape[0] < 2 || off.shape[1] != 2)
    throw Error("mol_offsets must be Nx2 matrix");
  const size_t natoms = ds["alpha"].d.size();
  for (int m = 0; m < 2; ++m) {
    const int s = off.i[2 * m], e = off.i[2 * m + 1];
    if (s < 0 || e < s || (size_t)e > natoms) throw Error("Molecule offset is past array boundaries");
    Molecule& M = p.mol[m];
    M.pos.assign(ds["pos"].d.begin() + 3 * s, ds["pos"].d.begin() + 3 * e);
    M.mass.assign(ds["mass"].d.begin() + s, ds["mass"].d.begin() + e);
    M.eps.assign(ds["eps"].d.begin() + s, ds["eps"].d.begin() + e);
    M.sigma.assign(ds["sigma"].d.begin() + s, ds["sigma"].d.begin() + e);
    M.alpha.assign(ds["alpha"].d.begin() + s, ds["alpha"].d.begin() + e);
  }
  return p;
}

// ---------------------------------------------------------------------------
// Geometry helpers (Geom/euler.hpp, Geom/transformation.hpp)
// ---------------------------------------------------------------------------
using Mat = std::array<double, 9>;

Mat euler_to_matrix(double a, double b, double g) {
  const double sa = std::sin(a), ca = std::cos(a), sb = std::sin(b), cb = std::cos(b), sg = std::sin(g),
               cg = std::cos(g);
  return {cg * ca, cg * sa, -sg, sb * sg * ca - cb * sa, sb * sg * sa + cb * ca, cg * sb,
          cb * sg * ca + sb * sa, cb * sg * sa - sb * ca, cg * cb};
}

void apply(const Mat& m, double* x) {
  double y[3];
  for (int i = 0; i < 3; ++i) {
    double acc = 0;
    for (int j = 0; j < 3; ++j) acc += m[3 * i + j] * x[j];
    y[i] = acc;
  }
  std::memcpy(x, y, sizeof y);
}

// loadRotationalGrid (Geom/rotational_grid.hpp:25-64)
std::vector<Mat> load_rotational_grid(const std::string& path) {
  std::ifstream in(path);
  if (!in.good()) throw Error("loadRotationalGrid(): Unable to open input file: " + path);
  int n = -1;
  in >> n;
  if (n <= 0) throw Error("loadRotationalGrid(): Unable to read the number of rotations from file: " + path);
  in.ignore(1000, '\n');
  const double deg = 4.0 * std::atan(1.0) / 180.;
  std::vector<Mat> out(n);
  for (int i = 0; i < n; ++i) {
    double a, b, g;
    in >> a >> b >> g;
    out[i] = euler_to_matrix(a * deg, b * deg, g * deg);
  }
  return out;
}

// ---------------------------------------------------------------------------
// Reference frame (docking_molstruct.hpp:199-281)
// ---------------------------------------------------------------------------
// Rigid transformation x -> m x + v (RotationTranslation, Geom/transformation.hpp:242-360).
struct RT {
  Mat m{1, 0, 0, 0, 1, 0, 0, 0, 1};
  double v[3] = {0, 0, 0};
};

// a * b (transformation.hpp: operator*(RotationTranslation))
RT compose(const RT& a, const RT& b) {
  RT r;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      double acc = 0;
      for (int k = 0; k < 3; ++k) acc += a.m[3 * i + k] * b.m[3 * k + j];
      r.m[3 * i + j] = acc;
    }
    double acc = 0;
    for (int k = 0; k < 3; ++k) acc += a.m[3 * i + k] * b.v[k];
    r.v[i] = acc + a.v[i];
  }
  return r;
}

// Rotation(m^T) * Translation(-v) (transformation.hpp: inverse())
RT inverse(const RT& a) {
  RT r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[3 * i + j] = a.m[3 * j + i];
  for (int i = 0; i < 3; ++i) {
    double acc = 0;
    for (int k = 0; k < 3; ++k) acc += r.m[3 * i + k] * (0 - a.v[k]);
    r.v[i] = acc;
  }
  return r;
}

// x -> m x + v in place
void apply(const RT& t, double* x) {
  double y[3];
  for (int i = 0; i < 3; ++i) {
    double acc = 0;
    for (int j = 0; j < 3; ++j) acc += t.m[3 * i + j] * x[j];
    y[i] = acc + t.v[i];
  }
  std::memcpy(x, y, sizeof y);
}

RT translation(double x, double y, double z) {
  RT r;
  r.v[0] = x;
  r.v[1] = y;
  r.v[2] = z;
  return r;
}

// Eu_xyz::rotMatrToEulerAng (Geom/euler.hpp:58-102): the angle triple of a rotation.
std::array<double, 3> matrix_to_euler(const Mat& m) {
  const double half_pi = 4.0 * std::atan(1.0) / 2.;
  double a0, a1, a2;
  if (m[2] < 1.0) {
    if (m[2] > -1.0) {
      a0 = std::atan2(-m[5], m[8]);
      a1 = std::asin(m[2]);
      a2 = std::atan2(-m[1], m[0]);
    } else {
      a0 = -std::atan2(m[3], m[4]);
      a1 = -half_pi;
      a2 = 0.0;
    }
  } else {
    a0 = std::atan2(m[3], m[4]);
    a1 = half_pi;
    a2 = 0.0;
  }
  return {-a2, -a0, -a1};
}

void check(pd_status s) {
  if (s != PD_OK) throw Error(std::string("libpdfft: ") + pd_last_error());
}

struct Frame {
  double box_lo[3], box_hi[3];
  double lig_size;
  RT tR, tL;  // original -> docking frame of receptor and ligand (MolStruct::trFromIni)
};

Frame prepare_frame(MolForceParams& p, const Options& opt) {
  Molecule& rec = p.mol[0];
  Molecule& lig = p.mol[1];
  // randomizeLigandOrientation (:248-266): spin about the ligand centre of mass
  const double spin = opt.num("startingSpin", 23.3734432);
  const Mat rs = euler_to_matrix(spin, spin * 3.7877, spin * -4.513432);
  double cm[3] = {0, 0, 0}, mt = 0;
  for (int j = 0; j < lig.n(); ++j) {
    for (int d = 0; d < 3; ++d) cm[d] += lig.mass[j] * lig.pos[3 * j + d];
    mt += lig.mass[j];
  }
  for (int d = 0; d < 3; ++d) cm[d] /= mt;
  for (int j = 0; j < lig.n(); ++j) {
    double* x = &lig.pos[3 * j];
    for (int d = 0; d < 3; ++d) x[d] -= cm[d];
    apply(rs, x);
    for (int d = 0; d < 3; ++d) x[d] += cm[d];
  }
  // receptor box: its approximate minimum-volume box (moveIntoMinBox,
  // :199-245); both molecules move by x -> A x - A c, A the box axes in rows
  // and c the box centre (Box::getTransformationToBoundCoordinates,
  // Geom/bounding.hpp:196-206)
  double axes[9], blo[3], bhi[3];
  check(pd_min_volume_box(rec.n(), rec.pos.data(), axes, blo, bhi));
  RT box;
  for (int i = 0; i < 9; ++i) box.m[i] = axes[i];
  // A c = the centre's coordinates along the axes, (lo + hi) / 2
  for (int k = 0; k < 3; ++k) box.v[k] = -0.5 * (blo[k] + bhi[k]);
  for (Molecule* m : {&rec, &lig})
    for (int a = 0; a < m->n(); ++a) apply(box, &m->pos[3 * a]);
  // ligand: centre of its diameter at the origin (:215-226), exact diameter
  int bi = 0, bj = 0;
  check(pd_diameter_pair(lig.n(), lig.pos.data(), &bi, &bj));
  double best = 0;
  for (int d = 0; d < 3; ++d) {
    const double t = lig.pos[3 * bi + d] - lig.pos[3 * bj + d];
    best += t * t;
  }
  double mid[3];
  for (int d = 0; d < 3; ++d) mid[d] = 0.5 * (lig.pos[3 * bi + d] + lig.pos[3 * bj + d]);
  for (int j = 0; j < lig.n(); ++j)
    for (int d = 0; d < 3; ++d) lig.pos[3 * j + d] -= mid[d];
  Frame f;
  // trFromIni (docking_molstruct.hpp:199-266): receptor moved by the box
  // transform; ligand spun about its centre of mass, moved by the box
  // transform, then its diameter centre moved to the origin
  f.tR = box;
  RT rspin;
  rspin.m = rs;
  const RT spin_cm = compose(compose(translation(cm[0], cm[1], cm[2]), rspin), translation(-cm[0], -cm[1], -cm[2]));
  f.tL = compose(translation(-mid[0], -mid[1], -mid[2]), compose(box, spin_cm));
  f.lig_size = std::sqrt(std::max(0.0, best));
  const double h = opt.num("gridStep", 1.5);
  const double pad = h * 2. + f.lig_size / 2. + opt.num("cutOffFft");  // :236
  // the receptor box in the docking frame: centred, half its extent per axis
  for (int k = 0; k < 3; ++k) {
    const double half = 0.5 * (bhi[k] - blo[k]);
    f.box_lo[k] = -half - pad;
    f.box_hi[k] = half + pad;
  }
  return f;
}

// FFTCorrelator::init grid size (docking_fft.hpp:541-553): the findBestSize
// edge as it is, odd edges included.
int grid_edge(double lo, double hi, double h) {
  const int start = (int)((hi - (lo - h)) / h) - 1;
  const int n = pd_find_best_size(start);
  if (n < 0) throw Error("no FFT size");
  return n;
}


// ---------------------------------------------------------------------------
// rot-scan task: Worker::init + Worker::run
// ---------------------------------------------------------------------------
void rot_scan(const Options& opt, MolForceParams& mfp) {
  // App::init (docking_app.hpp:56-67): ljComp (two terms) or ljAvg (one term)
  const std::string pot = opt.str("potentialName", "ljComp");
  if (pot != "ljComp" && pot != "ljAvg")
    throw Error("not_supported_error: Unknown 'potentialName' parameter value: " + pot);
  if (pot == "ljComp" && mfp.mix != 1) throw Error("ljComp requires mix == 1 (LJ_MIX_1)");
  const double h = opt.num("gridStep", 1.5);
  const double cut = opt.num("cutOffFft");
  const int maxNTrans = opt.integer("maxNTrans", 10);
  const int maxNTransInp = opt.integer("maxNTransInp", maxNTrans * 100);
  const double maxValCorr = opt.num("maxValCorr", 0.0);
  const bool rescore = opt.flag("rescore64", true);

  auto rots = load_rotational_grid(opt.str("anglesFile"));
  const int start = opt.integer("fft_rot_grid_start");
  int end = opt.integer("fft_rot_grid_end");
  if (!(start >= 0 && start < (int)rots.size())) throw Error("Rotation start index is out of bound");
  if (end < start) throw Error("rotation end index is out of bound");
  end = std::min<int>(end, (int)rots.size());

  Frame fr = prepare_frame(mfp, opt);
  pd_scan_desc d{};
  for (int a = 0; a < 3; ++a) d.n[a] = grid_edge(fr.box_lo[a], fr.box_hi[a], h);
  d.grid_step = h;
  d.n_terms = pot == "ljComp" ? 2 : 1;
  d.max_n_trans = maxNTrans;
  d.max_n_trans_inp = std::max(maxNTransInp, maxNTrans);
  d.max_val_corr = maxValCorr;
  d.batch = 64;
  d.device = opt.integer("device", 0);
  pd_scan* h_ = nullptr;
  check(pd_scan_create(&d, &h_));
  std::unique_ptr<pd_scan, pd_status (*)(pd_scan*)> hs(h_, pd_scan_destroy);

  const Molecule& rec = mfp.mol[0];
  const Molecule& lig = mfp.mol[1];
  std::vector<double> w;  // n_terms x ligand atoms
  if (pot == "ljComp") {
    // MolForceLJComp = MolForceLJRep + MolForceLJAttr (docking_molforce.hpp:455-602)
    std::vector<double> rep(2 * rec.n()), att(2 * rec.n()), wr(lig.n()), wa(lig.n());
    check(pd_ljcomp_params(rec.n(), rec.eps.data(), rec.sigma.data(), rec.alpha.data(), lig.n(),
                           lig.eps.data(), lig.sigma.data(), rep.data(), att.data(), wr.data(), wa.data()));
    pd_term terms[2] = {{PD_FIELD_LJ_REP, cut, -1.0, rep.data()}, {PD_FIELD_LJ_ATT, cut, -1.0, att.data()}};
    check(pd_scan_project_receptor(h_, rec.n(), rec.pos.data(), terms));
    w.insert(w.end(), wr.begin(), wr.end());
    w.insert(w.end(), wa.begin(), wa.end());
  } else {
    // MolForceLJ (docking_molforce.hpp:377-450)
    std::vector<double> par(5 * rec.n()), wl(lig.n());
    check(pd_ljavg_params(rec.n(), rec.eps.data(), rec.sigma.data(), rec.alpha.data(), lig.n(), lig.eps.data(),
                          lig.sigma.data(), mfp.mix, cut, par.data(), wl.data()));
    pd_term term = {PD_FIELD_LJ_FULL, cut, -1.0, par.data()};
    check(pd_scan_project_receptor(h_, rec.n(), rec.pos.data(), &term));
    w = wl;
  }
  check(pd_scan_set_ligand(h_, lig.n(), lig.pos.data(), w.data()));
  // Worker::init post-filters (docking_app.hpp:363-398): TranSymm with the
  // original-frame transformer and half the ligand's diameter, TranClust
  const bool clust = opt.flag("doClusterTranslations", false);
  const int n_multimer = opt.integer("nMultimer", 0);
  if (clust || n_multimer > 1) {
    pd_postfilter pf{};
    pf.n_multimer = n_multimer;
    pf.max_rmsd_symm = opt.num("maxRmsdSymm", 8.0);
    pf.ligand_radius = fr.lig_size / 2;
    const RT a = inverse(fr.tR);
    for (int i = 0; i < 9; ++i) {
      pf.frame_rec_inv[i] = a.m[i];
      pf.frame_lig[i] = fr.tL.m[i];
    }
    for (int i = 0; i < 3; ++i) {
      pf.frame_rec_inv[9 + i] = a.v[i];
      pf.frame_lig[9 + i] = fr.tL.v[i];
    }
    pf.cluster = clust ? 1 : 0;
    pf.cluster_radius = opt.num("clusterRadiusTrans", 5.0);
    check(pd_scan_set_postfilter(h_, &pf));
  }

  const int nr = end - start;
  std::vector<double> R(9 * (size_t)nr);
  for (int i = 0; i < nr; ++i) std::copy(rots[start + i].begin(), rots[start + i].end(), &R[9 * (size_t)i]);
  std::vector<pd_peak> peaks((size_t)nr * maxNTrans);
  std::vector<int> nout(nr);
  if (nr > 0) check(pd_scan_run(h_, nr, R.data(), start, peaks.data(), nout.data()));
  if (rescore && nr > 0) {
    std::vector<double> pr;
    std::vector<int> ps;
    for (int i = 0; i < nr; ++i)
      for (int k = 0; k < nout[i]; ++k) {
        pr.insert(pr.end(), &R[9 * (size_t)i], &R[9 * (size_t)i] + 9);
        const pd_peak& pk = peaks[(size_t)i * maxNTrans + k];
        ps.insert(ps.end(), pk.s, pk.s + 3);
      }
    std::vector<double> e(ps.size() / 3);
    if (!e.empty()) check(pd_scan_rescore64(h_, (int)e.size(), pr.data(), ps.data(), e.data()));
    size_t at = 0;
    for (int i = 0; i < nr; ++i)
      for (int k = 0; k < nout[i]; ++k) peaks[(size_t)i * maxNTrans + k].value64 = e[at++];
  }
  // RotFftScanIO_Bin::write_record per rotation (docking_io_bin.hpp:55-71)
  const std::string out = opt.str("fft_rot_scan_res");
  FILE* f = std::fopen(out.c_str(), "wb");
  if (!f) throw Error("Could not open output file: " + out);
  for (int i = 0; i < nr; ++i) {
    const int32_t sig = 2014011401;
    std::fwrite(&sig, 4, 1, f);
    std::fwrite(&R[9 * (size_t)i], 8, 9, f);
    const uint64_t n = (uint64_t)nout[i];
    std::fwrite(&n, 8, 1, f);
    for (int k = 0; k < nout[i]; ++k) {
      const pd_peak& pk = peaks[(size_t)i * maxNTrans + k];
      const double rec4[4] = {h * pk.s[0], h * pk.s[1], h * pk.s[2], rescore ? pk.value64 : (double)pk.value};
      std::fwrite(rec4, 8, 4, f);
    }
  }
  if (std::fclose(f) != 0) throw Error("Output failed");
  std::printf("rot-scan: grid %dx%dx%d h=%g rotations [%d,%d) -> %s\n", d.n[0], d.n[1], d.n[2], h, start, end,
              out.c_str());
}

// ---------------------------------------------------------------------------
// gather task: Foreman::run + writeResults (docking_app.hpp:174-281)
// ---------------------------------------------------------------------------
struct RotTranValue {
  RT t;
  double value;
};

// RotFftScanIO_Collector (docking_io_bin.hpp:81-116): records of the files named
// in a list, one file after another; the collection ends at the first file
// that yields no record.
class ScanCollector {
 public:
  explicit ScanCollector(const std::string& list) : list_(list) {
    if (!list_.good()) throw Error("Could not open list file: " + list);
  }
  bool next(Mat& rot, std::vector<std::array<double, 4>>& tv) {
    if (cur_ && read(rot, tv)) return true;
    std::string name;
    if (list_.good() && std::getline(list_, name)) {
      cur_.reset(new std::ifstream(name, std::ios::binary));
      if (!cur_->good()) throw Error("Could not open input file: " + name);
      return read(rot, tv);
    }
    return false;
  }

 private:
  // RotFftScanIO_Bin::read_record (docking_io_bin.hpp:37-53)
  bool read(Mat& rot, std::vector<std::array<double, 4>>& tv) {
    int32_t sig = 0;
    if (!cur_->read(reinterpret_cast<char*>(&sig), 4)) return false;
    if (sig != 2014011401) throw Error("Bad record signature in a rotation scan file");
    uint64_t n = 0;
    cur_->read(reinterpret_cast<char*>(rot.data()), 72);
    cur_->read(reinterpret_cast<char*>(&n), 8);
    tv.resize(n);
    if (n) cur_->read(reinterpret_cast<char*>(tv.data()), 32 * (std::streamsize)n);
    if (!*cur_) throw Error("Truncated record in a rotation scan file");
    return true;
  }
  std::ifstream list_;
  std::unique_ptr<std::ifstream> cur_;
};

void gather(const Options& opt, MolForceParams& mfp) {
  const int max_matches = opt.integer("maxNRigidMatches", 20000);
  if (max_matches < 1) throw Error("maxNRigidMatches must be positive");
  const double max_val = 1e16;  // rtvalQueue.init(maxNRigidMatches, 1e16)
  Frame fr = prepare_frame(mfp, opt);
  // BoundPriorityQueue (Common/queue.hpp:221-251) on the value alone
  // (cmp_rt_val, docking_app.hpp:131-145): a max-heap of at most max_matches
  // entries; a value enters only below the limit, and a full queue replaces its
  // top only with a strictly smaller value; sort_heap leaves them ascending.
  auto cmp = [](const RotTranValue& a, const RotTranValue& b) { return a.value < b.value; };
  std::vector<RotTranValue> q;
  q.reserve(max_matches);
  ScanCollector coll(opt.str("fft_rot_scan_list"));
  Mat rot;
  std::vector<std::array<double, 4>> tv;
  long long n_rec = 0, n_in = 0;
  while (coll.next(rot, tv)) {
    ++n_rec;
    RT r;
    r.m = rot;
    for (const auto& x : tv) {
      ++n_in;
      RotTranValue e;
      e.t = compose(translation(x[0], x[1], x[2]), r);  // tranVal.tran * doneRot
      e.value = x[3];
      if (!(e.value < max_val)) continue;
      if ((int)q.size() >= max_matches) {
        if (cmp(e, q.front())) {
          std::pop_heap(q.begin(), q.end(), cmp);
          q.pop_back();
        } else {
          continue;
        }
      }
      q.push_back(e);
      std::push_heap(q.begin(), q.end(), cmp);
    }
  }
  std::sort_heap(q.begin(), q.end(), cmp);
  // getResults: ToOriginalFrameTransformer, tR^-1 * t * tL (docking_molstruct.hpp:31-51)
  const RT tRi = inverse(fr.tR);
  // IORigid::writeCoords, format 'b' (IO/rigid.hpp:29-150): int32 signature,
  // int32 n, n x {float value, float angles[3], float displacement[3]}
  const std::string out = opt.str("fft_res");
  FILE* f = std::fopen(out.c_str(), "wb");
  if (!f) throw Error("Cannot open file for writing: " + out);
  const int32_t sig = 2005052501, n = (int32_t)q.size();
  std::fwrite(&sig, 4, 1, f);
  std::fwrite(&n, 4, 1, f);
  for (const RotTranValue& e : q) {
    const RT t = compose(compose(tRi, e.t), fr.tL);
    const auto ang = matrix_to_euler(t.m);
    const float rec[7] = {(float)e.value, (float)ang[0], (float)ang[1], (float)ang[2],
                          (float)t.v[0], (float)t.v[1], (float)t.v[2]};
    std::fwrite(rec, 4, 7, f);
  }
  if (std::fclose(f) != 0) throw Error("Error when writing to file: " + out);
  std::printf("gather: %lld records, %lld translations -> %d matches -> %s\n", n_rec, n_in, n, out.c_str());
}

void usage() {
  std::puts(
      "Options for FFT scan process:\n"
      "  --help                     produce help message\n"
      "  --options arg              file with options common for this run\n"
      "  --molforce-params arg      molforce parameters file\n"
      "  --task arg                 type of task to perform (rot-scan or gather)\n"
      "  --fft-rot-grid-start arg   start index in rot-grid\n"
      "  --fft-rot-grid-end arg     end index in rot-grid\n"
      "  --fft-rot-scan-res arg     output file of fft scan for one task\n"
      "  --fft-rot-scan-list arg    file with a list of fft rot scan tasks\n"
      "  --fft-res arg              output file for entire fft scan");
}

}  // namespace

int main(int argc, char** argv) {
  try {
    std::map<std::string, std::string> args;
    for (int i = 1; i < argc; ++i) {
      std::string a = argv[i];
      if (a == "--help") { usage(); return 0; }
      if (a.rfind("--", 0) != 0 || i + 1 >= argc) {
        std::fprintf(stderr, "error: bad argument %s\n", a.c_str());
        return -1;
      }
      std::string v = argv[++i];
      auto eq = a.find('=');
      args[a.substr(2)] = v;
      (void)eq;
    }
    Options opt;
    if (args.count("options")) JsonReader(read_file(args["options"])).parse(opt);
    if (!args.count("task")) throw Error("Option 'task' is required");
    const std::string task = args["task"];
    auto set = [&](const char* name) {
      if (!args.count(name)) throw Error(std::string("required option --") + name + " is missing");
      std::string k = name;
      std::replace(k.begin(), k.end(), '-', '_');
      opt.kv[k] = args[name];
    };
    if (task == "rot-scan") {
      set("fft-rot-grid-start");
      set("fft-rot-grid-end");
      set("fft-rot-scan-res");
    } else if (task == "gather") {
      set("fft-rot-scan-list");
      set("fft-res");
    } else {
      throw Error("Option 'task' has invalid value: " + task);
    }
    if (!args.count("molforce-params")) throw Error("required option --molforce-params is missing");
    MolForceParams mfp = load_molforce(args["molforce-params"]);
    if (task == "gather")
      gather(opt, mfp);
    else
      rot_scan(opt, mfp);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
