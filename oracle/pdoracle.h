/*
 * pdoracle.h -- FP64 CPU oracle for the FFT rigid-body correlation scan.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain CPU restatement of the
 * PRODDL2 docking worker's inner loop (include/PRODDL/docking_app.hpp:417-466)
 * and of every helper that loop calls.  It exists to check the HIP product
 * path; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product library (libpdfft.so) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Pinning of each
 * function (golden vectors, closed forms, brute force) lives in
 * tests/test_oracle_*.py.  Functions whose only pin is the mathematics are
 * marked there; none is "parity unpinned" except the exact libstdc++ order of
 * exactly-equal heap entries, which this oracle reproduces by calling the same
 * std::push_heap / pop_heap / sort_heap the reference calls.
 */
#ifndef PDORACLE_H
#define PDORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Radial field kinds; the numbering is shared with include/pdfft.h. */
enum {
  PDO_FIELD_LJ_REP = 0,   /* potentials_inc.hpp:311-314   params {e4s12, as6}           */
  PDO_FIELD_LJ_ATT = 1,   /* potentials_inc.hpp:349-352   params {e4s6, as6}            */
  PDO_FIELD_LJ_FULL = 2,  /* potentials_inc.hpp:214-219   params {e4s12,e4s6,as6,fsh,gsh}*/
  PDO_FIELD_TWOSTEP = 3,  /* new term: v_in for r2<=r_in2, v_out outside; {v_in,v_out,r_in2} */
  PDO_FIELD_COULOMB = 4,  /* new term: kq / r2 (distance-dependent dielectric); {kq}     */
  PDO_FIELD_DESOLV = 5    /* new term: d * (1 - r2/cut2)^2; {d}                          */
};

int pdo_field_nparams(int kind);

/* ---- integer helpers ---------------------------------------------------- */
/* Math/pfactors.hpp:29-57 */
int pdo_find_first_prime_factoring(int startN, int endN, const int* primes,
                                   int nPrimes, int* exps, int* resultN);
/* Math/fft.hpp:82-151 (primes {2,3,5,7,11,13}, e11+e13<=1) */
int pdo_find_best_size(int startN, int* resultN);

/* ---- geometry ----------------------------------------------------------- */
/* Grid/gridcoord.hpp:127-142,166-175: centred geometry for one axis. */
void pdo_geometry_centered(double h, int n, double* zero_h, double* zero_start_center);
/* Grid/gridcoord.hpp:114-121 + 185-192: geometry with explicit zero. */
int pdo_to_logical(double h, double zero_h, double x);
/* FFTCorrelator::init (docking_fft.hpp:541-553): grid edge from box diagonal. */
int pdo_grid_size_from_box(double lo, double hi, double h);

/* ---- rotations ---------------------------------------------------------- */
void pdo_euler_to_matrix(const double* ang, double* m);                 /* Geom/euler.hpp:38-52  */
void pdo_matrix_to_euler(const double* m, double* ang, int* unique);    /* Geom/euler.hpp:58-102 */
int pdo_load_rotational_grid(const char* path, double* mats, int cap);  /* Geom/rotational_grid.hpp:25-64 */
void pdo_rotate_points(const double* m, int n, const double* in, double* out); /* transformation.hpp:206-223, bztinyvec.hpp:52-65 */
void pdo_matmul3(const double* a, const double* b, double* c);          /* bztinyvec.hpp:70-85 */

/* ---- projection --------------------------------------------------------- */
/* ProjectorRadialField::projectFields (Grid/project.hpp:143-176); grid is
 * N0*N1*N2 C-order, zeroed first (zeroGrid=true). */
void pdo_project_receptor(const int* n, double h, int n_atoms, const double* pos,
                          int kind, const double* params, double cutoff,
                          double min_dist, double* grid);
/* MolForceRadial::projectLigand (docking_molforce.hpp:351-368) with the
 * bound check of Grid::operator() (grid.hpp:442-451).  Returns 0 or the
 * 1-based index of the first atom outside the grid. */
int pdo_project_ligand(const int* n, double h, int n_atoms, const double* pos,
                       const double* w, double* grid);
/* Cell index (i0,i1,i2) of each point (Grid/gridcoord.hpp:185-192). */
void pdo_cells(const int* n, double h, int n_atoms, const double* pos, int* cells);

/* ljComp / ljAvg per-atom parameters (docking_molforce.hpp:383-447, 461-581). */
void pdo_ljcomp_params(int n_rec, const double* eps_r, const double* sigma_r,
                       const double* alpha_r, int n_lig, const double* eps_l,
                       const double* sigma_l, double* rep_params /*n_rec*2*/,
                       double* att_params /*n_rec*2*/, double* w_rep, double* w_att);
void pdo_ljavg_params(int n_rec, const double* eps_r, const double* sigma_r,
                      const double* alpha_r, int n_lig, const double* eps_l,
                      const double* sigma_l, int mix, double cutoff,
                      double* params /*n_rec*5*/, double* w);

/* ---- FFT (own mixed-radix FP64 transform; FFTW sign conventions) -------- */
void pdo_fft1(int n, int sign, double* data /* n complex, in place */);
void pdo_rfft3(const int* n, const double* real, double* cplx /* N0*N1*(N2/2+1) complex */);
void pdo_irfft3(const int* n, const double* cplx, double* real /* unnormalised */);

/* ---- correlation -------------------------------------------------------- */
/* FFTCorrelators::correlate + MolForce::collectTotal
 * (docking_fft.hpp:622-633, 812-819; docking_molforce.hpp:248-269):
 * out = sum_t c2r(r2c(L_t) * conj(Rhat_t)) / N, summed in term order. */
void pdo_correlate(const int* n, int n_terms, const double* rec_spectra,
                   const double* lig_grids, double* out);

/* CorrelationProcessor::selectFromFFT + getTranValuesFilled
 * (docking_fft.hpp:316-345, 368-389, 429-464; Common/queue.hpp:221-251;
 * Math/fft.hpp:199-206).  Returns the number of translations written. */
int pdo_select(const int* n, double h, const double* grid, int max_inp,
               double max_val, int n_max, int* offsets, int* s, double* tran,
               double* values);

/* Energy by direct summation: E(s) = sum_t sum_j w_tj R_t((l_j + s) mod N). */
double pdo_direct_energy(const int* n, int n_terms, const double* rec_grids,
                         int n_lig, const int* cells, const double* w, const int* s);

/* One full rotation scan (Worker::scanRotation, docking_app.hpp:450-466) for
 * n_rot rotations, each on its own thread slice; K = maxNTrans outputs per
 * rotation.  Returns 0, or the 1-based rotation index whose ligand left the
 * grid. corr_out (optional, n_rot*N) receives the summed correlation grids. */
int pdo_scan(const int* n, double h, int n_terms, const double* rec_spectra,
             int n_lig, const double* lig_pos, const double* w, int n_rot,
             const double* rots, int max_inp, double max_val, int K,
             int n_threads, int* n_out, int* offsets, int* s, double* values,
             double* corr_out);

/* ---- post-filters of a rotation's translations --------------------------- */
/* A rigid transformation x -> m x + v is 12 doubles: m (3x3 row-major), v (3). */

/* TranSymm::process (docking_fft.hpp:146-230) with SymmetryCheckerCn
 * (Geom/symmetry.hpp:17-62) and ToOriginalFrameTransformer
 * (docking_molstruct.hpp:31-51): keeps, in order, the translations whose pose
 * frame_a * (Translation(tran) * rot) * frame_b applied n_symm times moves the
 * points radius*e1, radius*e2, radius*e3 and 0 by an RMSD below max_rmsd.
 * Compacts tran (n x 3), values and tag (n, carried along) in place; returns
 * the number kept.  rmsd (optional, n) receives every input's RMSD. */
int pdo_tran_symm(int n, double* tran, double* values, int* tag, const double* rot,
                  const double* frame_a, const double* frame_b, int n_symm, double mol_radius,
                  double max_rmsd, double* rmsd);

/* TranClust::process (docking_fft.hpp:30-140) with ClusterMatrix::cluster and
 * getClusters (Geom/cluster.hpp:120-420, M_hash = 3): leaders of a greedy
 * clustering with radius sqrt(3) * cluster_radius, visited in ascending order
 * of -value; the leaders, in that order, replace the first entries of tran /
 * values / tag.  Returns min(leaders, cap, max_out). */
int pdo_tran_clust(int n, double* tran, double* values, int* tag, double cluster_radius, int max_out,
                   int cap);

/* One rotation's selection and post-processing chain of Worker::findTranslationalMinima
 * + scanRotation (docking_app.hpp:417-466): the max_inp best translations below
 * max_val (CorrelationProcessor::selectFromFFT), then TranSymm when n_symm > 1,
 * then TranClust when cluster_radius > 0, then the first K
 * (getTranValuesFilled).  Writes offsets / s / values of the output; returns
 * their number.  rot (9), frames (12 each) as for pdo_tran_symm. */
int pdo_select_post(const int* n, double h, const double* grid, int max_inp, double max_val, int K,
                    const double* rot, const double* frame_a, const double* frame_b, int n_symm,
                    double mol_radius, double max_rmsd, double cluster_radius, int* offsets, int* s,
                    double* values);

/* RotFftScanIO_Bin (docking_io_bin.hpp:21-76). */
int pdo_write_scan_record(const char* path, int append, const double* rot,
                          int n, const double* tran, const double* values);
/* Reads up to cap records; returns record count, fills rot (9 each), counts and
 * up to maxn translations per record. */
int pdo_read_scan_records(const char* path, int cap, int maxn, double* rots,
                          int* counts, double* tran, double* values);

#ifdef __cplusplus
}
#endif
#endif
