/*
 * pdfft.h -- C ABI of libpdfft.so, the MI355X (gfx950) FFT rigid-body
 * correlation scan.
 *
 * The reference (andreyto/proddl2) has no C ABI: its scan is the header-only
 * template Docking<T>::Worker compiled into the proddl-dock-fft executable.
 * Each entry point below replaces one in-process seam of that executable; the
 * seam it replaces is cited next to it.  Plain pointers and sizes only; no
 * exceptions cross this boundary.  Every call returns a pd_status and, on
 * failure, pd_last_error() holds a message (thread-local).
 *
 * Ownership: the caller owns every buffer passed in or out.  A handle owns its
 * device memory and one HIP stream (default priority).  A handle is not
 * thread-safe; separate handles are independent.
 *
 * Conventions (all from the reference):
 *  - grids are N0 x N1 x N2, C order, last index fastest;
 *  - the geometry is centred: cell of x is int((x - zero_h)/h) - 1 with
 *    zero_h = -(h + h*(N-1))/2 - h  (Grid/gridcoord.hpp:127-142, 185-192);
 *  - energies: lower is better; a translation is kept only if value < maxValCorr;
 *  - displacement s of a correlation offset x: s = -x, wrapped into
 *    [-floor(N/2), N - floor(N/2) - 1] (Math/fft.hpp:199-206); translation = h*s.
 */
#ifndef PDFFT_H
#define PDFFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pd_status {
  PD_OK = 0,
  PD_ERR_INVALID_ARGUMENT = 1, /* dbg_exception from ATALWAYS (Common/debug.hpp:55,68-71) */
  PD_ERR_DOMAIN = 2,           /* grid_domain_error (Grid/grid.hpp:526): atom outside grid */
  PD_ERR_IO = 3,               /* file errors (docking_io_bin.hpp:27)                    */
  PD_ERR_DEVICE = 4,           /* HIP runtime failure (no reference counterpart)          */
  PD_ERR_UNSUPPORTED = 5,      /* not_supported_error (docking_app.hpp:65)                */
  PD_ERR_STATE = 6             /* call order violated (e.g. run before receptor set)      */
} pd_status;

/* Radial field kinds for receptor terms (MolForceRadialR plugins,
 * docking_molforce.hpp:287-374).  Per-atom parameter counts in brackets. */
enum {
  PD_FIELD_LJ_REP = 0,  /* SoftCoreLJRep::f2 (potentials_inc.hpp:311-314)  [e4s12, as6]   */
  PD_FIELD_LJ_ATT = 1,  /* SoftCoreLJAttr::f2 (potentials_inc.hpp:349-352) [e4s6, as6]    */
  PD_FIELD_LJ_FULL = 2, /* SoftCoreLJ::f2 (potentials_inc.hpp:214-219) [e4s12,e4s6,as6,fshift,gshift] */
  PD_FIELD_TWOSTEP = 3, /* new shape term: v_in if r2<=r_in2 else v_out   [v_in, v_out, r_in2] */
  PD_FIELD_COULOMB = 4, /* new: kq/r2, distance-dependent dielectric     [kq]                */
  PD_FIELD_DESOLV = 5   /* new: d*(1-r2/cut2)^2 short-range desolvation   [d]                 */
};

typedef struct pd_scan_desc {
  int n[3];               /* FFT grid edges (FFTCorrelator::init, docking_fft.hpp:545-553) */
  double grid_step;       /* option gridStep (docking_app.hpp:72)                           */
  int n_terms;            /* MolForce::nGrids() (docking_molforce.hpp:83)                   */
  int max_n_trans;        /* option maxNTrans: peaks kept per rotation (docking_app.hpp:84), 1..32 */
  int max_n_trans_inp;    /* option maxNTransInp (docking_app.hpp:374); >= max_n_trans       */
  double max_val_corr;    /* option maxValCorr (docking_app.hpp:363)                        */
  int batch;              /* rotations per kernel launch (0: 64)                             */
  int device;             /* HIP device ordinal                                              */
  int flags;              /* PD_SCAN_* bits, 0 by default                                    */
} pd_scan_desc;

/* pd_scan_desc.flags: run the plane passes as separate row and column kernels
 * even when the whole y-z plane fits on one CU (grids whose plane does not fit,
 * e.g. 256^3, always use them). */
#define PD_SCAN_SPLIT_PLANES 1
/* pd_scan_desc.flags: run all batches on the handle's stream (no overlap of
 * consecutive batches), e.g. to profile each kernel alone. */
#define PD_SCAN_ONE_LANE 2
/* pd_scan_desc.flags: with the row and column kernels, run the inverse y pass
 * and the inverse rows as two kernels also where one kernel holds the whole
 * x plane on chip (N1 = N2 = 256), e.g. to compare the two. */
#define PD_SCAN_SPLIT_INVERSE 4
/* pd_scan_desc.flags: likewise for the ligand plane pass (deposit, rows along
 * z and the y pass as separate kernels where one kernel would do them). */
#define PD_SCAN_SPLIT_FORWARD 8

typedef struct pd_term {
  int kind;               /* PD_FIELD_*                                                      */
  double cutoff;          /* field radius (option cutOffFft for ljComp)                      */
  double min_dist;        /* r is clamped to >= min_dist; < 0 means grid_step/4
                             (docking_molforce.hpp:307-310)                                  */
  const double* params;   /* n_atoms x nparams(kind), row-major                              */
} pd_term;

/* One selected translation of one rotation (TranValue, types.hpp:70-76, made explicit). */
typedef struct pd_peak {
  int32_t rot;            /* global rotation index = rot_index0 + position in the call     */
  int32_t rank;           /* 0-based rank within the rotation (ascending value)            */
  int32_t s[3];           /* displacement in cells                                          */
  int32_t offset;         /* row-major offset of the correlation cell                       */
  float value;            /* FP32 correlation energy                                        */
  float reserved;
  double value64;         /* FP64 rescored energy, NaN until pd_scan_rescore64 fills it     */
} pd_peak;

/* FFTCorrelators ctor + CorrelationProcessor::init (docking_fft.hpp:771-795, 283-312). */
pd_status pd_scan_create(const pd_scan_desc* desc, struct pd_scan** out);

/* MolForce::projectMol(iRec) + FFTCorrelators::preprocessReceptor
 * (docking_app.hpp:406-415; Grid/project.hpp:143-176; docking_fft.hpp:614-618):
 * FP64 projection of n_terms radial fields on the GPU, then FP32 r2c spectra. */
pd_status pd_scan_project_receptor(struct pd_scan* h, int n_atoms, const double* pos,
                                   const pd_term* terms);

/* Same as above from caller-projected FP64 grids (n_terms x N0*N1*N2). */
pd_status pd_scan_set_receptor_grids(struct pd_scan* h, const double* grids);

/* Copies the FP64 receptor grids back (n_terms x N0*N1*N2). */
pd_status pd_scan_get_receptor_grids(struct pd_scan* h, double* grids);

/* MolForceRadial ligand data (docking_molforce.hpp:351-372): reference-frame
 * coordinates (n x 3, FP64) and per-term weights (n_terms x n, FP64). */
pd_status pd_scan_set_ligand(struct pd_scan* h, int n_atoms, const double* pos,
                             const double* weights);

/* Worker::run over a rotation slice (docking_app.hpp:487-513 via 450-466, 417-447):
 * rots is n_rot x 9 row-major FP64 (host); peaks is n_rot x max_n_trans (host);
 * n_out[i] = peaks written for rotation i. */
pd_status pd_scan_run(struct pd_scan* h, int n_rot, const double* rots, int rot_index0,
                      pd_peak* peaks, int* n_out);

/* Same with device-resident rotations and outputs; asynchronous on the handle's
 * stream unless sync != 0.  Domain errors are reported by pd_scan_sync. */
pd_status pd_scan_run_device(struct pd_scan* h, int n_rot, const double* d_rots,
                             int rot_index0, pd_peak* d_peaks, int* d_n_out, int sync);
pd_status pd_scan_sync(struct pd_scan* h);

/* FP64 rescoring (new; no reference counterpart): E = sum_t sum_j w_tj R_t(l_j + s)
 * from the FP64 receptor grids, for n poses (rots n x 9, s n x 3, host). */
pd_status pd_scan_rescore64(struct pd_scan* h, int n, const double* rots, const int* s,
                            double* energies);

/* pd_scan_rescore64 without any allocation (no counterpart in the reference):
 * for callers that cannot allocate device memory, e.g. tight serving loops or
 * harnesses that forbid hipMalloc.  The poses are staged through a fixed
 * device-side area of the library (one per GPU, 736 KiB, held until the process
 * ends), so n is limited to 8192 poses per call (PD_ERR_INVALID_ARGUMENT above
 * that).  Same results as pd_scan_rescore64.  The area keeps the last call's
 * contents; calls on one GPU share it and are serialised by the library. */
pd_status pd_scan_rescore64_noalloc(struct pd_scan* h, int n, const double* rots, const int* s,
                                    double* energies);

/* The summed correlation grid of one rotation (FP32, N0*N1*N2), for parity tests
 * (the gridTot after collectTotal, docking_app.hpp:427). */
pd_status pd_scan_correlation_grid(struct pd_scan* h, const double* rot, float* out);

/* Per-kernel device time recorded with HIP events on the handle's stream.
 * enable != 0 starts recording; stats returns, for kernel id k
 * (0 bin, 1 ligand plane forward, 2 column, 3 inverse plane + top-K, 4 merge),
 * total milliseconds and launch count. */
pd_status pd_scan_set_timing(struct pd_scan* h, int enable);
pd_status pd_scan_kernel_stats(struct pd_scan* h, int k, double* total_ms, long long* launches);
void* pd_scan_stream(struct pd_scan* h);

/* The cells every rotated ligand atom can reach, [lo, lo + p) per axis (from
 * the largest atom radius; no reference counterpart), and the x planes per
 * (rotation, term) that the scan stores (p[0] rounded up to whole lane groups
 * of the column pass). */
pd_status pd_scan_ligand_box(struct pd_scan* h, int* lo, int* p, int* planes);

/* Frees the handle.  PD_ERR_STATE (and the handle stays valid) while a graph
 * built from it exists. */
pd_status pd_scan_destroy(struct pd_scan* h);
const char* pd_last_error(void);

/* Post-filters of each rotation's translations (Worker::init,
 * docking_app.hpp:363-398; applied in findTranslationalMinima, :417-447):
 * the scan keeps the max_n_trans_inp best translations (CorrelationProcessor),
 * then TranSymm drops those whose pose is not C_n-symmetric
 * (docking_fft.hpp:146-230), then TranClust keeps cluster leaders
 * (docking_fft.hpp:30-140), and the first max_n_trans of what is left are
 * reported.  Rigid transformations are 12 doubles: rotation (3x3 row-major),
 * then translation.  With TranClust the reported leaders come in the
 * reference's order: from the highest energy down (ClusterMatrix,
 * Geom/cluster.hpp:291-309). */
typedef struct pd_postfilter {
  int n_multimer;           /* option nMultimer: TranSymm when > 1                               */
  double max_rmsd_symm;     /* option maxRmsdSymm (reference default 8.0)                        */
  double ligand_radius;     /* MolStruct::getSizeLigand() / 2: radius of the symmetry test points */
  double frame_rec_inv[12]; /* tR^-1 of ToOriginalFrameTransformer (docking_molstruct.hpp:31-51)  */
  double frame_lig[12];     /* tL of ToOriginalFrameTransformer                                  */
  int cluster;              /* option doClusterTranslations: TranClust when != 0                  */
  double cluster_radius;    /* option clusterRadiusTrans (reference default 5.0)                  */
} pd_postfilter;

/* Sets (pf non-null with n_multimer > 1 or cluster != 0) or clears the
 * post-filters of pd_scan_run / pd_scan_run_device / graphs.  They need
 * max_n_trans_inp <= 4096; each batch then also keeps its real correlation
 * grids (batch x N0 N1 N2 floats per batch lane). */
pd_status pd_scan_set_postfilter(struct pd_scan* h, const pd_postfilter* pf);

/* The max_n_trans_inp best translations of each rotation before any
 * post-filter, ascending (CorrelationProcessor::selectFromFFT +
 * getTranValues, docking_fft.hpp:316-345, 409-414): peaks is
 * n_rot x max_n_trans_inp (host), n_out[i] the number found (fewer when fewer
 * cells lie below max_val_corr).  max_n_trans_inp <= 4096. */
pd_status pd_scan_candidates(struct pd_scan* h, int n_rot, const double* rots, int rot_index0,
                             pd_peak* peaks, int* n_out);

/* Rotations per kernel launch (pd_scan_desc.batch), changed after creation;
 * reallocates the per-batch buffers.  Results do not depend on it.
 * PD_ERR_STATE while a graph built from the handle exists (as for
 * pd_scan_set_ligand, which also reallocates them). */
pd_status pd_scan_set_batch(struct pd_scan* h, int batch);

/* Several scans in one HIP graph.  No reference counterpart: the reference
 * runs one rotation range per process (python/lib/PRODDL/dock.py:86-113).
 * Pair i scans n_rot[i] device rotations d_rots[i] into d_peaks[i] / d_n_out[i]
 * (rot_index0 = 0) on lane lane[i] < n_lanes; the pairs of a lane run in the given
 * order, lanes are independent branches.  The graph is built from kernel nodes
 * with explicit dependencies (no stream capture); within a pair, batches go round
 * the handle's batch buffers as in pd_scan_run_device.  pd_graph_launch replays
 * everything with one launch (the buffers named at creation are reused);
 * pd_graph_sync waits and reports domain errors of any pair.  A scan belongs to
 * at most one graph; while it does, pd_scan_set_batch, pd_scan_set_ligand and
 * pd_scan_destroy return PD_ERR_STATE, and work of its own (pd_scan_run...) must
 * not overlap a replay. */
pd_status pd_graph_create(int n, struct pd_scan* const* scans, const int* n_rot,
                          const double* const* d_rots, pd_peak* const* d_peaks,
                          int* const* d_n_out, const int* lane, int n_lanes,
                          struct pd_graph** out);
pd_status pd_graph_launch(struct pd_graph* g);
pd_status pd_graph_sync(struct pd_graph* g);
void* pd_graph_stream(struct pd_graph* g);
void pd_graph_destroy(struct pd_graph* g);

/* Launch choices of a scan for an autotuner (no counterpart in the reference,
 * whose worker has one fixed loop).  0 in a field keeps the library's rule.
 *   col_groups    rotation groups per column tile in the column pass: a tile's
 *                 groups share its receptor spectra through one XCD's L2
 *                 (1..batch; the library makes groups even-sized)
 *   col_mz_extra  column-pass variant: this many variants above the smallest
 *                 whose loads cover the ligand box (0..15; larger variants load
 *                 more zero planes)
 *   lig_pass      ligand pass: 1 the whole-plane kernel, 2 the row and column
 *                 kernels (whole-plane grids only; PD_ERR_UNSUPPORTED where the
 *                 choice does not exist for the grid)
 *   lanes         batch lanes, the streams consecutive batches go round (1..3)
 * Results do not depend on col_groups or lanes (same bits).  col_mz_extra and
 * lig_pass can select other transform arithmetic (a first DFT stage that skips
 * the zero planes, another order of the plane transforms): same FP32 bar, not
 * always the same bits.  Changing lig_pass after pd_scan_set_ligand re-lays the
 * ligand.  PD_ERR_STATE while a graph holds the scan. */
typedef struct pd_tuning {
  int col_groups;
  int col_mz_extra;
  int lig_pass;
  int lanes;
} pd_tuning;
pd_status pd_scan_set_tuning(struct pd_scan* h, const pd_tuning* t); /* t null: every rule */
/* The choices in effect for a launch of `batch` rotations (rules resolved;
 * after pd_scan_set_ligand, since the column pass follows the ligand box). */
pd_status pd_scan_get_tuning(struct pd_scan* h, int batch, pd_tuning* t);

/* FFTW_Size::findBestSize (Math/fft.hpp:141-151); returns -1 if none. */
int pd_find_best_size(int start);

/* Per-atom ljComp parameters: MolForceLJRep / MolForceLJAttr ctors
 * (docking_molforce.hpp:461-581).  rep/att: n_rec x 2; w_rep, w_att: n_lig. */
pd_status pd_ljcomp_params(int n_rec, const double* eps_r, const double* sigma_r,
                           const double* alpha_r, int n_lig, const double* eps_l,
                           const double* sigma_l, double* rep, double* att, double* w_rep,
                           double* w_att);

/* Per-atom ljAvg parameters: MolForceLJ ctor (docking_molforce.hpp:377-450) with
 * SoftCoreLJ and its shifted-force cutoff (potentials_inc.hpp:185-219): every
 * receptor atom's LJ is mixed with the ligand's average eps and sigma (mix 0:
 * arithmetic sigma, 1: geometric), params n_rec x 5 {e4s12, e4s6, as6, fshift,
 * gshift} for PD_FIELD_LJ_FULL; ligand weights w (n_lig) are all 1. */
pd_status pd_ljavg_params(int n_rec, const double* eps_r, const double* sigma_r, const double* alpha_r,
                          int n_lig, const double* eps_l, const double* sigma_l, int mix, double cutoff,
                          double* params, double* w);

/* Approximate minimum-volume bounding box of n points (pos: n x 3), the box
 * MolStruct::moveIntoMinBox orients the receptor along (Geom::Bounding::Box
 * with fixedDirections = false, Geom/bounding.hpp:135-183, docking_molstruct.hpp:
 * 199-245).  Barequet & Har-Peled's grid search (grid 5) with local improvement,
 * host code.  axes: 3 x 3, row k the unit direction of box axis k, a proper
 * rotation, rows ordered by decreasing extent; lo, hi: the points' extent along
 * each axis (lo[k] <= axes[k] . p <= hi[k]). */
pd_status pd_min_volume_box(int n, const double* pos, double* axes, double* lo, double* hi);

/* The two points farthest apart (Geom::Bounding::Diameter with eps = 0,
 * Geom/bounding.hpp:33-97): exact, the first such pair i < j. */
pd_status pd_diameter_pair(int n, const double* pos, int* i, int* j);

#ifdef __cplusplus
}
#endif
#endif /* PDFFT_H */
