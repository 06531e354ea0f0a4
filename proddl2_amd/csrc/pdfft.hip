// pdfft.hip -- host side of libpdfft.so: the C ABI declared in include/pdfft.h.
//
// A handle (pd_scan) stands for one FFTCorrelators + CorrelationProcessor pair
// of the reference worker (docking_fft.hpp:236-492, 767-883): the receptor
// spectra of every term stay resident in HBM, rotations stream through in
// batches, and each batch runs five kernels (scan_kernels.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/pdfft.h"
#include "scan_fast.hpp"
#include "scan_kernels.hpp"
#include "scan_post.hpp"
#include "scan_split.hpp"

namespace pd_detail {
// Message of the calling thread's last failed call, shared by every source of
// the library (pd_last_error).
std::string& last_error() {
  thread_local std::string msg;
  return msg;
}
}  // namespace pd_detail

namespace {

// Kernels specialised on a cubic grid edge (scan_fast.hpp); other grids run
// the generic runtime-plan kernels of scan_kernels.hpp.
using ColFn = void (*)(ScanGeom, const float2*, const float2*, const float4*, float2*, int, int, int);
using LanesFn = void (*)(ScanGeom, const float2*, float4*);
using LigFn = void (*)(ScanGeom, const float2*, const float2*, const float2*, const int*, int,
                       const float*, float2*);
using InvFn = void (*)(ScanGeom, const float2*, const float2*, const float2*, const float2*, float, int,
                       float*, int*, float*, unsigned*);
struct FastSet {
  int n;
  LigFn lig;
  InvFn inv;
  int threads;  // block size of both kernels (their template default, plane_threads)
};
#define PD_FAST(N) \
  FastSet { N, k_lig_plane_fwd_fast<N, N>, k_plane_inv_topk_fast<N, N>, plane_threads(N) }
const FastSet kFast[] = {PD_FAST(32),  PD_FAST(48),  PD_FAST(64),  PD_FAST(96),
                         PD_FAST(120), PD_FAST(128), PD_FAST(144), PD_FAST(160)};
#undef PD_FAST

// Column kernels depend on N0 only: they serve every grid whose N0 is listed,
// whatever its plane kernels (whole-plane, generic or split).  Each comes in
// sixteen variants by the number MZ of x positions per lane that may be
// nonzero (ceil(q A / 16), q = 1..16, so every MZ for A <= 16): the variant
// with the smallest MZ >= pp / B runs.
struct FastCol {
  int n;
  int col_lines;
  int lanes;  // B: lanes per column; L1 holds the ligand planes in whole groups of B
  int a;      // A = n / B positions per lane
  int ni;     // receptor value pairs per lane and term (k_rhat_lanes)
  LanesFn lanes_fn;
  int mz[16];
  ColFn col[16];
};
#define PD_MZ(N, Q) ((ColSplit<N>::A * (Q) + 15) / 16)
#define PD_MZK(N, Q) k_col_fast<N, PD_MZ(N, Q)>
#define PD_FCOL(N)                                                                                        \
  FastCol {                                                                                               \
    N, 256 / ColSplit<N>::B, ColSplit<N>::B, ColSplit<N>::A, ColLanes<N>::NI, k_rhat_lanes<N>,            \
        {PD_MZ(N, 1),  PD_MZ(N, 2),  PD_MZ(N, 3),  PD_MZ(N, 4),  PD_MZ(N, 5),  PD_MZ(N, 6),               \
         PD_MZ(N, 7),  PD_MZ(N, 8),  PD_MZ(N, 9),  PD_MZ(N, 10), PD_MZ(N, 11), PD_MZ(N, 12),              \
         PD_MZ(N, 13), PD_MZ(N, 14), PD_MZ(N, 15), PD_MZ(N, 16)},                                         \
        {PD_MZK(N, 1),  PD_MZK(N, 2),  PD_MZK(N, 3),  PD_MZK(N, 4),  PD_MZK(N, 5),  PD_MZK(N, 6),         \
         PD_MZK(N, 7),  PD_MZK(N, 8),  PD_MZK(N, 9),  PD_MZK(N, 10), PD_MZK(N, 11), PD_MZK(N, 12),        \
         PD_MZK(N, 13), PD_MZK(N, 14), PD_MZK(N, 15), PD_MZK(N, 16)}                                      \
  }
const FastCol kFastCol[] = {PD_FCOL(32),  PD_FCOL(48),  PD_FCOL(64),  PD_FCOL(96), PD_FCOL(120),
                            PD_FCOL(128), PD_FCOL(144), PD_FCOL(160), PD_FCOL(256)};
#undef PD_FCOL
#undef PD_MZ
#undef PD_MZK

// Row kernels of the split inverse specialised on N2.
using RowsFn = void (*)(ScanGeom, const float2*, const float2*, const float2*, float, int, float*, int*, float*,
                        unsigned*);
struct FastRows {
  int n2;
  int lds;  // bytes: (512 / B2) rows of LS complex
  RowsFn fn;
};
#define PD_FROWS(N) \
  FastRows { N, (512 / Split<N / 2>::B) * ((N / 2 + 1) | 1) * 8, k_inv_rows_topk_fast<N> }
const FastRows kFastRows[] = {PD_FROWS(48), PD_FROWS(64), PD_FROWS(96), PD_FROWS(128), PD_FROWS(256)};
#undef PD_FROWS

// y passes of the split path specialised on N1.
using YColFn = void (*)(ScanGeom, const float2*, float2*, int, int, int, int, int);
struct FastYCols {
  int n1;
  int cols;  // columns per block
  int lds;   // bytes: N1 rows of cols + 1 complex
  YColFn fwd, inv;
};
#define PD_FYCOL(N) \
  FastYCols { N, 256 / Split<N>::B, N * (256 / Split<N>::B + 1) * 8, k_cols_fast<N, -1>, k_cols_fast<N, +1> }
const FastYCols kFastYCols[] = {PD_FYCOL(32),  PD_FYCOL(48),  PD_FYCOL(64),  PD_FYCOL(96), PD_FYCOL(120),
                                PD_FYCOL(128), PD_FYCOL(144), PD_FYCOL(160), PD_FYCOL(256)};
#undef PD_FYCOL

const FastYCols* find_fast_ycols(int n1) {
  for (const FastYCols& f : kFastYCols)
    if (f.n1 == n1) return &f;
  return nullptr;
}

// Fused inverse y pass + rows + top-K of the split path (N1 = N2 = 256).
using InvSplitFn = void (*)(ScanGeom, const float2*, const float2*, const float2*, const float2*, float, int,
                           float*, int*, float*, unsigned*);
constexpr int kInvSplitN = 256;
constexpr int kInvSplitLds = 128 * (kInvSplitN / 2 + 1) * 8;  // 128 rows of the half plane
static_assert(kInvSplitLds >= 16 * 16 * 64 * 8, "y exchange fits the row buffer");
const InvSplitFn kInvSplit = k_inv_plane_topk_split<kInvSplitN, kInvSplitN>;
// ... and the one-kernel ligand plane pass (box rows up to kPlaneSplitRows).
using LigSplitFn = void (*)(ScanGeom, const float2*, const float2*, const int*, int, const float*,
                            float2*);
const LigSplitFn kLigSplit = k_lig_plane_split<kInvSplitN, kInvSplitN>;
static_assert(kInvSplitLds == kPlaneSplitRows * (kInvSplitN / 2 + 1) * 8, "box rows fill the row buffer");

const FastRows* find_fast_rows(int n2) {
  for (const FastRows& f : kFastRows)
    if (f.n2 == n2) return &f;
  return nullptr;
}

const FastCol* find_fast_col(int n0) {
  for (const FastCol& f : kFastCol)
    if (f.n == n0) return &f;
  return nullptr;
}

const FastSet* find_fast(int n0, int n1, int n2) {
  if (n0 != n1 || n1 != n2) return nullptr;
  for (const FastSet& f : kFast)
    if (f.n == n0) return &f;
  return nullptr;
}

pd_status fail(pd_status s, const std::string& msg) {
  pd_detail::last_error() = msg;
  return s;
}

#define PD_HIP(x)                                                                       \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess)                                                               \
      return fail(PD_ERR_DEVICE, std::string(#x) + " failed: " + hipGetErrorString(e_)); \
  } while (0)

// Device memory of one call, freed on every return path (hipFree waits for the
// work queued on it).
template <typename T>
struct CallBuf {
  T* p = nullptr;
  CallBuf() = default;
  CallBuf(const CallBuf&) = delete;
  CallBuf& operator=(const CallBuf&) = delete;
  ~CallBuf() {
    if (p) hipFree(p);
  }
};

constexpr int kPlaneThreads = 512;
constexpr int kColThreads = 256;
constexpr int kMaxLds = 160 * 1024;
constexpr int kMaxK = 32;

// Staging area of pd_scan_rescore64_noalloc: the poses' rotations, displacements
// and energies of one call, at most kNoAllocMaxPoses of them.  A file-scope
// device array, so each GPU holds its own copy from module load to process
// exit and no call allocates.  In host code `scratch` names a host-side stand-in:
// kernels and copies use the device address that hipGetSymbolAddress returns
// for the current device (scratch_ptr).
constexpr int kNoAllocMaxPoses = 8192;
constexpr size_t kNoAllocRotOff = 0;                                            // n x 9 f64
constexpr size_t kNoAllocShiftOff = kNoAllocRotOff + 72 * (size_t)kNoAllocMaxPoses;  // n x 3 i32
constexpr size_t kNoAllocEnergyOff = kNoAllocShiftOff + 12 * (size_t)kNoAllocMaxPoses;  // n x f64
constexpr size_t kNoAllocBytes = kNoAllocEnergyOff + 8 * (size_t)kNoAllocMaxPoses;
static_assert(kNoAllocShiftOff % 256 == 0 && kNoAllocEnergyOff % 256 == 0, "regions stay 256-byte aligned");

// Device address of `scratch` per device, set once per device while a handle is
// created there (pd_scan_create), and the lock that serialises the calls that
// use it on that device.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_setup;
void* g_scratch_ptr[kMaxDevices] = {};
std::string g_scratch_err[kMaxDevices];
std::mutex g_scratch_use[kMaxDevices];

int field_nparams(int kind) {
  switch (kind) {
    case PD_FIELD_LJ_REP: return 2;
    case PD_FIELD_LJ_ATT: return 2;
    case PD_FIELD_LJ_FULL: return 5;
    case PD_FIELD_TWOSTEP: return 3;
    case PD_FIELD_COULOMB: return 1;
    case PD_FIELD_DESOLV: return 1;
  }
  return -1;
}

// Centred geometry of one axis (Grid/gridcoord.hpp:127-142 then setup() 166-175).
void centred_axis(double h, int n, double* zero_h, double* zsc) {
  double upper = 0.0 + h + h * (n - 1);
  double lower = 0.0 + h * 0;
  double zero = -1. * (upper - lower) / 2;
  *zero_h = zero - h;
  *zsc = (zero - h * 0) + h / 2.0;
}

int to_logical_h(double x, double zero_h, double h) {
  double u = (x - zero_h) / h;
  if (!(u > -1.0e9 && u < 1.0e9)) return std::numeric_limits<int>::min();
  return static_cast<int>(u) - 1;
}

// Radix sequence for an FFT length: 16, 8, 4, 2 first, then odd primes <= 13.
bool radices_for(int n, std::vector<int>& r) {
  r.clear();
  while (n % 16 == 0) { r.push_back(16); n /= 16; }
  if (n % 8 == 0) { r.push_back(8); n /= 8; }
  if (n % 4 == 0) { r.push_back(4); n /= 4; }
  if (n % 2 == 0) { r.push_back(2); n /= 2; }
  for (int p : {3, 5, 7, 11, 13})
    while (n % p == 0) { r.push_back(p); n /= p; }
  return n == 1 && (int)r.size() <= PD_MAX_STAGES;
}

// Host copy of a plan's stages; the device copy is uploaded by upload_plan.
struct HostPlan {
  int n = 0;
  std::vector<int4> st;
};

bool make_plan(int n, HostPlan& P) {
  std::vector<int> r;
  if (n < 1 || !radices_for(n, r)) return false;
  P.n = n;
  P.st.clear();
  int ns = 1;
  for (size_t i = 0; i < r.size(); ++i) {
    P.st.push_back(make_int4(r[i], ns, n / (ns * r[i]), 0));
    ns *= r[i];
  }
  return true;
}

// Each thread keeps at most PD_MAXV/R butterflies of a radix-R stage in registers.
bool plan_fits(const HostPlan& P, int lines, int threads) {
  for (size_t s = 0; s < P.st.size(); ++s) {
    const int R = P.st[s].x;
    const int maxt = std::max(1, PD_MAXV / R);
    const long long tasks = (long long)lines * (P.n / R);
    if ((tasks + threads - 1) / threads > maxt) return false;
  }
  return true;
}

}  // namespace

// The no-alloc staging area (see kNoAllocBytes).
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

struct pd_graph;

struct pd_scan {
  pd_scan_desc d{};
  hipStream_t stream = nullptr;
  ScanGeom g{};
  int K = 10, B = 64, C = 32;
  long long N = 0;
  const FastSet* fast = nullptr;
  const FastCol* fcol = nullptr;  // column kernel specialised on N0, if any
  const FastRows* frows = nullptr;  // split inverse rows specialised on N2, if any
  const FastYCols* fycol = nullptr;  // split y passes specialised on N1, if any
  bool inv_fused = false;            // split inverse as one kernel (N1 = N2 = 256)
  bool lig_fused = false;            // split ligand pass as one kernel (N1 = N2 = 256)
  bool lig_rows = false;             // whole-plane grid whose ligand pass runs as row + column kernels
  bool rows_ok = false;              // the row + column ligand pass fits this whole-plane grid
  bool rows_auto = false;            // the library's rule for lig_rows
  pd_tuning tune{};                  // pd_scan_set_tuning (0: rule)
  std::vector<double> lig_pos_h, lig_w_h;  // the ligand as set, to re-lay it when lig_pass changes
  bool split = false;  // plane passes as row + column kernels (scan_split.hpp)
  HostPlan hx, hy, hz;
  FftPlan px{}, py{}, pz{};
  int4* d_stages = nullptr;  // stage tables of px, py, pz
  float2 *tw_x = nullptr, *tw_y = nullptr, *tw_z = nullptr, *tw2 = nullptr;
  double* R64 = nullptr;
  float2* Rhat = nullptr;
  float4* Rlanes = nullptr;  // Rhat in the fast column kernel's lane order
  int rhat_shift = 0;  // x shift of the receptor grid behind Rhat (g.pl when computed)
  bool receptor_ready = false;
  int n_lig = 0;
  double* q = nullptr;
  double* w64 = nullptr;
  float* w32 = nullptr;
  bool ligand_ready = false;
  // Per-batch buffers, one set per lane: consecutive batches go round the
  // lanes' streams so that one batch's kernels can share the CUs with the next
  // batches' (the column pass waits on memory, the plane passes on VALU).
  struct Lane {
    int* cells = nullptr;
    float2* L1 = nullptr;
    float2* S1 = nullptr;
    float* cval = nullptr;
    int* coff = nullptr;
    unsigned* thr = nullptr;  // per-rotation selection threshold of a batch
    float* grid = nullptr;    // the batch's real correlation grids (post-filters, candidate lists)
    hipStream_t s = nullptr;
    hipEvent_t done = nullptr;
  };
  // Up to three lanes (default_lanes picks two or three); four were slower
  // (with the handle's stream, three lanes fill the four hardware queues a
  // process gets by default).
  static constexpr int kLanes = 3;
  Lane lane[kLanes];
  hipEvent_t fork = nullptr;
  int* err = nullptr;
  double* d_rots = nullptr;
  size_t rots_cap = 0;
  PeakOut* d_peaks = nullptr;
  int* d_nout = nullptr;
  size_t out_cap = 0;
  bool timing = false;
  struct Pending { int k; hipEvent_t a, b; };
  std::vector<Pending> pending;
  double tot_ms[5] = {0, 0, 0, 0, 0};
  long long launches[5] = {0, 0, 0, 0, 0};
  // Post-filters (pd_scan_set_postfilter): when on, each batch selects
  // max_n_trans_inp candidates from its grids and filters them (k_select_post).
  bool post = false;
  PostGeom pg{};
  bool grids = false;  // the lanes carry grid buffers
  // The graph whose nodes name this scan's buffers, if any: while it exists the
  // buffers must not move (set_batch, set_ligand and destroy are refused).
  pd_graph* graph = nullptr;
};

namespace {

void free_batch(pd_scan* h) {
  for (auto& ln : h->lane) {
    hipFree(ln.cells); ln.cells = nullptr;
    hipFree(ln.L1); ln.L1 = nullptr;
    hipFree(ln.S1); ln.S1 = nullptr;
    hipFree(ln.cval); ln.cval = nullptr;
    hipFree(ln.coff); ln.coff = nullptr;
    hipFree(ln.thr); ln.thr = nullptr;
    hipFree(ln.grid); ln.grid = nullptr;
  }
}

// Per-batch buffers for h->B rotations (sizes follow the ligand box).
pd_status alloc_batch(pd_scan* h) {
  const ScanGeom& g = h->g;
  const size_t B = h->B;
  for (auto& ln : h->lane) {
    PD_HIP(hipMalloc(&ln.cells, sizeof(int) * B * h->n_lig));
    // planes of [pl, pl + pp) outside the ligand box are never written: zero them once
    const size_t l1 = sizeof(float2) * B * g.T * (size_t)g.pp * g.M;
    PD_HIP(hipMalloc(&ln.L1, l1));
    PD_HIP(hipMemset(ln.L1, 0, l1));
    PD_HIP(hipMalloc(&ln.S1, sizeof(float2) * B * (size_t)g.n0 * g.M));
    PD_HIP(hipMalloc(&ln.cval, sizeof(float) * B * g.n0 * h->K));
    PD_HIP(hipMalloc(&ln.coff, sizeof(int) * B * g.n0 * h->K));
    PD_HIP(hipMalloc(&ln.thr, sizeof(unsigned) * B));
    if (h->grids) PD_HIP(hipMalloc(&ln.grid, sizeof(float) * B * (size_t)h->N));
  }
  return PD_OK;
}

// Grid buffers for the post-filter and candidate paths, added to the lanes.
pd_status ensure_grids(pd_scan* h) {
  if (h->grids) return PD_OK;
  h->grids = true;
  if (!h->ligand_ready) return PD_OK;  // alloc_batch adds them with the ligand
  for (auto& ln : h->lane) PD_HIP(hipMalloc(&ln.grid, sizeof(float) * (size_t)h->B * (size_t)h->N));
  return PD_OK;
}

pd_status make_twiddles(pd_scan* h, int n, float2** out) {
  PD_HIP(hipMalloc(out, sizeof(float2) * (size_t)n));
  hipLaunchKernelGGL(k_twiddles, dim3((n + 255) / 256), dim3(256), 0, h->stream, n, *out);
  PD_HIP(hipGetLastError());
  return PD_OK;
}

pd_status set_lds(const void* fn, int bytes) {
  if (bytes > kMaxLds) return fail(PD_ERR_UNSUPPORTED, "LDS requirement exceeds 160 KiB");
  PD_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  return PD_OK;
}

struct Timer {
  pd_scan* h;
  int k;
  hipStream_t s;
  hipEvent_t a = nullptr, b = nullptr;
  Timer(pd_scan* h_, int k_, hipStream_t s_) : h(h_), k(k_), s(s_) {
    if (h->timing) {
      hipEventCreate(&a);
      hipEventCreate(&b);
      hipEventRecord(a, s);
    }
  }
  ~Timer() {
    if (h->timing) {
      hipEventRecord(b, s);
      h->pending.push_back({k, a, b});
    }
  }
};

// Where the kernels of a batch go: launched on a stream, or added as kernel
// nodes to a graph under construction, each node after the chain's previous
// node (the first after `deps`).  Arguments are converted to the kernel's
// parameter types and copied into the node (hipGraphAddKernelNode copies them).
struct Launcher {
  hipStream_t s = nullptr;
  hipGraph_t graph = nullptr;
  std::vector<hipGraphNode_t> deps;
  hipError_t err = hipSuccess;

  template <typename... P, typename... A>
  void operator()(void (*fn)(P...), dim3 grid, dim3 block, size_t shmem, A&&... args) {
    static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
    if (!graph) {
      hipLaunchKernelGGL(fn, grid, block, shmem, s, std::forward<A>(args)...);
      return;
    }
    std::tuple<std::decay_t<P>...> vals(std::forward<A>(args)...);
    add_node(reinterpret_cast<void*>(fn), grid, block, shmem, vals, std::index_sequence_for<P...>{});
  }

 private:
  template <typename Tup, size_t... I>
  void add_node(void* fn, dim3 grid, dim3 block, size_t shmem, Tup& vals, std::index_sequence<I...>) {
    void* ptrs[] = {static_cast<void*>(&std::get<I>(vals))...};
    hipKernelNodeParams p{};
    p.func = fn;
    p.gridDim = grid;
    p.blockDim = block;
    p.sharedMemBytes = (unsigned)shmem;
    p.kernelParams = ptrs;
    p.extra = nullptr;
    hipGraphNode_t node = nullptr;
    if (err != hipSuccess) return;
    err = hipGraphAddKernelNode(&node, graph, deps.empty() ? nullptr : deps.data(), deps.size(), &p);
    if (err == hipSuccess) deps.assign(1, node);
  }
};

pd_status resolve_timing(pd_scan* h) {
  if (h->pending.empty()) return PD_OK;
  PD_HIP(hipStreamSynchronize(h->stream));
  for (auto& p : h->pending) {
    float ms = 0;
    PD_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    h->tot_ms[p.k] += ms;
    h->launches[p.k] += 1;
    hipEventDestroy(p.a);
    hipEventDestroy(p.b);
  }
  h->pending.clear();
  return PD_OK;
}

int plane_lds(const pd_scan* h) { return h->g.n1 * h->g.ls * (int)sizeof(float2); }
int split_row_lds(const pd_scan* h) { return h->g.zrows * h->g.ls * (int)sizeof(float2); }
int split_col_lds(const pd_scan* h) { return h->g.n1 * kSplitColStride * (int)sizeof(float2); }
int col_lds(const pd_scan* h) { return h->g.n0 * h->C * (int)sizeof(float2); }

// Receptor forward transform of the FP64 grids already in h->R64.
pd_status receptor_spectra(pd_scan* h) {
  const ScanGeom& g = h->g;
  CallBuf<float> r32b;
  CallBuf<float2> tmpb;
  PD_HIP(hipMalloc(&r32b.p, sizeof(float) * h->N * g.T));
  PD_HIP(hipMalloc(&tmpb.p, sizeof(float2) * (size_t)g.n0 * g.M));
  float* const r32 = r32b.p;
  float2* const tmp = tmpb.p;
  const long long tot = h->N * g.T;
  hipLaunchKernelGGL(k_to_float, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, h->stream, h->R64,
                     r32, tot);
  PD_HIP(hipGetLastError());
  const int ntile = (g.M + h->C - 1) / h->C;
  for (int t = 0; t < g.T; ++t) {
    if (h->split) {
      hipLaunchKernelGGL(k_rec_rows, dim3(g.n0, (g.n1 + g.zrows - 1) / g.zrows), dim3(kSplitThreads),
                         split_row_lds(h), h->stream, g, h->pz, h->tw2, r32 + (size_t)t * h->N, tmp);
      if (h->fycol)
        hipLaunchKernelGGL(h->fycol->fwd, dim3(g.n0, (g.h2 + h->fycol->cols - 1) / h->fycol->cols), dim3(256),
                           h->fycol->lds, h->stream, g, h->tw_y, tmp, g.n0, g.n0, 0, 0, g.n1);
      else
        hipLaunchKernelGGL(k_cols_fft<-1>, dim3(g.n0, (g.h2 + kSplitCols - 1) / kSplitCols), dim3(kSplitThreads),
                           split_col_lds(h), h->stream, g, h->py, tmp, g.n0, g.n0, 0, 0, g.n1);
    } else {
      hipLaunchKernelGGL(k_rec_plane_fwd, dim3(g.n0), dim3(kPlaneThreads), plane_lds(h), h->stream, g,
                         h->pz, h->py, h->tw2, r32 + (size_t)t * h->N, tmp);
    }
    PD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_col_fwd, dim3(ntile), dim3(kColThreads), col_lds(h), h->stream, g, h->px,
                       tmp, h->Rhat + (size_t)t * g.n0 * g.M, h->C, g.pl);
    PD_HIP(hipGetLastError());
  }
  if (h->fcol) {
    const long long ntile = (g.M + h->fcol->col_lines - 1) / h->fcol->col_lines;
    const long long tot = (long long)g.T * ntile * 4 * h->fcol->ni * 64;
    hipLaunchKernelGGL(h->fcol->lanes_fn, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, h->stream, g, h->Rhat,
                       h->Rlanes);
    PD_HIP(hipGetLastError());
  }
  PD_HIP(hipStreamSynchronize(h->stream));
  h->rhat_shift = g.pl;
  h->receptor_ready = true;
  return PD_OK;
}

// Launch shape of the column pass for a batch of nb rotations: the kernel
// (specialised or generic), its tiles, rotation groups per tile (nyb) of bpb
// rotations, and the specialised kernel's MZ variant.
struct ColShape {
  const FastCol* fc;
  int ntile, nyb, bpb, v;
};

ColShape col_shape(const pd_scan* h, int nb) {
  const ScanGeom& g = h->g;
  // The specialised column kernel addresses a block group's ligand planes and
  // outputs through 32-bit buffer offsets: cap the rotations per group.
  const FastCol* fc = h->fcol;
  const long long lig_bytes = (long long)g.T * g.pp * g.M * 8, out_bytes = (long long)g.n0 * g.M * 8;
  const long long max_bpb = std::min(0xFFFFFFFFLL / lig_bytes, 0xFFFFFFFFLL / out_bytes);
  if (max_bpb < 1) fc = nullptr;
  const int C = fc ? fc->col_lines : h->C;
  const int ntile = (g.M + C - 1) / C;
  int nyb = std::max(1, std::min(nb, 4096 / std::max(1, ntile)));
  if (fc) {
    // About 4,096 blocks: the rotation groups of a tile run together on one
    // XCD and share each read of the tile's receptor spectra through L2.
    // Measured with the direct-summation ligand pass (two blocks per CU) and
    // the default lanes: 128^3 (260 tiles) 16 groups of 4 rotations 96.7k
    // rotations/s against 89.4-96.0k with 4 groups of 16 (the lower figure
    // with three lanes where two of them shared a hardware queue); 96^3
    // (147 tiles) 16 groups 158k against 147k with the earlier 1,024-block
    // rule; 160^3 (810 tiles) 5 groups 21.4k against 20.8k with 1.  256^3
    // (2,064 tiles) gets 2 groups of 32, the fewest its 32-bit offsets allow
    // (column pass 525 ms per 10,000 rotations against 607 with 16 groups
    // of 4).
    nyb = std::max(1, std::min((nb + 1) / 2, (4096 + ntile / 2) / ntile));
  }
  if (h->tune.col_groups > 0) nyb = std::min(nb, h->tune.col_groups);
  if (fc) nyb = std::max(nyb, (int)((nb + max_bpb - 1) / max_bpb));
  int bpb = (nb + nyb - 1) / nyb;
  // the fast column pass takes rotations in pairs: an even group size
  // leaves no phantom rotation except at an odd batch end
  if (fc && bpb % 2 && bpb < max_bpb) ++bpb;
  nyb = (nb + bpb - 1) / bpb;
  ColShape cs{fc, ntile, nyb, bpb, 0};
  if (fc) {
    const int mz = (g.pp + fc->lanes - 1) / fc->lanes;  // positions per lane that hold ligand planes
    int v = 0;
    while (v < 15 && fc->mz[v] < mz) ++v;
    cs.v = std::min(15, v + h->tune.col_mz_extra);
  }
  return cs;
}

// The four FFT-path kernels of one batch of nb rotations (bin, ligand plane
// pass, column pass, inverse plane pass + top-K).  grid_out (optional)
// receives the scaled correlation grids.
pd_status launch_fft_path(pd_scan* h, const pd_scan::Lane& ln, int nb, const double* d_rots,
                          float* grid_out, Launcher& L) {
  hipStream_t stream = L.s;
  const ScanGeom& g = h->g;
  const float scale = (float)(1.0 / (double)h->N);
  {
    Timer tm(h, 0, stream);
    const long long tot = (long long)nb * h->n_lig;
    L(k_bin, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, g, d_rots, nb,
                       h->q, h->n_lig, ln.cells, h->err, ln.thr);
  }
  const FastSet* f = h->fast;
  {
    Timer tm(h, 1, stream);
    if (h->split && h->lig_fused && g.p[1] <= kPlaneSplitRows) {
      L(kLigSplit, dim3((g.p[0] + 15) / 16 * 16, nb, g.T), dim3(1024), kInvSplitLds, g, h->tw_y, h->tw2,
                         ln.cells, h->n_lig, h->w32, ln.L1);
    } else if (h->split || h->lig_rows) {
      L(k_lig_rows, dim3(g.p[0], nb, (g.p[1] + g.zrows - 1) / g.zrows),
                         dim3(kSplitThreads), split_row_lds(h), g, h->pz, h->tw2, ln.cells, h->n_lig,
                         h->w32, ln.L1);
      if (h->fycol)
        L(h->fycol->fwd, dim3(nb * g.T * g.p[0], (g.h2 + h->fycol->cols - 1) / h->fycol->cols),
                           dim3(256), h->fycol->lds, g, h->tw_y, ln.L1, g.p[0], g.pp, g.lo[0] - g.pl,
                           g.lo[1], g.lo[1] + g.p[1]);
      else
        L(k_cols_fft<-1>, dim3(nb * g.T * g.p[0], (g.h2 + kSplitCols - 1) / kSplitCols),
                           dim3(kSplitThreads), split_col_lds(h), g, h->py, ln.L1, g.p[0], g.pp,
                           g.lo[0] - g.pl, g.lo[1], g.lo[1] + g.p[1]);
    } else if (f)
      L(f->lig, dim3((g.p[0] + 15) / 16 * 16, nb), dim3(f->threads), plane_lds(h), g,
                         h->tw_y, h->tw_z, h->tw2, ln.cells, h->n_lig, h->w32, ln.L1);
    else
      L(k_lig_plane_fwd, dim3(g.p[0], nb), dim3(kPlaneThreads), plane_lds(h),
                         g, h->pz, h->py, h->tw2, ln.cells, h->n_lig, h->w32, ln.L1);
  }
  {
    Timer tm(h, 2, stream);
    const ColShape cs = col_shape(h, nb);
    const FastCol* fc = cs.fc;
    const int ntile = cs.ntile, nyb = cs.nyb, bpb = cs.bpb;
    if (fc) {
      const int nblk = 8 * ((ntile + 7) / 8) * nyb;
      L(fc->col[cs.v], dim3(nblk), dim3(kColThreads), 0, g, h->tw_x, ln.L1, h->Rlanes,
                         ln.S1, bpb, nb, nyb);
    } else
      L(k_col, dim3(ntile, nyb), dim3(kColThreads), col_lds(h), g, h->px,
                         ln.L1, h->Rhat, ln.S1, h->C, bpb, nb);
  }
  {
    Timer tm(h, 3, stream);
    if (h->split && h->inv_fused) {
      L(kInvSplit, dim3(nb, g.n0), dim3(1024), kInvSplitLds, g, h->tw_y, h->tw_z, h->tw2,
                         ln.S1, scale, h->K, ln.cval, ln.coff, grid_out, ln.thr);
    } else if (h->split) {
      if (h->fycol)
        L(h->fycol->inv, dim3(nb * g.n0, (g.h2 + h->fycol->cols - 1) / h->fycol->cols), dim3(256),
                           h->fycol->lds, g, h->tw_y, ln.S1, g.n0, g.n0, 0, 0, g.n1);
      else
        L(k_cols_fft<+1>, dim3(nb * g.n0, (g.h2 + kSplitCols - 1) / kSplitCols),
                           dim3(kSplitThreads), split_col_lds(h), g, h->py, ln.S1, g.n0, g.n0, 0, 0, g.n1);
      if (h->frows)
        L(h->frows->fn, dim3(nb, g.n0), dim3(512), h->frows->lds, g, h->tw_z, h->tw2, ln.S1,
                           scale, h->K, ln.cval, ln.coff, grid_out, ln.thr);
      else
        L(k_inv_rows_topk, dim3(g.n0, nb), dim3(kSplitThreads), split_row_lds(h), g,
                           h->pz, h->tw2, ln.S1, scale, h->K, ln.cval, ln.coff, grid_out);
    } else if (f)
      L(f->inv, dim3(nb, g.n0), dim3(f->threads), plane_lds(h), g, h->tw_y,
                         h->tw_z, h->tw2, ln.S1, scale, h->K, ln.cval, ln.coff, grid_out, ln.thr);
    else
      L(k_plane_inv_topk, dim3(g.n0, nb), dim3(kPlaneThreads), plane_lds(h),
                         g, h->pz, h->py, h->tw2, ln.S1, scale, h->K, ln.cval, ln.coff, grid_out);
  }
  PD_HIP(hipGetLastError());
  return PD_OK;
}

// One batch of nb rotations starting at rotation b0 of the call: the FFT path
// and the per-rotation merge into the call's outputs.
// Output of a batch: the first K peaks (merge of the per-plane lists, or the
// post-filters' output), or the whole candidate list.
enum class BatchOut { kPeaks, kCandidates };

pd_status launch_batch(pd_scan* h, const pd_scan::Lane& ln, int b0, int nb, const double* d_rots, int rot0,
                       PeakOut* d_peaks, int* d_nout, Launcher& L, BatchOut what = BatchOut::kPeaks) {
  const ScanGeom& g = h->g;
  const bool select = h->post || what == BatchOut::kCandidates;
  pd_status st = launch_fft_path(h, ln, nb, d_rots + 9 * (size_t)b0, select ? ln.grid : nullptr, L);
  if (st != PD_OK) return st;
  if (select) {
    Timer tm(h, 4, L.s);
    PostGeom pg = h->pg;
    pg.candidates = what == BatchOut::kCandidates;
    const int stride = pg.candidates ? pg.k_inp : h->K;
    L(k_select_post, dim3(nb), dim3(kSelThreads), 0, g, pg, ln.grid, d_rots + 9 * (size_t)b0, rot0 + b0,
      d_peaks + (size_t)b0 * stride, d_nout + b0);
  } else {
    Timer tm(h, 4, L.s);
    if (g.n0 <= 64 * kMergePlanes)
      L(k_merge_wave, dim3(nb), dim3(256), (size_t)g.n0 * h->K * 8, g, h->K, ln.cval, ln.coff,
        h->d.max_val_corr, rot0 + b0, d_peaks + (size_t)b0 * h->K, d_nout + b0);
    else
      L(k_merge, dim3(nb), dim3(256), (size_t)g.n0 * h->K * 8, g, h->K, ln.cval, ln.coff, h->d.max_val_corr,
        rot0 + b0, d_peaks + (size_t)b0 * h->K, d_nout + b0);
  }
  if (L.err != hipSuccess) return fail(PD_ERR_DEVICE, std::string("graph node: ") + hipGetErrorString(L.err));
  PD_HIP(hipGetLastError());
  return PD_OK;
}

pd_status check_run(pd_scan* h, int n_rot) {
  if (!h->receptor_ready) return fail(PD_ERR_STATE, "receptor not set");
  if (!h->ligand_ready) return fail(PD_ERR_STATE, "ligand not set");
  if (n_rot < 0) return fail(PD_ERR_INVALID_ARGUMENT, "n_rot < 0");
  return PD_OK;
}

// Batch lanes by default: two up to N0 = 128, three above.  The lanes and the
// handle's stream take three of the four hardware queues a process gets, so a
// caller's own stream (torch's, for one) never shares a queue with a lane;
// with three lanes at 128^3 that sharing took the scan from 96.0k to 89.4k
// rotations/s, two lanes ran 96.7k either way (96^3: 158k against 152k).
// Grids above 128 keep three (160^3: 21.6k against 21.4k; 256^3: 9.9k
// against 9.7k).
int default_lanes(const pd_scan* h) { return h->g.n0 > 128 ? pd_scan::kLanes : 2; }

// Lanes a scan of nbatch batches uses: one while kernels are timed (so that
// each recorded duration belongs to one kernel alone), with PD_SCAN_ONE_LANE,
// or for a single batch.
int lanes_for(const pd_scan* h, int nbatch) {
  const bool single = h->timing || nbatch <= 1 || (h->d.flags & PD_SCAN_ONE_LANE) != 0;
  return single ? 1 : (h->tune.lanes > 0 ? h->tune.lanes : default_lanes(h));
}

// Eager scan on the handle's stream: batches go round the lanes' streams,
// which fork from and join back into the handle's stream.
pd_status run_batches(pd_scan* h, int n_rot, const double* d_rots, int rot0, PeakOut* d_peaks,
                      int* d_nout, BatchOut what = BatchOut::kPeaks) {
  pd_status st = check_run(h, n_rot);
  if (st != PD_OK) return st;
  const int nbatch = (n_rot + h->B - 1) / h->B;
  if (nbatch == 0) return PD_OK;
  const int nl = lanes_for(h, nbatch);
  pd_scan::Lane own = h->lane[0];
  own.s = h->stream;
  if (nl > 1) {
    PD_HIP(hipEventRecord(h->fork, h->stream));
    for (int l = 0; l < nl; ++l) PD_HIP(hipStreamWaitEvent(h->lane[l].s, h->fork, 0));
  }
  for (int i = 0; i < nbatch && st == PD_OK; ++i) {
    const pd_scan::Lane& ln = nl == 1 ? own : h->lane[i % nl];
    Launcher L;
    L.s = ln.s;
    const int b0 = i * h->B;
    st = launch_batch(h, ln, b0, std::min(h->B, n_rot - b0), d_rots, rot0, d_peaks, d_nout, L, what);
  }
  // the lanes join the handle's stream also after a failed launch, so that
  // later work on that stream still follows everything already queued
  if (nl > 1)
    for (int l = 0; l < nl; ++l) {
      PD_HIP(hipEventRecord(h->lane[l].done, h->lane[l].s));
      PD_HIP(hipStreamWaitEvent(h->stream, h->lane[l].done, 0));
    }
  return st;
}

// The same scan as kernel nodes of a graph: batch i on lane i % nl waits for
// the previous batch of its lane (which used the same buffers), the first batch
// of each lane for `start`.  `tail` receives the last node of every lane.
pd_status graph_batches(pd_scan* h, hipGraph_t graph, const std::vector<hipGraphNode_t>& start, int n_rot,
                        const double* d_rots, PeakOut* d_peaks, int* d_nout, std::vector<hipGraphNode_t>& tail) {
  pd_status st = check_run(h, n_rot);
  if (st != PD_OK) return st;
  const int nbatch = (n_rot + h->B - 1) / h->B;
  const int nl = lanes_for(h, nbatch);
  std::vector<std::vector<hipGraphNode_t>> last(nl, start);
  for (int i = 0; i < nbatch; ++i) {
    Launcher L;
    L.graph = graph;
    L.deps = last[i % nl];
    const int b0 = i * h->B;
    st = launch_batch(h, h->lane[i % nl], b0, std::min(h->B, n_rot - b0), d_rots, 0, d_peaks, d_nout, L);
    if (st != PD_OK) return st;
    last[i % nl] = L.deps;
  }
  tail.clear();
  for (int l = 0; l < nl; ++l)
    if (nl == 1 || l < nbatch) tail.insert(tail.end(), last[l].begin(), last[l].end());
  if (tail.empty()) tail = start;
  return PD_OK;
}

pd_status check_err(pd_scan* h) {
  int e = 0;
  PD_HIP(hipMemcpyAsync(&e, h->err, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  PD_HIP(hipStreamSynchronize(h->stream));
  if (e) {
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(1), 0, h->stream, h->err, 1, 0);
    return fail(PD_ERR_DOMAIN, "a rotated ligand atom fell outside the FFT grid");
  }
  return PD_OK;
}

pd_status ensure_io(pd_scan* h, int n_rot) {
  // a failed allocation leaves no freed pointer behind a nonzero capacity
  if ((size_t)n_rot > h->rots_cap) {
    hipFree(h->d_rots);
    h->d_rots = nullptr;
    h->rots_cap = 0;
    PD_HIP(hipMalloc(&h->d_rots, sizeof(double) * 9 * (size_t)n_rot));
    h->rots_cap = n_rot;
  }
  if ((size_t)n_rot > h->out_cap) {
    hipFree(h->d_peaks);
    hipFree(h->d_nout);
    h->d_peaks = nullptr;
    h->d_nout = nullptr;
    h->out_cap = 0;
    PD_HIP(hipMalloc(&h->d_peaks, sizeof(PeakOut) * (size_t)n_rot * h->K));
    PD_HIP(hipMalloc(&h->d_nout, sizeof(int) * (size_t)n_rot));
    h->out_cap = n_rot;
  }
  return PD_OK;
}

}  // namespace

extern "C" {

const char* pd_last_error(void) { return pd_detail::last_error().c_str(); }

int pd_find_best_size(int start) {
  // FFTW_Size::findBestSize (Math/fft.hpp:82-151): first n >= max(2, start)
  // factoring over {2,3,5,7,11,13} with e11 + e13 <= 1.
  for (long long n = std::max(2, start); n < std::numeric_limits<int>::max(); ++n) {
    long long m = n;
    int e[6] = {0, 0, 0, 0, 0, 0};
    const int pr[6] = {2, 3, 5, 7, 11, 13};
    for (int i = 0; i < 6; ++i)
      while (m % pr[i] == 0) { m /= pr[i]; ++e[i]; }
    if (m == 1 && e[4] + e[5] <= 1) return (int)n;
  }
  return -1;
}

pd_status pd_ljcomp_params(int n_rec, const double* eps_r, const double* sigma_r,
                           const double* alpha_r, int n_lig, const double* eps_l,
                           const double* sigma_l, double* rep, double* att, double* w_rep,
                           double* w_att) {
  if (n_rec < 0 || n_lig <= 0) return fail(PD_ERR_INVALID_ARGUMENT, "atom counts");
  double sl = 0;
  for (int j = 0; j < n_lig; ++j) sl += sigma_l[j];
  const double sig_avg = sl / n_lig;
  for (int a = 0; a < n_rec; ++a) {
    const double sg = std::sqrt(sigma_r[a]), sa = std::sqrt(sigma_r[a] * sig_avg),
                 ep = std::sqrt(eps_r[a]);
    const double t1 = sg * sg * sg, s6 = t1 * t1;
    const double t2 = sa * sa * sa, sa6 = t2 * t2;
    const double as6 = alpha_r[a] * sa6;
    const double e4s6 = 4. * ep * s6;
    rep[2 * a] = e4s6 * s6;
    rep[2 * a + 1] = as6;
    att[2 * a] = e4s6;
    att[2 * a + 1] = as6;
  }
  for (int j = 0; j < n_lig; ++j) {
    const double s = sigma_l[j], e = std::sqrt(eps_l[j]);
    const double t3 = s * s * s;
    w_rep[j] = e * (t3 * t3);
    w_att[j] = e * t3;
  }
  return PD_OK;
}

pd_status pd_ljavg_params(int n_rec, const double* eps_r, const double* sigma_r, const double* alpha_r,
                          int n_lig, const double* eps_l, const double* sigma_l, int mix, double cutoff,
                          double* params, double* w) {
  if (n_rec < 0 || n_lig <= 0) return fail(PD_ERR_INVALID_ARGUMENT, "atom counts");
  if (!(cutoff > 0)) return fail(PD_ERR_INVALID_ARGUMENT, "cutoff must be > 0");
  // ligand averages (blitz::mean, docking_molforce.hpp:395-396)
  double se = 0, ss = 0;
  for (int j = 0; j < n_lig; ++j) {
    se += eps_l[j];
    ss += sigma_l[j];
  }
  const double eps_avg = se / n_lig, sig_avg = ss / n_lig;
  const double cut2 = cutoff * cutoff;
  for (int a = 0; a < n_rec; ++a) {
    // mixed parameters of atom a with the average ligand atom (:406-421)
    const double sigma = mix == 0 ? (sigma_r[a] + sig_avg) / 2. : std::sqrt(sigma_r[a] * sig_avg);
    const double eps = std::sqrt(eps_r[a] * eps_avg);
    // SoftCoreLJ with the shifted-force cutoff (potentials_inc.hpp:185-208)
    const double t = sigma * sigma * sigma, s6 = t * t;
    const double as6 = alpha_r[a] * s6, e4s6 = 4. * eps * s6, e4s12 = e4s6 * s6;
    // force at the cutoff, g2(cut^2, (cut, 0, 0)) before any shift (:221-228)
    const double r6 = cut2 * cut2 * cut2, A = r6 + as6, Bq = s6 / (r6 + as6) - 1.;
    const double Cq = -6. * (cut2 * cut2) * (e4s6 * A * Bq + e4s12) / (A * A * A);
    const double g_cut = cutoff * Cq + cutoff * (0.0 / std::sqrt(cut2));
    // energy at the cutoff, f2(cut^2) before any shift (:214-219)
    const double f_cut = (e4s12 / A - e4s6) / A + 0.0 + 0.0 * std::sqrt(cut2);
    const double gshift = -g_cut;
    const double fshift = -f_cut - gshift * cutoff;
    double* p = params + 5 * (size_t)a;
    p[0] = e4s12;
    p[1] = e4s6;
    p[2] = as6;
    p[3] = fshift;
    p[4] = gshift;
  }
  for (int j = 0; j < n_lig; ++j) w[j] = 1;  // (:439-446)
  return PD_OK;
}

pd_status pd_scan_create(const pd_scan_desc* desc, pd_scan** out) {
  if (!desc || !out) return fail(PD_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  const pd_scan_desc& d = *desc;
  for (int i = 0; i < 3; ++i)
    if (d.n[i] < 2 || d.n[i] > 4096) return fail(PD_ERR_INVALID_ARGUMENT, "grid edge out of range");
  {
    std::vector<int> r;
    const int zlen = d.n[2] % 2 ? d.n[2] : d.n[2] / 2;  // length of the z transform
    if (!radices_for(d.n[0], r) || !radices_for(d.n[1], r) || !radices_for(zlen, r))
      return fail(PD_ERR_UNSUPPORTED, "grid edges must factor over {2,3,5,7,11,13}");
  }
  if (!(d.grid_step > 0)) return fail(PD_ERR_INVALID_ARGUMENT, "grid_step must be > 0");
  if (d.n_terms < 1 || d.n_terms > 16) return fail(PD_ERR_INVALID_ARGUMENT, "n_terms must be 1..16");
  if (d.max_n_trans < 1 || d.max_n_trans > kMaxK)
    return fail(PD_ERR_UNSUPPORTED, "max_n_trans must be 1..32");
  if (d.max_n_trans_inp < d.max_n_trans)
    return fail(PD_ERR_INVALID_ARGUMENT, "max_n_trans_inp < max_n_trans");
  if (d.batch < 0) return fail(PD_ERR_INVALID_ARGUMENT, "batch < 0");
  int ndev = 0;
  PD_HIP(hipGetDeviceCount(&ndev));
  if (d.device < 0 || d.device >= ndev) return fail(PD_ERR_DEVICE, "no such HIP device");
  PD_HIP(hipSetDevice(d.device));
  if (d.device < kMaxDevices) {
    // once per device, with that device current: the device address of the
    // no-alloc staging array (a failure only disables pd_scan_rescore64_noalloc,
    // which then reports it)
    std::lock_guard<std::mutex> lk(g_scratch_setup);
    if (!g_scratch_ptr[d.device] && g_scratch_err[d.device].empty()) {
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e == hipSuccess && p)
        g_scratch_ptr[d.device] = p;
      else
        g_scratch_err[d.device] = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
    }
  }

  pd_scan* h = new pd_scan();
  h->d = d;
  h->K = d.max_n_trans;
  h->B = d.batch > 0 ? d.batch : 64;
  ScanGeom& g = h->g;
  g.n0 = d.n[0];
  g.n1 = d.n[1];
  g.n2 = d.n[2];
  g.h2 = g.n2 / 2 + 1;
  g.odd2 = g.n2 % 2;
  g.m2 = g.odd2 ? g.n2 : g.n2 / 2;  // the z transform: full complex length for an odd n2
  g.ls = g.odd2 ? g.n2 : (g.h2 | 1);
  g.M = g.n1 * g.h2;
  g.T = d.n_terms;
  g.h = d.grid_step;
  // measured: 256^3 column pass 702 -> 606 ms per 10,000 rotations; 128^3
  // with 16 rotations per block 201.5 -> 189.2 ms per 30,000 (89.2k -> 92.4k
  // rotations/s); 120^3 104.4k -> 105.6k rotations/s, 96^3 163.5k -> 164.2k
  g.col_sync = g.n0 >= 96;
  for (int a = 0; a < 3; ++a) {
    centred_axis(d.grid_step, d.n[a], &g.zero_h[a], &g.zsc[a]);
    g.lo[a] = 0;
    g.p[a] = d.n[a];
  }
  g.pl = 0;
  g.pp = g.n0;
  g.il = 0;
  g.ilsh = 0;
  h->N = (long long)g.n0 * g.n1 * g.n2;
  h->pg.K = d.max_n_trans;
  h->pg.k_inp = d.max_n_trans_inp;
  h->pg.max_val = d.max_val_corr;
  h->C = std::max(1, std::min(64, 4096 / g.n0));
  auto bad = [&](pd_status s, const std::string& m) {
    pd_scan_destroy(h);
    return fail(s, m);
  };
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
    return bad(PD_ERR_DEVICE, "hipStreamCreate failed");
  for (auto& ln : h->lane)
    if (hipStreamCreateWithFlags(&ln.s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ln.done, hipEventDisableTiming) != hipSuccess)
      return bad(PD_ERR_DEVICE, "lane stream creation failed");
  if (hipEventCreateWithFlags(&h->fork, hipEventDisableTiming) != hipSuccess)
    return bad(PD_ERR_DEVICE, "hipEventCreate failed");
  pd_status st;
  if ((st = make_twiddles(h, g.n0, &h->tw_x)) != PD_OK) return bad(st, pd_detail::last_error());
  if ((st = make_twiddles(h, g.n1, &h->tw_y)) != PD_OK) return bad(st, pd_detail::last_error());
  if ((st = make_twiddles(h, g.m2, &h->tw_z)) != PD_OK) return bad(st, pd_detail::last_error());
  if ((st = make_twiddles(h, g.n2, &h->tw2)) != PD_OK) return bad(st, pd_detail::last_error());
  if (!make_plan(g.n0, h->hx) || !make_plan(g.n1, h->hy) || !make_plan(g.m2, h->hz))
    return bad(PD_ERR_UNSUPPORTED, "grid edges must factor over {2,3,5,7,11,13}");
  // rows per block of the split row kernels: 32, halved while the z plan does
  // not fit the threads' registers (long odd z transforms such as 243 = 3^5)
  g.zrows = kSplitRows;
  while (g.zrows > 1 && !plan_fits(h->hz, g.zrows, kSplitThreads)) g.zrows /= 2;
  h->split = (d.flags & PD_SCAN_SPLIT_PLANES) != 0 || plane_lds(h) > kMaxLds ||
             !plan_fits(h->hy, g.h2, kPlaneThreads) || !plan_fits(h->hz, g.n1, kPlaneThreads);
  if (!plan_fits(h->hx, h->C, kColThreads) || g.n0 * h->C > PD_COL_PER * kColThreads)
    return bad(PD_ERR_UNSUPPORTED, "grid edge too large for the column register plan");
  if (h->split && (!plan_fits(h->hz, g.zrows, kSplitThreads) || !plan_fits(h->hy, kSplitCols, kSplitThreads) ||
                   split_row_lds(h) > kMaxLds || split_col_lds(h) > kMaxLds))
    return bad(PD_ERR_UNSUPPORTED, "grid edge too large for the split plane kernels");
  {
    std::vector<int4> all;
    all.insert(all.end(), h->hx.st.begin(), h->hx.st.end());
    all.insert(all.end(), h->hy.st.begin(), h->hy.st.end());
    all.insert(all.end(), h->hz.st.begin(), h->hz.st.end());
    if (hipMalloc(&h->d_stages, sizeof(int4) * all.size()) != hipSuccess ||
        hipMemcpy(h->d_stages, all.data(), sizeof(int4) * all.size(), hipMemcpyHostToDevice) != hipSuccess)
      return bad(PD_ERR_DEVICE, "stage table upload failed");
    const int nx = (int)h->hx.st.size(), ny = (int)h->hy.st.size(), nz = (int)h->hz.st.size();
    h->px = FftPlan{g.n0, nx, h->d_stages, h->tw_x};
    h->py = FftPlan{g.n1, ny, h->d_stages + nx, h->tw_y};
    h->pz = FftPlan{g.m2, nz, h->d_stages + nx + ny, h->tw_z};
  }
  // A whole-plane grid whose plane takes more than half of LDS (one block per
  // CU) runs the ligand pass as the split row and column kernels when it has
  // no specialised whole-plane kernel: they beat the generic deposit + FFT
  // kernel (160^3: 32.9 -> 15.4 us per rotation).  The specialised kernels,
  // which sum each row's z spectrum directly, beat the split kernels at 144^3
  // (10.2 against 15.3 us) and 160^3 (8.2 against 16.4 us; pipeline 24.1k ->
  // 30.2k rotations/s) and run there.  The inverse keeps the whole-plane kernel.
  h->rows_ok = !h->split && plan_fits(h->hz, g.zrows, kSplitThreads) && plan_fits(h->hy, kSplitCols, kSplitThreads) &&
               split_row_lds(h) <= kMaxLds && split_col_lds(h) <= kMaxLds;
  h->rows_auto = h->rows_ok && 2 * plane_lds(h) > kMaxLds && !find_fast(g.n0, g.n1, g.n2);
  h->lig_rows = h->rows_auto;
  if (h->rows_ok) {
    if ((st = set_lds((const void*)k_lig_rows, split_row_lds(h))) != PD_OK ||
        (st = set_lds((const void*)k_cols_fft<-1>, split_col_lds(h))) != PD_OK)
      return bad(st, pd_detail::last_error());
    h->fycol = find_fast_ycols(g.n1);
    if (h->fycol && (st = set_lds((const void*)h->fycol->fwd, h->fycol->lds)) != PD_OK)
      return bad(st, pd_detail::last_error());
  }
  if (h->split) {
    const void* rows[] = {(const void*)k_lig_rows, (const void*)k_rec_rows, (const void*)k_inv_rows_topk};
    for (const void* fn : rows)
      if ((st = set_lds(fn, split_row_lds(h))) != PD_OK) return bad(st, pd_detail::last_error());
    const void* cols[] = {(const void*)k_cols_fft<-1>, (const void*)k_cols_fft<+1>};
    for (const void* fn : cols)
      if ((st = set_lds(fn, split_col_lds(h))) != PD_OK) return bad(st, pd_detail::last_error());
  } else {
    if ((st = set_lds((const void*)k_lig_plane_fwd, plane_lds(h))) != PD_OK) return bad(st, pd_detail::last_error());
    if ((st = set_lds((const void*)k_plane_inv_topk, plane_lds(h))) != PD_OK) return bad(st, pd_detail::last_error());
    if ((st = set_lds((const void*)k_rec_plane_fwd, plane_lds(h))) != PD_OK) return bad(st, pd_detail::last_error());
  }
  if ((st = set_lds((const void*)k_col, col_lds(h))) != PD_OK) return bad(st, pd_detail::last_error());
  if ((st = set_lds((const void*)k_col_fwd, col_lds(h))) != PD_OK) return bad(st, pd_detail::last_error());
  if ((st = set_lds((const void*)k_merge, g.n0 * h->K * 8)) != PD_OK) return bad(st, pd_detail::last_error());
  if ((st = set_lds((const void*)k_merge_wave, g.n0 * h->K * 8)) != PD_OK) return bad(st, pd_detail::last_error());
  h->fast = h->split ? nullptr : find_fast(g.n0, g.n1, g.n2);
  h->fcol = find_fast_col(g.n0);
  if (h->fcol) {
    // the lane-order receptor spectra are addressed with 32-bit offsets
    const long long ntile = (g.M + h->fcol->col_lines - 1) / h->fcol->col_lines;
    if ((long long)g.T * ntile * 4 * h->fcol->ni * 1024 > 0xFFFFFFFFLL) h->fcol = nullptr;
  }
  if (h->split) {
    h->frows = find_fast_rows(g.n2);
    if (h->frows && (st = set_lds((const void*)h->frows->fn, h->frows->lds)) != PD_OK)
      return bad(st, pd_detail::last_error());
    h->fycol = find_fast_ycols(g.n1);
    if (h->fycol && ((st = set_lds((const void*)h->fycol->fwd, h->fycol->lds)) != PD_OK ||
                     (st = set_lds((const void*)h->fycol->inv, h->fycol->lds)) != PD_OK))
      return bad(st, pd_detail::last_error());
    h->inv_fused = g.n1 == kInvSplitN && g.n2 == kInvSplitN && !(h->d.flags & PD_SCAN_SPLIT_INVERSE);
    if (h->inv_fused && (st = set_lds((const void*)kInvSplit, kInvSplitLds)) != PD_OK)
      return bad(st, pd_detail::last_error());
    h->lig_fused = g.n1 == kInvSplitN && g.n2 == kInvSplitN && !(h->d.flags & PD_SCAN_SPLIT_FORWARD);
    if (h->lig_fused && (st = set_lds((const void*)kLigSplit, kInvSplitLds)) != PD_OK)
      return bad(st, pd_detail::last_error());
  }
  if (h->fast) {
    if ((st = set_lds((const void*)h->fast->lig, plane_lds(h))) != PD_OK) return bad(st, pd_detail::last_error());
    if ((st = set_lds((const void*)h->fast->inv, plane_lds(h))) != PD_OK) return bad(st, pd_detail::last_error());
  }
  if (hipMalloc(&h->R64, sizeof(double) * h->N * g.T) != hipSuccess ||
      hipMalloc(&h->Rhat, sizeof(float2) * (size_t)g.T * g.n0 * g.M) != hipSuccess ||
      hipMalloc(&h->err, sizeof(int)) != hipSuccess)
    return bad(PD_ERR_DEVICE, "device allocation failed");
  if (h->fcol) {
    const size_t ntile = (g.M + h->fcol->col_lines - 1) / h->fcol->col_lines;
    if (hipMalloc(&h->Rlanes, sizeof(float4) * g.T * ntile * 4 * h->fcol->ni * 64) != hipSuccess)
      return bad(PD_ERR_DEVICE, "device allocation failed");
  }
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(1), 0, h->stream, h->err, 1, 0);
  if (hipStreamSynchronize(h->stream) != hipSuccess) return bad(PD_ERR_DEVICE, "stream sync failed");
  *out = h;
  return PD_OK;
}

pd_status pd_scan_destroy(pd_scan* h) {
  if (!h) return PD_OK;
  if (h->graph) return fail(PD_ERR_STATE, "the scan belongs to a graph: destroy the graph first");
  if (h->stream) hipStreamSynchronize(h->stream);
  for (auto& ln : h->lane)
    if (ln.s) hipStreamSynchronize(ln.s);
  for (auto& p : h->pending) { hipEventDestroy(p.a); hipEventDestroy(p.b); }
  free_batch(h);
  hipFree(h->tw_x); hipFree(h->tw_y); hipFree(h->tw_z); hipFree(h->tw2); hipFree(h->d_stages);
  hipFree(h->R64); hipFree(h->Rhat); hipFree(h->Rlanes);
  hipFree(h->q); hipFree(h->w64); hipFree(h->w32);
  hipFree(h->err); hipFree(h->d_rots); hipFree(h->d_peaks); hipFree(h->d_nout);
  for (auto& ln : h->lane) {
    if (ln.s) hipStreamDestroy(ln.s);
    if (ln.done) hipEventDestroy(ln.done);
  }
  if (h->fork) hipEventDestroy(h->fork);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return PD_OK;
}

pd_status pd_scan_project_receptor(pd_scan* h, int n_atoms, const double* pos, const pd_term* terms) {
  if (!h || !pos || !terms || n_atoms < 1) return fail(PD_ERR_INVALID_ARGUMENT, "bad receptor arguments");
  PD_HIP(hipSetDevice(h->d.device));
  const ScanGeom& g = h->g;
  double* dpos = nullptr;
  int* dreg = nullptr;
  double* dpar = nullptr;
  PD_HIP(hipMalloc(&dpos, sizeof(double) * 3 * n_atoms));
  PD_HIP(hipMalloc(&dreg, sizeof(int) * 6 * n_atoms));
  PD_HIP(hipMalloc(&dpar, sizeof(double) * 5 * n_atoms));
  PD_HIP(hipMemcpy(dpos, pos, sizeof(double) * 3 * n_atoms, hipMemcpyHostToDevice));
  std::vector<int> reg(6 * (size_t)n_atoms);
  const int nb = ((g.n0 + 3) / 4) * ((g.n1 + 7) / 8) * ((g.n2 + 7) / 8);
  pd_status st = PD_OK;
  for (int t = 0; t < g.T && st == PD_OK; ++t) {
    const pd_term& tm = terms[t];
    const int np = field_nparams(tm.kind);
    if (np < 0 || !tm.params || !(tm.cutoff > 0)) { st = fail(PD_ERR_INVALID_ARGUMENT, "bad term"); break; }
    const double cut = tm.cutoff;
    const double md = tm.min_dist < 0 ? g.h / 4. : tm.min_dist;
    // GridRegionProjectionHelper (Grid/project.hpp:89-98) + cropSubDomain (grid.hpp:613-620)
    for (int a = 0; a < n_atoms; ++a) {
      for (int dd = 0; dd < 3; ++dd) {
        const double c = pos[3 * a + dd];
        int lo = to_logical_h(c - cut, g.zero_h[dd], g.h);
        int hi = to_logical_h(c + cut, g.zero_h[dd], g.h);
        lo = std::max(0, lo);
        hi = std::min(h->d.n[dd] - 1, hi);
        reg[6 * a + 2 * dd] = lo;
        reg[6 * a + 2 * dd + 1] = hi;
      }
    }
    if (hipMemcpy(dreg, reg.data(), sizeof(int) * reg.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dpar, tm.params, sizeof(double) * np * (size_t)n_atoms, hipMemcpyHostToDevice) !=
            hipSuccess) {
      st = fail(PD_ERR_DEVICE, "receptor upload failed");
      break;
    }
    hipLaunchKernelGGL(k_project, dim3(nb), dim3(256), 0, h->stream, g, n_atoms, dpos, dreg, tm.kind,
                       dpar, np, cut * cut, md * md, h->R64 + (size_t)t * h->N);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess)
      st = fail(PD_ERR_DEVICE, "receptor projection failed");
  }
  hipFree(dpos);
  hipFree(dreg);
  hipFree(dpar);
  if (st != PD_OK) return st;
  return receptor_spectra(h);
}

pd_status pd_scan_set_receptor_grids(pd_scan* h, const double* grids) {
  if (!h || !grids) return fail(PD_ERR_INVALID_ARGUMENT, "null argument");
  PD_HIP(hipSetDevice(h->d.device));
  PD_HIP(hipMemcpy(h->R64, grids, sizeof(double) * h->N * h->g.T, hipMemcpyHostToDevice));
  return receptor_spectra(h);
}

pd_status pd_scan_get_receptor_grids(pd_scan* h, double* grids) {
  if (!h || !grids) return fail(PD_ERR_INVALID_ARGUMENT, "null argument");
  if (!h->receptor_ready) return fail(PD_ERR_STATE, "receptor not set");
  PD_HIP(hipSetDevice(h->d.device));
  PD_HIP(hipMemcpy(grids, h->R64, sizeof(double) * h->N * h->g.T, hipMemcpyDeviceToHost));
  return PD_OK;
}

pd_status pd_scan_set_ligand(pd_scan* h, int n_atoms, const double* pos, const double* weights) {
  if (!h || !pos || !weights || n_atoms < 1) return fail(PD_ERR_INVALID_ARGUMENT, "bad ligand arguments");
  if (h->graph) return fail(PD_ERR_STATE, "the scan belongs to a graph: its buffers cannot change");
  PD_HIP(hipSetDevice(h->d.device));
  ScanGeom& g = h->g;
  // Rotation-invariant box of the ligand: every rotated atom lies within
  // r_max of the origin, so its cells lie in [cell(-r), cell(r)] per axis.
  double rmax = 0;
  for (int j = 0; j < n_atoms; ++j) {
    const double* x = pos + 3 * j;
    double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    if (!(r2 == r2) || std::isinf(r2)) return fail(PD_ERR_INVALID_ARGUMENT, "non-finite ligand coordinate");
    rmax = std::max(rmax, std::sqrt(r2));
  }
  const double r = rmax * (1.0 + 1e-12) + 1e-9;
  for (int a = 0; a < 3; ++a) {
    int lo = to_logical_h(-r, g.zero_h[a], g.h), hi = to_logical_h(r, g.zero_h[a], g.h);
    lo = std::max(0, lo);
    hi = std::min(h->d.n[a] - 1, hi);
    if (hi < lo) return fail(PD_ERR_DOMAIN, "ligand does not overlap the grid");
    g.lo[a] = lo;
    g.p[a] = hi - lo + 1;
  }
  // L1 planes: the x box from its first plane, rounded up to whole groups of
  // the column kernel's lanes; the receptor spectra follow the shift.
  const int al = h->fcol ? h->fcol->lanes : 1;
  g.pl = g.lo[0];
  g.pp = std::min(g.n0, (g.p[0] + al - 1) / al * al);
  // Interleaved L1 order for the fast column pass, where one kernel writes
  // each ligand plane (the whole-plane ligand kernel, or the one-kernel
  // ligand pass of the split path; the row and column kernels transform L1
  // planes in place): needs whole 64/B column blocks and a group that its
  // 32-bit offsets cover.
  g.il = 0;
  g.ilsh = 0;
  const bool one_writer =
      h->split ? h->lig_fused && g.p[1] <= kPlaneSplitRows : !h->lig_rows;
  if (h->fcol && one_writer && g.M % (64 / al) == 0 && (long long)g.pp * g.M * 8 <= 0xFFFFFFFFLL) {
    g.il = al;
    while ((1 << g.ilsh) < 64 / al) ++g.ilsh;
  }
  if (!h->split && !plan_fits(h->hz, g.p[1], kPlaneThreads))
    return fail(PD_ERR_UNSUPPORTED, "ligand box too large for the register plan");
  hipFree(h->q); hipFree(h->w64); hipFree(h->w32);
  h->q = nullptr; h->w64 = nullptr; h->w32 = nullptr;
  free_batch(h);
  h->ligand_ready = false;
  const size_t nw = (size_t)g.T * n_atoms;
  std::vector<float> w32(nw);
  for (size_t i = 0; i < nw; ++i) w32[i] = (float)weights[i];
  PD_HIP(hipMalloc(&h->q, sizeof(double) * 3 * n_atoms));
  PD_HIP(hipMalloc(&h->w64, sizeof(double) * nw));
  PD_HIP(hipMalloc(&h->w32, sizeof(float) * nw));
  PD_HIP(hipMemcpy(h->q, pos, sizeof(double) * 3 * n_atoms, hipMemcpyHostToDevice));
  PD_HIP(hipMemcpy(h->w64, weights, sizeof(double) * nw, hipMemcpyHostToDevice));
  PD_HIP(hipMemcpy(h->w32, w32.data(), sizeof(float) * nw, hipMemcpyHostToDevice));
  h->n_lig = n_atoms;
  if (h->lig_pos_h.data() != pos) {
    h->lig_pos_h.assign(pos, pos + 3 * (size_t)n_atoms);
    h->lig_w_h.assign(weights, weights + nw);
  }
  pd_status st = alloc_batch(h);
  if (st != PD_OK) return st;
  if (h->receptor_ready && h->rhat_shift != g.pl) {
    h->receptor_ready = false;
    if ((st = receptor_spectra(h)) != PD_OK) return st;
  }
  h->ligand_ready = true;
  return PD_OK;
}

pd_status pd_scan_set_batch(pd_scan* h, int batch) {
  if (!h) return fail(PD_ERR_INVALID_ARGUMENT, "null handle");
  if (batch < 1 || batch > 4096) return fail(PD_ERR_INVALID_ARGUMENT, "batch must be 1..4096");
  if (h->graph) return fail(PD_ERR_STATE, "the scan belongs to a graph: its buffers cannot change");
  PD_HIP(hipSetDevice(h->d.device));
  PD_HIP(hipStreamSynchronize(h->stream));
  for (auto& ln : h->lane) PD_HIP(hipStreamSynchronize(ln.s));
  if (batch == h->B) return PD_OK;
  free_batch(h);
  h->B = batch;
  if (h->ligand_ready) {
    pd_status st = alloc_batch(h);
    if (st != PD_OK) {
      h->ligand_ready = false;
      return st;
    }
  }
  return PD_OK;
}

pd_status pd_scan_set_tuning(pd_scan* h, const pd_tuning* t) {
  if (!h) return fail(PD_ERR_INVALID_ARGUMENT, "null handle");
  if (h->graph) return fail(PD_ERR_STATE, "the scan belongs to a graph: its launches are fixed");
  const pd_tuning z{};
  const pd_tuning& u = t ? *t : z;
  if (u.col_groups < 0 || u.col_mz_extra < 0 || u.col_mz_extra > 15 || u.lig_pass < 0 || u.lig_pass > 2 ||
      u.lanes < 0 || u.lanes > pd_scan::kLanes)
    return fail(PD_ERR_INVALID_ARGUMENT, "tuning value out of range");
  if (u.lig_pass == 2 && !h->rows_ok)
    return fail(PD_ERR_UNSUPPORTED, "this grid has no row and column ligand pass");
  if (u.lig_pass == 1 && h->split) return fail(PD_ERR_UNSUPPORTED, "this grid has no whole-plane ligand pass");
  PD_HIP(hipSetDevice(h->d.device));
  PD_HIP(hipStreamSynchronize(h->stream));
  for (auto& ln : h->lane) PD_HIP(hipStreamSynchronize(ln.s));
  const pd_tuning old = h->tune;
  const bool old_rows = h->lig_rows;
  h->tune = u;
  const bool rows = u.lig_pass == 0 ? h->rows_auto : u.lig_pass == 2;
  if (!h->split && rows != h->lig_rows) {
    h->lig_rows = rows;
    // the ligand's L1 layout depends on the ligand pass: lay it out again
    if (h->ligand_ready) {
      const std::vector<double> pos = h->lig_pos_h, w = h->lig_w_h;
      const int n = (int)(pos.size() / 3);
      pd_status st = pd_scan_set_ligand(h, n, pos.data(), w.data());
      if (st != PD_OK) {
        // keep the handle usable: the previous choice and its layout
        const std::string msg = pd_detail::last_error();
        h->tune = old;
        h->lig_rows = old_rows;
        pd_scan_set_ligand(h, n, pos.data(), w.data());
        return fail(st, msg);
      }
    }
  }
  return PD_OK;
}

pd_status pd_scan_get_tuning(pd_scan* h, int batch, pd_tuning* t) {
  if (!h || !t || batch < 1) return fail(PD_ERR_INVALID_ARGUMENT, "bad tuning query");
  if (!h->ligand_ready) return fail(PD_ERR_STATE, "the launch shape follows the ligand box: set the ligand first");
  const ColShape cs = col_shape(h, batch);
  t->col_groups = cs.nyb;
  t->col_mz_extra = h->tune.col_mz_extra;
  t->lig_pass = h->split ? 0 : (h->lig_rows ? 2 : 1);
  t->lanes = lanes_for(h, 2);
  return PD_OK;
}

pd_status pd_scan_set_postfilter(pd_scan* h, const pd_postfilter* pf) {
  if (!h) return fail(PD_ERR_INVALID_ARGUMENT, "null handle");
  if (h->graph) return fail(PD_ERR_STATE, "the scan belongs to a graph: its buffers cannot change");
  const bool symm = pf && pf->n_multimer > 1, clust = pf && pf->cluster != 0;
  if (!symm && !clust) {
    h->post = false;
    return PD_OK;
  }
  if (h->pg.k_inp > kCandMax)
    return fail(PD_ERR_UNSUPPORTED, "post-filters take max_n_trans_inp <= " + std::to_string(kCandMax));
  if (symm && !(pf->ligand_radius >= 0 && pf->max_rmsd_symm == pf->max_rmsd_symm))
    return fail(PD_ERR_INVALID_ARGUMENT, "bad symmetry parameters");
  if (clust && !(pf->cluster_radius > 0)) return fail(PD_ERR_INVALID_ARGUMENT, "cluster_radius must be > 0");
  PD_HIP(hipSetDevice(h->d.device));
  PD_HIP(hipStreamSynchronize(h->stream));
  for (auto& ln : h->lane) PD_HIP(hipStreamSynchronize(ln.s));
  pd_status st = ensure_grids(h);
  if (st != PD_OK) return st;
  PostGeom& pg = h->pg;
  pg.n_symm = symm ? pf->n_multimer : 0;
  pg.cluster = clust ? 1 : 0;
  pg.mol_radius = symm ? pf->ligand_radius : 0;
  pg.max_rmsd = symm ? pf->max_rmsd_symm : 0;
  // the neighbour test of ClusterMatrix: dist2 <= cutoff * cutoff with
  // cutoff = sqrt(3 r^2) (Geom/cluster.hpp:160-164, partpoints.hpp:551-584)
  const double r2 = clust ? pf->cluster_radius * pf->cluster_radius : 0;
  const double cutoff = std::sqrt(r2 * 3);
  pg.clust_rad2 = cutoff * cutoff;
  for (int i = 0; i < 12; ++i) {
    pg.fa[i] = symm ? pf->frame_rec_inv[i] : (i % 4 == 0 && i < 9 ? 1 : 0);
    pg.fb[i] = symm ? pf->frame_lig[i] : (i % 4 == 0 && i < 9 ? 1 : 0);
  }
  h->post = true;
  return PD_OK;
}

pd_status pd_scan_candidates(pd_scan* h, int n_rot, const double* rots, int rot_index0, pd_peak* peaks,
                             int* n_out) {
  if (!h || n_rot < 0 || (n_rot > 0 && (!rots || !peaks || !n_out)))
    return fail(PD_ERR_INVALID_ARGUMENT, "null argument");
  if (h->pg.k_inp > kCandMax)
    return fail(PD_ERR_UNSUPPORTED, "candidate lists take max_n_trans_inp <= " + std::to_string(kCandMax));
  if (h->graph) return fail(PD_ERR_STATE, "the scan belongs to a graph: its buffers cannot change");
  if (n_rot == 0) return PD_OK;
  PD_HIP(hipSetDevice(h->d.device));
  pd_status st = ensure_grids(h);
  if (st != PD_OK) return st;
  st = ensure_io(h, n_rot);
  if (st != PD_OK) return st;
  CallBuf<PeakOut> d_cand;
  const size_t nc = (size_t)n_rot * h->pg.k_inp;
  PD_HIP(hipMalloc(&d_cand.p, sizeof(PeakOut) * nc));
  PD_HIP(hipMemcpyAsync(h->d_rots, rots, sizeof(double) * 9 * (size_t)n_rot, hipMemcpyHostToDevice, h->stream));
  st = run_batches(h, n_rot, h->d_rots, rot_index0, d_cand.p, h->d_nout, BatchOut::kCandidates);
  if (st == PD_OK) {
    if (hipMemcpyAsync(peaks, d_cand.p, sizeof(PeakOut) * nc, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipMemcpyAsync(n_out, h->d_nout, sizeof(int) * (size_t)n_rot, hipMemcpyDeviceToHost, h->stream) !=
            hipSuccess)
      st = fail(PD_ERR_DEVICE, "candidate copy failed");
    else
      st = pd_scan_sync(h);
  }
  hipStreamSynchronize(h->stream);
  return st;
}

pd_status pd_scan_run_device(pd_scan* h, int n_rot, const double* d_rots, int rot_index0,
                             pd_peak* d_peaks, int* d_n_out, int sync) {
  if (!h || (n_rot > 0 && (!d_rots || !d_peaks || !d_n_out)))
    return fail(PD_ERR_INVALID_ARGUMENT, "null argument");
  PD_HIP(hipSetDevice(h->d.device));
  pd_status st = run_batches(h, n_rot, d_rots, rot_index0, reinterpret_cast<PeakOut*>(d_peaks), d_n_out);
  if (st != PD_OK) return st;
  if (sync) return pd_scan_sync(h);
  return PD_OK;
}

pd_status pd_scan_sync(pd_scan* h) {
  if (!h) return fail(PD_ERR_INVALID_ARGUMENT, "null handle");
  PD_HIP(hipSetDevice(h->d.device));
  pd_status st = resolve_timing(h);
  if (st != PD_OK) return st;
  return check_err(h);
}

pd_status pd_scan_run(pd_scan* h, int n_rot, const double* rots, int rot_index0, pd_peak* peaks,
                      int* n_out) {
  if (!h || (n_rot > 0 && (!rots || !peaks || !n_out))) return fail(PD_ERR_INVALID_ARGUMENT, "null argument");
  if (n_rot == 0) return PD_OK;
  PD_HIP(hipSetDevice(h->d.device));
  pd_status st = ensure_io(h, n_rot);
  if (st != PD_OK) return st;
  PD_HIP(hipMemcpyAsync(h->d_rots, rots, sizeof(double) * 9 * (size_t)n_rot, hipMemcpyHostToDevice,
                        h->stream));
  st = run_batches(h, n_rot, h->d_rots, rot_index0, h->d_peaks, h->d_nout);
  if (st != PD_OK) return st;
  PD_HIP(hipMemcpyAsync(peaks, h->d_peaks, sizeof(PeakOut) * (size_t)n_rot * h->K,
                        hipMemcpyDeviceToHost, h->stream));
  PD_HIP(hipMemcpyAsync(n_out, h->d_nout, sizeof(int) * (size_t)n_rot, hipMemcpyDeviceToHost, h->stream));
  return pd_scan_sync(h);
}

pd_status pd_scan_rescore64(pd_scan* h, int n, const double* rots, const int* s, double* energies) {
  if (!h || (n > 0 && (!rots || !s || !energies))) return fail(PD_ERR_INVALID_ARGUMENT, "null argument");
  if (!h->receptor_ready || !h->ligand_ready) return fail(PD_ERR_STATE, "receptor or ligand not set");
  if (n == 0) return PD_OK;
  PD_HIP(hipSetDevice(h->d.device));
  CallBuf<double> dr, de;
  CallBuf<int> ds;
  PD_HIP(hipMalloc(&dr.p, sizeof(double) * 9 * (size_t)n));
  PD_HIP(hipMalloc(&ds.p, sizeof(int) * 3 * (size_t)n));
  PD_HIP(hipMalloc(&de.p, sizeof(double) * (size_t)n));
  PD_HIP(hipMemcpy(dr.p, rots, sizeof(double) * 9 * (size_t)n, hipMemcpyHostToDevice));
  PD_HIP(hipMemcpy(ds.p, s, sizeof(int) * 3 * (size_t)n, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_rescore, dim3(n), dim3(256), 0, h->stream, h->g, dr.p, ds.p, h->q, h->n_lig, h->w64,
                     h->R64, de.p);
  PD_HIP(hipGetLastError());
  PD_HIP(hipMemcpyAsync(energies, de.p, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, h->stream));
  PD_HIP(hipStreamSynchronize(h->stream));
  return PD_OK;
}

pd_status pd_scan_rescore64_noalloc(pd_scan* h, int n, const double* rots, const int* s, double* energies) {
  if (!h || n < 0 || (n > 0 && (!rots || !s || !energies))) return fail(PD_ERR_INVALID_ARGUMENT, "null argument");
  if (n > kNoAllocMaxPoses)
    return fail(PD_ERR_INVALID_ARGUMENT, "pd_scan_rescore64_noalloc takes at most " +
                                             std::to_string(kNoAllocMaxPoses) + " poses per call (got " +
                                             std::to_string(n) + "); use pd_scan_rescore64");
  if (!h->receptor_ready || !h->ligand_ready) return fail(PD_ERR_STATE, "receptor or ligand not set");
  if (n == 0) return PD_OK;
  const int dev = h->d.device;
  if (dev >= kMaxDevices) return fail(PD_ERR_UNSUPPORTED, "device ordinal beyond the staging table");
  if (!g_scratch_ptr[dev]) return fail(PD_ERR_DEVICE, "no staging area on this device: " + g_scratch_err[dev]);
  PD_HIP(hipSetDevice(dev));
  unsigned char* base = static_cast<unsigned char*>(g_scratch_ptr[dev]);
  double* dr = reinterpret_cast<double*>(base + kNoAllocRotOff);
  int* ds = reinterpret_cast<int*>(base + kNoAllocShiftOff);
  double* de = reinterpret_cast<double*>(base + kNoAllocEnergyOff);
  // one call at a time per device: the staging area is shared by every handle there
  std::lock_guard<std::mutex> lk(g_scratch_use[dev]);
  PD_HIP(hipMemcpyAsync(dr, rots, sizeof(double) * 9 * (size_t)n, hipMemcpyHostToDevice, h->stream));
  PD_HIP(hipMemcpyAsync(ds, s, sizeof(int) * 3 * (size_t)n, hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(k_rescore, dim3(n), dim3(256), 0, h->stream, h->g, dr, ds, h->q, h->n_lig, h->w64, h->R64, de);
  PD_HIP(hipGetLastError());
  PD_HIP(hipMemcpyAsync(energies, de, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, h->stream));
  PD_HIP(hipStreamSynchronize(h->stream));
  return PD_OK;
}

pd_status pd_scan_correlation_grid(pd_scan* h, const double* rot, float* out) {
  if (!h || !rot || !out) return fail(PD_ERR_INVALID_ARGUMENT, "null argument");
  if (!h->receptor_ready || !h->ligand_ready) return fail(PD_ERR_STATE, "receptor or ligand not set");
  PD_HIP(hipSetDevice(h->d.device));
  pd_status st = ensure_io(h, 1);
  if (st != PD_OK) return st;
  CallBuf<float> dgrid;
  PD_HIP(hipMalloc(&dgrid.p, sizeof(float) * h->N));
  PD_HIP(hipMemcpy(h->d_rots, rot, sizeof(double) * 9, hipMemcpyHostToDevice));
  pd_scan::Lane ln = h->lane[0];
  ln.s = h->stream;
  Launcher L;
  L.s = h->stream;
  st = launch_fft_path(h, ln, 1, h->d_rots, dgrid.p, L);
  if (st != PD_OK) return st;
  PD_HIP(hipMemcpyAsync(out, dgrid.p, sizeof(float) * h->N, hipMemcpyDeviceToHost, h->stream));
  PD_HIP(hipStreamSynchronize(h->stream));
  return check_err(h);
}

pd_status pd_scan_set_timing(pd_scan* h, int enable) {
  if (!h) return fail(PD_ERR_INVALID_ARGUMENT, "null handle");
  pd_status st = resolve_timing(h);
  if (st != PD_OK) return st;
  h->timing = enable != 0;
  for (int k = 0; k < 5; ++k) { h->tot_ms[k] = 0; h->launches[k] = 0; }
  return PD_OK;
}

pd_status pd_scan_kernel_stats(pd_scan* h, int k, double* total_ms, long long* launches) {
  if (!h || k < 0 || k >= 5 || !total_ms || !launches) return fail(PD_ERR_INVALID_ARGUMENT, "bad argument");
  pd_status st = resolve_timing(h);
  if (st != PD_OK) return st;
  *total_ms = h->tot_ms[k];
  *launches = h->launches[k];
  return PD_OK;
}

void* pd_scan_stream(pd_scan* h) { return h ? (void*)h->stream : nullptr; }

pd_status pd_scan_ligand_box(pd_scan* h, int* lo, int* p, int* planes) {
  if (!h || !lo || !p || !planes) return fail(PD_ERR_INVALID_ARGUMENT, "null argument");
  if (!h->ligand_ready) return fail(PD_ERR_STATE, "ligand not set");
  for (int a = 0; a < 3; ++a) {
    lo[a] = h->g.lo[a];
    p[a] = h->g.p[a];
  }
  *planes = h->g.pp;
  return PD_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Several scans in one graph.  The graph is built node by node (no stream
// capture): each scan's batches become kernel nodes with explicit edges, as in
// graph_batches; pairs on one graph lane follow each other, lanes are
// independent branches from one root node.
// ---------------------------------------------------------------------------
struct pd_graph {
  int device = 0;
  hipStream_t origin = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<pd_scan*> scans;
};

namespace {
void graph_free(pd_graph* g) {
  if (!g) return;
  for (pd_scan* h : g->scans)
    if (h && h->graph == g) h->graph = nullptr;
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  if (g->origin) hipStreamDestroy(g->origin);
  delete g;
}
}  // namespace

extern "C" {

pd_status pd_graph_create(int n, pd_scan* const* scans, const int* n_rot, const double* const* d_rots,
                          pd_peak* const* d_peaks, int* const* d_n_out, const int* lane, int n_lanes,
                          pd_graph** out) {
  if (!out || n < 1 || !scans || !n_rot || !d_rots || !d_peaks || !d_n_out || !lane || n_lanes < 1)
    return fail(PD_ERR_INVALID_ARGUMENT, "bad graph arguments");
  *out = nullptr;
  for (int i = 0; i < n; ++i)
    if (!scans[i]) return fail(PD_ERR_INVALID_ARGUMENT, "null scan handle");
  const int dev = scans[0]->d.device;
  for (int i = 0; i < n; ++i) {
    pd_scan* h = scans[i];
    if (h->d.device != dev) return fail(PD_ERR_INVALID_ARGUMENT, "scans must share one device");
    if (h->graph) return fail(PD_ERR_STATE, "a scan already belongs to a graph");
    if (!h->receptor_ready || !h->ligand_ready) return fail(PD_ERR_STATE, "a scan has no receptor or ligand");
    if (h->timing) return fail(PD_ERR_STATE, "kernel timing must be off while a graph holds the scan");
    if (lane[i] < 0 || lane[i] >= n_lanes) return fail(PD_ERR_INVALID_ARGUMENT, "lane out of range");
    if (n_rot[i] < 1 || !d_rots[i] || !d_peaks[i] || !d_n_out[i]) return fail(PD_ERR_INVALID_ARGUMENT, "bad pair");
    for (int j = 0; j < i; ++j)
      if (scans[j] == h) return fail(PD_ERR_INVALID_ARGUMENT, "a scan appears twice");
  }
  PD_HIP(hipSetDevice(dev));
  pd_graph* g = new pd_graph();
  g->device = dev;
  auto bad = [&](pd_status s) {
    graph_free(g);
    return s;
  };
  if (hipStreamCreateWithFlags(&g->origin, hipStreamNonBlocking) != hipSuccess)
    return bad(fail(PD_ERR_DEVICE, "hipStreamCreate failed"));
  for (int i = 0; i < n; ++i)
    if (hipStreamSynchronize(scans[i]->stream) != hipSuccess) return bad(fail(PD_ERR_DEVICE, "sync failed"));
  if (hipGraphCreate(&g->graph, 0) != hipSuccess) return bad(fail(PD_ERR_DEVICE, "hipGraphCreate failed"));
  hipGraphNode_t root = nullptr;
  if (hipGraphAddEmptyNode(&root, g->graph, nullptr, 0) != hipSuccess)
    return bad(fail(PD_ERR_DEVICE, "hipGraphAddEmptyNode failed"));
  std::vector<std::vector<hipGraphNode_t>> lane_tail(n_lanes, std::vector<hipGraphNode_t>(1, root));
  for (int i = 0; i < n; ++i) {
    std::vector<hipGraphNode_t> tail;
    pd_status st = graph_batches(scans[i], g->graph, lane_tail[lane[i]], n_rot[i], d_rots[i],
                                 reinterpret_cast<PeakOut*>(d_peaks[i]), d_n_out[i], tail);
    if (st != PD_OK) return bad(st);
    lane_tail[lane[i]] = tail;
  }
  if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
    return bad(fail(PD_ERR_DEVICE, "hipGraphInstantiate failed"));
  g->scans.assign(scans, scans + n);
  for (pd_scan* h : g->scans) h->graph = g;
  *out = g;
  return PD_OK;
}

pd_status pd_graph_launch(pd_graph* g) {
  if (!g) return fail(PD_ERR_INVALID_ARGUMENT, "null graph");
  PD_HIP(hipSetDevice(g->device));
  PD_HIP(hipGraphLaunch(g->exec, g->origin));
  return PD_OK;
}

pd_status pd_graph_sync(pd_graph* g) {
  if (!g) return fail(PD_ERR_INVALID_ARGUMENT, "null graph");
  PD_HIP(hipSetDevice(g->device));
  PD_HIP(hipStreamSynchronize(g->origin));
  for (pd_scan* h : g->scans) {
    pd_status st = check_err(h);
    if (st != PD_OK) return st;
  }
  return PD_OK;
}

void* pd_graph_stream(pd_graph* g) { return g ? (void*)g->origin : nullptr; }

void pd_graph_destroy(pd_graph* g) {
  if (!g) return;
  hipSetDevice(g->device);
  if (g->origin) hipStreamSynchronize(g->origin);
  graph_free(g);
}

}  // extern "C"
