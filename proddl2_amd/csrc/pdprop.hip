// pdprop.hip -- host side of the learned rotation proposal (include/pdprop.h).
//
// One handle holds the network (FP32 master weights, Adam state), the maps of
// the current complex, and activation buffers for up to max_rot rotations.
// Every call enqueues kernels (prop_kernels.hpp) on the handle's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pdprop.h"
#include "prop_kernels.hpp"

namespace pd_detail {
std::string& last_error();
}

using namespace prop;

namespace {

pd_status fail(pd_status s, const std::string& msg) {
  pd_detail::last_error() = msg;
  return s;
}

#define PP_HIP(x)                                                                       \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess)                                                               \
      return fail(PD_ERR_DEVICE, std::string(#x) + " failed: " + hipGetErrorString(e_)); \
  } while (0)

#define PP_LAUNCH(...)                 \
  do {                                 \
    hipLaunchKernelGGL(__VA_ARGS__);   \
    PP_HIP(hipGetLastError());         \
  } while (0)

constexpr int kPattBlocks = 256;
constexpr int kRfLds = C * NG * (int)sizeof(float);  // 128 KiB
constexpr int kRfBwdChunk = 32;                       // rotations per block in k_rf_bwd_rec

int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

// splitmix64 -> uniform [0, 1)
struct Rng {
  uint64_t s;
  double next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  }
};

}  // namespace

struct pd_prop {
  int device = 0;
  int max_rot = 0;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  bool have_complex = false;
  // network
  float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
  long long step = 0;
  // maps and conv activations, [molecule][...]: 0 receptor, 1 ligand
  float* P[2] = {nullptr, nullptr};   // CIN x NG
  float* h1[2] = {nullptr, nullptr};  // C x NG (after ReLU)
  float* h2[2] = {nullptr, nullptr};
  float* F[2] = {nullptr, nullptr};   // C x NG (conv3 + input)
  float* dF[2] = {nullptr, nullptr};
  float* dy2[2] = {nullptr, nullptr};
  float* dy1[2] = {nullptr, nullptr};
  float* patt_part = nullptr;         // kPattBlocks x CIN x NG
  float* surf = nullptr;
  double* atoms = nullptr;
  int atoms_cap = 0;
  // rotation functions
  int nv = 0;
  Shells sh{};
  int* vlist = nullptr;
  float *FRv = nullptr, *dFRv = nullptr, *rf_part = nullptr, *lig_part = nullptr, *dFL_cell = nullptr;
  int rf_part_blocks = 0, lig_part_blocks = 0;
  // per-rotation activations
  float *rf = nullptr, *X0 = nullptr, *H1 = nullptr, *H2 = nullptr, *out = nullptr;
  float *y = nullptr, *energy = nullptr, *dout = nullptr, *dZ2 = nullptr, *dZ1 = nullptr, *dX0 = nullptr, *drf = nullptr;
  float* stats = nullptr;   // 2 x NF (features), then 2 (target)
  float* loss = nullptr;
  double* rots = nullptr;   // staging for host-side calls
};

namespace {

pd_status alloc(float** p, long long n) {
  PP_HIP(hipMalloc(p, sizeof(float) * (size_t)n));
  return PD_OK;
}

#define PP_TRY(x)                        \
  do {                                   \
    pd_status s_ = (x);                  \
    if (s_ != PD_OK) return s_;          \
  } while (0)

void init_params(std::vector<float>& p, unsigned seed) {
  // torch's default: weights and biases uniform in +-1/sqrt(fan_in); the last
  // convolution of each stack starts at zero so that the stack starts as the
  // identity on the input maps.
  Rng rng{0x5DEECE66Dull ^ ((uint64_t)seed << 17)};
  p.assign(NPARAM, 0.0f);
  auto fill = [&](int off, int n, int fan_in) {
    const double b = 1.0 / std::sqrt((double)fan_in);
    for (int i = 0; i < n; ++i) p[off + i] = (float)((2.0 * rng.next() - 1.0) * b);
  };
  for (int m = 0; m < 2; ++m) {
    const int o = m * CONV_STACK;
    fill(o, CW1, CIN * 27);
    fill(o + CW1, CB1, CIN * 27);
    fill(o + CW1 + CB1, CW2, C * 27);
    fill(o + CW1 + CB1 + CW2, CB2, C * 27);
    // w3, b3 stay 0
  }
  fill(OFF_W1, HID * NIN, NIN);
  fill(OFF_B1, HID, NIN);
  fill(OFF_W2, HID * HID, HID);
  fill(OFF_B2, HID, HID);
  fill(OFF_W3, HID, HID);
  fill(OFF_B3, 1, HID);
}

struct StackPtr {
  const float *w1, *b1, *w2, *b2, *w3, *b3;
};
struct StackGrad {
  float *w1, *b1, *w2, *b2, *w3, *b3;
};
StackPtr stack_params(const float* base, int m) {
  const float* o = base + m * CONV_STACK;
  return {o, o + CW1, o + CW1 + CB1, o + CW1 + CB1 + CW2, o + CW1 + CB1 + CW2 + CB2,
          o + CW1 + CB1 + CW2 + CB2 + CW3};
}
StackGrad stack_grads(float* base, int m) {
  float* o = base + m * CONV_STACK;
  return {o, o + CW1, o + CW1 + CB1, o + CW1 + CB1 + CW2, o + CW1 + CB1 + CW2 + CB2, o + CW1 + CB1 + CW2 + CB2 + CW3};
}

pd_status conv_forward(pd_prop* h) {
  StackPtr s[2] = {stack_params(h->params, 0), stack_params(h->params, 1)};
  ConvFwd a[2];
  for (int m = 0; m < 2; ++m) a[m] = {h->P[m], s[m].w1, s[m].b1, nullptr, h->h1[m]};
  PP_LAUNCH(k_conv_fwd, dim3(NG / 256, C, 2), dim3(256), 0, h->stream, a[0], a[1], CIN, 1);
  for (int m = 0; m < 2; ++m) a[m] = {h->h1[m], s[m].w2, s[m].b2, nullptr, h->h2[m]};
  PP_LAUNCH(k_conv_fwd, dim3(NG / 256, C, 2), dim3(256), 0, h->stream, a[0], a[1], C, 1);
  for (int m = 0; m < 2; ++m) a[m] = {h->h2[m], s[m].w3, s[m].b3, h->P[m], h->F[m]};
  PP_LAUNCH(k_conv_fwd, dim3(NG / 256, C, 2), dim3(256), 0, h->stream, a[0], a[1], C, 0);
  return PD_OK;
}

template <int EPI>
pd_status gemm(pd_prop* h, int M, int N, int K, const float* A, long long sam, long long sak, const float* B,
               long long sbk, long long sbn, float* Cm, long long ldc, const float* bias, const float* mask,
               long long ldm) {
  PP_LAUNCH(k_gemm_bf16<EPI>, dim3(cdiv(N, GT), cdiv(M, GT)), dim3(256), 0, h->stream, M, N, K, A, sam, sak, B, sbk,
            sbn, Cm, ldc, bias, mask, ldm);
  return PD_OK;
}

// Forward from the conv stacks to the scores (h->out) for n rotations.
pd_status forward(pd_prop* h, int n, const double* d_rots) {
  PP_TRY(conv_forward(h));
  PP_LAUNCH(k_rf_pack, dim3(cdiv(h->nv * C, 256)), dim3(256), 0, h->stream, h->nv, h->vlist, h->F[0], h->FRv);
  const int blocks = std::min(256, cdiv(n, RF_WAVES));  // one block per CU
  PP_LAUNCH(k_rf_fwd, dim3(blocks), dim3(RF_THREADS), kRfLds, h->stream, n, d_rots, h->F[1], h->FRv, h->vlist, h->sh,
            h->rf);
  PP_LAUNCH(k_colstats, dim3(NF), dim3(256), 0, h->stream, n, h->rf, NF, 1e-6f, h->stats);
  PP_LAUNCH(k_build_input, dim3(cdiv((long long)n * NINP, 256)), dim3(256), 0, h->stream, n, h->rf, h->stats, d_rots,
            h->X0);
  const float* W1 = h->params + OFF_W1;
  const float* W2 = h->params + OFF_W2;
  // H1 = relu(X0 W1^T + b1): A(m=r, k=i) = X0[r][i], B(k=i, n=o) = W1[o][i]
  PP_TRY(gemm<1>(h, n, HID, NIN, h->X0, NINP, 1, W1, 1, NIN, h->H1, HID, h->params + OFF_B1, nullptr, 0));
  PP_TRY(gemm<1>(h, n, HID, HID, h->H1, HID, 1, W2, 1, HID, h->H2, HID, h->params + OFF_B2, nullptr, 0));
  PP_LAUNCH(k_out_fwd, dim3(cdiv(n, 4)), dim3(256), 0, h->stream, n, h->H2, h->params + OFF_W3, h->params + OFF_B3,
            h->out);
  return PD_OK;
}

// Loss and gradients into h->grads (targets in h->energy).
pd_status backward(pd_prop* h, int n, const double* d_rots) {
  // standardised target
  PP_LAUNCH(k_colstats, dim3(1), dim3(256), 0, h->stream, n, h->energy, 1, 0.0f, h->stats + 2 * NF);
  PP_LAUNCH(k_standardize, dim3(cdiv(n, 256)), dim3(256), 0, h->stream, n, h->energy, h->stats + 2 * NF, h->y);
  PP_LAUNCH(k_mse, dim3(1), dim3(1024), 0, h->stream, n, h->out, h->y, h->dout, h->loss);
  float* g = h->grads;
  PP_LAUNCH(k_fill, dim3(cdiv(NPARAM, 256)), dim3(256), 0, h->stream, g, (long long)NPARAM, 0.0f);
  // output layer
  PP_LAUNCH(k_colsum, dim3(cdiv(HID, 64)), dim3(256), 0, h->stream, n, HID, h->H2, HID, h->dout, g + OFF_W3);
  PP_LAUNCH(k_sum, dim3(1), dim3(1024), 0, h->stream, n, h->dout, g + OFF_B3);
  PP_LAUNCH(k_out_bwd, dim3(cdiv((long long)n * HID, 256)), dim3(256), 0, h->stream, n, h->dout, h->params + OFF_W3,
            h->H2, h->dZ2);
  // layer 2: dW2 = dZ2^T H1; db2 = colsum dZ2; dZ1 = (dZ2 W2) * (H1 > 0)
  PP_TRY(gemm<0>(h, HID, HID, n, h->dZ2, 1, HID, h->H1, HID, 1, g + OFF_W2, HID, nullptr, nullptr, 0));
  PP_LAUNCH(k_colsum, dim3(cdiv(HID, 64)), dim3(256), 0, h->stream, n, HID, h->dZ2, HID, nullptr, g + OFF_B2);
  PP_TRY(gemm<2>(h, n, HID, HID, h->dZ2, HID, 1, h->params + OFF_W2, HID, 1, h->dZ1, HID, nullptr, h->H1, HID));
  // layer 1: dW1 = dZ1^T X0 (30 columns); db1; dX0 = dZ1 W1
  PP_TRY(gemm<0>(h, HID, NIN, n, h->dZ1, 1, HID, h->X0, NINP, 1, g + OFF_W1, NIN, nullptr, nullptr, 0));
  PP_LAUNCH(k_colsum, dim3(cdiv(HID, 64)), dim3(256), 0, h->stream, n, HID, h->dZ1, HID, nullptr, g + OFF_B1);
  PP_TRY(gemm<0>(h, n, NIN, HID, h->dZ1, HID, 1, h->params + OFF_W1, NIN, 1, h->dX0, NINP, nullptr, nullptr, 0));
  // standardisation
  PP_LAUNCH(k_zscore_bwd, dim3(NF), dim3(256), 0, h->stream, n, h->rf, h->stats, 1e-6f, h->dX0, h->drf);
  // rotation functions -> dF (receptor listed vectors, ligand cells)
  const int rblocks = std::min(h->rf_part_blocks, cdiv(n, kRfBwdChunk));
  const int chunk = cdiv(n, rblocks);
  PP_LAUNCH(k_rf_bwd_rec, dim3(rblocks), dim3(RF_THREADS), kRfLds, h->stream, n, d_rots, h->F[1], h->vlist, h->nv,
            h->sh, h->drf, chunk, h->rf_part);
  PP_LAUNCH(k_sum_partials, dim3(cdiv(h->nv * C, 256)), dim3(256), 0, h->stream, rblocks, h->nv * C, h->rf_part,
            h->dFRv);
  PP_LAUNCH(k_fill, dim3(cdiv(C * NG, 256)), dim3(256), 0, h->stream, h->dF[0], (long long)C * NG, 0.0f);
  PP_LAUNCH(k_rf_unpack, dim3(cdiv(h->nv * C, 256)), dim3(256), 0, h->stream, h->nv, h->vlist, h->dFRv, h->dF[0]);
  const int lblocks = std::min(h->lig_part_blocks, cdiv(n, RF_WAVES));
  PP_LAUNCH(k_rf_bwd_lig, dim3(lblocks), dim3(RF_THREADS), kRfLds, h->stream, n, d_rots, h->FRv, h->vlist, h->sh,
            h->drf, h->lig_part);
  PP_LAUNCH(k_sum_partials, dim3(cdiv(C * NG, 256)), dim3(256), 0, h->stream, lblocks, C * NG, h->lig_part,
            h->dFL_cell);
  PP_LAUNCH(k_cell_to_channel, dim3(cdiv(C * NG, 256)), dim3(256), 0, h->stream, h->dFL_cell, h->dF[1]);
  // conv stacks (F = input + conv3(relu(conv2(relu(conv1(P))))))
  StackPtr s[2] = {stack_params(h->params, 0), stack_params(h->params, 1)};
  StackGrad gs[2] = {stack_grads(g, 0), stack_grads(g, 1)};
  ConvBwdW bw[2];
  ConvBwdData bd[2];
  for (int m = 0; m < 2; ++m) bw[m] = {h->dF[m], h->h2[m], gs[m].w3, gs[m].b3};
  PP_LAUNCH(k_conv_bwd_w, dim3(27, C, 2 * C), dim3(256), 0, h->stream, bw[0], bw[1], C, C);
  for (int m = 0; m < 2; ++m) bd[m] = {h->dF[m], s[m].w3, h->h2[m], h->dy2[m]};
  PP_LAUNCH(k_conv_bwd_data, dim3(NG / 256, C, 2), dim3(256), 0, h->stream, bd[0], bd[1], C, C);
  for (int m = 0; m < 2; ++m) bw[m] = {h->dy2[m], h->h1[m], gs[m].w2, gs[m].b2};
  PP_LAUNCH(k_conv_bwd_w, dim3(27, C, 2 * C), dim3(256), 0, h->stream, bw[0], bw[1], C, C);
  for (int m = 0; m < 2; ++m) bd[m] = {h->dy2[m], s[m].w2, h->h1[m], h->dy1[m]};
  PP_LAUNCH(k_conv_bwd_data, dim3(NG / 256, C, 2), dim3(256), 0, h->stream, bd[0], bd[1], C, C);
  for (int m = 0; m < 2; ++m) bw[m] = {h->dy1[m], h->P[m], gs[m].w1, gs[m].b1};
  PP_LAUNCH(k_conv_bwd_w, dim3(27, CIN, 2 * C), dim3(256), 0, h->stream, bw[0], bw[1], CIN, C);
  return PD_OK;
}

pd_status check_n(pd_prop* h, int n) {
  if (!h) return fail(PD_ERR_INVALID_ARGUMENT, "null handle");
  if (!h->have_complex) return fail(PD_ERR_STATE, "pd_prop_set_complex has not been called");
  if (n < 2 || n > h->max_rot) return fail(PD_ERR_INVALID_ARGUMENT, "n_rot must be 2..max_rot");
  return PD_OK;
}

pd_status build_maps(pd_prop* h, int m, int n, const double* pos) {
  if (n > h->atoms_cap) {
    hipFree(h->atoms);
    hipFree(h->surf);
    h->atoms = nullptr;
    h->surf = nullptr;
    h->atoms_cap = 0;  // a failed allocation leaves no capacity behind
    PP_HIP(hipMalloc(&h->atoms, sizeof(double) * 3 * (size_t)n));
    PP_HIP(hipMalloc(&h->surf, sizeof(float) * (size_t)n));
    h->atoms_cap = n;
  }
  PP_HIP(hipMemcpyAsync(h->atoms, pos, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, h->stream));
  PP_LAUNCH(k_patt_surface, dim3(cdiv(n, 256)), dim3(256), 0, h->stream, n, h->atoms, h->surf);
  const int blocks = std::min(kPattBlocks, n);
  PP_LAUNCH(k_patt_splat, dim3(blocks), dim3(256), 0, h->stream, n, h->atoms, h->surf, h->patt_part);
  PP_LAUNCH(k_patt_reduce, dim3(cdiv(CIN * NG, 256)), dim3(256), 0, h->stream, blocks, h->patt_part, 1.0f / n,
            h->P[m]);
  // the host buffer of pos must stay valid until the copy has run
  PP_HIP(hipStreamSynchronize(h->stream));
  return PD_OK;
}

}  // namespace

extern "C" {

int pd_prop_n_params(void) { return NPARAM; }

pd_status pd_prop_create(const pd_prop_desc* d, pd_prop** out) {
  if (!d || !out) return fail(PD_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (d->max_rot < 2) return fail(PD_ERR_INVALID_ARGUMENT, "max_rot < 2");
  PP_HIP(hipSetDevice(d->device));
  pd_prop* h = new pd_prop();
  h->device = d->device;
  h->max_rot = d->max_rot;
  auto bad = [&](pd_status s) {
    pd_prop_destroy(h);
    return s;
  };
  if (hipStreamCreateWithFlags(&h->own, hipStreamNonBlocking) != hipSuccess)
    return bad(fail(PD_ERR_DEVICE, "hipStreamCreate failed"));
  h->stream = h->own;
  const long long R = d->max_rot;
  pd_status st = PD_OK;
#define PP_A(p, n) \
  if ((st = alloc(&(p), (n))) != PD_OK) return bad(st)
  PP_A(h->params, NPARAM);
  PP_A(h->grads, NPARAM);
  PP_A(h->adam_m, NPARAM);
  PP_A(h->adam_v, NPARAM);
  for (int m = 0; m < 2; ++m) {
    PP_A(h->P[m], CIN * NG);
    PP_A(h->h1[m], C * NG);
    PP_A(h->h2[m], C * NG);
    PP_A(h->F[m], C * NG);
    PP_A(h->dF[m], C * NG);
    PP_A(h->dy2[m], C * NG);
    PP_A(h->dy1[m], C * NG);
  }
  PP_A(h->patt_part, (long long)kPattBlocks * CIN * NG);
  // listed vectors, in shell order
  std::vector<int> vl;
  for (int s = 0; s < NSH; ++s) {
    h->sh.off[s] = (int)vl.size();
    for (int p = 0; p < NG; ++p) {
      const double x = ((p >> 8) - 7.5) * DX, y = (((p >> 4) & 15) - 7.5) * DX, z = ((p & 15) - 7.5) * DX;
      const double r = std::sqrt(x * x + y * y + z * z);
      if (r >= SHELL_R[s] && r < SHELL_R[s + 1]) vl.push_back(p);
    }
  }
  h->sh.off[NSH] = (int)vl.size();
  h->nv = (int)vl.size();
  if (hipMalloc(&h->vlist, sizeof(int) * vl.size()) != hipSuccess) return bad(fail(PD_ERR_DEVICE, "hipMalloc failed"));
  if (hipMemcpy(h->vlist, vl.data(), sizeof(int) * vl.size(), hipMemcpyHostToDevice) != hipSuccess)
    return bad(fail(PD_ERR_DEVICE, "hipMemcpy failed"));
  h->rf_part_blocks = std::min(256, cdiv(R, kRfBwdChunk));
  h->lig_part_blocks = std::min(256, cdiv(R, RF_WAVES));
  PP_A(h->FRv, (long long)h->nv * C);
  PP_A(h->dFRv, (long long)h->nv * C);
  PP_A(h->rf_part, (long long)h->rf_part_blocks * h->nv * C);
  PP_A(h->lig_part, (long long)h->lig_part_blocks * C * NG);
  PP_A(h->dFL_cell, C * NG);
  PP_A(h->rf, R * NF);
  PP_A(h->X0, R * NINP);
  PP_A(h->H1, R * HID);
  PP_A(h->H2, R * HID);
  PP_A(h->out, R);
  PP_A(h->y, R);
  PP_A(h->energy, R);
  PP_A(h->dout, R);
  PP_A(h->dZ2, R * HID);
  PP_A(h->dZ1, R * HID);
  PP_A(h->dX0, R * NINP);
  PP_A(h->drf, R * NF);
  PP_A(h->stats, 2 * NF + 2);
  PP_A(h->loss, 1);
#undef PP_A
  if (hipMalloc(&h->rots, sizeof(double) * 9 * (size_t)R) != hipSuccess) return bad(fail(PD_ERR_DEVICE, "hipMalloc failed"));
  const void* big[] = {(const void*)k_rf_fwd, (const void*)k_rf_bwd_rec, (const void*)k_rf_bwd_lig};
  for (const void* fn : big)
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, kRfLds) != hipSuccess)
      return bad(fail(PD_ERR_DEVICE, "hipFuncSetAttribute failed"));
  std::vector<float> p;
  init_params(p, d->seed);
  if (hipMemcpy(h->params, p.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice) != hipSuccess)
    return bad(fail(PD_ERR_DEVICE, "hipMemcpy failed"));
  hipLaunchKernelGGL(k_fill, dim3(cdiv(NPARAM, 256)), dim3(256), 0, h->stream, h->adam_m, (long long)NPARAM, 0.0f);
  hipLaunchKernelGGL(k_fill, dim3(cdiv(NPARAM, 256)), dim3(256), 0, h->stream, h->adam_v, (long long)NPARAM, 0.0f);
  if (hipStreamSynchronize(h->stream) != hipSuccess) return bad(fail(PD_ERR_DEVICE, "init failed"));
  *out = h;
  return PD_OK;
}

void pd_prop_destroy(pd_prop* h) {
  if (!h) return;
  hipSetDevice(h->device);
  if (h->stream) hipStreamSynchronize(h->stream);
  float* fs[] = {h->params, h->grads, h->adam_m, h->adam_v, h->patt_part, h->surf, h->FRv, h->dFRv, h->rf_part,
                 h->lig_part, h->dFL_cell, h->rf, h->X0, h->H1, h->H2, h->out, h->y, h->energy, h->dout, h->dZ2,
                 h->dZ1, h->dX0, h->drf, h->stats, h->loss};
  for (float* f : fs) hipFree(f);
  for (int m = 0; m < 2; ++m) {
    hipFree(h->P[m]);
    hipFree(h->h1[m]);
    hipFree(h->h2[m]);
    hipFree(h->F[m]);
    hipFree(h->dF[m]);
    hipFree(h->dy2[m]);
    hipFree(h->dy1[m]);
  }
  hipFree(h->atoms);
  hipFree(h->vlist);
  hipFree(h->rots);
  if (h->own) hipStreamDestroy(h->own);
  delete h;
}

pd_status pd_prop_set_stream(pd_prop* h, void* stream) {
  if (!h) return fail(PD_ERR_INVALID_ARGUMENT, "null handle");
  PP_HIP(hipStreamSynchronize(h->stream));
  h->stream = stream ? (hipStream_t)stream : h->own;
  return PD_OK;
}

void* pd_prop_stream(pd_prop* h) { return h ? (void*)h->stream : nullptr; }

pd_status pd_prop_sync(pd_prop* h) {
  if (!h) return fail(PD_ERR_INVALID_ARGUMENT, "null handle");
  PP_HIP(hipStreamSynchronize(h->stream));
  return PD_OK;
}

pd_status pd_prop_set_complex(pd_prop* h, int n_rec, const double* rec_pos, int n_lig, const double* lig_pos) {
  if (!h || !rec_pos || !lig_pos) return fail(PD_ERR_INVALID_ARGUMENT, "null argument");
  if (n_rec < 2 || n_lig < 2) return fail(PD_ERR_INVALID_ARGUMENT, "a molecule needs at least 2 atoms");
  PP_HIP(hipSetDevice(h->device));
  PP_TRY(build_maps(h, 0, n_rec, rec_pos));
  PP_TRY(build_maps(h, 1, n_lig, lig_pos));
  h->have_complex = true;
  return PD_OK;
}

pd_status pd_prop_get_maps(pd_prop* h, float* rec, float* lig) {
  if (!h || !rec || !lig) return fail(PD_ERR_INVALID_ARGUMENT, "null argument");
  if (!h->have_complex) return fail(PD_ERR_STATE, "pd_prop_set_complex has not been called");
  PP_HIP(hipSetDevice(h->device));
  PP_HIP(hipMemcpyAsync(rec, h->P[0], sizeof(float) * CIN * NG, hipMemcpyDeviceToHost, h->stream));
  PP_HIP(hipMemcpyAsync(lig, h->P[1], sizeof(float) * CIN * NG, hipMemcpyDeviceToHost, h->stream));
  PP_HIP(hipStreamSynchronize(h->stream));
  return PD_OK;
}

pd_status pd_prop_predict_device(pd_prop* h, int n_rot, const double* d_rots, float* d_scores) {
  PP_TRY(check_n(h, n_rot));
  if (!d_rots || !d_scores) return fail(PD_ERR_INVALID_ARGUMENT, "null argument");
  PP_HIP(hipSetDevice(h->device));
  PP_TRY(forward(h, n_rot, d_rots));
  PP_HIP(hipMemcpyAsync(d_scores, h->out, sizeof(float) * n_rot, hipMemcpyDeviceToDevice, h->stream));
  return PD_OK;
}

pd_status pd_prop_train_step_device(pd_prop* h, int n_rot, const double* d_rots, const float* d_energy, float lr,
                                    float* loss) {
  PP_TRY(check_n(h, n_rot));
  if (!d_rots || !d_energy) return fail(PD_ERR_INVALID_ARGUMENT, "null argument");
  if (!(lr > 0.0f)) return fail(PD_ERR_INVALID_ARGUMENT, "lr must be > 0");
  PP_HIP(hipSetDevice(h->device));
  PP_HIP(hipMemcpyAsync(h->energy, d_energy, sizeof(float) * n_rot, hipMemcpyDeviceToDevice, h->stream));
  PP_TRY(forward(h, n_rot, d_rots));
  PP_TRY(backward(h, n_rot, d_rots));
  h->step += 1;
  const double b1 = 0.9, b2 = 0.999;
  const double bc1 = 1.0 - std::pow(b1, (double)h->step), bc2 = 1.0 - std::pow(b2, (double)h->step);
  PP_LAUNCH(k_adam, dim3(cdiv(NPARAM, 256)), dim3(256), 0, h->stream, NPARAM, h->params, h->grads, h->adam_m,
            h->adam_v, (float)(lr / bc1), (float)b1, (float)b2, 1e-8f, (float)std::sqrt(bc2));
  if (loss) {
    PP_HIP(hipMemcpyAsync(loss, h->loss, sizeof(float), hipMemcpyDeviceToHost, h->stream));
    PP_HIP(hipStreamSynchronize(h->stream));
  }
  return PD_OK;
}

pd_status pd_prop_gradients(pd_prop* h, int n_rot, const double* rots, const float* energy, float* grads,
                            float* loss) {
  PP_TRY(check_n(h, n_rot));
  if (!rots || !energy || !grads) return fail(PD_ERR_INVALID_ARGUMENT, "null argument");
  PP_HIP(hipSetDevice(h->device));
  PP_HIP(hipMemcpyAsync(h->rots, rots, sizeof(double) * 9 * (size_t)n_rot, hipMemcpyHostToDevice, h->stream));
  PP_HIP(hipMemcpyAsync(h->energy, energy, sizeof(float) * n_rot, hipMemcpyHostToDevice, h->stream));
  PP_TRY(forward(h, n_rot, h->rots));
  PP_TRY(backward(h, n_rot, h->rots));
  PP_HIP(hipMemcpyAsync(grads, h->grads, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, h->stream));
  if (loss) PP_HIP(hipMemcpyAsync(loss, h->loss, sizeof(float), hipMemcpyDeviceToHost, h->stream));
  PP_HIP(hipStreamSynchronize(h->stream));
  return PD_OK;
}

pd_status pd_prop_get_params(pd_prop* h, float* params) {
  if (!h || !params) return fail(PD_ERR_INVALID_ARGUMENT, "null argument");
  PP_HIP(hipSetDevice(h->device));
  PP_HIP(hipMemcpyAsync(params, h->params, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, h->stream));
  PP_HIP(hipStreamSynchronize(h->stream));
  return PD_OK;
}

pd_status pd_prop_set_params(pd_prop* h, const float* params) {
  if (!h || !params) return fail(PD_ERR_INVALID_ARGUMENT, "null argument");
  PP_HIP(hipSetDevice(h->device));
  PP_HIP(hipMemcpyAsync(h->params, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice, h->stream));
  PP_LAUNCH(k_fill, dim3(cdiv(NPARAM, 256)), dim3(256), 0, h->stream, h->adam_m, (long long)NPARAM, 0.0f);
  PP_LAUNCH(k_fill, dim3(cdiv(NPARAM, 256)), dim3(256), 0, h->stream, h->adam_v, (long long)NPARAM, 0.0f);
  PP_HIP(hipStreamSynchronize(h->stream));
  h->step = 0;
  return PD_OK;
}

pd_status pd_prop_select_device(pd_prop* h, int n_rot, const float* d_scores, int n_keep, int* d_idx) {
  if (!h || !d_scores || !d_idx) return fail(PD_ERR_INVALID_ARGUMENT, "null argument");
  if (n_rot < 1 || n_keep < 1 || n_keep > n_rot) return fail(PD_ERR_INVALID_ARGUMENT, "need 1 <= n_keep <= n_rot");
  PP_HIP(hipSetDevice(h->device));
  PP_LAUNCH(k_select_lowest, dim3(1), dim3(SEL_THREADS), 0, h->stream, n_rot, d_scores, n_keep, d_idx);
  return PD_OK;
}

pd_status pd_prop_gather_rotations(pd_prop* h, int n, const int* d_idx, const double* d_in, double* d_out) {
  if (!h || !d_idx || !d_in || !d_out) return fail(PD_ERR_INVALID_ARGUMENT, "null argument");
  if (n < 1) return fail(PD_ERR_INVALID_ARGUMENT, "n < 1");
  PP_HIP(hipSetDevice(h->device));
  PP_LAUNCH(k_gather_rots, dim3(cdiv(9LL * n, 256)), dim3(256), 0, h->stream, n, d_idx, d_in, d_out);
  return PD_OK;
}

}  // extern "C"

namespace prop {
__global__ void k_remap_peaks(int n, int K, pd_peak* __restrict__ peaks, const int* __restrict__ n_out,
                              const int* __restrict__ idx) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= n * K) return;
  const int i = e / K, k = e % K;
  if (k >= n_out[i]) return;
  const int r = peaks[e].rot;
  if (r >= 0 && r < n) peaks[e].rot = idx[r];
}
}  // namespace prop

extern "C" pd_status pd_prop_remap_peaks(pd_prop* h, int n, int K, pd_peak* d_peaks, const int* d_n_out,
                                         const int* d_idx) {
  if (!h || !d_peaks || !d_n_out || !d_idx) return fail(PD_ERR_INVALID_ARGUMENT, "null argument");
  if (n < 1 || K < 1) return fail(PD_ERR_INVALID_ARGUMENT, "n and K must be >= 1");
  PP_HIP(hipSetDevice(h->device));
  PP_LAUNCH(k_remap_peaks, dim3(cdiv((long long)n * K, 256)), dim3(256), 0, h->stream, n, K, d_peaks, d_n_out, d_idx);
  return PD_OK;
}
