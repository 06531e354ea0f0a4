/*
 * pdprop.h -- C ABI of the learned rotation proposal in libpdfft.so.
 *
 * The reference has no counterpart: its worker scans every rotation of the
 * rotational grid (docking_app.hpp:487-513).  This network ranks the
 * rotations of one complex before the scan so that only a top fraction gets
 * the full FFT correlation (pd_scan_run_device on the kept rotations).  The
 * exhaustive scan stays available as fallback and as the measure of recall.
 *
 * Model (fixed shapes):
 *  - per molecule, two 16^3 interatomic-vector (Patterson) maps on a grid of
 *    0.9 A centred on the zero vector: channel 0 over all atom pairs,
 *    channel 1 over pairs of under-coordinated (surface) atoms; each pair
 *    vector v contributes exp(-|g - v|^2 / (2 * 0.5^2)) within 2 A per axis,
 *    pairs closer than 7.2 A, summed and divided by the atom count;
 *  - per molecule, a 3-layer 3x3x3 convolution stack 2 -> 8 -> 8 -> 8 (ReLU
 *    after the first two), added to the input maps (channel c += map c % 2);
 *  - for each rotation R and channel c, three rotation functions
 *    RF[c][s] = sum over grid vectors v in shell s of Frec[c](v) * Flig[c](R^T v)
 *    (trilinear), shells 2.0-3.6, 3.6-4.8, 4.8-7.2 A;
 *  - the 24 rotation functions, standardised over the rotations of the call,
 *    and the first two columns of R feed an MLP 30 -> 384 -> 384 -> 1 (ReLU);
 *    the output predicts the standardised best energy of the rotation
 *    (lower = better).
 * Convolution and linear weights are used in BF16 (round to nearest even);
 * the linear layers run on MFMA with FP32 accumulation; master weights and
 * Adam state are FP32.
 *
 * Rotations are row-major 3x3 FP64 matrices, the same as pd_scan_run.
 * All *_device calls are asynchronous on the handle's stream.
 */
#ifndef PDPROP_H
#define PDPROP_H

#include "pdfft.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pd_prop_desc {
  int max_rot;    /* most rotations in one predict / train call            */
  int device;     /* HIP device ordinal                                     */
  unsigned seed;  /* initial weights (uniform +-1/sqrt(fan_in); last conv 0) */
} pd_prop_desc;

pd_status pd_prop_create(const pd_prop_desc* desc, struct pd_prop** out);
void pd_prop_destroy(struct pd_prop* h);

/* Runs the handle's work on an external stream (for example pd_scan_stream of
 * the scan that consumes the selection), or on its own stream when NULL. */
pd_status pd_prop_set_stream(struct pd_prop* h, void* stream);
void* pd_prop_stream(struct pd_prop* h);
pd_status pd_prop_sync(struct pd_prop* h);

/* Builds the receptor and ligand maps of one complex on the GPU (atoms in the
 * scan's frames: receptor centred, ligand in its reference frame). */
pd_status pd_prop_set_complex(struct pd_prop* h, int n_rec, const double* rec_pos, int n_lig,
                              const double* lig_pos);
/* Copies the current maps back: rec and lig are 2 x 16^3 floats each. */
pd_status pd_prop_get_maps(struct pd_prop* h, float* rec, float* lig);

/* Scores of n_rot rotations (device pointers); lower = better. */
pd_status pd_prop_predict_device(struct pd_prop* h, int n_rot, const double* d_rots, float* d_scores);

/* One Adam step (lr; betas 0.9, 0.999; eps 1e-8) on the mean squared error between
 * the prediction and the standardised energies d_energy (the best exhaustive
 * energy of each rotation).  *loss (host, may be NULL) receives the loss and
 * makes the call synchronous. */
pd_status pd_prop_train_step_device(struct pd_prop* h, int n_rot, const double* d_rots,
                                    const float* d_energy, float lr, float* loss);

/* Loss and gradient of every parameter without an update (host buffers; grads
 * has pd_prop_n_params() floats in the pd_prop_get_params layout). */
pd_status pd_prop_gradients(struct pd_prop* h, int n_rot, const double* rots, const float* energy,
                            float* grads, float* loss);

/* Flat FP32 parameters: receptor conv stack, ligand conv stack (each w1[8][2][27],
 * b1[8], w2[8][8][27], b2[8], w3[8][8][27], b3[8]), then W1[384][30], b1[384],
 * W2[384][384], b2[384], W3[384], b3[1]. */
int pd_prop_n_params(void);
pd_status pd_prop_get_params(struct pd_prop* h, float* params);
pd_status pd_prop_set_params(struct pd_prop* h, const float* params);

/* Indices (ascending) of the n_keep lowest scores; ties at the cut go to the
 * lower index. */
pd_status pd_prop_select_device(struct pd_prop* h, int n_rot, const float* d_scores, int n_keep,
                                int* d_idx);
/* d_out[i] = d_in[d_idx[i]] for i < n (9 doubles each). */
pd_status pd_prop_gather_rotations(struct pd_prop* h, int n, const int* d_idx, const double* d_in,
                                   double* d_out);
/* After a scan of gathered rotations with rot_index0 = 0: peak.rot = d_idx[peak.rot]
 * for the first n_out[i] peaks of each of the n rotations (K peaks per rotation). */
pd_status pd_prop_remap_peaks(struct pd_prop* h, int n, int K, pd_peak* d_peaks, const int* d_n_out,
                              const int* d_idx);

#ifdef __cplusplus
}
#endif
#endif /* PDPROP_H */
