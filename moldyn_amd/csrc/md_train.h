// md_train.h -- device state of the surrogate trainer (md_mlp_train.hip) and
// its launchers.
#pragma once

#include "md_device.h"

namespace md {

struct TrainState {
    // data set: n_frames frames of n_atoms atoms
    const double* pos;    // n_frames x n_atoms x 3
    const double* box;    // n_frames x 3
    const float* Eref;    // n_frames x n_atoms  per-atom energy labels
    const float* Fref;    // n_frames x n_atoms x 3
    float* G;             // n_frames x n_atoms x 32 features
    int32_t n_atoms, n_frames;
    int64_t n_total;      // atoms the loss is averaged over
    // pairs within the cutoff, built once: atom (f, i) has ncnt[f n + i]
    // neighbours nbr[(f n + i) cap + q], ascending j
    int32_t* nbr;
    int32_t* ncnt;
    int32_t cap;
    // frames of the current minibatch: batch slot b holds frame bf[b]
    const int32_t* bf;
    // basis and normalisation
    float mu[kMlpK], g_mean[kMlpK], g_inv_std[kMlpK];
    float rc, inv_2w2, inv_w2;
    float e_scale, e_shift, w_e, w_f;
    // parameters and optimiser
    float* theta;         // kMlpNP
    float* grad;          // kMlpNP
    float* adam_m;
    float* adam_v;
    // minibatch work arrays (rows = atoms of the minibatch)
    float *X, *Gam, *U;                                  // M x 32
    float *H1, *S1, *S1p, *H2, *S2, *S2p;                // M x 128
    float *D2, *T, *D1, *V1, *PP, *V2, *R1, *W3T;        // M x 128
    float *E, *dy;                                       // M
    float *F, *phi;                                      // M x 3
    float* gpart;         // split-K / column-sum partials
    int32_t splits, colsum_blocks;
};

hipError_t tr_count_pairs(const TrainState& t, int f0, int nf, int32_t* max_cnt, hipStream_t st);
hipError_t tr_fill_pairs(const TrainState& t, int f0, int nf, hipStream_t st);
hipError_t tr_features(const TrainState& t, int f0, int nf, hipStream_t st);
hipError_t tr_feature_stats(const TrainState& t, int64_t rows, double* part, int blocks, hipStream_t st);
// minibatch of nf frames t.bf[0..nf): M = nf n_atoms rows
hipError_t tr_forward_backward(const TrainState& t, int nf, hipStream_t st);
hipError_t tr_forces(const TrainState& t, int nf, double* loss_part, hipStream_t st);
hipError_t tr_loss_sum(const TrainState& t, int nf, const double* loss_part, double* out, hipStream_t st);
hipError_t tr_gradient(const TrainState& t, int nf, hipStream_t st);
hipError_t tr_adam(const TrainState& t, float lr, int step, hipStream_t st);
hipError_t tr_gemm(bool ta, bool tb, int M, int N, int K, float alpha, const float* A, int lda, const float* B,
                   int ldb, float beta, float* C, int ldc, int splits, hipStream_t st);

}  // namespace md
