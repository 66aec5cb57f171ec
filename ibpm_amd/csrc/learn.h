// learn.h -- descriptors and launchers of the learned CG warm start.
#pragma once
#include "common.h"

namespace ibpmg {

constexpr int kNnIn = 13;                                   // features per point
constexpr int kNnHid = 32;                                  // hidden units (tanh)
constexpr int kNnParams = kNnHid * kNnIn + kNnHid + 2 * kNnHid + 2;   // W1 b1 W2 b2
constexpr int kNnScratch = 2 * kNnHid + 3;                  // h, d2, d1, loss per sample

struct NnOps {
    int nb, substep;
    const double *fa, *fb, *fc, *fd, *fe;   // f_{k-1}, f_{k-S}, f_{k-S-1}, f_{k-2S}, f_{k-2S-1}
    const double *b_now, *b_prev;           // constraint of this and of the previous substep
    double* X;                              // nb x kNnIn features
    double* scale;                          // rms(f_{k-1})
    const double* target;                   // f_k after CG (sample collection) or null
    double* data_x;                         // data set (rows of kNnIn) or null
    double* data_y;                         // targets (rows of 2)
    long data_off;
    const double* params;
    double* x0;                             // CG unknown receiving the prediction
};

void launch_nn_features(const NnOps* d, int nflows, hipStream_t s);
void launch_nn_infer(const NnOps* d, int nflows, int max_nb, hipStream_t s);
void launch_nn_grad(const double* P, const double* X, const double* Y, long n, double* S, double* part,
                    int nchunks, double* m, double* v, double* grad, double lr, double b1t, double b2t,
                    double* params_out, hipStream_t s);

}  // namespace ibpmg
