// learn.hip -- the learned CG warm start: a small network applied to every
// boundary point, trained and evaluated on the GPU (FP64).
//
// For substep k of a moving-body run the CG unknown is the body force f_k
// (ConjugateGradientSolver.cc:35-65).  The reference warm-starts from the
// previous substep's force f_{k-1} (ProjectionSolver.cc:72), but the RK3
// substeps use different beta and M = C A^-1 beta B scales with beta, so that
// start is no better than zero.  The network is anchored on the same substep
// of the previous step, f_{k-S} (same beta), and predicts the one-step
// increment from the one-step increments of the history:
//   s      = rms(f_{k-S}) over the flow's points, sb = rms(b_k - b_{k-1})
//   x(p)   = [(f_{k-1}-f_{k-S-1})/s, (f_{k-S}-f_{k-2S})/s, (f_{k-S-1}-f_{k-2S-1})/s,
//             (f_{k-1}-f_{k-S})/s, (b_k-b_{k-1})/sb, one-hot(substep)]   at point p
//   y(p)   = W2 tanh(W1 x + b1) + b2          (13 -> 32 -> 2)
//   f0(p)  = f_{k-S}(p) + s y(p)              (CG start; CG then runs to the
//                                              same absolute tolerance)
// Components of point p are (p, nb + p) as in BoundaryVector.h:65-69.  With
// W2 = 0 (the initial parameters) the start is f_{k-S}.
#include "learn.h"

namespace ibpmg {

__device__ inline double rms_of(const double* v, int n, double* red) {
    double s = 0;
    for (int k = threadIdx.x; k < n; k += kThreads) s += v[k] * v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    __syncthreads();
    if (lane == 0) red[wave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < kThreads / 64; ++w) t += red[w];
        red[kThreads / 64] = sqrt(t / (n > 0 ? n : 1)) + 1e-300;
    }
    __syncthreads();
    return red[kThreads / 64];
}

// Features of every point of one flow (one workgroup per flow).  With a
// non-null target the sample (x, (f_k - f_{k-1})/s) is appended to the data set.
__global__ void __launch_bounds__(kThreads) k_nn_features(const NnOps* __restrict__ ops) {
    const NnOps& O = ops[blockIdx.x];
    __shared__ double red[kThreads / 64 + 1];
    const int nb = O.nb, n = 2 * nb;
    const double s = rms_of(O.fb, n, red);
    double sb = 0;
    {
        double t = 0;
        for (int k = threadIdx.x; k < n; k += kThreads) {
            const double d = O.b_now[k] - O.b_prev[k];
            t += d * d;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
        __syncthreads();
        if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = t;
        __syncthreads();
        if (threadIdx.x == 0) {
            double u = 0;
            for (int w = 0; w < kThreads / 64; ++w) u += red[w];
            red[kThreads / 64] = sqrt(u / (n > 0 ? n : 1)) + 1e-300;
        }
        __syncthreads();
        sb = red[kThreads / 64];
    }
    for (int p = threadIdx.x; p < nb; p += kThreads) {
        double* x = O.X + (long)p * kNnIn;
        const int cx = p, cy = nb + p;
        const double* hp[4] = {O.fa, O.fb, O.fc, O.fa};
        const double* hm[4] = {O.fc, O.fd, O.fe, O.fb};
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            x[2 * h] = (hp[h][cx] - hm[h][cx]) / s;
            x[2 * h + 1] = (hp[h][cy] - hm[h][cy]) / s;
        }
        x[8] = (O.b_now[cx] - O.b_prev[cx]) / sb;
        x[9] = (O.b_now[cy] - O.b_prev[cy]) / sb;
        x[10] = O.substep == 0 ? 1.0 : 0.0;
        x[11] = O.substep == 1 ? 1.0 : 0.0;
        x[12] = O.substep >= 2 ? 1.0 : 0.0;
        if (O.scale) O.scale[0] = s;
        if (O.target && O.data_x) {
            const long row = O.data_off + p;
            for (int j = 0; j < kNnIn; ++j) O.data_x[row * kNnIn + j] = x[j];
            O.data_y[row * 2] = (O.target[cx] - O.fb[cx]) / s;
            O.data_y[row * 2 + 1] = (O.target[cy] - O.fb[cy]) / s;
        }
    }
}

__device__ inline void nn_forward(const double* __restrict__ P, const double* x, double* h, double* y) {
    const double* W1 = P;                          // [H][In]
    const double* b1 = P + kNnHid * kNnIn;         // [H]
    const double* W2 = b1 + kNnHid;                // [2][H]
    const double* b2 = W2 + 2 * kNnHid;            // [2]
    double xr[kNnIn];
#pragma unroll
    for (int i = 0; i < kNnIn; ++i) xr[i] = x[i];
#pragma unroll
    for (int j = 0; j < kNnHid; ++j) {
        double a = b1[j];
#pragma unroll
        for (int i = 0; i < kNnIn; ++i) a += W1[j * kNnIn + i] * xr[i];
        h[j] = tanh(a);
    }
#pragma unroll
    for (int o = 0; o < 2; ++o) {
        double a = b2[o];
#pragma unroll
        for (int j = 0; j < kNnHid; ++j) a += W2[o * kNnHid + j] * h[j];
        y[o] = a;
    }
}

// f0 = f_{k-S} + s * net(x) written into the CG unknown.
__global__ void __launch_bounds__(kThreads) k_nn_infer(const NnOps* __restrict__ ops) {
    const NnOps& O = ops[blockIdx.y];
    const int p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= O.nb) return;
    double h[kNnHid], y[2];
    nn_forward(O.params, O.X + (long)p * kNnIn, h, y);
    const double s = O.scale[0];
    O.x0[p] = O.fb[p] + s * y[0];
    O.x0[O.nb + p] = O.fb[O.nb + p] + s * y[1];
}

// Per-sample forward and backward of the mean squared error: stores the
// hidden activations and the two back-propagated error vectors.
__global__ void __launch_bounds__(kThreads) k_nn_backprop(const double* __restrict__ P, const double* __restrict__ X,
                                                          const double* __restrict__ Y, long n, double* __restrict__ S) {
    const long s = (long)blockIdx.x * kThreads + threadIdx.x;
    if (s >= n) return;
    double h[kNnHid], y[2];
    nn_forward(P, X + s * kNnIn, h, y);
    const double* W2 = P + kNnHid * kNnIn + kNnHid;
    double* out = S + s * kNnScratch;   // [h (H), d2 (2), d1 (H), loss]
    const double inv = 1.0 / (double)n;
    double d2[2], loss = 0;
    for (int o = 0; o < 2; ++o) {
        const double e = y[o] - Y[s * 2 + o];
        loss += e * e;
        d2[o] = 2.0 * e * inv;   // d(mean sum e^2)/dy
    }
#pragma unroll
    for (int j = 0; j < kNnHid; ++j) out[j] = h[j];
    out[kNnHid] = d2[0];
    out[kNnHid + 1] = d2[1];
#pragma unroll
    for (int j = 0; j < kNnHid; ++j) {
        const double g = W2[j] * d2[0] + W2[kNnHid + j] * d2[1];
        out[kNnHid + 2 + j] = g * (1 - h[j] * h[j]);
    }
    out[2 * kNnHid + 2] = loss * inv;
}

// Gradient of every parameter over a chunk of samples (one thread per
// parameter, samples in order): partial sums per chunk, combined in fixed
// order by k_nn_adam, so training is deterministic.
__global__ void __launch_bounds__(kThreads) k_nn_grad_partial(const double* __restrict__ X, const double* __restrict__ S,
                                                              long n, long chunk, double* __restrict__ part) {
    const int q = blockIdx.x * kThreads + threadIdx.x;
    if (q > kNnParams) return;   // q == kNnParams: loss
    const long s0 = (long)blockIdx.y * chunk;
    const long s1 = s0 + chunk < n ? s0 + chunk : n;
    double g = 0;
    if (q < kNnHid * kNnIn) {                       // W1[j][i] = sum d1_j x_i
        const int j = q / kNnIn, i = q % kNnIn;
        for (long s = s0; s < s1; ++s) g += S[s * kNnScratch + kNnHid + 2 + j] * X[s * kNnIn + i];
    } else if (q < kNnHid * kNnIn + kNnHid) {       // b1[j] = sum d1_j
        const int j = q - kNnHid * kNnIn;
        for (long s = s0; s < s1; ++s) g += S[s * kNnScratch + kNnHid + 2 + j];
    } else if (q < kNnHid * kNnIn + kNnHid + 2 * kNnHid) {   // W2[o][j] = sum d2_o h_j
        const int r = q - kNnHid * kNnIn - kNnHid;
        const int o = r / kNnHid, j = r % kNnHid;
        for (long s = s0; s < s1; ++s) g += S[s * kNnScratch + kNnHid + o] * S[s * kNnScratch + j];
    } else if (q < kNnParams) {                     // b2[o] = sum d2_o
        const int o = q - (kNnHid * kNnIn + 3 * kNnHid);
        for (long s = s0; s < s1; ++s) g += S[s * kNnScratch + kNnHid + o];
    } else {
        for (long s = s0; s < s1; ++s) g += S[s * kNnScratch + 2 * kNnHid + 2];
    }
    part[(long)blockIdx.y * (kNnParams + 1) + q] = g;
}

// Sums the chunk partials in chunk order into grad (and loss); then, when
// lr > 0, one Adam step (Kingma & Ba 2015) on the parameters.
__global__ void __launch_bounds__(kThreads) k_nn_adam(double* __restrict__ P, double* __restrict__ m, double* __restrict__ v,
                                                      const double* __restrict__ part, int nchunks, double* __restrict__ grad,
                                                      double lr, double b1t, double b2t) {
    const int q = blockIdx.x * kThreads + threadIdx.x;
    if (q > kNnParams) return;
    double g = 0;
    for (int c = 0; c < nchunks; ++c) g += part[(long)c * (kNnParams + 1) + q];
    grad[q] = g;
    if (q == kNnParams || lr <= 0) return;
    const double beta1 = 0.9, beta2 = 0.999, eps = 1e-12;
    m[q] = beta1 * m[q] + (1 - beta1) * g;
    v[q] = beta2 * v[q] + (1 - beta2) * g * g;
    const double mh = m[q] / (1 - b1t), vh = v[q] / (1 - b2t);
    P[q] -= lr * mh / (sqrt(vh) + eps);
}

void launch_nn_features(const NnOps* d, int nflows, hipStream_t s) {
    hipLaunchKernelGGL(k_nn_features, dim3(nflows), dim3(kThreads), 0, s, d);
}
void launch_nn_infer(const NnOps* d, int nflows, int max_nb, hipStream_t s) {
    hipLaunchKernelGGL(k_nn_infer, dim3((max_nb + kThreads - 1) / kThreads, nflows), dim3(kThreads), 0, s, d);
}
void launch_nn_grad(const double* P, const double* X, const double* Y, long n, double* S, double* part,
                    int nchunks, double* m, double* v, double* grad, double lr, double b1t, double b2t,
                    double* params_out, hipStream_t s) {
    (void)params_out;
    hipLaunchKernelGGL(k_nn_backprop, dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads), 0, s,
                       P, X, Y, n, S);
    const long chunk = (n + nchunks - 1) / nchunks;
    hipLaunchKernelGGL(k_nn_grad_partial, dim3((kNnParams + kThreads) / kThreads, nchunks), dim3(kThreads), 0, s,
                       X, S, n, chunk, part);
    hipLaunchKernelGGL(k_nn_adam, dim3((kNnParams + kThreads) / kThreads), dim3(kThreads), 0, s,
                       const_cast<double*>(P), m, v, part, nchunks, grad, lr, b1t, b2t);
}

}  // namespace ibpmg
