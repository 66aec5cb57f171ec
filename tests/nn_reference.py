"""CPU FP64 reference of the warm-start network of ibpm_amd/csrc/learn.hip
(13 -> 32 tanh -> 2, mean squared error), used to check the GPU loss and
gradients.  Parameter layout: W1 [32][13], b1 [32], W2 [2][32], b2 [2]."""
import numpy as np

NI, NH = 13, 32
NP = NH * NI + NH + 2 * NH + 2


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    W1 = p[:NH * NI].reshape(NH, NI)
    b1 = p[NH * NI:NH * NI + NH]
    W2 = p[NH * NI + NH:NH * NI + 3 * NH].reshape(2, NH)
    b2 = p[NH * NI + 3 * NH:NP]
    return W1, b1, W2, b2


def forward(p, X):
    W1, b1, W2, b2 = unpack(p)
    H = np.tanh(X @ W1.T + b1)
    return H, H @ W2.T + b2


def loss_grad(p, X, Y):
    """Loss = mean over samples of sum over the 2 outputs of (y - Y)^2."""
    W1, b1, W2, b2 = unpack(p)
    n = X.shape[0]
    H, Yp = forward(p, X)
    E = Yp - Y
    loss = np.sum(E * E) / n
    D2 = 2.0 * E / n
    gW2 = D2.T @ H
    gb2 = D2.sum(0)
    D1 = (D2 @ W2) * (1 - H * H)
    gW1 = D1.T @ X
    gb1 = D1.sum(0)
    return loss, np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2])
