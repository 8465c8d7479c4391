"""FP64 CPU restatement of the row-nnz predictor -- TEST INFRASTRUCTURE ONLY.

Checks the library's HIP forward / backward / feature kernels
(rtspmspm_amd/csrc/k_learned.hip).  The reference has no learned component;
this restates the network the build defines: 8 inputs -> 32 ReLU -> 32 ReLU
-> 1, output = network + 32 * x[6] (residual on log2(1 + min(P, cols))),
loss = mean (y_hat - y)^2.  Parameter layout: W1[32][8], b1[32], W2[32][32],
b2[32], W3[32], b3.

Pinned by tests/test_learned_cpu.py: the analytic gradient below is checked
against central finite differences in FP64 (so a dropped ReLU mask or a
transposed weight fails), and features against a pure-Python loop.
"""
import numpy as np

NF, H = 8, 32
OW1, OB1 = 0, H * NF
OW2 = OB1 + H
OB2 = OW2 + H * H
OW3 = OB2 + H
OB3 = OW3 + H
NPARAM = OB3 + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OW1:OB1].reshape(H, NF), p[OB1:OW2], p[OW2:OB2].reshape(H, H), p[OB2:OW3],
            p[OW3:OB3], p[OB3])


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    z1 = X @ W1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0)
    out = h2 @ W3 + b3 + 32.0 * X[:, 6]
    return out, (X, z1, h1, z2, h2)


def loss_grad(p, X, y):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    out, (X, z1, h1, z2, h2) = forward(p, X)
    y = np.asarray(y, np.float64)
    n = y.size
    d = out - y
    loss = float((d * d).mean())
    g = 2.0 * d / n
    gW3 = h2.T @ g
    gb3 = g.sum()
    d2 = np.outer(g, W3) * (z2 > 0)
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (z1 > 0)
    gW1 = d1.T @ X
    gb1 = d1.sum(0)
    grad = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return loss, grad


def features(A, B, prod=None):
    """Per-row features in the kernel's definition (FP64, cast to float32)."""
    m = A.shape[0]
    ncols = B.shape[1]
    blen = np.diff(B.row_ptr)
    F = np.zeros((m, NF), np.float32)
    for i in range(m):
        a0, a1 = int(A.row_ptr[i]), int(A.row_ptr[i + 1])
        nA = a1 - a0
        ks = A.col[a0:a1]
        lens = blen[ks] if nA else np.zeros(0, np.int64)
        P = float(lens.sum()) if prod is None else float(prod[i])
        maxlen = int(lens.max()) if nA else 0
        mask = 0
        ns = 0
        for l in range(min(16, nA)):
            e = a0 + (l if nA <= 16 else (l * nA) // 16)
            k = int(A.col[e])
            q0 = int(B.row_ptr[k])
            L = int(B.row_ptr[k + 1]) - q0
            t_n = min(8, L)
            for t in range(t_n):
                c = int(B.col[q0 + (t * L) // t_n])
                mask |= 1 << (((c * 0x9E3779B1) & 0xFFFFFFFF) >> 26)
            ns += t_n
        U = min(P, float(ncols))
        F[i] = [np.log2(1 + nA) / 16, np.log2(1 + P) / 32, np.log2(1 + maxlen) / 24,
                np.log2(1 + P / max(1.0, nA)) / 24, bin(mask).count("1") / 64.0,
                np.log2(1 + ns) / 8, np.log2(1 + U) / 32, np.log2(1 + ncols) / 32]
    return F
