"""TEST INFRASTRUCTURE: FP64 restatement of the learned-ordering pair scorer
(c2a_amd/csrc/scorer.h, mlp_dev.h) used to check the matrix-core kernels.

The network is 32 -> 64 -> 64 -> 1 with ReLU.  With bf16=True the same
values are rounded to BF16 as on the GPU: the inputs, W1, W2 and the first
hidden layer in the forward pass, and the two hidden-layer gradients where
they feed the matrix cores in the backward pass; everything else is FP64.
The backward pass itself is pinned by central finite differences of the
unrounded network (tests/test_scorer.py)."""
import numpy as np

NF, NH = 32, 64
P_W1 = 0
P_B1 = P_W1 + NF * NH
P_W2 = P_B1 + NH
P_B2 = P_W2 + NH * NH
P_W3 = P_B2 + NH
P_B3 = P_W3 + NH
NPARAM = P_B3 + 1


def bf16(x):
    """Round-to-nearest-even to BF16, returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[P_W1:P_B1].reshape(NF, NH), p[P_B1:P_W2], p[P_W2:P_B2].reshape(NH, NH),
            p[P_B2:P_W3], p[P_W3:P_B3], p[P_B3])


def forward(p, X, bf=True):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    r = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    Xr = r(X)
    h1 = np.maximum(Xr @ r(W1) + b1, 0.0)
    h1r = r(h1)
    h2 = np.maximum(h1r @ r(W2) + b2, 0.0)
    s = h2 @ w3 + b3
    return s, (Xr, h1r, h2)


def loss_grad(p, X, y, bf=True):
    """Mean squared error and its gradient with respect to the 6337 parameters."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    r = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    n = len(y)
    s, (Xr, h1r, h2) = forward(p, X, bf)
    e = s - y
    loss = float((e * e).sum() / n)
    dy = e * (2.0 / n)
    g = np.zeros(NPARAM)
    g[P_W3:P_B3] = h2.T @ dy
    g[P_B3] = dy.sum()
    dh2 = np.outer(dy, w3) * (h2 > 0)
    g[P_B2:P_W3] = dh2.sum(0)
    dh2r = r(dh2)
    g[P_W2:P_B2] = (h1r.T @ dh2r).reshape(-1)
    dh1 = (dh2r @ r(W2).T) * (h1r > 0)
    g[P_B1:P_W2] = dh1.sum(0)
    g[P_W1:P_B1] = (Xr.T @ r(dh1)).reshape(-1)
    return loss, g
