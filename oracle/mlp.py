"""TEST INFRASTRUCTURE ONLY: FP64 CPU restatement of the bracket surrogate's network.

The MLP 24 -> 256 -> 256 -> 256 -> 1 (GELU in its tanh form) of machinig-sim_amd/csrc/mlp.hip,
its loss sum((y - t)^2) / (2n) and the gradient of that loss.  With quant=True the BF16
rounding points of the device (inputs, weights, activations, dY, dZ; round to nearest even)
are reproduced, so the device gradients differ from this one only by FP32 accumulation;
with quant=False it is the exact FP64 gradient of the same network.  No reference
counterpart exists (the surrogate is new); the gradient is pinned by central finite
differences in tests/test_oracle_mlp.py.
"""
import numpy as np

IN, HID = 24, 256
NPARAM = IN * HID + HID + 2 * (HID * HID + HID) + HID + 1


def bf16(x):
    """round to BF16 (nearest, ties to even) through FP32, returned as float64"""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    o = 0
    out = []
    for shape in ((IN, HID), (HID,), (HID, HID), (HID,), (HID, HID), (HID,), (HID,), (1,)):
        k = int(np.prod(shape))
        out.append(P[o:o + k].reshape(shape))
        o += k
    return out


def gelu(z):
    c = np.sqrt(2.0 / np.pi)
    return 0.5 * z * (1.0 + np.tanh(c * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    c = np.sqrt(2.0 / np.pi)
    t = np.tanh(c * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * c * (1.0 + 3.0 * 0.044715 * z * z)


def forward(P, X, quant=True):
    """X: scaled inputs (n x 24).  Returns y (n,) and the cache of the backward pass."""
    q = bf16 if quant else (lambda v: np.asarray(v, dtype=np.float64))
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(P)
    H0 = q(X)
    Ws = [q(W1), q(W2), q(W3)]
    bs = [b1, b2, b3]
    Hs, Zs = [H0], []
    for l in range(3):
        Z = Hs[-1] @ Ws[l] + bs[l]
        Zs.append(Z)
        Hs.append(q(gelu(Z)))
    w4 = q(W4)
    y = Hs[-1] @ w4 + b4[0]
    return y, (Hs, Zs, Ws, w4)


def grads(P, X, t, quant=True):
    """flat gradient (parameter order W1 b1 W2 b2 W3 b3 W4 b4) and the loss"""
    q = bf16 if quant else (lambda v: np.asarray(v, dtype=np.float64))
    n = len(t)
    y, (Hs, Zs, Ws, w4) = forward(P, X, quant)
    d = y - np.asarray(t, dtype=np.float64)
    loss = float(np.sum(d * d) / (2 * n))
    dY = d / n
    dYq = q(dY)
    g = [None] * 8
    g[7] = np.array([dY.sum()])
    g[6] = Hs[3].T @ dYq
    dZ = q(np.outer(dYq, w4) * gelu_grad(Zs[2]))
    for l in (2, 1, 0):
        g[2 * l] = Hs[l].T @ dZ
        g[2 * l + 1] = dZ.sum(axis=0)
        if l:
            dZ = q((dZ @ Ws[l].T) * gelu_grad(Zs[l - 1]))
    return np.concatenate([x.ravel() for x in g]), loss
