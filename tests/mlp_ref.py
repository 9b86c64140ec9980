"""CPU (numpy, float64) restatement of the prescreen network for checking the GPU kernels.

With emulate=True the values the GPU rounds to BF16 are rounded the same way (inputs, weight
copies, stored activations and activation gradients; round to nearest even), so the remaining
differences come only from FP32 accumulation order. With emulate=False it is the plain float64
network, whose gradients are pinned against central finite differences (tests/test_mlp_cpu.py).
Parameter layout: W1[64][256], b1[256], W2[256][256], b2[256], w3[256], b3.
"""
import numpy as np

IN, HID = 64, 256
SIZES = [("W1", IN * HID), ("b1", HID), ("W2", HID * HID), ("b2", HID), ("w3", HID), ("b3", 1)]
N_PARAMS = sum(n for _, n in SIZES)


def bf16(x):
    """round float values to the nearest BF16 (ties to even), returned as float64"""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7fff + ((u >> 16) & 1)) & 0xffff0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = {}, 0
    for name, n in SIZES:
        out[name] = p[o:o + n]
        o += n
    out["W1"] = out["W1"].reshape(IN, HID)
    out["W2"] = out["W2"].reshape(HID, HID)
    return out


def forward(p, X, emulate=True):
    q = unpack(p)
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    Xb = r(X)
    H1 = r(np.maximum(Xb @ r(q["W1"]) + q["b1"], 0))
    H2 = r(np.maximum(H1 @ r(q["W2"]) + q["b2"], 0))
    y = H2 @ q["w3"] + q["b3"][0]
    return y, (Xb, H1, H2)


def loss(p, X, t, emulate=True):
    y, _ = forward(p, X, emulate)
    return float(np.mean(np.maximum(y, 0) - y * t + np.log1p(np.exp(-np.abs(y)))))


def grads(p, X, t, emulate=True):
    """gradient of the mean BCE-with-logits loss, flat, in the parameter layout"""
    q = unpack(p)
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    y, (Xb, H1, H2) = forward(p, X, emulate)
    B = len(t)
    dy = (1 / (1 + np.exp(-y)) - t) / B
    dW3 = H2.T @ dy
    db3 = dy.sum()
    dH2 = r(dy[:, None] * q["w3"][None, :]) * (H2 > 0)
    dW2 = H1.T @ dH2
    db2 = dH2.sum(axis=0)
    dH1 = r(dH2 @ r(q["W2"]).T) * (H1 > 0)
    dW1 = Xb.T @ dH1
    db1 = dH1.sum(axis=0)
    return np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3, [db3]])


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    ph = p - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return ph, m, v
