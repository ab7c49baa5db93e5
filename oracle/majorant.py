"""CPU restatement of the learned rate majorant -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py, as the checker. The majorant has no
counterpart in the reference (no file:line to follow); this module states the network, its loss, the
backward pass and Adam in numpy, at the precision points of the BF16-MFMA kernels:

* inputs standardized in FP32 ((x - mean) * (1 / scale)), then rounded to BF16 (round to nearest even);
* W1 and W2 rounded to BF16; products accumulated exactly here (float64), FP32 on the GPU;
* SiLU(A1) rounded to BF16 (the layer-2 operand); layer 3 in FP32 from SiLU(A2) unrounded;
* backward: dA2 rounded to BF16 for dH1 = dA2 W2^T and dW2 = H1^T dA2; dA1 rounded to BF16 for
  dW1 = X^T dA1; bias and w3 gradients from the unrounded FP32 values;
* loss of one sample (ln softplus(z) - ln t)^2, averaged over the batch;
* Adam (Kingma and Ba 2015) with bias correction, p -= lr * mhat / (sqrt(vhat) + eps), in FP32.

With exact=True every rounding is skipped and all arithmetic is float64: that form is pinned by finite
differences and by torch's autograd and Adam on the CPU (tests/test_oracle_majorant.py); the rounded
form differs from it only by the rounding calls.
"""
from __future__ import annotations

import numpy as np

F, H = 8, 64
P_W1 = 0
P_B1 = P_W1 + F*H
P_W2 = P_B1 + H
P_B2 = P_W2 + H*H
P_W3 = P_B2 + H
P_B3 = P_W3 + H
P_COUNT = P_B3 + 1


def bf16(x) -> np.ndarray:
    """Round float32 values to the nearest BF16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def unpack(p: np.ndarray) -> dict:
    p = np.asarray(p)
    return {"W1": p[P_W1:P_B1].reshape(F, H), "b1": p[P_B1:P_W2], "W2": p[P_W2:P_B2].reshape(H, H),
            "b2": p[P_B2:P_W3], "w3": p[P_W3:P_B3], "b3": p[P_B3]}


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.ravel(d["W1"]), d["b1"], np.ravel(d["W2"]), d["b2"], d["w3"], [d["b3"]]])


def _sigmoid(x):
    return 1.0/(1.0 + np.exp(-x))


def _silu(x):
    return x*_sigmoid(x)


def _silu_grad(x):
    s = _sigmoid(x)
    return s*(1.0 + x*(1.0 - s))


def _softplus(z):
    return np.maximum(z, 0.0) + np.log1p(np.exp(-np.abs(z)))


def forward(params, norm, X, exact: bool = False) -> dict:
    """norm = {mean[8], scale[8]}. Returns the intermediates and y = softplus(z)."""
    q = unpack(params)
    if exact:
        f = np.float64
        xs = (np.asarray(X, f) - np.asarray(norm[:F], f))/np.asarray(norm[F:], f)
        W1, W2 = q["W1"].astype(f), q["W2"].astype(f)
        A1 = xs @ W1 + q["b1"].astype(f)
        H1 = _silu(A1)
        A2 = H1 @ W2 + q["b2"].astype(f)
        H2 = _silu(A2)
        z = H2 @ q["w3"].astype(f) + f(q["b3"])
    else:
        f32 = np.float32
        inv = (f32(1.0)/np.asarray(norm[F:], f32)).astype(f32)
        xs = bf16((np.asarray(X, f32) - np.asarray(norm[:F], f32))*inv).astype(np.float64)
        W1, W2 = bf16(q["W1"]).astype(np.float64), bf16(q["W2"]).astype(np.float64)
        A1 = (xs @ W1 + q["b1"].astype(np.float64)).astype(f32)
        H1 = bf16(_silu(A1)).astype(np.float64)
        A2 = (H1 @ W2 + q["b2"].astype(np.float64)).astype(f32)
        H2 = _silu(A2)
        z = (H2.astype(np.float64) @ q["w3"].astype(np.float64) + float(q["b3"])).astype(f32)
    return {"xs": xs, "A1": A1, "H1": H1, "A2": A2, "H2": H2, "z": z, "y": _softplus(z), "W1": W1, "W2": W2}


def predict(params, norm, X, exact: bool = False) -> np.ndarray:
    return forward(params, norm, X, exact)["y"]


def loss_grad(params, norm, X, log_target, exact: bool = False) -> tuple[np.ndarray, float]:
    """Gradient of the mean loss over the rows of X and the loss value."""
    fw = forward(params, norm, X, exact)
    q = unpack(params)
    n = len(X)
    f = np.float64
    rnd = (lambda a: a) if exact else (lambda a: bf16(a).astype(f))
    z = np.asarray(fw["z"], f)
    y = _softplus(z)
    d = np.log(y) - np.asarray(log_target, f)
    loss = float(np.sum(d*d)/n)
    dz = 2.0*d*_sigmoid(z)/y/n
    A2, A1 = np.asarray(fw["A2"], f), np.asarray(fw["A1"], f)
    H2 = np.asarray(fw["H2"], f)
    g = {}
    g["w3"] = H2.T @ dz
    g["b3"] = np.sum(dz)
    dA2 = dz[:, None]*q["w3"].astype(f)[None, :]*_silu_grad(A2)
    g["b2"] = dA2.sum(axis=0)
    dA2r = rnd(dA2)
    dH1 = dA2r @ fw["W2"].T
    dA1 = dH1*_silu_grad(A1)
    g["b1"] = dA1.sum(axis=0)
    g["W2"] = np.asarray(fw["H1"], f).T @ dA2r
    g["W1"] = np.asarray(fw["xs"], f).T @ rnd(dA1)
    return pack(g), loss


def loss_value(params, norm, X, log_target, exact: bool = True) -> float:
    y = predict(params, norm, X, exact)
    d = np.log(np.asarray(y, np.float64)) - np.asarray(log_target, np.float64)
    return float(np.mean(d*d))


def adam_step(p, m, v, g, step: int, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update in FP32; returns new (p, m, v)."""
    f32 = np.float32
    p, m, v, g = (np.asarray(a, f32) for a in (p, m, v, g))
    b1, b2 = f32(beta1), f32(beta2)
    m = b1*m + (f32(1.0) - b1)*g
    v = b2*v + (f32(1.0) - b2)*(g*g)
    bc1 = f32(1.0 - beta1**step)
    bc2 = f32(1.0 - beta2**step)
    p = p - f32(lr)*(m/bc1)/(np.sqrt(v/bc2) + f32(eps))
    return p.astype(f32), m, v


def shuffle_order(n: int, state: int) -> tuple[np.ndarray, int]:
    """The training sample order: Fisher-Yates over splitmix64 draws (same as the library)."""
    M = (1 << 64) - 1
    perm = np.arange(n, dtype=np.int64)
    s = state
    for i in range(n - 1, 0, -1):
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30))*0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27))*0x94D049BB133111EB) & M
        z ^= z >> 31
        k = z % (i + 1)
        perm[i], perm[k] = perm[k], perm[i]
    return perm, s


def train(params, norm, X, log_target, epochs: int, batch: int, lr: float, seed: int):
    """Minibatch Adam as gdx_col_majorant_train (rounded form); returns params and epoch losses."""
    p = np.asarray(params, np.float32).copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    n = len(X)
    s = seed
    step = 0
    losses = []
    for _ in range(epochs):
        perm, s = shuffle_order(n, s)
        tot = 0.0
        for b in range(n//batch):
            idx = perm[b*batch:(b + 1)*batch]
            g, loss = loss_grad(p, norm, X[idx], log_target[idx])
            step += 1
            p, m, v = adam_step(p, m, v, g, step, lr)
            tot += loss
        losses.append(tot/(n//batch))
    return p, np.array(losses)


def init_params(seed: int) -> np.ndarray:
    """The library's initial parameters: uniform(+-1/sqrt(fan_in)) from splitmix64, b3 = softplus^-1(1)."""
    M = (1 << 64) - 1
    s = seed
    out = np.zeros(P_COUNT, np.float32)

    def unit():
        nonlocal s
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30))*0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27))*0x94D049BB133111EB) & M
        z ^= z >> 31
        return (z >> 11)*2.0**-53

    for off, count, fan in ((P_W1, F*H, F), (P_B1, H, F), (P_W2, H*H, H), (P_B2, H, H), (P_W3, H, H)):
        bound = 1.0/np.sqrt(fan)
        for i in range(count):
            out[off + i] = np.float32((2.0*unit() - 1.0)*bound)
    out[P_B3] = np.float32(0.54132485)
    return out
