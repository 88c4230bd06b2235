"""CPU oracle for the format selector -- TEST INFRASTRUCTURE ONLY.

The selector has no reference counterpart (the reference picks no format), so
this module restates the library's own definitions in numpy float64:

* ``features``: the 64 structure features of a CSR matrix, straight from
  their definitions (row lengths from ``np.diff``, histograms with
  ``np.bincount``), not from the device algorithm (grid-stride counters,
  binary search per run of entries).
* ``forward`` / ``loss_grad``: the MLP 64 -> 128 -> 128 -> 8 with ReLU and a
  softmax cross-entropy against target distributions, and its analytic
  gradient.  ``numeric_grad`` differentiates ``loss`` by central differences;
  the tests use it to pin ``loss_grad`` before trusting it as the checker.
* ``adam``: one Adam step (Kingma & Ba 2015, bias-corrected).

Parameters are packed W1[64][128] b1 W2[128][128] b2 W3[128][8] b3.
"""
from __future__ import annotations

import numpy as np

NI, NH, NO = 64, 128, 8
SHAPES = [("W1", (NI, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH, NO)), ("b3", (NO,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(theta):
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(theta[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def pack(parts):
    return np.concatenate([np.asarray(parts[name], np.float64).ravel() for name, _ in SHAPES])


def forward(theta, X):
    p = unpack(theta)
    X = np.asarray(X, np.float64)
    H1 = np.maximum(X @ p["W1"] + p["b1"], 0.0)
    H2 = np.maximum(H1 @ p["W2"] + p["b2"], 0.0)
    Z = H2 @ p["W3"] + p["b3"]
    return H1, H2, Z


def softmax(Z):
    Z = Z - Z.max(axis=1, keepdims=True)
    E = np.exp(Z)
    return E / E.sum(axis=1, keepdims=True)


def loss(theta, X, P):
    _, _, Z = forward(theta, X)
    Zs = Z - Z.max(axis=1, keepdims=True)
    logq = Zs - np.log(np.exp(Zs).sum(axis=1, keepdims=True))
    return float(-(np.asarray(P, np.float64) * logq).sum() / X.shape[0])


def loss_grad(theta, X, P):
    """(loss, gradient) by backpropagation."""
    p = unpack(theta)
    X = np.asarray(X, np.float64)
    P = np.asarray(P, np.float64)
    n = X.shape[0]
    H1, H2, Z = forward(theta, X)
    Q = softmax(Z)
    dZ = (Q - P) / n
    g = {"W3": H2.T @ dZ, "b3": dZ.sum(0)}
    dH2 = (dZ @ p["W3"].T) * (H2 > 0)
    g["W2"], g["b2"] = H1.T @ dH2, dH2.sum(0)
    dH1 = (dH2 @ p["W2"].T) * (H1 > 0)
    g["W1"], g["b1"] = X.T @ dH1, dH1.sum(0)
    return loss(theta, X, P), pack(g)


def numeric_grad(theta, X, P, idx, h=1e-6):
    """Central differences of ``loss`` at the parameter indices ``idx``."""
    theta = np.asarray(theta, np.float64).copy()
    out = np.empty(len(idx))
    for j, i in enumerate(idx):
        t0 = theta[i]
        theta[i] = t0 + h
        lp = loss(theta, X, P)
        theta[i] = t0 - h
        lm = loss(theta, X, P)
        theta[i] = t0
        out[j] = (lp - lm) / (2 * h)
    return out


def adam(theta, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v


def features(row_ptr, columns, rows, cols):
    """The 64 structure features, from their definitions."""
    rp = np.asarray(row_ptr, np.int64)
    ci = np.asarray(columns, np.int64)
    lens = np.diff(rp)
    nnz = int(rp[-1])
    rows_f, cols_f = float(rows), float(cols)
    nz = float(nnz) if nnz else 1.0
    mean = nnz / rows_f
    # the sum of squared lengths is an exact 64-bit integer sum
    sumsq = int((lens.astype(np.uint64) ** 2).sum(dtype=np.uint64))
    var = max(0.0, float(sumsq) / rows_f - mean * mean)
    sd = np.sqrt(var)
    cv = sd / mean if mean > 0 else 0.0
    mx = int(lens.max()) if lens.size else 0
    f = np.zeros(64)
    f[0] = np.log2(rows_f + 1) / 32
    f[1] = np.log2(cols_f + 1) / 32
    f[2] = np.log2(nnz + 1.0) / 40
    f[3] = np.log2(1 + mean) / 16
    f[4] = np.log2(1.0 + mx) / 16
    f[5] = min(cv, 32.0) / 32
    f[6] = np.count_nonzero(lens == 0) / rows_f
    f[7] = np.log2(1 + sd) / 16
    ne = lens > 0
    lb = np.zeros(lens.size, np.int64)
    # exact floor(log2 len): len = m 2^e with m in [0.5, 1)
    lb[ne] = np.frexp(lens[ne].astype(np.float64))[1] - 1
    bins = np.where(ne, np.minimum(23, lb + 1), 0)
    f[8:32] = np.bincount(bins, minlength=24)[:24] / rows_f
    f[32:48] = np.bincount(np.minimum(15, lb[ne]), weights=lens[ne], minlength=16)[:16] / nz
    if nnz:
        r = np.repeat(np.arange(rows, dtype=np.int64), lens)
        a = ci.astype(np.uint64) * np.uint64(rows)
        b = r.astype(np.uint64) * np.uint64(cols)
        d = np.where(a > b, a - b, b - a).astype(np.float64) / (rows_f * cols_f)
        th = 1e-4 * 4.0 ** np.arange(7)
        k = np.searchsorted(th, d, side="right")
        f[48:56] = np.bincount(k, minlength=8)[:8] / nz
        f[56] = float(np.floor(d * 16777216.0).astype(np.uint64).sum(dtype=np.uint64)) / 16777216.0 / nz
        f[57] = np.count_nonzero(ci == r) / nz
        first, last = ci[rp[:-1][ne]], ci[rp[1:][ne] - 1]
        f[58] = float((np.abs(last - first) + 1).sum()) / np.count_nonzero(ne) / cols_f
        f[59] = np.log2(rows_f * mx / nnz) / 16
    f[60] = lens[lens > 4096].sum() / nz
    f[61] = lens[ne & (lens <= 4)].sum() / nz
    f[62] = np.count_nonzero(lens > 64) / rows_f
    f[63] = 1.0
    return f


def bf16_round(a):
    """Round float32 values to the nearest bfloat16 (ties to even), returned
    as float32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def predict_bf16(theta, X):
    """Inference with operands rounded to BF16 and exact (float64) products
    and sums, bias / ReLU / softmax in float64."""
    p = unpack(np.asarray(theta, np.float32))
    h = np.asarray(X, np.float32)
    for i, (w, b) in enumerate((("W1", "b1"), ("W2", "b2"), ("W3", "b3"))):
        z = bf16_round(h).astype(np.float64) @ bf16_round(p[w].astype(np.float32)).astype(np.float64)
        z = z + p[b]
        h = (np.maximum(z, 0.0) if i < 2 else z).astype(np.float32)
    return softmax(h.astype(np.float64))
