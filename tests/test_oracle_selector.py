"""Pins of the selector oracle (features and MLP) on the CPU: closed forms of
the features on matrices whose structure is known, central differences for
the backpropagated gradient, and the softmax cross-entropy at uniform logits."""
import math

import numpy as np
import pytest

from oracle import selector as S


def _stencil2d(n):
    """5-point stencil CSR on an n x n grid, ascending columns."""
    rp, ci = [0], []
    for i in range(n):
        for j in range(n):
            r = i * n + j
            nb = [r - n] * (i > 0) + [r - 1] * (j > 0) + [r] + [r + 1] * (j < n - 1) + [r + n] * (i < n - 1)
            ci += nb
            rp.append(len(ci))
    return np.array(rp, np.uint32), np.array(ci, np.uint32)


def test_param_count():
    assert S.NPARAM == 25864 == 64 * 128 + 128 + 128 * 128 + 128 + 128 * 8 + 8


def test_identity_features():
    n = 1000
    f = S.features(np.arange(n + 1), np.arange(n), n, n)
    assert f[3] == pytest.approx(math.log2(2) / 16)
    assert f[5] == 0 and f[6] == 0 and f[7] == 0
    assert f[9] == 1.0 and f[8:32].sum() == 1.0  # every row has length 1 -> bin 1
    assert f[32] == 1.0                           # all entries in rows of floor(log2 len) = 0
    assert f[48] == 1.0 and f[56] == 0.0 and f[57] == 1.0
    assert f[58] == pytest.approx(1 / n)
    assert f[59] == 0.0                           # no ELL padding
    assert f[61] == 1.0 and f[60] == 0 and f[62] == 0 and f[63] == 1


def test_stencil_features_closed_form():
    n = 40
    rp, ci = _stencil2d(n)
    N = n * n
    nnz = 5 * N - 4 * n
    assert rp[-1] == nnz
    f = S.features(rp, ci, N, N)
    # corners have 3 entries, other edge rows 4, interior rows 5
    assert f[8 + 2] == pytest.approx(4 / N)
    assert f[8 + 3] == pytest.approx((N - 4) / N)
    assert f[32 + 1] == pytest.approx(12 / nnz)
    assert f[32 + 2] == pytest.approx((nnz - 12) / nnz)
    mean = nnz / N
    sumsq = 4 * 9 + 4 * (n - 2) * 16 + (n - 2) ** 2 * 25
    sd = math.sqrt(sumsq / N - mean * mean)
    assert f[5] == pytest.approx(sd / mean / 32)
    assert f[7] == pytest.approx(math.log2(1 + sd) / 16)
    assert f[57] == pytest.approx(N / nnz)
    # |c - r| is 0, 1 or n; relative to cols: 0, 1/N, n/N
    d1, dn = 1 / N, n / N
    th = 1e-4 * 4.0 ** np.arange(7)
    k1, kn = int(np.searchsorted(th, d1, "right")), int(np.searchsorted(th, dn, "right"))
    n1, nn = 2 * n * (n - 1), 2 * n * (n - 1)  # horizontal / vertical neighbour pairs, both directions
    exp = np.zeros(8)
    exp[0] += N
    exp[k1] += n1
    exp[kn] += nn
    assert np.allclose(f[48:56], exp / nnz)
    # spans: last - first + 1 per row, from the explicit rows
    spans = [int(ci[rp[r + 1] - 1]) - int(ci[rp[r]]) + 1 for r in range(N)]
    assert f[58] == pytest.approx(np.mean(spans) / N)
    assert f[59] == pytest.approx(math.log2(N * 5 / nnz) / 16)


def test_empty_rows_and_long_rows():
    # 8 rows: lengths 0, 0, 5000, 1, 3, 64, 65, 0 over 10000 columns
    lens = [0, 0, 5000, 1, 3, 64, 65, 0]
    rp = np.concatenate([[0], np.cumsum(lens)])
    ci = np.concatenate([np.arange(L) for L in lens]).astype(np.int64)
    f = S.features(rp, ci, 8, 10000)
    nnz = sum(lens)
    assert f[6] == 3 / 8 and f[8] == 3 / 8
    assert f[8 + 13] == 1 / 8           # 5000: floor(log2) = 12 -> bin 13
    assert f[32 + 12] == pytest.approx(5000 / nnz)
    assert f[60] == pytest.approx(5000 / nnz)
    assert f[61] == pytest.approx(4 / nnz)
    assert f[62] == 2 / 8               # 5000 and 65
    assert f[4] == pytest.approx(math.log2(5001) / 16)


def _tiny_problem(seed, n=5):
    rng = np.random.default_rng(seed)
    theta = rng.normal(0, 0.2, S.NPARAM)
    X = rng.uniform(0, 1, (n, 64))
    P = rng.dirichlet(np.ones(8), n)
    return theta, X, P


@pytest.mark.parametrize("seed", [0, 1])
def test_backprop_matches_central_differences(seed):
    theta, X, P = _tiny_problem(seed)
    _, g = S.loss_grad(theta, X, P)
    rng = np.random.default_rng(10 + seed)
    # indices in every parameter block
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in S.SHAPES])
    idx = np.concatenate([rng.integers(offs[i], offs[i + 1], 6) for i in range(len(S.SHAPES))])
    num = S.numeric_grad(theta, X, P, idx)
    assert np.allclose(g[idx], num, rtol=1e-5, atol=1e-8)


def test_a_dropped_relu_mask_is_caught():
    """The check above must fail for a plausible backprop mistake."""
    theta, X, P = _tiny_problem(3)
    p = S.unpack(theta)
    H1, H2, Z = S.forward(theta, X)
    dZ = (S.softmax(Z) - P) / X.shape[0]
    dH2 = dZ @ p["W3"].T  # mask missing
    wrong_W2 = H1.T @ dH2
    off = 64 * 128 + 128
    idx = np.arange(off, off + 50)
    num = S.numeric_grad(theta, X, P, idx)
    assert not np.allclose(wrong_W2.ravel()[:50], num, rtol=1e-5, atol=1e-8)


def test_uniform_logits_closed_form():
    theta = np.zeros(S.NPARAM)
    X = np.random.default_rng(0).uniform(0, 1, (7, 64))
    P = np.eye(8)[np.arange(7)]
    loss, g = S.loss_grad(theta, X, P)
    assert loss == pytest.approx(math.log(8))
    gb3 = g[-8:]
    assert np.allclose(gb3, (1 / 8 - P).mean(axis=0))
    assert np.all(g[:-8] == 0)  # zero weights: nothing flows below the output bias


def test_adam_first_step_is_sign_times_lr():
    g = np.array([0.3, -2.0, 1e-3])
    t, m, v = S.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 0.01, eps=0.0)
    assert np.allclose(t, -0.01 * np.sign(g))


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    a = np.concatenate([rng.normal(0, 10, 10000), rng.uniform(-1, 1, 1000) * 1e-38,
                        [0.0, -0.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 3.0e38]]).astype(np.float32)
    ref = torch.from_numpy(a).to(torch.bfloat16).float().numpy()
    assert np.array_equal(S.bf16_round(a).view(np.uint32), ref.view(np.uint32))
