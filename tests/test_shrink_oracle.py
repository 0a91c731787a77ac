"""Pins of the learned-shrinkage restatement (oracle/shrink_oracle.py), CPU only.

What each pin would catch: a transposed DCT or a wrong Haar layout / DC index
(Wiener reproduction), a wrong tap direction or a dropped term in the backward
pass (finite differences), a wrong rounding mode (torch's BF16 conversion), a
wrong bias correction in Adam (its closed-form first step)."""
import numpy as np
import pytest

import oracle
from oracle import shrink_oracle as so


def test_parameter_layout_size():
    assert so.NPARAM == 49600
    from gbm3d_amd._abi import GBM3D_SHRINK_PARAMS
    from gbm3d_amd import shrink

    assert so.NPARAM == GBM3D_SHRINK_PARAMS
    for l in range(so.NL):
        assert shrink._off_w(l) == so.off_w(l) and shrink._off_b(l) == so.off_b(l)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 4000), rng.normal(0, 1e4, 4000), [0.0, -0.0, 1.0, 3.0]]).astype(np.float32)
    # exact ties: low 16 bits 0x8000 with even and odd kept parts
    ties = (np.arange(64, dtype=np.uint32) << 16 | 0x3F800000 | 0x8000).view(np.float32)
    x = np.concatenate([x, ties, -ties])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(so.bf16(x), ref)


def test_transform_is_orthonormal_and_invertible():
    rng = np.random.default_rng(1)
    V = rng.normal(0, 30, (8, 8, 16, 5))
    c = so.coefficients(V)
    assert c.shape == (5, 64, 16)
    assert np.allclose((c ** 2).sum(axis=(1, 2)), (V ** 2).sum(axis=(0, 1, 2)), rtol=1e-12)
    assert np.allclose(so.inverse(c), V, atol=1e-10)
    # a constant group has only the DC approximation (t = 0, k = 14, 15)
    cc = so.coefficients(np.full((8, 8, 16, 1), 2.0))
    nz = np.argwhere(np.abs(cc[0]) > 1e-9)
    assert set(map(tuple, nz)) <= {(0, 14), (0, 15)}


@pytest.mark.parametrize("sigma", [5.0, 25.0])
def test_transform_reproduces_wiener(sigma):
    """Wiener gains applied with these coefficients give oracle.wiener()."""
    rng = np.random.default_rng(2)
    V1 = rng.normal(100, 40, (8, 8, 16, 6))
    V2 = V1 + rng.normal(0, sigma, V1.shape)
    c1, c2 = so.coefficients(V1), so.coefficients(V2)
    f = c1 ** 2 / (c1 ** 2 + sigma ** 2)
    f[:, 0, 14:] = 1.0
    out = np.maximum(so.inverse(c2 * f), 0)
    ref, g2 = oracle.wiener(V1, V2, sigma, precision=64)
    assert np.allclose(out, ref.reshape(out.shape, order="F"), atol=1e-9)
    assert np.allclose((f ** 2).sum(axis=(1, 2)), np.ravel(g2, order="F"), rtol=1e-12)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = so.init_params(4).astype(np.float64) * 0.7
    G = 3
    X = rng.normal(0, 2, (G, 16, 64))
    T = rng.uniform(0, 1, (G, 16, 64))
    Wt = rng.uniform(0.5, 3, (G, 16, 64))
    sig = np.array([10.0, 25.0, 40.0])
    _, g = so.loss_and_grad(p, X, T, Wt, sig, exact=True)
    idx = np.concatenate([rng.choice(so.KW * so.CH, 10, replace=False) + so.off_w(l) for l in range(so.NL)]
                         + [np.arange(so.OFF_WSIG, so.OFF_WSIG + 192, 37), so.off_b(0) + np.arange(0, 64, 21),
                            so.off_b(3) + np.arange(0, 64, 17)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = so.loss_and_grad(pp, X, T, Wt, sig, exact=True)
        lm, _ = so.loss_and_grad(pm, X, T, Wt, sig, exact=True)
        num = (lp - lm) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(num), (i, num, g[i])


def test_rounded_gradient_close_to_exact():
    rng = np.random.default_rng(5)
    p = so.init_params(6)
    X = rng.normal(0, 2, (8, 16, 64)).astype(np.float32)
    T = rng.uniform(0, 1, X.shape).astype(np.float32)
    Wt = np.ones_like(X)
    sig = np.full(8, 20.0, np.float32)
    l1, g1 = so.loss_and_grad(p, X, T, Wt, sig)
    l2, g2 = so.loss_and_grad(p.astype(np.float64), X, T, Wt, sig, exact=True)
    assert abs(l1 - l2) < 2e-2 * l2
    assert np.linalg.norm(g1 - g2) < 5e-2 * np.linalg.norm(g2)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(7)
    p = rng.normal(0, 1, 100).astype(np.float32)
    g = rng.normal(0, 1e-3, 100).astype(np.float32)
    z = np.zeros_like(p)
    p1, m, v = so.adam(p, z, z, g, 1, 1e-2)
    # bias-corrected first step: m = (1-b1) g, v = (1-b2) g^2, so the update is
    # -lr g / (|g| + eps / sqrt(1 - b2))
    assert np.allclose(p1, p - 1e-2 * g / (np.abs(g) + 1e-8 / np.sqrt(1e-3)), rtol=1e-5, atol=1e-7)


def test_features_targets_and_weights():
    rng = np.random.default_rng(8)
    Vc = rng.uniform(0, 255, (8, 8, 16, 4))
    Vn = Vc + rng.normal(0, 25, Vc.shape)
    X, T, Wt = so.features(Vc, Vn, 25.0)
    cc = so.coefficients(Vc).transpose(0, 2, 1)
    cn = so.coefficients(Vn).transpose(0, 2, 1)
    assert np.allclose(X, cn / 25.0, rtol=1e-5, atol=1e-5)
    assert np.allclose(T, cc ** 2 / (cc ** 2 + 625.0), rtol=1e-5, atol=1e-6)
    assert np.all(Wt[:, 14:, 0] == 0) and np.all(Wt[:, :14, :] >= 1) and Wt.max() <= 100
    # energy-matched targets: |T y| = |c| where |c| <= |y|, else 1; weight y^2/sigma^2
    X, T, Wt = so.features(Vc, Vn, 25.0, target="energy")
    y = cn
    assert np.allclose(T, np.minimum(np.abs(cc) / np.abs(y), 1), rtol=1e-4, atol=1e-6)
    assert np.allclose(Wt[:, :14, :], np.minimum(y[:, :14, :] ** 2 / 625.0, 100), rtol=1e-4, atol=1e-6)
    assert np.all(Wt[:, 14:, 0] == 0)
