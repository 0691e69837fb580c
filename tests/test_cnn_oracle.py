"""The CPU oracle of the learned grid-1 predictor, pinned without the GPU:
its loss gradient against central finite differences of its own FP64 loss
(every layer's weights and biases), its forward pass against a closed form
(centre-tap copy kernels carry input pixel (8y, 8x) through all six layers),
and its bfloat16-storage mode against the FP64 mode within bfloat16
rounding."""
import numpy as np
import pytest

ARCH = [(2, 32, 2), (32, 64, 2), (64, 64, 2), (64, 64, 1), (64, 64, 1), (64, 2, 1)]


def offsets():
    o, out = 0, []
    for cin, cout, _ in ARCH:
        out.append((o, o + cout * 9 * cin, cout * 9 * cin, cout))
        o += cout * 9 * cin + cout
    return out


def he_params(orc, rng, scale=1.0):
    p = np.zeros(orc.cnn_param_count())
    for (cin, cout, _), (ow, ob, nw, nb) in zip(ARCH, offsets()):
        p[ow:ow + nw] = rng.normal(0.0, scale * np.sqrt(2.0 / (9 * cin)), nw)
        p[ob:ob + nb] = rng.normal(0.0, 0.05, nb)
    return p


def test_param_count(orc):
    assert orc.cnn_param_count() == sum(cout * 9 * cin + cout for cin, cout, _ in ARCH) == 131042


def test_cnn_gradient_matches_finite_differences(orc):
    rng = np.random.default_rng(11)
    p = he_params(orc, rng)
    x = rng.normal(0.0, 1.0, (2, 32, 32, 2))
    y = rng.normal(0.0, 2.0, (2, 2, 16))
    loss, g = orc.cnn_grad(p, x, y)
    assert loss > 0
    h = 1e-6
    worst = 0.0
    for ow, ob, nw, nb in offsets():
        for i in list(rng.choice(nw, 6, replace=False) + ow) + list(rng.choice(nb, 2, replace=False) + ob):
            q = p.copy()
            q[i] += h
            lp, _ = orc.cnn_grad(q, x, y)
            q[i] -= 2 * h
            lm, _ = orc.cnn_grad(q, x, y)
            fd = (lp - lm) / (2 * h)
            err = abs(fd - g[i]) / max(abs(fd), abs(g[i]), 1e-6)
            worst = max(worst, err)
            assert err < 1e-4, (i, fd, g[i])
    print(f"[cnn oracle] worst relative finite-difference error {worst:.2e}")


def test_cnn_forward_centre_tap_copy(orc):
    """Weight 1 on the centre tap from channel 0 to channel 0 in every
    layer (and to channel 1 in the last), zero elsewhere: the output is input
    pixel (8y, 8x) of channel 0, times 0.1^5 where it is negative (five leaky
    ReLUs)."""
    rng = np.random.default_rng(3)
    p = np.zeros(orc.cnn_param_count())
    for li, ((cin, cout, _), (ow, ob, nw, nb)) in enumerate(zip(ARCH, offsets())):
        w = np.zeros((cout, 9, cin))
        w[0, 4, 0] = 1.0
        if li == 5:
            w[1, 4, 0] = 1.0
        p[ow:ow + nw] = w.ravel()
    x = rng.normal(0.0, 1.0, (2, 64, 48, 2))
    out = orc.cnn_forward(p, x)
    src = x[:, ::8, ::8, 0].reshape(2, -1)
    want = np.where(src > 0, src, src * 0.1 ** 5)
    np.testing.assert_allclose(out[:, 0], want, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(out[:, 1], want, rtol=1e-12, atol=1e-15)


def test_cnn_bf16_mode_is_within_bf16_rounding(orc):
    rng = np.random.default_rng(5)
    p = he_params(orc, rng)
    x = rng.normal(0.0, 1.0, (2, 32, 64, 2))
    y = rng.normal(0.0, 2.0, (2, 2, 32))
    o64 = orc.cnn_forward(p, x)
    o16 = orc.cnn_forward(p, x, bf16=True)
    rel = np.linalg.norm(o16 - o64) / np.linalg.norm(o64)
    l64, g64 = orc.cnn_grad(p, x, y)
    l16, g16 = orc.cnn_grad(p, x, y, bf16=True)
    grel = np.linalg.norm(g16 - g64) / np.linalg.norm(g64)
    print(f"[cnn oracle] bf16 storage vs FP64: output {rel:.2e}, gradient {grel:.2e}")
    assert 1e-4 < rel < 2e-2  # bfloat16 keeps 8 significant bits
    assert grel < 5e-2
    assert abs(l16 - l64) / l64 < 2e-2
    with pytest.raises(AssertionError):
        orc.cnn_forward(p, x[:, :30])  # spatial sizes must be multiples of 8
