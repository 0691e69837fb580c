"""The learned grid-1 predictor on the GPU: network kernels against the CPU
oracle (forward pass and parameter gradient, the oracle emulating the GPU's
bfloat16 storage; the oracle itself is pinned by finite differences in
test_cnn_oracle.py), the Adam step against numpy, the device renderer of
synthetic pairs, training on GPU-labelled pairs and what the trained network
gains over an untrained one, and the outlier-fraction fallback to the plain
grid-1 passes.

Bars (per assertion): forward outputs within 1e-2 relative L2 of the
bfloat16-storage oracle (accumulation order differs, and a value close to a
bfloat16 rounding boundary can round the other way); each layer's gradient
within 3e-2 relative L2 and cosine similarity >= 0.999; the Adam step within
1e-6 relative (plus 1e-5 of the learning rate) of numpy's FP32 update.
"""
import numpy as np
import pytest

from test_cnn_oracle import ARCH, he_params, offsets

pytestmark = pytest.mark.gpu


def _report(name, **kw):
    print(f"[learned] {name}: " + ", ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _bf16_round(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("h,w", [(64, 128), (64, 256), (256, 512)])
def test_cnn_forward_and_gradient_match_oracle(pivx, orc, h, w):
    """Forward pass and parameter gradient against the FP64 oracle; w = 256
    gives the stride-2 data-gradient launches four 16-pixel tiles per wave
    (layer 1) and two (layer 2), w = 128 two and one; 256 x 512 takes the
    LDS-staged weight gradients of the hidden layers (whole 64-pixel blocks
    of an output row)."""
    rng = np.random.default_rng(21)
    p = he_params(orc, rng).astype(np.float32)
    x = rng.normal(0.0, 1.0, (2, h, w, 2)).astype(np.float32)
    y = rng.normal(0.0, 3.0, (2, 2, (h // 8) * (w // 8))).astype(np.float32)
    with pivx.Cnn(0) as net:
        net.set_params(p)
        np.testing.assert_array_equal(net.params(), p)
        out = net.forward(x)
        loss, g = net.gradient(x, y)
    p64 = p.astype(np.float64)
    ref_out = orc.cnn_forward(p64, x, bf16=True)
    rel_out = float(np.linalg.norm(out - ref_out) / np.linalg.norm(ref_out))
    ref_loss, ref_g = orc.cnn_grad(p64, x, y, bf16=True)
    _report("forward", rel_l2=f"{rel_out:.2e}", max_abs=f"{float(np.abs(out - ref_out).max()):.2e}")
    _report("loss", gpu=f"{loss:.6f}", oracle=f"{ref_loss:.6f}")
    assert rel_out < 1e-2
    assert abs(loss - ref_loss) / ref_loss < 1e-2
    for li, (ow, ob, nw, nb) in enumerate(offsets()):
        for name, sl in (("W", slice(ow, ow + nw)), ("b", slice(ob, ob + nb))):
            a, r = g[sl].astype(np.float64), ref_g[sl]
            rel = float(np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-30))
            cos = float(a @ r / max(np.linalg.norm(a) * np.linalg.norm(r), 1e-30))
            _report(f"gradient layer {li} {name}", rel_l2=f"{rel:.2e}", cosine=f"{cos:.6f}")
            assert rel < 3e-2, (li, name)
            assert cos > 0.999, (li, name)


def test_cnn_forward_closed_form(pivx, orc):
    """Centre-tap copy kernels: the output is input pixel (8y, 8x) of channel
    0 (times 0.1^5 where negative), after bfloat16 rounding of the input."""
    n = pivx.Cnn(0).n
    p = np.zeros(n, np.float32)
    for li, ((cin, cout, _), (ow, ob, nw, nb)) in enumerate(zip(ARCH, offsets())):
        w = np.zeros((cout, 9, cin), np.float32)
        w[0, 4, 0] = 1.0
        if li == 5:
            w[1, 4, 0] = 1.0
        p[ow:ow + nw] = w.ravel()
    rng = np.random.default_rng(4)
    x = rng.normal(0.0, 1.0, (1, 32, 256, 2)).astype(np.float32)
    with pivx.Cnn(0) as net:
        net.set_params(p)
        out = net.forward(x)
    src = _bf16_round(x)[:, ::8, ::8, 0].reshape(1, -1).astype(np.float64)
    want = np.where(src > 0, src, src * 0.1 ** 5)
    # positive values pass through exactly; negative ones are scaled and rounded five times
    pos = src > 0
    np.testing.assert_array_equal(out[:, 0][pos], want[pos].astype(np.float32))
    np.testing.assert_allclose(out[:, 0][~pos], want[~pos], rtol=2e-2)
    np.testing.assert_array_equal(out[:, 0], out[:, 1])


def test_adam_step_matches_numpy(pivx, orc):
    rng = np.random.default_rng(8)
    p = he_params(orc, rng).astype(np.float32)
    x = rng.normal(0.0, 1.0, (1, 32, 128, 2)).astype(np.float32)
    y = rng.normal(0.0, 3.0, (1, 2, 4 * 16)).astype(np.float32)
    lr = np.float32(1e-3)
    with pivx.Cnn(0) as net:
        net.set_params(p)
        m = np.zeros_like(p)
        v = np.zeros_like(p)
        q = p.copy()
        for t in (1, 2):
            _, g = net.gradient(x, y)
            net.adam_step(float(lr))
            m = np.float32(0.9) * m + np.float32(0.1) * g
            v = np.float32(0.999) * v + np.float32(0.001) * g * g
            c1 = np.float32(1 - 0.9 ** t)
            c2 = np.float32(1 - 0.999 ** t)
            q = q - lr * (m / c1) / (np.sqrt(v / c2) + np.float32(1e-8))
            got = net.params()
            # FP32 rounding of the update itself: 1e-5 of a step of size lr
            np.testing.assert_allclose(got, q, rtol=1e-6, atol=1e-5 * float(lr))
        assert np.abs(got - p).max() > 1e-4  # the parameters moved


def test_device_synthetic_pair_follows_its_flow(pivx):
    """Frames rendered on the GPU carry the analytic flow: the plain two-grid
    pipeline recovers it (random Fourier modes, max 12 px), two renders are
    identical, and the intensity statistics match the host generator's."""
    nx, ny = 1024, 1024
    sp = pivx.synth_params(nx, ny, seed=4242, flow=2, a0=12.0)
    p = pivx.init_params(nx=nx, ny=ny, ws_x1=64, ws_y1=64, spacing_x1=32, spacing_y1=32, npass1=3, ws_x2=32,
                         ws_y2=32, spacing_x2=8, spacing_y2=8, npass2=3)
    fp = pivx.FilterParams(2, 5, 2.0, 2.0, 1)
    host_a, host_b = pivx.synth_pair(sp)
    with pivx.Xcorr2dMulti(p) as x:
        x.synth_pair_device(sp)
        x.run_pair(fp, fp)
        fx, fy, fv = x.filtered(2)
        gx, gy = x.grid_coords(2)
        r1 = x.raw(1)
        x.synth_pair_device(sp)
        x.run_pair(fp, fp)
        r2 = x.raw(1)
    for a, b in zip(r1, r2):
        np.testing.assert_array_equal(a, b)
    u, v = pivx.synth_flow(sp, gx, gy)
    inner = (gx > 48) & (gx < nx - 48) & (gy > 48) & (gy < ny - 48)
    rms = float(np.sqrt(np.mean((fx[inner] - u[inner]) ** 2 + (fy[inner] - v[inner]) ** 2)))
    _report("device synthetic pair", rms_px=f"{rms:.4f}", valid=f"{float(fv[inner].mean()):.4f}",
            host_mean=f"{float(host_a.mean()):.2f}")
    assert rms < 0.1
    assert float(fv[inner].mean()) > 0.99


def _learned_setup(pivx, nx=2048, ny=2048):
    p = pivx.init_params(nx=nx, ny=ny, ws_x1=64, ws_y1=64, spacing_x1=32, spacing_y1=32, npass1=3, ws_x2=32,
                         ws_y2=32, spacing_x2=8, spacing_y2=8, npass2=3)
    fp = pivx.FilterParams(2, 5, 2.0, 2.0, 1)
    return p, fp


def _rms(pivx, sp, x):
    fx, fy, _ = x.filtered(2)
    gx, gy = x.grid_coords(2)
    u, v = pivx.synth_flow(sp, gx, gy)
    inner = (gx > 48) & (gx < sp.nx - 48) & (gy > 48) & (gy < sp.ny - 48)
    return float(np.sqrt(np.mean((fx[inner] - u[inner]) ** 2 + (fy[inner] - v[inner]) ** 2)))


def test_training_gains_over_untrained_network(pivx):
    """A short training run (512 GPU-rendered 2048x2048 pairs, Fourier flows up
    to 12 px, labels from the plain grid-1 passes, 8 epochs) lowers the loss,
    and on 6 held-out pairs the trained network's grid-1 field is far closer
    to the plain grid-1 field than an untrained network's; with the fallback
    off, the learned pipeline's final grid-2 field then stays within 0.02 px
    RMS of the plain pipeline's error against the analytic flow."""
    p, fp = _learned_setup(pivx)
    synth = pivx.synth_params(2048, 2048, seed=100000, flow=2, a0=12.0)
    tp = pivx.CnnTrainParams(npairs=512, epochs=8, batch=8, lr=2e-3, shuffle_seed=7, synth=synth, grid1=fp)
    with pivx.Xcorr2dMulti(p) as x, pivx.Cnn(0, seed=1) as net, pivx.Cnn(0, seed=1) as raw:
        losses, rep = net.train(x, tp)
        _report("training", pairs=tp.npairs, epochs=tp.epochs, gen_s=f"{rep.gen_seconds:.2f}",
                train_s=f"{rep.train_seconds:.2f}", losses=[round(float(v), 3) for v in losses])
        assert losses[-1] < 0.5 * losses[0]
        err_t, err_u, d_final = [], [], []
        for i in range(6):
            sp = pivx.synth_params(2048, 2048, seed=900000 + i, flow=2, a0=12.0)
            x.synth_pair_device(sp)
            x.run_pair(fp, fp)
            plain1 = x.filtered(1)
            rms_plain = _rms(pivx, sp, x)
            for k, (n, errs) in enumerate(((net, err_t), (raw, err_u))):
                x.run_pair_learned(n, fp, fp, max_outlier_fraction=1.0)
                nx_, ny_, _ = x.filtered(1)
                # grid-1 field after the pair = the network field (no fallback)
                errs.append(float(np.sqrt(np.mean((nx_ - plain1[0]) ** 2 + (ny_ - plain1[1]) ** 2))))
                if k == 0:
                    d_final.append(_rms(pivx, sp, x) - rms_plain)
        _report("held-out", trained_vs_plain_grid1_rms=[round(v, 3) for v in err_t],
                untrained_vs_plain_grid1_rms=[round(v, 3) for v in err_u],
                final_rms_minus_plain=[round(v, 4) for v in d_final])
        assert np.mean(err_t) < 0.5 * np.mean(err_u)
        assert max(d_final) <= 0.02


def test_fallback_runs_plain_grid1(pivx, orc):
    """A network with random weights gives a noisy field that fails the
    normalised median test in more than 0.5% of the vectors: the plain grid-1
    passes run instead and the pair's results are bit-identical to
    pivx_run_pair, eagerly and as a replayed graph; with the threshold at 100%
    the network field is kept (no fallback)."""
    p, fp = _learned_setup(pivx, 2048, 1024)
    sp = pivx.synth_params(2048, 1024, seed=31337, flow=2, a0=8.0)
    with pivx.Xcorr2dMulti(p) as x, pivx.Cnn(0) as net:
        net.set_params(he_params(orc, np.random.default_rng(12)).astype(np.float32))
        x.synth_pair_device(sp)
        x.run_pair(fp, fp)
        want = [x.raw(1), x.filtered(1), x.raw(2), x.filtered(2), x.filtered(3)]
        for graph in (False, True):
            x.set_graph_mode(graph)
            for _ in range(2):
                x.run_pair_learned(net, fp, fp, max_outlier_fraction=0.005)
                outl, fb = x.learned_stats()
                got = [x.raw(1), x.filtered(1), x.raw(2), x.filtered(2), x.filtered(3)]
                _report("fallback", graph=graph, outliers=outl, fallback=fb)
                assert fb == 1 and outl > 0.005 * 64 * 32
                for A, B in zip(want, got):
                    for a, b in zip(A, B):
                        np.testing.assert_array_equal(a, b)
            x.run_pair_learned(net, fp, fp, max_outlier_fraction=1.0)
            outl, fb = x.learned_stats()
            assert fb == 0
            assert not np.array_equal(x.filtered(1)[0], want[1][0])  # the network field was kept
            # the gate's count (set by the median launch's last block) equals the
            # vectors the FP32 median test rejects in the network field
            fx, fy, fv = x.filtered(1)
            _, _, dv, _ = orc.median_pass((fx, fy, fv), (fx, fy, fv), 1, (64, 32), (p.wx1_median, p.wy1_median),
                                          fp.min_nneigh, fp.insert_thr, prec=32)
            _report("gate count", graph=graph, gpu=outl, oracle=int((dv == 0).sum()))
            assert outl == int((dv == 0).sum())


def _net_input(img):
    """The network input plane of one frame on the host: 4x4 block means,
    normalised as (mean - 30) / 40 in FP32, rounded to bfloat16."""
    ny, nx = img.shape
    s = img.reshape(ny // 4, 4, nx // 4, 4).astype(np.int64).sum(axis=(1, 3))
    f = (s.astype(np.float32) * np.float32(1.0 / 16.0) - np.float32(30.0)) * np.float32(1.0 / 40.0)
    return _bf16_round(f)


def test_network_input_formed_from_frames(pivx):
    """The learned pipeline's first layer forms the network input from the
    frames as it reads them (no downsampled plane), and for one pair the last
    three layers run as one launch (k_conv_tail): the grid-1 field it produces
    equals pivx_cnn_forward on the same input formed on the host (4x4 block
    means of both frames, normalised, rounded to bfloat16), bit for bit, both
    for that one input (the same fused path) and for a batch of two copies of
    it (every layer its own launch)."""
    p, fp = _learned_setup(pivx, 2048, 1024)
    sp = pivx.synth_params(2048, 1024, seed=4242, flow=2, a0=8.0)
    a, b = pivx.synth_pair(sp)
    with pivx.Xcorr2dMulti(p) as x, pivx.Cnn(0, seed=3) as net:
        x.prep_img_data(a, b)
        x.run_pair_learned(net, fp, fp, max_outlier_fraction=1.0)
        fx, fy, fv = x.filtered(1)
        x0 = np.stack([_net_input(a), _net_input(b)], axis=-1)[None]
        out = net.forward(x0)
        out2 = net.forward(np.concatenate([x0, x0]))
    _report("network input from frames", max_abs_dx=float(np.abs(fx - out[0, 0]).max()),
            max_abs_dy=float(np.abs(fy - out[0, 1]).max()),
            layerwise_max_abs=float(np.abs(out2[0] - out[0]).max()))
    np.testing.assert_array_equal(fx, out[0, 0])
    np.testing.assert_array_equal(fy, out[0, 1])
    np.testing.assert_array_equal(out2[0], out[0])
    np.testing.assert_array_equal(out2[1], out[0])
    assert fv.all()
