"""The learned gather on the GPU: the MFMA GEMM against a float64 product, the MLP's
GPU gradients against the CPU oracle's back-propagation, training, and the
learned+fallback bake against the plain 1024-ray bake on held-out scenes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bf16(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def _from_bf16(b):
    return (b.astype(np.uint32) << 16).view(np.float32)


@pytest.mark.parametrize("shape", [(16, 16, 16), (48, 32, 64), (64, 16, 128)])
def test_mfma_gemm_matches_float64(gpu, shape):
    from lightsprintsdk_amd import native as N
    from lightsprintsdk_amd.collider import DeviceBuffer
    m, n, k = shape
    rng = np.random.default_rng(m * 7 + n)
    a = _bf16(rng.normal(size=(m, k)))
    b = _bf16(rng.normal(size=(k, n)))
    da, db, dc = DeviceBuffer.from_array(a), DeviceBuffer.from_array(b), DeviceBuffer(4 * m * n)
    N.check(N.lib().lsg_mlp_gemm_test(da.ptr, db.ptr, dc.ptr, m, n, k), "gemm")
    c = dc.download(np.float32, (m, n))
    ref = _from_bf16(a).astype(np.float64) @ _from_bf16(b).astype(np.float64)
    np.testing.assert_allclose(c, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())


def test_mlp_gradients_match_oracle(gpu, oracle):
    """One backward pass on the GPU (BF16 MFMA, FP32 accumulation) against the oracle's
    binary64 back-propagation rounded to BF16 at the same points."""
    from lightsprintsdk_amd.learned import Surrogate
    from lightsprintsdk_amd.collider import DeviceBuffer
    rng = np.random.default_rng(5)
    net = Surrogate(max_batch=256, seed=3)
    p = net.init_params(3)
    p[-16 * 65:] += (rng.normal(size=16 * 65) * 0.05).astype(np.float32)   # make the last layer non-trivial
    net.set_params(p)
    p = net.get_params()
    n = 96
    x = rng.normal(size=(n, 32)).astype(np.float32)
    t = rng.normal(size=(n, 3)).astype(np.float32)
    dx, dt = DeviceBuffer.from_array(x), DeviceBuffer.from_array(t)
    loss = net.step(dx, dt, n, lr=0.0, want_loss=True)
    g = net.get_grads()
    lo, y, go = oracle.mlp_loss_grad(p, x, t, lam=0.5, bf16=True)
    assert abs(loss - lo) <= 2e-3 * abs(lo)
    np.testing.assert_allclose(net.predict(x), y, rtol=2e-2, atol=2e-3)
    rel = np.linalg.norm(g - go) / np.linalg.norm(go)
    cos = float(g @ go / (np.linalg.norm(g) * np.linalg.norm(go)))
    print(f"grad rel diff {rel:.2e}, cosine {cos:.6f}", flush=True)
    assert rel < 2e-2 and cos > 0.9995
    # per layer as well (each weight block and bias)
    o = 0
    for (i, j) in Surrogate.layer_shapes():
        for blk in (i * j, j):
            gg, oo = g[o:o + blk], go[o:o + blk]
            if np.linalg.norm(oo) > 0:
                assert np.linalg.norm(gg - oo) / np.linalg.norm(oo) < 5e-2
            o += blk


def test_mlp_training_converges(gpu):
    from lightsprintsdk_amd.learned import Surrogate
    rng = np.random.default_rng(9)
    n = 65536
    x = rng.uniform(-1, 1, size=(n, 32)).astype(np.float32)
    x[:, 31] = 1
    t = np.stack([np.sin(2 * x[:, 0]) + x[:, 1] ** 2, np.cos(3 * x[:, 2]) * x[:, 3], x[:, 4] * x[:, 5]], 1).astype(np.float32)
    net = Surrogate(max_batch=8192, seed=1)
    hist = net.train(x, t, epochs=30, batch=1024, lr=3e-3)
    assert hist[-1] < 0.15 * hist[0], hist


def test_learned_gather_heldout(gpu):
    """Train on seeded box-room scenes, pick the error bound on one more scene,
    then on a held-out scene the learned+fallback bake stays within 2% RMSE of the
    plain 1024-ray bake (filtered lightmaps, as updateLightmaps writes them) while
    re-gathering only part of the texels, and the network alone does better than the
    cheap gather it starts from (the same pipeline without the network)."""
    from lightsprintsdk_amd import learned as Lr
    X, T = Lr.build_dataset(range(3000, 3004))
    net = Lr.Surrogate(max_batch=16384, seed=0)
    hist = net.train(X, T, epochs=6, batch=8192, lr=2e-3)
    del X, T
    print("train loss", [round(h, 5) for h in hist], flush=True)
    assert hist[-1] < 0.2 * hist[0]
    cal = [Lr.prepare_eval(3100), Lr.prepare_eval(3101)]
    target = Lr.calibrate_target(net, cal, target_rmse=0.015)
    ls = Lr.prepare_eval(3200)
    ls.cheap_only()
    r_cheap = Lr.rmse_rel(ls.final_rgb(), ls.ref_final)
    r_net, f_unrel = Lr.evaluate(ls, net, tau=float("inf"))
    r_final, frac = Lr.evaluate(ls, net, target=target)
    print(f"held-out: target {target} (tau {ls.last_tau:.4g}, estimated network rmse {ls.last_est_rmse:.4f}) "
          f"cheap {r_cheap:.4f} network {r_net:.4f} (unreliable {f_unrel:.3f}) "
          f"learned+fallback {r_final:.4f} fallback {frac:.3f}", flush=True)
    assert r_net < 0.8 * r_cheap
    assert r_final <= 0.02
    assert frac < 0.5


def test_committed_surrogates_heldout(gpu):
    """The committed surrogates on a held-out scene of their family each (the box room
    and a 1/16-area city): learned+fallback within 2% RMSE of the plain bake, and the
    network alone better than the cheap gather."""
    from lightsprintsdk_amd import learned as Lr
    for family, seed in (("box_room", 2004), ("city", 6001)):
        got = Lr.load_weights(Lr.weights_path(family))
        assert got is not None, f"no committed surrogate for {family}"
        net, bound, meta = got
        ls = Lr.prepare_eval(seed, family)
        ls.cheap_only()
        r_cheap = Lr.rmse_rel(ls.final_rgb(), ls.ref_final)
        r_net, _ = Lr.evaluate(ls, net, tau=float("inf"))
        r, f = Lr.evaluate(ls, net, target=bound["target"])
        print(f"{family} {seed}: {ls.n} texels, cheap {r_cheap:.4f} network {r_net:.4f} learned+fallback {r:.4f} "
              f"fallback {f:.3f}", flush=True)
        assert r <= 0.02 and r_net < r_cheap
        ls.close()
        del ls


@pytest.mark.parametrize("n", [1, 15, 16, 100, 513, 70000])
def test_fused_inference_matches_oracle(gpu, oracle, n):
    """The one-launch inference kernel (weights in LDS, 16-row tiles per wave) on ragged
    row counts against the oracle's forward pass with BF16 rounding at the same points."""
    from lightsprintsdk_amd.learned import Surrogate
    rng = np.random.default_rng(n)
    net = Surrogate(max_batch=256, seed=7)
    p = net.init_params(7)
    p[-16 * 65:] += (rng.normal(size=16 * 65) * 0.05).astype(np.float32)
    net.set_params(p)
    x = rng.normal(size=(n, 32)).astype(np.float32)
    t = np.zeros((n, 3), np.float32)
    _, y_ref, _ = oracle.mlp_loss_grad(net.get_params(), x, t, lam=0.5, bf16=True, want_grad=False)
    y = net.predict(x)
    assert y.shape == (n, 4) and np.all(np.isfinite(y))
    np.testing.assert_allclose(y, y_ref, rtol=2e-2, atol=2e-3)
