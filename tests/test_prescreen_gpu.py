"""GPU checks of the learned prescreen: the feature kernel against its CPU restatement, the
BF16 MFMA network (inference, gradients, Adam) against the float64 oracle with BF16 rounding
emulated, subset runs of a plan, and a small end-to-end train / calibrate / screen run whose
recall is measured against the exact DP.

Tolerances: features rtol 1e-5 (float32 divisions and log2); logits |d| <= 2e-3 * (1 + |y|);
gradients: max |g_gpu - g_cpu| <= 2e-2 * max |g_cpu| per tensor (BF16-stored activation
gradients may round differently by one ulp when the FP32 and float64 sums straddle a rounding
boundary); Adam: rtol 1e-5 on the updated parameters.
"""
import ctypes as C

import numpy as np
import pytest

import mlp_ref
import nsffi
from feature_ref import features as ref_features
from protlocus_amd import capi, prescreen, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu(product):
    capi.init_gpu(0)
    return capi


def _tables(oracle):
    L = oracle.lib
    return (list((C.c_uint8 * 256).in_dll(L, "or_nt4")), list((C.c_uint8 * 256).in_dll(L, "or_aa20")),
            list((C.c_uint8 * 64).in_dll(L, "or_codon")))


def _cands(seed, n):
    prots, dnas, _ = prescreen.candidates(n, seed, min_len=5, max_len=200)
    rng = np.random.default_rng(seed)
    # edge cases: tiny proteins, a window with N runs, lower-case and numeric DNA, empty window
    prots += [b"M", b"MW", b"MWK", b"ACDEFGHIKLMNPQRSTVWY" * 3]
    dnas += [b"ATG", b"", b"atgtggaaannnnACGT", bytes(rng.integers(0, 5, size=400).astype(np.uint8))]
    return prots, dnas


def test_features_match_cpu(oracle, gpu):
    nt4, aa20, codon = _tables(oracle)
    prots, dnas = _cands(1, 40)
    got = gpu.features(prots, dnas)
    for k, (p, d) in enumerate(zip(prots, dnas)):
        exp = ref_features(p, d, nt4, aa20, codon, oracle.blosum62)
        assert np.allclose(got[k], exp, rtol=1e-5, atol=1e-6), (k, np.nonzero(~np.isclose(got[k], exp, 1e-5, 1e-6)))


def test_plan_features_equal_standalone(gpu):
    prots, dnas = _cands(2, 30)
    prots, dnas = prots[:-3], dnas[:-3]   # plans need al >= 1 and the same inputs
    plan = gpu.Plan(dnas, prots, gpu.opt(**prescreen.LABEL_KW))
    try:
        assert np.array_equal(plan.features_host(), gpu.features(prots, dnas))
    finally:
        plan.close()


def test_run_subset_and_global_mode(oracle, gpu):
    prots, dnas, _ = prescreen.candidates(40, 3, min_len=20, max_len=300)
    plan = gpu.Plan(dnas, prots, gpu.opt(**prescreen.LABEL_KW))
    try:
        plan.run()
        full = plan.scores()
        cells_full = plan.cells(1)
        # global mode equals ns_global_gs16b without CIGAR or EXT flags (every row: no x-drop)
        for k in range(0, len(prots), 7):
            exp = oracle.global_dp(dnas[k], prots[k], oracle.opt(**prescreen.LABEL_KW), mode=16)[2]
            assert full[k] == exp, k
        keep = np.arange(len(prots)) % 3 == 0
        plan.run(keep=keep)
        sub = plan.scores()
        assert np.array_equal(sub[keep], full[keep])
        assert (sub[~keep] == -2 ** 31).all()
        assert 0 < plan.cells(1) < cells_full
    finally:
        plan.close()


def test_mlp_inference_matches_oracle(gpu):
    rng = np.random.default_rng(4)
    p = prescreen.init_params(rng)
    p[64 * 256:64 * 256 + 256] = rng.normal(0, 0.1, 256)          # b1
    p[-257 - 256:-257] = rng.normal(0, 0.1, 256)                   # b2
    p[-1] = 0.3
    X = rng.normal(0, 1, size=(1000, 64)).astype(np.float32)       # not a multiple of 64: a ragged last tile
    net = gpu.MLP(p)
    try:
        got = net.infer(X)
    finally:
        net.close()
    exp, _ = mlp_ref.forward(p, X)
    assert np.all(np.abs(got - exp) <= 2e-3 * (1 + np.abs(exp))), np.max(np.abs(got - exp))


def test_mlp_inference_many_blocks_standardised(gpu):
    # 40,001 samples: each work-group of the persistent inference kernel takes several 64-sample
    # blocks in turn, the last one ragged; the device standardises the raw features. Against the
    # oracle on the standardised inputs, and bit for bit against the same rows inferred alone
    rng = np.random.default_rng(5)
    p = prescreen.init_params(rng)
    p[64 * 256:64 * 256 + 256] = rng.normal(0, 0.1, 256)
    p[-257 - 256:-257] = rng.normal(0, 0.1, 256)
    X = rng.normal(3, 2, size=(40001, 64)).astype(np.float32)
    mu = rng.normal(3, 0.5, 64).astype(np.float32)
    sd = rng.uniform(1, 3, 64).astype(np.float32)
    net = gpu.MLP(p, mu, sd)
    try:
        got = net.infer(X)
        head = net.infer(X[:1000])
        tail = net.infer(X[-777:])
    finally:
        net.close()
    Xn = (X - mu) * (np.float32(1) / sd)          # float32, as the device computes it
    exp, _ = mlp_ref.forward(p, Xn)
    # FP32 MFMA sums and the oracle's float64 sums can round a hidden activation to neighbouring
    # BF16 values; over 40,001 samples a few such flips move a logit by a few 1e-3
    err = np.abs(got - exp) / (1 + np.abs(exp))
    assert np.mean(err > 2e-3) <= 1e-3 and err.max() <= 2e-2, (int(np.sum(err > 2e-3)), err.max())
    assert np.array_equal(got[:1000], head) and np.array_equal(got[-777:], tail)


def test_mlp_gradients_match_oracle(gpu):
    rng = np.random.default_rng(5)
    p = prescreen.init_params(rng)
    p[-1] = -0.2
    X = rng.normal(0, 1, size=(256, 64)).astype(np.float32)
    t = (rng.random(256) < 0.4).astype(np.float32)
    net = gpu.MLP(p)
    try:
        loss = net.train_step(X, t, 0.0)
        p_after, g = net.params(grads=True)
    finally:
        net.close()
    assert np.array_equal(p_after, p)            # lr = 0: no update
    exp = mlp_ref.grads(p, X, t)
    assert abs(loss - mlp_ref.loss(p, X, t)) < 1e-3
    o = 0
    for name, n in mlp_ref.SIZES:
        d = np.max(np.abs(g[o:o + n] - exp[o:o + n]))
        assert d <= 2e-2 * np.max(np.abs(exp[o:o + n])) + 1e-7, (name, d, np.max(np.abs(exp[o:o + n])))
        o += n


def test_mlp_adam_step(gpu):
    rng = np.random.default_rng(6)
    p = prescreen.init_params(rng)
    X = rng.normal(0, 1, size=(128, 64)).astype(np.float32)
    t = (rng.random(128) < 0.5).astype(np.float32)
    net = gpu.MLP(p)
    try:
        net.train_step(X, t, 0.0)
        _, g = net.params(grads=True)
        net.train_step(X, t, 1e-3)                  # same batch, same parameters: same gradients
        p1 = net.params()
    finally:
        net.close()
    exp, _, _ = mlp_ref.adam(p.astype(np.float64), g.astype(np.float64), 0.0, 0.0, 1, 1e-3)
    assert np.allclose(p1, exp, rtol=1e-5, atol=1e-7)


def test_training_separates_classes(gpu):
    rng = np.random.default_rng(7)
    X = rng.normal(0, 1, size=(8192, 64)).astype(np.float32)
    w = rng.normal(0, 1, size=64)
    y = ((X @ w + 0.5 * np.sin(3 * X[:, 0])) > 0).astype(np.float32)
    params, mu, sd = prescreen.train(X, y, epochs=6, batch=512, lr=3e-3)
    net = gpu.MLP(params, mu, sd)
    try:
        acc = ((net.infer(X) > 0) == (y > 0)).mean()
    finally:
        net.close()
    assert acc > 0.95, acc


def test_prescreen_end_to_end_small(gpu):
    res = prescreen.evaluate(n_train=2048, n_val=1024, n_eval=1024, epochs=8, log=None)
    assert res["recall"] >= 0.99, res
    assert res["kept_frac"] < 0.8, res
    assert res["cells_saved"] > 0.2, res
