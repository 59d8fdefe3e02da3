"""Device MLP (bf16 MFMA, fp32 accumulation) against the torch CPU reference.

Tolerances: bf16 operands with fp32 accumulation in a different order than
torch's; a rounding boundary crossed by an activation moves it by one bf16
ulp (2^-8 relative).  Forward outputs agree within 1e-2 of the output scale,
gradients within 3e-2 relative (L2, per tensor), the Adam update within
1e-6 absolute.
"""
import sys

import numpy as np
import pytest

import gz_native as gz
import mlp_oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def device():
    gz.check(gz.lib().gz_set_device(0))


def random_codes(rng, n):
    c = rng.integers(0, 21, (n, 128)).astype(np.uint8)
    c[rng.random((n, 128)) < 0.05] = 20
    return c


@pytest.mark.parametrize("n", [1, 63, 64, 1000, 70001])
def test_forward_matches_reference(device, n):
    say("forward n", n)
    rng = np.random.default_rng(100 + n)
    flat = gz.Mlp.init_weights(3)
    flat[-1] = 0.5
    flat[21 * 8 + 256 * 1024:21 * 8 + 256 * 1024 + 256] = rng.normal(0, 0.1, 256)
    mlp = gz.Mlp(flat)
    codes = random_codes(rng, n)
    got = mlp.predict(codes)
    want, _, _ = mlp_oracle.forward_backward(flat, codes)
    scale = np.abs(want).max() + 1e-3
    err = np.abs(got - want).max()
    say("  max err", err, "scale", scale)
    assert err <= 1e-2 * scale


def test_gradients_match_reference(device):
    rng = np.random.default_rng(7)
    flat = gz.Mlp.init_weights(5)
    mlp = gz.Mlp(flat)
    n = 300
    codes = random_codes(rng, n)
    targets = rng.normal(8, 2, n).astype(np.float32)
    loss, grad = mlp.train_step(codes, targets, 0.0, want_grad=True)
    _, rloss, rgrad = mlp_oracle.forward_backward(flat, codes, targets)
    say("loss", loss, rloss)
    assert abs(loss - rloss) <= 1e-3 * abs(rloss)
    o = 0
    for name, shape in mlp_oracle.SIZES:
        k = int(np.prod(shape))
        g, r = grad[o:o + k], rgrad[o:o + k]
        rel = np.linalg.norm(g - r) / (np.linalg.norm(r) + 1e-12)
        say("  grad", name, "rel", rel)
        assert rel <= 3e-2, name
        o += k
    # lr == 0 leaves the weights alone
    assert np.array_equal(mlp.weights(), flat)


def test_adam_step_matches_reference(device):
    rng = np.random.default_rng(8)
    flat = gz.Mlp.init_weights(9)
    mlp = gz.Mlp(flat)
    codes = random_codes(rng, 128)
    targets = rng.normal(6, 1, 128).astype(np.float32)
    m = np.zeros_like(flat)
    v = np.zeros_like(flat)
    w = flat.copy()
    for step in (1, 2, 3):
        _, g = mlp.train_step(codes, targets, 1e-3, want_grad=True)
        w, m, v = mlp_oracle.adam_step(w, g, m, v, step, 1e-3)
        err = np.abs(mlp.weights() - w).max()
        say("adam step", step, "max err", err)
        assert err <= 1e-6


def test_training_fits_a_learnable_target(device):
    rng = np.random.default_rng(10)
    mlp = gz.Mlp(gz.Mlp.init_weights(11))
    codes = random_codes(rng, 4096)
    # a smooth function of the first window's composition
    t = ((codes[:, :32] < 5).sum(axis=1) * 0.25 + 3).astype(np.float32)
    mu = np.float32(t.mean())  # trained on centred targets, as the cost model is
    first = None
    for it in range(600):
        idx = rng.integers(0, len(codes), 256)
        loss = mlp.train_step(codes[idx], t[idx] - mu, 1e-3)
        first = loss if first is None else first
    pred = mlp.predict(codes) + mu
    mae = np.abs(pred - t).mean()
    base = np.abs(t - t.mean()).mean()
    say("loss", first, "->", loss, "mae", mae, "constant-predictor mae", base)
    assert mae < 0.5 * base


def test_device_resident_training_matches_per_step_training(device):
    """gz_mlp_train_epochs runs the steps of gz_mlp_train_step without host
    round trips (runs of 16 steps replayed from one captured graph, here five
    replays and a remainder of 13 steps run directly): the same batches give
    the same weights up to the order of float atomics (Adam's bias
    corrections are computed on the device), and a batch that does not
    divide the order leaves the remainder out."""
    rng = np.random.default_rng(12)
    codes = random_codes(rng, 3000)
    t = ((codes[:, :32] < 5).sum(axis=1) * 0.25).astype(np.float32)
    t -= t.mean()
    order = np.concatenate([rng.permutation(3000) for _ in range(4)]).astype(np.uint32)
    batch = 128
    a = gz.Mlp(gz.Mlp.init_weights(13))
    losses = []
    for b in range(0, len(order) - batch + 1, batch):
        idx = order[b:b + batch]
        losses.append(a.train_step(codes[idx], t[idx], 1e-3))
    b_ = gz.Mlp(gz.Mlp.init_weights(13))
    mean_loss, ms = b_.train_epochs(codes, t, order, batch, 1e-3)
    wa, wb = a.weights(), b_.weights()
    say("steps", len(losses), "mean loss", np.mean(losses), mean_loss, "device ms", ms)
    assert len(losses) == len(order) // batch
    # the embedding gradient's float atomics sum in any order, and Adam turns
    # a near-zero gradient's noise into whole lr-sized steps of either sign:
    # the weight vectors agree to 1e-3 relative (L2), and no weight moves by
    # more than ten steps of lr
    assert np.linalg.norm(wa - wb) <= 1e-3 * np.linalg.norm(wa)
    assert np.abs(wa - wb).max() <= 10 * 1e-3
    assert abs(mean_loss - np.mean(losses)) <= 1e-3 * max(1.0, abs(np.mean(losses)))
    assert ms > 0
    with pytest.raises(gz.GzError):
        b_.train_epochs(codes, t, np.array([0, 3000], np.uint32), 1, 1e-3)


def test_device_features_match_host_features(device):
    """Feature codes gathered on the device from packed sequences equal the
    host's (the codes the network is trained on)."""
    from helpers import related_pair_set
    rng = np.random.default_rng(12)
    q, qo, d, do, qpos, dpos = related_pair_set(rng, 3000, min_len=5, max_len=90)
    _, pq = gz.pack_on_device(q)
    _, pd = gz.pack_on_device(d)
    for reverse in (0, 1):
        a, b = (qpos - 1, dpos - 1) if reverse else (qpos, dpos)
        a = np.ascontiguousarray(a, np.uint32)
        b = np.ascontiguousarray(b, np.uint32)
        want = gz.cost_features(q, d, a, b, reverse)
        da, db_ = gz.DeviceBuffer.from_array(a), gz.DeviceBuffer.from_array(b)
        out = gz.DeviceBuffer(a.size * 128)
        gz.check(gz.lib().gz_cost_features_packed_async(pq.ptr, pd.ptr, da.ptr, db_.ptr, a.size,
                                                         reverse, out.ptr, None))
        got = out.to_array(np.uint8, a.size * 128).reshape(-1, 128)
        say("features reverse", reverse, "seeds", a.size)
        assert np.array_equal(got, want)


def test_swapped_rows_predict_like_swapped_codes(device):
    """Predictions read with each row's halves swapped (how the aligner
    predicts a hit's forward side from its reverse side's features) equal
    predictions of rows swapped on the host, bit for bit."""
    rng = np.random.default_rng(14)
    mlp = gz.Mlp(gz.Mlp.init_weights(5))
    codes = rng.integers(0, 21, (5000, 128)).astype(np.uint8)
    want = mlp.predict(np.concatenate([codes[:, 64:], codes[:, :64]], axis=1))
    got = np.zeros(len(codes), np.float32)
    gz.check(gz.lib().gz_mlp_predict_swapped(mlp.handle, gz.ptr(codes), len(codes), gz.ptr(got)))
    say("swapped rows", len(codes))
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def wave_cost(cells, order, lanes=64):
    """Lane-steps a wave-per-64-jobs launch spends: each wave runs as long as
    its most expensive job."""
    c = cells[order]
    pad = (-len(c)) % lanes
    c = np.concatenate([c, np.zeros(pad, c.dtype)]).reshape(-1, lanes)
    return int(c.max(axis=1).sum()) * lanes


def test_trained_network_gain_over_plain_order(device):
    """What the learned part buys: trained on the plain run's extension traces
    of one read set, the network predicts the one-sided extension costs of
    another read set's seeds far better than a constant, and ordering each
    side's jobs by it packs the waves of a launch with fewer idle lane-steps
    than the plain order."""
    import synthetic
    rng = np.random.default_rng(21)
    (names, res, offs), (dnames, dres, doffs) = synthetic.translated_workload(22, 20000, 4000)
    db = gz.Database(dnames, dres, doffs)
    half = len(names) // 2

    def traced(lo, hi):
        q = gz.Queries(names[lo:hi], res[int(offs[lo]):int(offs[hi])],
                       offs[lo:hi + 1] - offs[lo], translated=True)
        al = gz.Aligner(db, record_traces=1)
        al.align(q)
        tr = al.traces()
        concat = al.last_concatenated_queries()
        seq = db.chunk(0)["sequence"]
        feats = np.zeros((len(tr), 128), np.uint8)
        for rev in (0, 1):
            sel = np.nonzero(tr[:, 3] == rev)[0]
            feats[sel] = gz.cost_features(concat, seq, tr[sel, 0], tr[sel, 1], rev)
        return tr, feats, np.log2(np.maximum(tr[:, 4], 1)).astype(np.float32)

    tr_a, f_a, t_a = traced(0, half)
    tr_b, f_b, t_b = traced(half, len(names))
    say("training extensions", len(t_a), "evaluation extensions", len(t_b))
    assert len(t_a) > 10000 and len(t_b) > 10000
    mu = float(t_a.mean())
    mlp = gz.Mlp(gz.Mlp.init_weights(23))
    for ep in range(12):
        order = rng.permutation(len(t_a))
        for b in range(0, len(order) - 255, 256):
            mlp.train_step(f_a[order[b:b + 256]], t_a[order[b:b + 256]] - mu, 1e-3)
    pred = mlp.predict(f_b) + mu
    mae = float(np.abs(pred - t_b).mean())
    base = float(np.abs(t_b - mu).mean())
    gains = []
    for rev in (1, 0):
        sel = np.nonzero(tr_b[:, 3] == rev)[0]  # this side's jobs in the plain order
        cells = tr_b[sel, 4].astype(np.int64)
        learned = np.argsort(-pred[sel], kind="stable")
        gains.append(wave_cost(cells, np.arange(len(sel))) / wave_cost(cells, learned))
    say("mae", mae, "constant-predictor mae", base, "wave-cost gain (reverse, forward)", gains)
    assert mae < 0.5 * base
    assert min(gains) > 1.1
