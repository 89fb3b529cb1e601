"""The learned radiance cache on the GPU: its BF16 MFMA forward and gradients
against the FP64 CPU oracle, and what the trained network gains over the
plain multi-bounce method at equal frames.

Tolerances: BF16 operands are identical on both sides and products are exact
in FP32; the GPU accumulates in FP32 and rounds inter-layer activations to
BF16 from FP32 where the oracle rounds from FP64, so an activation can differ
by one BF16 ulp (2^-8 relative). Outputs and gradients are therefore checked
to 2% relative (per layer, in norm), not bit for bit.
"""
import numpy as np
import pytest

from conftest import gpu_tracer, make_scene

pytestmark = pytest.mark.gpu


def scene_of(g):
    b = g.bounds()
    extent = max(b[3] - b[0] + 1, b[4] - b[1] + 1, b[5] - b[2] + 1) * 16
    return np.array([b[0] * 16, b[1] * 16, b[2] * 16, 1.0 / extent], np.float32)


def random_queries(vgi, setup, n, rng):
    faces = vgi.exposed_faces(setup["grid"])
    pick = faces[rng.choice(len(faces), n, replace=False)]
    q = np.zeros((n, 4), np.uint32)
    q[:, 0:3] = pick["voxel"].astype(np.uint32)
    mats = setup["grid"][pick["voxel"][:, 2], pick["voxel"][:, 1], pick["voxel"][:, 0]].astype(np.uint32)
    q[:, 3] = pick["face"] | (mats << 8)
    return q


LAYERS = [(0, 2048), (2048, 6144), (6144, 10240), (10240, 14336), (14336, 14528)]


@pytest.fixture(params=[0, 1], ids=["hash", "frequency"])
def encoding(request, orc):
    """Both position encodings: the hash grid and the frequency-only one."""
    orc.mlp_encoding(request.param)
    yield request.param
    orc.mlp_encoding(0)


def test_mlp_forward_matches_fp64_oracle(vgi, orc, encoding):
    setup = make_scene(vgi, orc, 2, 0x5EED0002, edge=64)
    g = gpu_tracer(vgi, setup)
    g.mlp_enable(seed=7)
    g.mlp_encoding(encoding)
    rng = np.random.default_rng(3)
    q = random_queries(vgi, setup, 4096, rng)
    q[::97, 3] = 0xFFFFFFFF  # invalid queries answer 0
    w = g.mlp_weights()
    # A hash table far from its initialisation, so the grid features count.
    table = rng.uniform(-0.5, 0.5, vgi.MLP_TABLE).astype(np.float32)
    g.mlp_table(table)
    assert np.array_equal(g.mlp_table(), table)
    gpu = g.mlp_infer(q)
    cpu = orc.mlp_forward(w, table, q, setup["diffuse"], scene_of(g))
    scale = np.abs(cpu).mean()
    assert scale > 1e-3
    err = np.abs(gpu - cpu).max() / scale
    print("forward max error / mean |out| = %.3g" % err, flush=True)
    assert err < 2e-2
    assert np.all(gpu[::97] == 0.0)
    g.close()


def test_mlp_gradient_matches_fp64_oracle(vgi, orc, encoding):
    setup = make_scene(vgi, orc, 2, 0x5EED0002, edge=64)
    g = gpu_tracer(vgi, setup)
    g.mlp_enable(seed=11)
    g.mlp_encoding(encoding)
    rng = np.random.default_rng(4)
    q = random_queries(vgi, setup, 512, rng)
    targets = np.zeros((512, 4), np.float32)
    targets[:, :3] = rng.random((512, 3)).astype(np.float32) * 2.0
    targets[:, 3] = (rng.random(512) > 0.1).astype(np.float32)
    w = g.mlp_weights()
    table = rng.uniform(-0.5, 0.5, vgi.MLP_TABLE).astype(np.float32)
    g.mlp_table(table)
    grad, loss = g.mlp_train(q, targets, update=False)
    table_grad = g.mlp_table_grad()
    ref, ref_table, ref_loss = orc.mlp_gradient(w, table, q, targets, setup["diffuse"], scene_of(g))
    print("loss gpu %.6f oracle %.6f" % (loss, ref_loss), flush=True)
    assert abs(loss - ref_loss) <= 1e-3 * abs(ref_loss) + 1e-6
    for a, b in LAYERS:
        e = np.linalg.norm(grad[a:b] - ref[a:b]) / max(np.linalg.norm(ref[a:b]), 1e-30)
        cos = np.dot(grad[a:b], ref[a:b]) / (np.linalg.norm(grad[a:b]) * np.linalg.norm(ref[a:b]) + 1e-30)
        print("layer [%d,%d): relative error %.3g cosine %.6f" % (a, b, e, cos), flush=True)
        assert e < 2e-2 and cos > 0.9995
    # The hash table's gradient, per level (none with the frequency encoding).
    size = 1 << 23  # entries x features per level
    if encoding == 1:
        assert not table_grad.any() and not ref_table.any()
    for level in range(8 if encoding == 0 else 0):
        a, b = table_grad[level * size:(level + 1) * size], ref_table[level * size:(level + 1) * size]
        e = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        cos = np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30)
        print("hash level %d: relative error %.3g cosine %.6f, %d entries" % (level, e, cos, int((b != 0).sum())), flush=True)
        assert e < 2e-2 and cos > 0.9995
        assert np.all(a[b == 0] == 0), "only the entries the queries touch get gradient"
    # A second identical call returns the same gradient (the buffers are cleared).
    grad2, _ = g.mlp_train(q, targets, update=False)
    assert np.allclose(grad, grad2, rtol=1e-5, atol=1e-9)
    assert np.allclose(g.mlp_table_grad(), table_grad, rtol=1e-4, atol=1e-9)
    g.close()


def test_adam_steps_reduce_the_loss(vgi, orc):
    setup = make_scene(vgi, orc, 2, 0x5EED0002, edge=64)
    g = gpu_tracer(vgi, setup)
    g.mlp_enable(seed=5, lr=1e-2)
    rng = np.random.default_rng(9)
    q = random_queries(vgi, setup, 1024, rng)
    targets = np.zeros((1024, 4), np.float32)
    targets[:, :3] = (0.2 + 0.5 * (q[:, 1:2] % 7) / 7.0).astype(np.float32)  # a learnable field of the height
    targets[:, 3] = 1.0
    _, first = g.mlp_train(q, targets, update=True)
    for _ in range(200):
        _, last = g.mlp_train(q, targets, update=True)
    print("loss %.4f -> %.4f" % (first, last), flush=True)
    assert last < 0.2 * first
    g.close()


def test_learned_cache_gains_over_plain(vgi, orc):
    """What the trained network gains over the plain multi-bounce tracer on a
    128^3 noise scene (4 bounces, 4 samples per face per frame): after 200
    training-only frames of 64k path-traced samples, the first frames of face
    irradiance are closer to the FP64 long-horizon reference (relative RMSE
    with the reference's own noise removed) at equal frame counts, while each
    learned frame traces only to the first hit.

    With one-path targets and a short warm-up the learned estimate still
    carries the network's fit error, which the unbiased plain estimate
    eventually passes; the gain asserted is the early one, with margin (the
    convergence to 2% is measured by the bench on the 512^3 cave)."""
    setup = make_scene(vgi, orc, 2, 0x5EED0002, edge=128)
    faces = vgi.exposed_faces(setup["grid"])
    sub = faces[::8]
    ref_t = gpu_tracer(vgi, setup, resident=True)
    ref_t.set_transport(bounces=4, samples=4, plain=True)
    ref_t.bind_faces(sub)
    sums = ref_t.converged(512)
    ref_t.run_frame()
    rec, _ = ref_t.read_bound()
    reference, var = vgi.converged_irradiance(sums, rec[:, 16:19])
    ref_t.close()
    curves, rays = {}, {}
    for mode in ("plain", "learned"):
        g = gpu_tracer(vgi, setup, resident=True)
        g.set_transport(bounces=4, samples=4, plain=True)
        g.bind_faces(faces)
        if mode == "learned":
            g.mlp_enable(seed=1, train_per_frame=65536, steps=8, lr=1e-3)
            g.mlp_pretrain(200)
            print("pretrained: loss %.4f" % g.mlp_last_loss(), flush=True)
        curve = []
        for f in range(8):
            g.run_frame(flags=vgi.COUNT_RAYS if f == 0 else 0)
            g.sync()
            rec, counters = g.read_bound()
            if f == 0:
                rays[mode] = int(counters[0])
            est = rec[::8, 0:3] + rec[::8, 16:19]
            curve.append(vgi.relative_rmse(est, reference, var))
        curves[mode] = curve
        print(mode, "rays/frame", rays[mode], ["%.3f" % c for c in curve], flush=True)
        g.close()
    for f in (0, 3, 7):
        assert curves["learned"][f] < 0.92 * curves["plain"][f]
    # The learned frame stops at the first hit: fewer rays than plain paths.
    assert rays["learned"] < 0.8 * rays["plain"]


def test_frozen_learned_frame_matches_oracle(vgi, orc):
    """One learned frame with frozen weights against the oracle: the oracle's
    learned paths (first hit, its MIS-weighted emission and direct light, the
    receiver's emitter sample; bit-exact FP32 path arithmetic) plus its FP64
    forward of the same weights and hash table at the same queries, combined
    per face as value = partial + factor * max(prediction, 0).

    Bar: known sample counts exact; faces none of whose samples query the
    network equal the oracle to FP32 summation order (1e-5 relative); every
    face within the forward's BF16 tolerance (2% of the mean |prediction|,
    times the face's mean |factor|) of the oracle."""
    setup = make_scene(vgi, orc, 2, 0x5EED0002, edge=64)
    env = (0.3, 0.5, 0.8)
    faces = vgi.exposed_faces(setup["grid"])[::3][:8192]
    g = gpu_tracer(vgi, setup)
    g.set_transport(bounces=3, samples=2, environment=env, plain=True)
    g.bind_faces(faces)
    g.mlp_enable(seed=5, train_per_frame=4096, steps=4, lr=3e-3)
    g.mlp_train_samples(4)
    g.mlp_pretrain(30)
    g.mlp_training(0, 1, 3e-3)  # frozen: the frame only infers
    w = g.mlp_weights()
    table = g.mlp_table()
    g.run_frame()
    g.sync()
    rec, _ = g.read_bound()
    # A frozen network trains nothing during the frame.
    assert np.array_equal(g.mlp_weights(), w)
    tasks = g.bound_tasks()
    scene = scene_of(g)
    g.close()

    t = setup["tracer"]
    t.set_transport(bounces=3, samples=2, environment=env)
    partial, factor, queries = t.trace_learned(tasks)
    pred = orc.mlp_forward(w, table, queries, setup["diffuse"], scene)
    known = partial[:, 3] > 0
    value = partial[:, 0:3].astype(np.float64) + factor.astype(np.float64) * np.maximum(pred, 0.0)
    value[~known] = 0.0
    n = len(faces)
    spp = 2
    k = known.reshape(n, spp).sum(axis=1)
    mean = value.reshape(n, spp, 3).sum(axis=1) / np.maximum(k, 1)[:, None]
    assert np.array_equal(rec[:, 5].astype(np.int64), k), "known sample counts"

    queried = (queries[:, 3] != 0xFFFFFFFF).reshape(n, spp).any(axis=1)
    term = np.abs(factor * np.maximum(pred, 0.0)).sum(axis=1).mean()
    print("faces %d, queried %d, mean |factor * prediction| %.4f, mean |partial| %.4f" % (
        n, int(queried.sum()), term, np.abs(partial[:, 0:3]).sum(axis=1).mean()), flush=True)
    assert queried.sum() > 0.5 * n and (~queried).sum() > 0
    # The prediction term matters: a wrong factor or query could not hide.
    assert term > 0.1 * np.abs(partial[:, 0:3]).sum(axis=1).mean()
    plain_faces = ~queried
    assert np.allclose(rec[plain_faces, 0:3], mean[plain_faces], rtol=1e-5, atol=1e-6)
    scale = np.abs(pred[queries[:, 3] != 0xFFFFFFFF]).mean()
    fmag = np.abs(factor).reshape(n, spp, 3).mean(axis=1)
    bound = 2e-2 * scale * fmag + 1e-5 * np.abs(mean) + 1e-6
    err = np.abs(rec[:, 0:3] - mean)
    worst = float((err / bound).max())
    print("worst error / bound %.3f" % worst, flush=True)
    assert worst <= 1.0


def test_adam_steps_follow_the_update_rule(vgi, orc):
    """Two Adam steps against the rule restated in numpy (Kingma & Ba 2015;
    beta1 0.9, beta2 0.99; eps 1e-10 for the network, 1e-15 for the table;
    bias correction by the global step): the network's weights everywhere,
    the hash table lazily -- only entries whose gradient is non-zero in a step
    move; an entry the second batch does not touch keeps its value exactly
    (no momentum-only step), and entries neither batch touched keep their
    initial values. Each step's gradient is read first with a gradient-only
    call at the same weights.

    Tolerance: the update divides the gradient by its own RMS, so values whose
    gradient is within FP32 atomic-order noise of 0 may take either sign; the
    check covers entries with |g| above 1e-4 of the largest and allows at most
    0.1% of them to differ by more than 1e-5 relative."""
    setup = make_scene(vgi, orc, 2, 0x5EED0002, edge=64)
    g = gpu_tracer(vgi, setup)
    lr = 1e-3
    g.mlp_enable(seed=13, lr=lr)
    rng = np.random.default_rng(6)
    table = rng.uniform(-0.5, 0.5, vgi.MLP_TABLE).astype(np.float32)
    g.mlp_table(table)
    w = g.mlp_weights().astype(np.float64)
    t = table.astype(np.float64)
    mw, vw = np.zeros_like(w), np.zeros_like(w)
    mt, vt = np.zeros_like(t), np.zeros_like(t)
    ever = np.zeros(len(t), bool)
    for step in range(2):
        q = random_queries(vgi, setup, 1024 if step == 0 else 256, rng)
        targets = np.zeros((len(q), 4), np.float32)
        targets[:, :3] = rng.random((len(q), 3)).astype(np.float32)
        targets[:, 3] = 1.0
        gw, _ = g.mlp_train(q, targets, update=False)
        gw = gw.astype(np.float64)
        gt = g.mlp_table_grad().astype(np.float64)
        before = g.mlp_table()
        g.mlp_train(q, targets, update=True)
        c1, c2 = 1.0 - 0.9 ** (step + 1), 1.0 - 0.99 ** (step + 1)
        mw = 0.9 * mw + 0.1 * gw
        vw = 0.99 * vw + 0.01 * gw * gw
        w = w - lr * (mw / c1) / (np.sqrt(vw / c2) + 1e-10)
        live = gt != 0.0
        ever |= live
        mt[live] = 0.9 * mt[live] + 0.1 * gt[live]
        vt[live] = 0.99 * vt[live] + 0.01 * gt[live] ** 2
        t[live] = t[live] - 10.0 * lr * (mt[live] / c1) / (np.sqrt(vt[live] / c2) + 1e-15)
        got_w = g.mlp_weights().astype(np.float64)
        got_t = g.mlp_table()
        for name, got, want, grad in (("network", got_w, w, gw), ("table", got_t.astype(np.float64), t, gt)):
            sel = np.abs(grad) > 1e-4 * np.abs(grad).max()
            bad = np.abs(got[sel] - want[sel]) > 1e-5 * np.maximum(np.abs(want[sel]), 1e-2)
            print("step %d %s: %d checked, %d differ" % (step + 1, name, int(sel.sum()), int(bad.sum())), flush=True)
            assert sel.sum() > 100 and bad.sum() <= 1e-3 * sel.sum()
        assert np.array_equal(got_t[~live], before[~live]), "entries the step did not touch keep their values"
        if step == 1:
            assert (ever & ~live).sum() > 1000, "some entries were touched by the first batch only"
        assert np.array_equal(got_t[~ever], table[~ever])
        # Continue from the library's FP32 state.
        w, t = got_w, got_t.astype(np.float64)
    g.close()
