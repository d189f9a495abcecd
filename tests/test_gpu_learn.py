"""GPU tests of the learned-proposal network: the MFMA GEMM (which also pins the
v_mfma_f32_16x16x32_bf16 lane layout the kernels assume), the forward pass and the InfoNCE
gradient against the float64 CPU oracle, and training."""
import numpy as np
import pytest

import mlp_oracle as MO
import resynthesizer_amd as R

pytestmark = pytest.mark.gpu


def _bf16(rng, shape, scale=1.0):
    return MO.bf16_bits(rng.normal(0, scale, size=shape).astype(np.float32))


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_mfma_gemm_matches_float64(ta, tb):
    rng = np.random.RandomState(10 + 2 * ta + tb)
    M, N, K = 72, 200, 96
    A = _bf16(rng, (K, M) if ta else (M, K))
    B = _bf16(rng, (N, K) if tb else (K, N))
    C = R.gemm_bf16(A, B, ta=bool(ta), tb=bool(tb))
    a = MO.from_bits(A)
    b = MO.from_bits(B)
    ref = (a.T if ta else a) @ (b.T if tb else b)
    assert np.abs(C - ref).max() <= 1e-4 * np.abs(ref).max() + 1e-5



@pytest.mark.parametrize("P", [1, 2])
def test_query_routing_matches_brute_force(P):
    """The index's query routing (MFMA score tiles over the query's super-cluster, per-lane
    lists, merge): the P sub-centroids of the query's super-cluster with the largest dot
    products, best first, as cluster ids s*Q + c; a negative super-cluster counts as 0."""
    rng = np.random.RandomState(40 + P)
    S, Q, nq = 9, 128, 777
    sub = _bf16(rng, (S * Q, 64))
    subf = MO.from_bits(sub)
    qb = _bf16(rng, (nq, 64))
    qf = MO.from_bits(qb)
    sup = rng.randint(0, S, nq).astype(np.int32)
    sup[::50] = -1
    got = R.route(qb, sup, sub, S, Q, P)
    for i in range(nq):
        s = max(int(sup[i]), 0)
        scores = subf[s * Q:(s + 1) * Q] @ qf[i]
        order = np.argsort(-scores, kind="stable")
        assert (got[i] >= s * Q).all() and (got[i] < (s + 1) * Q).all(), i
        gs = scores[got[i] - s * Q]
        assert np.allclose(gs, scores[order[:P]], atol=1e-5), (i, gs, scores[order[:P]])

def _check_fine_top16(qb, kb, cl, start, pts):
    """Every query's 16 best keys of its clusters against brute force, best first, padded
    with (-1, -1) when its clusters hold fewer than 16 keys; ties may come in any order."""
    qf, kf = MO.from_bits(qb), MO.from_bits(kb)
    out = R.fine_top16(qb, cl, kb, start, pts)
    for i in range(len(qb)):
        cand = np.concatenate([np.arange(start[c], start[c + 1]) for c in cl[i] if c >= 0] + [np.zeros(0, int)])
        cand = cand.astype(np.int64)
        scores = kf[cand] @ qf[i]
        want = min(16, len(cand))
        got = out[i]
        assert (got[want:] == -1).all(), i
        ids = (got[:want, 0] // 2) + (got[:want, 1] // 2) * 1000
        assert (got[:want] % 2 == 0).all() and len(set(ids.tolist())) == want, i
        assert set(ids.tolist()) <= set(cand.tolist()), i
        s = kf[ids] @ qf[i]
        assert (np.diff(s) <= 1e-5).all(), (i, s)
        if want:
            kth = np.sort(scores)[::-1][want - 1]
            assert s.min() >= kth - 1e-5, (i, s.min(), kth)


def _clusters(rng, sizes, nq, P):
    nc = len(sizes)
    start = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    cl = np.stack([rng.permutation(nc)[:P] for _ in range(nq)]).astype(np.int32)
    cl[rng.rand(nq) < 0.05, 0] = -1
    if P == 2:
        cl[rng.rand(nq) < 0.05, 1] = -1
    nk = int(start[-1])
    pts = np.stack([np.arange(nk) % 1000, np.arange(nk) // 1000], 1).astype(np.int32)
    return start, cl, pts


@pytest.mark.parametrize("P", [1, 2])
def test_fine_search_returns_exact_top16(P):
    """The index's fine search (16 x 16 MFMA score tiles, per-lane sorted lists, a merge per
    query) against brute force.  Cluster
    sizes cover empty, partial key tiles and multi-tile clusters; tiles of 16 sorted queries
    straddle cluster boundaries."""
    rng = np.random.RandomState(20 + P)
    sizes = [0, 1, 5, 15, 16, 17, 31, 33, 240, 700, 64, 3] * 3
    start, cl, pts = _clusters(rng, sizes, 523, P)
    _check_fine_top16(_bf16(rng, (523, 64)), _bf16(rng, (int(start[-1]), 64)), cl, start, pts)


@pytest.mark.parametrize("distinct", [1, 3, 40])
def test_fine_search_exact_under_ties(distinct):
    """Keys drawn from a few distinct vectors tie in their products (with one, every key of a
    cluster ties): the top 16 are still 16 distinct keys of the query's clusters, best first,
    none below the 16th best product."""
    rng = np.random.RandomState(50 + distinct)
    sizes = [240, 700, 16, 90, 33, 300] * 2
    start, cl, pts = _clusters(rng, sizes, 301, 2)
    base = _bf16(rng, (distinct, 64))
    kb = base[rng.randint(0, distinct, size=int(start[-1]))]
    _check_fine_top16(_bf16(rng, (301, 64)), kb, cl, start, pts)


def _features(n, seed):
    rng = np.random.RandomState(seed)
    x = np.zeros((n, 128), np.float32)
    x[:, :120] = rng.uniform(0, 1, size=(n, 120))
    pres = rng.randint(0, 2, size=(n, 30)).astype(np.float32)
    x[:, 3:120:4] = pres
    for c in range(3):
        x[:, c:120:4] *= pres
    return x


def _oracle_params(model):
    """The model's parameters as the GPU uses them: weights rounded to bf16, biases fp32."""
    p = model.params().astype(np.float64)
    d = MO.unpack(p)
    for k in ("W1", "W2", "W3"):
        d[k] = MO.bf16_round(d[k])
    return MO.pack(d)


def test_forward_matches_oracle():
    m = R.Model(seed=7)
    x = _features(300, 1)
    xb = MO.bf16_bits(x)
    z = m.embed(xb)
    zr, _ = MO.forward(_oracle_params(m), MO.from_bits(xb))
    assert z.shape == (300, 64)
    assert np.allclose((z * z).sum(1), 1.0, atol=1e-4)
    cos = (z * zr).sum(1)
    assert cos.min() > 0.995, cos.min()
    assert np.abs(z - zr).max() < 0.05


def test_infonce_gradient_matches_oracle():
    """Loss within 1% and every parameter block's gradient within 5% (relative norm) of the
    float64 oracle at bf16 inputs and weights; tau 0.1, 128 pairs."""
    m = R.Model(seed=3)
    xq = _features(128, 2)
    xk = np.clip(xq + np.random.RandomState(3).normal(0, 0.03, size=xq.shape).astype(np.float32), 0, 1)
    xk[:, 120:] = 0
    qb, kb = MO.bf16_bits(xq), MO.bf16_bits(xk)
    loss, g = m.infonce(qb, kb, 0.1)
    lr, gr = MO.infonce(_oracle_params(m), MO.from_bits(qb), MO.from_bits(kb), 0.1)
    assert abs(loss - lr) <= 0.01 * abs(lr), (loss, lr)
    o = 0
    for name, shp in MO.SHAPES:
        n = int(np.prod(shp))
        a, b = g[o:o + n].astype(np.float64), gr[o:o + n]
        o += n
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert rel < 0.05, (name, rel)


def test_training_reduces_loss():
    """Adam on 4096 noisy pairs: the loss of the last 20 steps is well below the first."""
    m = R.Model(seed=11)
    xq = _features(4096, 5)
    xk = np.clip(xq + np.random.RandomState(6).normal(0, 0.05, size=xq.shape).astype(np.float32), 0, 1)
    xk[:, 120:] = 0
    losses = m.train(MO.bf16_bits(xq), MO.bf16_bits(xk), steps=300, batch=256, lr=2e-3, tau=0.1, seed=1)
    assert np.isfinite(losses).all()
    first, last = losses[:20].mean(), losses[-20:].mean()
    print("loss %.3f -> %.3f" % (first, last))
    assert last < 0.5 * first, (first, last)


def _trained_model(n_tex=6, steps=150):
    from resynthesizer_amd import workloads as W
    m = R.Model(seed=1)
    for i, (img, mask) in enumerate(W.training_textures(n_tex, 256)):
        err, out = m.log_pairs(img, mask, R.T_RGB, R.default_params(), max_pairs=2048, seed=i + 1)
        assert err == 0
        sel = mask != 0
        assert np.array_equal(out[~sel], img[~sel])
    assert m.pair_count() == n_tex * 2048
    losses = m.fit(steps, batch=512, lr=1e-3, tau=0.1, seed=2)
    assert losses[-10:].mean() < losses[:10].mean()
    return m


@pytest.mark.parametrize("refresh", [0, 0b10])
def test_learned_proposals_beat_plain(refresh):
    """The batched engine with learned proposals (per-chunk, or refreshed once at pass 1)
    on a 768x768 value-noise texture with a 192x192 hole: the context is untouched, every
    target takes a corpus colour, the index is built, proposals are used, and the mean
    match error is well below the plain method's with 200 probes (and not above the same
    proposals from an untrained network by more than noise)."""
    from resynthesizer_amd import workloads as W
    m = _trained_model()
    img, mask = W.config5(size=768)
    sel = mask != 0
    corpus = {tuple(c) for c in img[~sel]}
    p = R.default_params()
    kw = dict(learned_refresh=refresh, learned_clusters=1, learned_probes=1, learned_probes_later=16,
              max_passes=3) if refresh else dict(learned_probes=32)
    rl, out = R.image_synth(img, mask, R.T_RGB, p, mode=R.MODE_BATCHED, seed=1, model=m, **kw)
    assert rl.err == 0
    st = rl.stats
    assert st["learnedKeys"] > 100000 and st["learnedClusters"] >= 128 and st["learnedLaunches"] > 0, st
    assert np.array_equal(out[~sel], img[~sel])
    assert all(tuple(c) in corpus for c in out[sel][::11])
    rp, _ = R.image_synth(img, mask, R.T_RGB, p, mode=R.MODE_BATCHED, seed=1)
    untrained = R.Model(seed=1)
    ru, _ = R.image_synth(img, mask, R.T_RGB, p, mode=R.MODE_BATCHED, seed=1, model=untrained, **kw)
    print("mean error: learned %.0f untrained %.0f plain %.0f" % (rl.meanError, ru.meanError, rp.meanError))
    assert rl.meanError < 0.6 * rp.meanError
    assert rl.meanError < 1.10 * ru.meanError
