"""The configured workloads at full size on the GPU.

  config 2: 2048 x 2048 RGBA heal, 531 632 targets.  The serial mode reproduces the CPU
            oracle's serial engine byte for byte (tests/golden/config_serial.json, made by
            tests/golden/make_config_goldens.py); the batched mode's mean match error over
            four probe seeds is within 2% of the oracle's serial mean over eight seeds.
  config 3: the 96 uneven requests in one multi-request call; every request keeps its
            context, and on every fourth request (three probe seeds each) the mean match error,
            weighted by target count, is within 2% of the oracle's serial engine (four seeds).
  config 5: the 4096 x 4096 learned heal (1 048 576 targets) with a trained network: the
            context is kept, the index covers the corpus, and the error is below the plain
            method's; training lowers the error by more than its seed-to-seed spread.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import resynthesizer_amd as R
from resynthesizer_amd import workloads as W

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config_serial.json")


@pytest.fixture(scope="module")
def serial():
    with open(GOLDEN) as f:
        return json.load(f)


def test_config2_serial_equals_the_oracle_at_full_size(serial):
    img, mask = W.config2()
    g0 = next(r for r in serial["config2"] if r["seed"] == 0)
    print("config 2 serial run, %d targets" % g0["nTargets"], flush=True)
    r, out = R.image_synth(img, mask, R.T_RGBA, R.default_params(), mode=R.MODE_SERIAL)
    assert r.err == 0 and r.nTargets == g0["nTargets"]
    assert r.passes == g0["passes"] and list(r.betters)[:r.passes] == g0["betters"]
    assert hashlib.sha256(out.tobytes()).hexdigest() == g0["sha256"]
    assert abs(r.meanError - g0["meanError"]) <= 1e-9 * g0["meanError"]
    # the run built only the part of the sorted offsets its scans can reach
    assert r.nOffsets < (2 * 2048 - 1) ** 2 // 10, r.nOffsets


def test_config2_batched_within_two_percent_of_serial(serial):
    img, mask = W.config2()
    sel = mask != 0
    serial_mean = float(np.mean([r["meanError"] for r in serial["config2"]]))
    errs = []
    for seed in range(1, 5):
        r, out = R.image_synth(img, mask, R.T_RGBA, R.default_params(), mode=R.MODE_BATCHED, seed=seed)
        assert r.err == 0 and r.nTargets == serial["config2"][0]["nTargets"]
        assert np.array_equal(out[~sel], img[~sel]) and np.array_equal(out[..., 3], img[..., 3])
        errs.append(r.meanError)
    ratio = np.mean(errs) / serial_mean
    print("config 2 batched / serial mean match error: %.4f (%s vs %.0f)" % (ratio, np.round(errs), serial_mean))
    assert ratio <= 1.02


def test_config3_all_requests_in_one_call(serial):
    imgs, masks, patches = W.config3()
    params = [R.default_params(patchSize=p) for p in patches]
    sub = {}
    for r in serial["config3"]:
        sub.setdefault(r["request"], []).append(r)
    errs = {k: [] for k in sub}
    for seed in (1, 2, 3):
        print("config 3, 96 requests, seed", seed, flush=True)
        rc, res, st, outs = R.image_synth_batch(imgs, masks, [R.T_RGB] * len(imgs), params, seed=seed)
        assert rc == 0 and st["requests"] == len(imgs) and st["graphLaunches"] == 1 and st["waitsExpired"] == 0
        for i, (img, mask, out) in enumerate(zip(imgs, masks, outs)):
            assert res[i]["error"] == 0 and res[i]["passes"] >= 1
            assert np.array_equal(out[mask == 0], img[mask == 0]), i
            if i in sub:
                assert res[i]["nTargets"] == sub[i][0]["nTargets"]
                errs[i].append(res[i]["meanError"])
    num = sum(sub[k][0]["nTargets"] * np.mean(errs[k]) for k in sub)
    den = sum(sub[k][0]["nTargets"] * np.mean([r["meanError"] for r in sub[k]]) for k in sub)
    per = {k: round(np.mean(errs[k]) / np.mean([r["meanError"] for r in sub[k]]), 3) for k in sub}
    print("config 3 multi-request / serial error (target-weighted over %d requests): %.4f; per request %s"
          % (len(sub), num / den, per))
    assert num / den <= 1.02


# ---------------------------------------------------------------- config 5

# the benchmark's learned setting (chunks capped at 262 144 targets)
LEARNED = dict(learned_refresh=0b10, learned_clusters=1, learned_probes=1, learned_probes_later=16, max_passes=2,
               ramp_permille=300, batch=262144)


@pytest.fixture(scope="module")
def trained():
    """The network as the benchmark trains it: pairs logged by the plain engine on 200 seeded
    512 x 512 textures, 1500 Adam steps of 1024 pairs (about 17 s)."""
    m = R.Model(seed=1)
    for i, (img, mask) in enumerate(W.training_textures(200, 512)):
        err, _ = m.log_pairs(img, mask, R.T_RGB, R.default_params(), max_pairs=4096, seed=i + 1)
        assert err == 0
    losses = m.fit(1500, batch=1024, lr=1e-3, tau=0.1, seed=3)
    assert losses[-20:].mean() < losses[:20].mean()
    return m


def test_config5_learned_heal_at_full_size(trained):
    img, mask = W.config5()
    sel = mask != 0
    print("config 5 learned heal", flush=True)
    r, out = R.image_synth(img, mask, R.T_RGB, R.default_params(), seed=1, model=trained, **LEARNED)
    assert r.err == 0 and r.nTargets == 1024 * 1024
    st = r.stats
    # keys: every corpus point with even coordinates (4096^2 - 1024^2 pixels, a quarter of them)
    assert st["learnedKeys"] == (4096 * 4096 - 1024 * 1024) // 4
    assert st["learnedClusters"] == 128 * 128 and st["learnedLaunches"] > 0 and st["learnedKeysScored"] > 0
    assert np.array_equal(out[~sel], img[~sel])
    code = lambda a: (a[..., 0].astype(np.int64) << 16) | (a[..., 1].astype(np.int64) << 8) | a[..., 2]
    assert np.isin(code(out[sel]), np.unique(code(img[~sel]))).all()   # every target copies a corpus colour
    rp, _ = R.image_synth(img, mask, R.T_RGB, R.default_params(), seed=1)
    print("config 5: learned %.0f, plain 200 probes %.0f" % (r.meanError, rp.meanError))
    assert r.meanError < 0.5 * rp.meanError


@pytest.mark.parametrize("setting", ["bench", "per-chunk proposals, 32 probes"])
def test_training_lowers_the_error_beyond_the_seed_spread(trained, setting):
    """On config 5's heal (4096 x 4096, 1024 x 1024 hole) over four probe seeds, the trained
    network's mean match error is below the untrained network's (the same initial weights)
    by more than either's seed-to-seed standard deviation and by more than 3%, in the
    benchmark's setting and with proposals computed per chunk beside 32 random probes (the
    configured setting).  Runs with the same probe seed are paired: at least three of the
    four must improve."""
    img, mask = W.config5()
    kw = LEARNED if setting == "bench" else dict(learned_probes=32)
    untrained = R.Model(seed=1)
    e_t, e_u = [], []
    for seed in (1, 2, 3, 4):
        rt, _ = R.image_synth(img, mask, R.T_RGB, R.default_params(), seed=seed, model=trained, **kw)
        ru, _ = R.image_synth(img, mask, R.T_RGB, R.default_params(), seed=seed, model=untrained, **kw)
        assert rt.err == 0 and ru.err == 0
        e_t.append(rt.meanError)
        e_u.append(ru.meanError)
    gain = np.mean(e_u) - np.mean(e_t)
    spread = max(np.std(e_t), np.std(e_u))
    print("%s: trained %s untrained %s: gain %.0f (%.1f%%), spread %.0f"
          % (setting, np.round(e_t), np.round(e_u), gain, 100 * gain / np.mean(e_u), spread))
    # paired by probe seed: most seeds improve, and the mean gain exceeds the spread
    assert sum(t < u for t, u in zip(e_t, e_u)) >= 3
    assert gain > spread and gain > 0.03 * np.mean(e_u)
