"""Writes tests/golden/config_serial.json: the CPU oracle's serial engine (GRand, the
reference's GIMP build) run on full-size workloads, for the GPU tests that need a serial
result too slow to recompute per test run.

  config2 : the 2048 x 2048 RGBA heal (resynthesizer_amd.workloads.config2), the engine's
            seed and seven other seeds: sha256 of the synthesized image, mean best-match
            error, passes and betters per pass.  The GPU serial mode must reproduce the
            engine-seed image byte for byte; the batched mode's mean error is judged
            against the mean over the eight seeds.
  config3 : every fourth request of the 96 uneven requests (workloads.config3_request),
            four seeds each: mean best-match error and target count.

Usage:  python tests/golden/make_config_goldens.py   (about 10 minutes on 8 cores)
"""
import hashlib
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

CONFIG2_SEEDS = [0, 11, 12, 13, 14, 15, 16, 17]   # 0: the engine's seed 1198472 (lib/engine.c:643)
CONFIG3_REQUESTS = list(range(0, 96, 4))
CONFIG3_SEEDS = [0, 21, 22, 23]


def _params(patch=30):
    return dict(tileH=0, tileV=0, matchContextType=1, mapWeight=0.5, sensitivity=0.117, patchSize=patch,
                maxProbeCount=200)


def config2_run(seed):
    import oracle_py as O
    from resynthesizer_amd import workloads as W
    img, mask = W.config2()
    t0 = time.time()
    r, out = O.image_synth(img, mask, 1, _params(), seed=seed, want_maps=True)   # T_RGBA
    assert r.err == 0
    return dict(seed=seed, sha256=hashlib.sha256(out.tobytes()).hexdigest(), meanError=r.meanError,
                passes=r.passes, betters=list(r.betters)[:r.passes], nTargets=r.nTargets,
                seconds=round(time.time() - t0, 1))


def config3_run(args):
    k, seed = args
    import oracle_py as O
    from resynthesizer_amd import workloads as W
    img, mask, patch = W.config3_request(k)
    r, _ = O.image_synth(img, mask, 0, _params(patch), seed=seed, want_maps=True)   # T_RGB
    assert r.err == 0
    return dict(request=k, seed=seed, meanError=r.meanError, nTargets=r.nTargets, patch=patch)


def main():
    with ProcessPoolExecutor(8) as ex:
        c2 = list(ex.map(config2_run, CONFIG2_SEEDS))
        c3 = list(ex.map(config3_run, [(k, s) for k in CONFIG3_REQUESTS for s in CONFIG3_SEEDS]))
    out = dict(config2=c2, config3=c3)
    with open(os.path.join(HERE, "config_serial.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("config2 serial mean over seeds:", sum(r["meanError"] for r in c2) / len(c2))


if __name__ == "__main__":
    main()
