"""Benchmark of the synthesis hot path on one MI355X.

One step = one full engine run (imageSynth on device-resident buffers: adaptation, point
lists, sorted offsets, target ordering, every refinement pass, write-back).

Default workload (config 5, learned proposals vs the plain method): a 4096x4096 RGB
value-noise texture with a centred 1024x1024 hole (1 048 576 targets), patchSize 30,
sensitivity 0.117, batched mode.  The proposal network is first trained on pairs the plain
engine logs on 200 seeded textures (not timed); `value` is the learned mode's target px/s;
the plain method (200 probes) is timed the same way.  At equal error: the fastest learned
setting whose mean match error is at most the plain method's, timed against the plain
method; plain runs with more probes bound the plain time at the learned error; the
configured setting (proposals per chunk + 32 random probes, the refiner's whole schedule)
is timed beside them.

--workload config2: the plain method on a 2048x2048 RGBA value-noise texture with 12 seeded
ellipses as the hole (531 632 targets), against the exact serial engine's error.
--workload config3: 96 uneven heal requests in one multi-request launch.
--workload config4: map-guided transfer through engine().

Prints one JSON line.  `value` is target pixels per second of the whole job with the inputs
already in HBM.  `roofline` prices the dominant kernels (the large-chunk visit kernels) by
their algorithmic bytes: a corpus pixel word per patch neighbour per scored candidate, plus
the per-visit neighbour scan (9 B per offset read), staging (12 B per neighbour) and
write-back (13 B).  `cpu_baseline` is the CPU oracle in the reference's threaded mode
(8 threads, lib/refinerThreaded.h) on the whole workload, with a config-2 leg.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload W] [--no-cpu] [--no-quality]
                       [--no-ladder] [--save-model F.npy | --load-model F.npy]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "target px/s at equal mean match error: learned-proposal vs plain, 1 MI355X"
HBM_PEAK_GBS = 8000.0


def cpu_baseline(threads=8):
    """The CPU oracle in the reference's threaded mode (8 threads) on the whole config-2 heal:
    px/s of one engine run."""
    O = _oracle()
    from resynthesizer_amd import workloads as W
    img, mask = W.config2()
    t = time.time()
    r, _ = O.image_synth(img, mask, 1, O.default_params(), threads=threads)
    dt = time.time() - t
    out = {"value": round(r.nTargets / dt, 1), "unit": "target px/s", "cores": threads, "kind": "port",
           "sample": "the whole config-2 heal (%d target px), one engine run of %.1f s, 8-thread reference "
                     "mode of the CPU oracle" % (r.nTargets, dt)}
    out.update(_cpu_info())
    return out


def bench_config3(args, np, torch, R, W):
    """96 uneven heal requests (sides U[128,1536], grating+noise RGB, 1-6 ellipse masks of
    3-25% area, patchSize from {10,30,50}) in one multi-request launch per step, on
    device-resident images; the kernel shape comes from the autotuner run on the first 8
    requests."""
    imgs, masks, patches = W.config3()
    d_src = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in imgs]
    d_img = [torch.empty_like(t) for t in d_src]
    d_mask = [torch.from_numpy(m).cuda() for m in masks]
    dev_img = [(t.data_ptr(), a.shape[1], a.shape[0], a.strides[0]) for t, a in zip(d_img, imgs)]
    dev_mask = [(t.data_ptr(), m.shape[1], m.shape[0], m.strides[0]) for t, m in zip(d_mask, masks)]
    fmts = [R.T_RGB] * len(imgs)
    params = [R.default_params(patchSize=p) for p in patches]
    for a, b in zip(d_img, d_src):
        a.copy_(b)
    torch.cuda.synchronize()
    t = time.time()
    rc, best, ms = R.autotune_batch(dev_img[:8], dev_mask[:8], fmts[:8], params[:8], device=True)
    assert rc == 0
    tune_s = time.time() - t

    def step():
        for a, b in zip(d_img, d_src):
            a.copy_(b)
        rc, res, st, _ = R.image_synth_batch(dev_img, dev_mask, fmts, params, device=True)
        assert rc == 0 and all(r["error"] == 0 for r in res), [r["error"] for r in res]
        return res, st

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res, st = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    n_t = sum(r["nTargets"] for r in res)
    alg_bytes = st["patchTerms"] * 4
    achieved = alg_bytes / (st["synthKernelMs"] / 1e3) / 1e9
    out = {
        "metric": METRIC, "value": round(n_t * args.steps / elapsed, 1), "unit": "target px/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (seeded grating + noise images, seeded ellipse masks)",
        "config": {"workload": "config 3: 96 uneven heal requests, %d target px, one multi-request launch"
                               % n_t, "target_px": n_t, "requests": len(imgs),
                   "kernel_config": st["config"], "autotune_s": round(tune_s, 1),
                   "autotune_ms": [round(float(v), 2) for v in ms], "prep_ms": round(st["prepMs"], 1),
                   "synth_kernel_ms": round(st["synthKernelMs"], 2),
                   "mean_match_error": round(float(np.mean([r["meanError"] for r in res])), 1)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_synth_many",
                     "launches": st["graphLaunches"],
                     "avg_launch_ms": round(st["synthKernelMs"] / max(st["graphLaunches"], 1), 3),
                     "algorithmic_bytes_per_launch": int(alg_bytes / max(st["graphLaunches"], 1))},
    }
    print(json.dumps(out), flush=True)


# Learned-mode settings of the config-5 line: proposals of every target computed once at
# the start of pass 1 from the pass-0 result (1 inverted-file cluster searched, top 16),
# 1 random probe beside them, 16 random probes in pass 0, whose chunks grow by 30% of the
# visited targets up to 262 144 targets (a quarter of the hole; the library's default cap
# is nT/32), 2 passes.
LEARNED_BATCH = 262144
LEARNED = dict(learned_refresh=0b10, learned_clusters=1, learned_probes=1, learned_probes_later=16, max_passes=2,
               ramp_permille=300, batch=LEARNED_BATCH)
# The same with 32 random probes beside the proposals (the probe count of the configured
# setting).
LEARNED32 = dict(LEARNED, learned_probes=32)
# The configured learned setting: proposals for every chunk from its targets' current
# neighbourhoods (top 16 of 2 inverted-file clusters) beside 32 random probes, in every pass
# of the refiner's own schedule (up to 6 passes, 10% stop rule).
CONFIGURED = dict(learned_probes=32)
# Cheaper learned settings (fewer random probes in pass 0, and in pass 1 beside the
# proposals), searched for the fastest one whose error still reaches the plain method's.
LEARNED_LADDER = [dict(LEARNED, first_chunk=256, ramp_permille=500), dict(LEARNED, first_chunk=1024, ramp_permille=500),
                  dict(LEARNED, first_chunk=1024, ramp_permille=500, learned_probes_later=4),
                  dict(LEARNED, first_chunk=4096, ramp_permille=500, learned_probes_later=1, learned_probes=0)]
# bytes priced per unit of visit-kernel work: a corpus pixel word per patch term of a
# scored candidate (priced at its word width), 9 B per patch offset the neighbour scan
# reads (an 8 B offset and the target's presence byte), 12 B per staged neighbour (offset,
# pixel, source) and 13 B per visit's write-back (colour, source, best)
SCAN_B, STAGE_B, VISIT_B = 9, 12, 13


def visit_bytes(st):
    """Algorithmic bytes of the large-chunk visit kernels of one run (ResynthStats)."""
    return (st.wavePatchTerms * st.pixelWordBytes + st.waveScanned * SCAN_B + st.waveStaged * STAGE_B
            + st.waveVisits * VISIT_B)


def _profile_json(name):
    path = os.path.join(ROOT, "profiles", name)
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def _roofline(st, kernel, traffic_file=None, counters_file=None):
    """roofline object of the large-chunk visit kernels from one timed run's statistics:
    algorithmic bytes per launch over the mean launch time (HIP events recorded on the
    library's stream around every launch)."""
    launches = max(st.waveLaunches, 1)
    kern_ms = st.waveKernelMs
    alg = visit_bytes(st)
    achieved = alg / (kern_ms / 1e3) / 1e9 if kern_ms > 0 else 0.0
    tr = _profile_json(traffic_file) if traffic_file else None
    ct = _profile_json(counters_file) if counters_file else None
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4),
           "traffic": tr.get("hbm_bytes_per_launch") if tr else None,
           "kernel": kernel, "launches": st.waveLaunches, "avg_launch_ms": round(kern_ms / launches, 4),
           "algorithmic_bytes_per_launch": int(alg / launches),
           "algorithmic_bytes_split": {"gathers": st.wavePatchTerms * st.pixelWordBytes,
                                       "scan": st.waveScanned * SCAN_B, "staging": st.waveStaged * STAGE_B,
                                       "write_back": st.waveVisits * VISIT_B},
           "candidate_evals_per_s": round(st.candidates / (st.synthKernelMs / 1e3), 1) if st.synthKernelMs else None,
           "valu_busy": ct.get("valu_busy") if ct else None,
           "valu_insts_per_visit": ct.get("valu_insts_per_visit") if ct else None}
    return out


def _train_or_load(args, R, W, np):
    """The proposal network: trained on pairs the plain engine logs on seeded textures, or
    loaded from --load-model (a .npy of its parameters written by --save-model)."""
    model = R.Model(seed=1)
    info = {}
    t = time.time()
    if args.load_model:
        model.set_params(np.load(args.load_model))
        info["loaded"] = os.path.basename(args.load_model)
    else:
        for i, (img, mask) in enumerate(W.training_textures(args.train_textures, 512)):
            err, _ = model.log_pairs(img, mask, R.T_RGB, R.default_params(), max_pairs=4096, seed=i + 1)
            assert err == 0, err
        losses = model.fit(args.train_steps, batch=1024, lr=1e-3, tau=0.1, seed=3)
        info.update(train_pairs=int(model.pair_count()), train_textures=args.train_textures,
                    train_steps=args.train_steps,
                    train_loss_first_last=[round(float(losses[:20].mean()), 3), round(float(losses[-20:].mean()), 3)])
    info["train_s"] = round(time.time() - t, 1)
    if args.save_model:
        np.save(args.save_model, model.params())
    return model, info


def bench_config5(args, np, torch, R, W):
    """Learned proposals vs the plain method on a 4096x4096 RGB value-noise + grating texture
    with a centred 1024x1024 hole (1 048 576 targets).  The network is trained first (pairs
    logged by the plain engine on 200 seeded 512x512 textures of config 2's generator,
    InfoNCE, Adam); training is reported, not timed.  Timed steps run the learned mode; the
    plain method (200 probes) is timed the same way.  Equal error: the fastest learned
    setting whose mean match error is at most the plain method's is timed against it; plain
    runs with more probes bound the plain method's time at the learned error."""
    model, train = _train_or_load(args, R, W, np)
    img, mask = W.config5(size=args.size or 4096)
    h, w, c = img.shape
    d_src = torch.from_numpy(img).cuda()
    d_img = torch.empty_like(d_src)
    d_mask = torch.from_numpy(mask).cuda()

    def run(probes, learned, seed=0, timing=False, plain=None):
        d_img.copy_(d_src)
        torch.cuda.synchronize()
        kw = dict(learned, model=model) if learned is not None else dict(plain or {})
        r = R.image_synth_device(d_img.data_ptr(), w, h, w * c, d_mask.data_ptr(), w, R.T_RGB,
                                 R.default_params(maxProbeCount=probes), mode=R.MODE_BATCHED, seed=seed,
                                 timing=timing, **kw)
        assert r.err == 0, r.err
        return r

    def timed(probes, learned, steps, warmup, plain=None):
        for _ in range(warmup):
            run(probes, learned, plain=plain)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        errs = []
        for k in range(steps):
            errs.append(run(probes, learned, seed=k + 1, plain=plain).meanError)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps, float(np.mean(errs))

    sec_l, err_l = timed(200, LEARNED, args.steps, args.warmup)
    sec_p, err_p = timed(200, None, max(2, args.steps // 2), 1)
    # the plain method with the learned line's chunk cap (faster, at a higher error): reported
    sec_pc, err_pc = timed(200, None, 2, 1, plain=dict(batch=LEARNED_BATCH))
    rl = run(200, LEARNED, timing=True)
    rp = run(200, None, timing=True)
    n_t = rl.nTargets
    ladder, learned_ladder, configured, probes32 = [], [], None, None
    if not args.no_ladder:
        # the plain method with more probes, until its error reaches the learned error or the
        # probe budget ends: its time there bounds the plain time at equal error from below
        for probes in (1000, 4000):
            sec, err = timed(probes, None, 1, 0)
            ladder.append({"probes": probes, "ms": round(sec * 1e3, 1), "mean_match_error": round(err, 1)})
            if err <= err_l:
                break
        # cheaper learned settings, while their error stays at or below the plain method's
        for st in LEARNED_LADDER:
            sec, err = timed(200, st, 3, 1)
            learned_ladder.append({"setting": {k: v for k, v in st.items() if LEARNED.get(k) != v},
                                   "ms": round(sec * 1e3, 2), "mean_match_error": round(err, 1)})
            if err > err_p:
                break
        sec_32, err_32 = timed(200, LEARNED32, 5, 1)
        probes32 = {"setting": "the bench setting with 32 random probes beside the proposals",
                    "ms": round(sec_32 * 1e3, 2), "px_per_s": round(n_t / sec_32, 1),
                    "mean_match_error": round(err_32, 1), "speedup_vs_plain200": round(sec_p / sec_32, 3),
                    "error_ratio_vs_plain200": round(err_32 / err_p, 4)}
        sec_c, err_c = timed(200, CONFIGURED, 3, 1)
        rc = run(200, CONFIGURED, timing=True)
        configured = {"setting": "proposals per chunk (top 16 of 2 clusters) + 32 random probes, every pass, "
                                 "the refiner's schedule", "ms": round(sec_c * 1e3, 2),
                      "px_per_s": round(n_t / sec_c, 1), "mean_match_error": round(err_c, 1),
                      "passes": rc.passes, "speedup_vs_plain200": round(sec_p / sec_c, 3),
                      "error_ratio_vs_plain200": round(err_c / err_p, 4)}
    # equal error: the fastest learned setting (bench setting or a cheaper one) whose error
    # is at most the plain 200-probe error
    eq = [(sec_l * 1e3, err_l, "bench setting")] + [(e["ms"], e["mean_match_error"], e["setting"])
                                                     for e in learned_ladder]
    if probes32:
        eq.append((probes32["ms"], probes32["mean_match_error"], "32 random probes"))
    eq = [e for e in eq if e[1] <= err_p]
    best_eq = min(eq, key=lambda e: e[0])
    reached = bool(ladder) and ladder[-1]["mean_match_error"] <= err_l
    out = {
        "metric": METRIC, "value": round(n_t / sec_l, 1), "unit": "target px/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sec_l * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32+bf16",
        "data": "synthetic (seeded value noise + grating; training textures from seeds 10000..10199)",
        "config": {"workload": "config 5: learned proposals, 4096x4096 RGB, 1024x1024 hole, %d target px" % n_t,
                   "target_px": n_t, "learned": LEARNED,
                   "learned_mean_match_error": round(err_l, 1),
                   "plain200_px_per_s": round(n_t / sec_p, 1), "plain200_ms": round(sec_p * 1e3, 2),
                   "plain200_mean_match_error": round(err_p, 1),
                   "plain200_at_learned_chunk_cap": {"batch": LEARNED_BATCH, "ms": round(sec_pc * 1e3, 2),
                                                     "mean_match_error": round(err_pc, 1)},
                   "speedup_vs_plain200": round(sec_p / sec_l, 3),
                   "error_ratio_learned_vs_plain200": round(err_l / err_p, 4),
                   "speedup_at_equal_error": round(sec_p * 1e3 / best_eq[0], 3),
                   "equal_error_basis": "plain 200-probe time over the fastest learned setting whose mean match "
                                        "error is at most the plain error (%s: %.1f ms, error %.0f)"
                                        % (best_eq[2], best_eq[0], best_eq[1]),
                   "learned_ladder": learned_ladder,
                   "with_32_probes": probes32,
                   "configured_setting": configured,
                   "plain_probe_ladder": ladder,
                   "plain_time_at_learned_error_over_learned": (round(ladder[-1]["ms"] / (sec_l * 1e3), 2)
                                                                if reached else None),
                   "plain_time_at_learned_error_over_learned_lower_bound": (
                       round(ladder[-1]["ms"] / (sec_l * 1e3), 2) if ladder else None),
                   "train": train,
                   "index_ms": round(rl.learnedIndexMs, 2), "keys": rl.learnedKeys,
                   "clusters": rl.learnedClusters, "propose_ms": round(rl.learnedProposeMs, 2),
                   "fine_search_keys_scored": rl.learnedKeysScored,
                   "synth_kernels_ms": round(rl.synthKernelMs, 2), "passes": rl.passes,
                   "plain_synth_kernels_ms": round(rp.synthKernelMs, 2)},
        "roofline": _roofline(rl, "k_synth_seg", "synth_seg_config5_traffic.json",
                              "synth_seg_config5_counters.json"),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_config5(np, torch, R, W)
    print(json.dumps(out), flush=True)


def bench_config4(args, np, torch, R, W):
    """Map-guided texture transfer through engine(): a 768x768 RGB value-noise corpus with its
    luminance as the map, and a whole 1536x1536 target canvas (2 359 296 targets) guided by a
    radial-gradient + value-noise map; mapWeight 0.5, batched mode.  Host-buffer entry (the
    engine() contract takes host Maps), so the line includes the PCIe copies."""
    tp, cp = W.config4()
    p = R.default_params(mapWeight=0.5)
    ix = R.prepare_indices(3, 1, False, False, True)
    for _ in range(args.warmup):
        R.engine(p, ix, tp, cp, mode=R.MODE_BATCHED)
    t0 = time.perf_counter()
    errs = []
    for k in range(args.steps):
        r, _ = R.engine(p, ix, tp, cp, mode=R.MODE_BATCHED, seed=k + 1)
        assert r.err == 0
        errs.append(r.meanError)
    elapsed = time.perf_counter() - t0
    rt, _ = R.engine(p, ix, tp, cp, mode=R.MODE_BATCHED, timing=True)
    n_t = rt.nTargets
    out = {
        "metric": METRIC, "value": round(n_t * args.steps / elapsed, 1), "unit": "target px/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (seeded value noise, luminance map, radial-gradient target map)",
        "config": {"workload": "config 4: map-guided transfer, corpus 768x768 RGB + map, target 1536x1536, "
                               "%d target px, mapWeight 0.5, batched mode, host buffers" % n_t,
                   "target_px": n_t, "passes": rt.passes, "mean_match_error": round(float(np.mean(errs)), 1),
                   "synth_kernels_ms": round(rt.synthKernelMs, 2), "prep_ms": round(rt.prepMs, 2),
                   "total_ms": round(rt.totalMs, 2)},
        "roofline": _roofline(rt, "k_synth_seg"),
    }
    print(json.dumps(out), flush=True)


def _cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O
    return O


def cpu_baseline_config5(np, torch, R, W, threads=8):
    """The CPU oracle in the reference's threaded mode (8 threads, lib/refinerThreaded.h,
    lib/buildSwitches.h:89) on the whole config-5 heal (the plain method, 200 probes; the
    reference has no learned proposals): px/s of one engine run.  Beside it, the config-2
    leg: the oracle on the whole 2048x2048 config-2 heal against the plain GPU batched mode
    on the same input."""
    O = _oracle()
    img, mask = W.config5()
    t = time.time()
    r, _ = O.image_synth(img, mask, 0, O.default_params(), threads=threads)
    dt5 = time.time() - t
    out = {"value": round(r.nTargets / dt5, 1), "unit": "target px/s", "cores": threads, "kind": "port",
           "sample": "the whole config-5 heal (4096x4096 RGB, 1024x1024 hole, %d target px), plain method, "
                     "one engine run of %.1f s, 8-thread reference mode of the CPU oracle" % (r.nTargets, dt5)}
    out.update(_cpu_info())
    # config-2 leg
    img2, mask2 = W.config2()
    t = time.time()
    r2, _ = O.image_synth(img2, mask2, 1, O.default_params(), threads=threads)
    dt2 = time.time() - t
    h, w, c = img2.shape
    d_src = torch.from_numpy(img2).cuda()
    d_img = torch.empty_like(d_src)
    d_mask = torch.from_numpy(mask2).cuda()
    secs = []
    for k in range(4):
        d_img.copy_(d_src)
        torch.cuda.synchronize()
        t = time.perf_counter()
        g = R.image_synth_device(d_img.data_ptr(), w, h, w * c, d_mask.data_ptr(), w, R.T_RGBA, R.default_params(),
                                 mode=R.MODE_BATCHED, seed=k + 1)
        torch.cuda.synchronize()
        assert g.err == 0
        if k:
            secs.append(time.perf_counter() - t)
    gpu2 = g.nTargets / float(np.median(secs))
    out["config2"] = {"cpu_px_per_s": round(r2.nTargets / dt2, 1), "cpu_s": round(dt2, 1),
                      "gpu_plain_px_per_s": round(gpu2, 1), "gpu_over_cpu": round(gpu2 / (r2.nTargets / dt2), 1),
                      "target_px": r2.nTargets}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=0, help="image side (default 2048 for config2, 4096 for config5)")
    ap.add_argument("--batch", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-quality", action="store_true")
    ap.add_argument("--train-textures", type=int, default=200, help="config5: training textures (profiling runs)")
    ap.add_argument("--train-steps", type=int, default=1500, help="config5: Adam steps (profiling runs)")
    ap.add_argument("--no-ladder", action="store_true", help="config5: skip the probe ladders and the "
                    "configured-setting run")
    ap.add_argument("--save-model", default="", help="config5: write the trained parameters here (.npy)")
    ap.add_argument("--load-model", default="", help="config5: load parameters from --save-model instead of "
                    "training (profiling runs)")
    ap.add_argument("--workload", default="config5", choices=["config2", "config3", "config4", "config5"],
                    help="config5 (default): learned proposals vs the plain method on a 4096x4096 heal; "
                         "config2: the plain method on a 2048x2048 heal; config3: 96 uneven heal "
                         "requests in one multi-request launch")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (calibrates HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the engine runs on one GPU (the sharded path is a design only); --gpus %d is not "
              "supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    if args.copy_probe:
        n = args.copy_probe << 20
        a = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n}), flush=True)
        return
    import resynthesizer_amd as R
    from resynthesizer_amd import workloads as W
    if args.workload == "config3":
        bench_config3(args, np, torch, R, W)
        return
    if args.workload == "config5":
        bench_config5(args, np, torch, R, W)
        return
    if args.workload == "config4":
        bench_config4(args, np, torch, R, W)
        return

    img, mask = W.config2(size=args.size or 2048)
    h, w, c = img.shape
    d_src = torch.from_numpy(img).cuda()
    d_img = torch.empty_like(d_src)
    d_mask = torch.from_numpy(mask).cuda()
    params = R.default_params()
    torch.cuda.synchronize()

    def step(timing=False):
        d_img.copy_(d_src)
        torch.cuda.synchronize()
        r = R.image_synth_device(d_img.data_ptr(), w, h, w * c, d_mask.data_ptr(), w, R.T_RGBA, params,
                                 mode=R.MODE_BATCHED, batch=args.batch, timing=timing)
        assert r.err == 0, r.err
        return r

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    n_t = r.nTargets
    value = n_t * args.steps / elapsed

    # the dominant kernels (the large-chunk visit kernels): HIP events recorded by the
    # library around each of their launches on its stream, and the work they tallied
    rt = step(timing=True)
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "target px/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u%d" % (8 * rt.pixelWordBytes),
        "data": "synthetic (seeded value noise + grating, 12 seeded ellipses)",
        "config": {"workload": "config 2: plain heal, %dx%d RGBA, %d target px, patch 30, 200 probes, "
                               "sensitivity 0.117, batched mode" % (w, h, n_t),
                   "proposal": "plain (200 random probes)", "target_px": n_t,
                   "passes": rt.passes, "visits": rt.visits, "candidates": rt.candidates,
                   "mean_match_error": round(rt.meanError, 1)},
        "roofline": dict(_roofline(rt, "k_synth_seg", "synth_seg_config2_traffic.json"),
                         all_synth_kernels_ms=round(rt.synthKernelMs, 3), all_synth_launches=rt.launches),
    }
    if not args.no_quality:
        # The batched mode's mean match error over four probe seeds (a single run varies by a
        # few percent with the seed and with which same-chunk writes land first), against the
        # exact serial engine on the same input (the reference's own result, engine seed) and
        # against the mean of eight serial seeds that the CPU oracle recorded for the tests.
        errs = []
        for seed in range(4):
            d_img.copy_(d_src)
            torch.cuda.synchronize()
            rb = R.image_synth_device(d_img.data_ptr(), w, h, w * c, d_mask.data_ptr(), w, R.T_RGBA, params,
                                      mode=R.MODE_BATCHED, batch=args.batch, seed=seed)
            assert rb.err == 0, rb.err
            errs.append(rb.meanError)
        d_img.copy_(d_src)
        torch.cuda.synchronize()
        t = time.time()
        rs = R.image_synth_device(d_img.data_ptr(), w, h, w * c, d_mask.data_ptr(), w, R.T_RGBA, params,
                                  mode=R.MODE_SERIAL)
        batched = float(np.mean(errs))
        out["config"]["batched_mean_match_error_4_seeds"] = [round(e, 1) for e in errs]
        out["config"]["serial_mean_match_error"] = round(rs.meanError, 1)
        out["config"]["error_ratio_batched_vs_serial"] = round(batched / rs.meanError, 4)
        golden = os.path.join(ROOT, "tests", "golden", "config_serial.json")
        if os.path.exists(golden) and (w, h) == (2048, 2048):
            with open(golden) as f:
                ser8 = float(np.mean([r["meanError"] for r in json.load(f)["config2"]]))
            out["config"]["serial_8_seed_mean_match_error"] = round(ser8, 1)
            out["config"]["error_ratio_batched_vs_serial_8_seeds"] = round(batched / ser8, 4)
        out["config"]["serial_seconds"] = round(time.time() - t, 2)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
