"""The bound behind the engine's cut offset table (launch_reach_radius): every target's
neighbour scan gathers its patch before it passes an offset farther than reach + 21, where
reach is the largest distance from a tile holding a target to the nearest full context tile
(8x8 pixels, all holding a value from the start).  Checked by brute force on random heals:
the distance to each target's (L-1)-th nearest context pixel, over every L the engine
admits.  The device kernels restated here (tile maps, tile-box radius) follow
resynthesizer_amd/csrc/resynth_kernels.hip."""
import numpy as np

K = 8   # tile side


def _tile_maps(ctx, tgt):
    H, W = ctx.shape
    tH, tW = (H + K - 1) // K, (W + K - 1) // K
    full = np.zeros((tH, tW), bool)
    has = np.zeros((tH, tW), bool)
    for ty in range(tH):
        for tx in range(tW):
            blk = ctx[ty * K:ty * K + K, tx * K:tx * K + K]
            full[ty, tx] = blk.shape == (K, K) and bool(blk.all())
            has[ty, tx] = bool(tgt[ty * K:ty * K + K, tx * K:tx * K + K].any())
    return full, has


def _box_radius(mark):
    """floor of the least Euclidean distance between each tile's box and a marked tile's box
    (boxes t tiles apart in a row or column are (t - 1) * 8 + 1 pixels apart)."""
    pts = np.argwhere(mark)
    tH, tW = mark.shape
    ty, tx = np.mgrid[:tH, :tW]

    def gap(t):
        return np.where(t <= 0, 0, (t - 1) * K + 1)

    best = np.full(mark.shape, np.iinfo(np.int64).max, np.int64)
    for y, x in pts:
        best = np.minimum(best, gap(np.abs(tx - x)) ** 2 + gap(np.abs(ty - y)) ** 2)
    r = np.floor(np.sqrt(best)).astype(np.int64)
    r -= (r * r > best)
    r += ((r + 1) * (r + 1) <= best)
    return r


def test_scan_reach_bound_holds_on_random_heals():
    rng = np.random.default_rng(7)
    checked = 0
    worst = 0.0
    for trial in range(60):
        H, W = int(rng.integers(16, 64)), int(rng.integers(16, 64))
        yy, xx = np.mgrid[:H, :W]
        tgt = np.zeros((H, W), bool)
        for _ in range(int(rng.integers(1, 4))):
            cy, cx = rng.integers(0, H), rng.integers(0, W)
            ry, rx = rng.integers(2, H), rng.integers(2, W)
            tgt |= ((yy - cy) / ry) ** 2 + ((xx - cx) / rx) ** 2 <= 1
        ctx = ~tgt
        if trial % 3 == 0:
            ctx &= rng.random((H, W)) < 0.9   # pixels with no value outside the hole
        full, has = _tile_maps(ctx, tgt)
        if not full.any() or not has.any():
            continue
        reach = int(_box_radius(full)[has].max())
        cp = np.argwhere(ctx)
        for y, x in np.argwhere(tgt)[::7]:
            d2 = np.sort((cp[:, 0] - y) ** 2 + (cp[:, 1] - x) ** 2)
            for L1 in (1, 2, 29, 63):   # L - 1 neighbours besides the target, L <= 64
                if L1 <= len(d2):
                    assert d2[L1 - 1] <= (reach + 21) ** 2, (trial, y, x, L1, reach)
                    worst = max(worst, float(np.sqrt(d2[L1 - 1])) - reach)
                    checked += 1
    assert checked > 1000
    assert worst < 21
