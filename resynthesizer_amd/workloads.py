"""Seeded synthetic inputs of the benchmark configurations (no datasets: every image is
generated from a seed with numpy's MT19937-based RandomState).

  grating_noise  : sum of 3 sinusoidal gratings per channel (period U[6,40] px, random
                   orientation and phase, amplitude 40) + integer noise U{-16..16} + 128
  value_noise    : per-channel 4-octave value noise (cells 64/32/16/8 px, persistence 0.5,
                   smoothstep lattice) + one grating, min-max scaled to 0..255
  ellipse_mask   : union of seeded rotated ellipses
"""
import numpy as np


def grating_noise(h, w, channels, seed):
    rng = np.random.RandomState(seed)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    img = np.zeros((h, w, channels))
    for c in range(channels):
        for _ in range(3):
            th, per, ph = rng.uniform(0, np.pi), rng.uniform(6, 40), rng.uniform(0, 2 * np.pi)
            img[..., c] += 40 * np.sin(2 * np.pi * (x * np.cos(th) + y * np.sin(th)) / per + ph)
    img += rng.randint(-16, 17, size=img.shape) + 128
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def _value_octave(h, w, cell, rng):
    gh, gw = h // cell + 2, w // cell + 2
    lattice = rng.uniform(0, 1, size=(gh, gw))
    ys = np.arange(h) / cell
    xs = np.arange(w) / cell
    y0 = np.floor(ys).astype(int)
    x0 = np.floor(xs).astype(int)
    fy = ys - y0
    fx = xs - x0
    sy = fy * fy * (3 - 2 * fy)
    sx = fx * fx * (3 - 2 * fx)
    a = lattice[y0][:, x0]
    b = lattice[y0][:, x0 + 1]
    c = lattice[y0 + 1][:, x0]
    d = lattice[y0 + 1][:, x0 + 1]
    top = a + (b - a) * sx[None, :]
    bot = c + (d - c) * sx[None, :]
    return top + (bot - top) * sy[:, None]


def value_noise(h, w, channels, seed, octaves=4, base_cell=64):
    rng = np.random.RandomState(seed)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    img = np.zeros((h, w, channels))
    for c in range(channels):
        acc = np.zeros((h, w))
        amp, cell = 1.0, base_cell
        for _ in range(octaves):
            acc += amp * _value_octave(h, w, max(cell, 1), rng)
            amp *= 0.5
            cell //= 2
        th, per, ph = rng.uniform(0, np.pi), rng.uniform(6, 40), rng.uniform(0, 2 * np.pi)
        acc += 0.25 * np.sin(2 * np.pi * (x * np.cos(th) + y * np.sin(th)) / per + ph)
        lo, hi = acc.min(), acc.max()
        img[..., c] = (acc - lo) / (hi - lo) * 255.0
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def ellipse_mask(h, w, count, seed, axes=(60, 260)):
    rng = np.random.RandomState(seed)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    m = np.zeros((h, w), bool)
    for _ in range(count):
        cx, cy = rng.uniform(0, w), rng.uniform(0, h)
        a, b = rng.uniform(*axes), rng.uniform(*axes)
        t = rng.uniform(0, np.pi)
        dx, dy = x - cx, y - cy
        u = dx * np.cos(t) + dy * np.sin(t)
        v = -dx * np.sin(t) + dy * np.cos(t)
        m |= (u / a) ** 2 + (v / b) ** 2 <= 1.0
    return m


def config1(seed=1):
    """256x256 RGB grating+noise, centred 64x64 hole (4096 targets)."""
    img = grating_noise(256, 256, 3, seed)
    mask = np.zeros((256, 256), np.uint8)
    mask[96:160, 96:160] = 255
    return img, mask


# Mask seed of config 2: the union of 12 ellipses with semi-axes U[60,260] covers ~19% of
# the image on average; seed 1269 is the one among 1000..1299 closest to the configured
# ~12% (531,632 targets at 2048^2).
CONFIG2_MASK_SEED = 1269


def config2(seed=2, size=2048, mask_seed=CONFIG2_MASK_SEED):
    """size^2 RGBA value noise + grating (alpha 255), union of 12 seeded ellipses with
    semi-axes U[60,260] px (scaled with size)."""
    rgb = value_noise(size, size, 3, seed)
    img = np.concatenate([rgb, np.full((size, size, 1), 255, np.uint8)], axis=2)
    s = size / 2048.0
    m = ellipse_mask(size, size, 12, mask_seed, axes=(60 * s, 260 * s))
    return img, (m * 255).astype(np.uint8)


def luminance(rgb):
    """8-bit luminance (299 R + 587 G + 114 B + 500) / 1000."""
    r, g, b = (rgb[..., i].astype(np.int64) for i in range(3))
    return ((299 * r + 587 * g + 114 * b + 500) // 1000).astype(np.uint8)


# Format indices of the map-guided transfer: pixels [mask, R, G, B, map] (colours 1..3,
# one map channel at 4; lib/imageFormatIndicies.h:46-58).
CONFIG4_INDICES = dict(colorEndBip=4, alpha_bip=4, map_start_bip=4, map_end_bip=5, img_match_bpp=3,
                       map_match_bpp=1, total_bpp=5, isAlphaTarget=0, isAlphaSource=0)


def config4(seed=4, corpus_size=768, target_size=1536):
    """Map-guided texture transfer in the engine's pixmap layout [mask, R, G, B, map]:
    corpus = value-noise RGB texture with its luminance as the map, every pixel selected;
    target = a whole canvas (every pixel a target, no context) whose map is a radial
    gradient blended with 2-octave value noise.  Returns (target, corpus) uint8 pixmaps."""
    rgb = value_noise(corpus_size, corpus_size, 3, seed)
    corpus = np.concatenate([np.full((corpus_size, corpus_size, 1), 255, np.uint8), rgb,
                             luminance(rgb)[..., None]], axis=2)
    n = target_size
    y, x = np.mgrid[0:n, 0:n].astype(np.float64)
    r = np.sqrt((x - n / 2) ** 2 + (y - n / 2) ** 2) / (n / np.sqrt(2))
    noise = value_noise(n, n, 1, seed + 1, octaves=2, base_cell=max(n // 12, 2))[..., 0] / 255.0
    tmap = np.clip(np.rint(255.0 * (0.7 * (1.0 - r) + 0.3 * noise)), 0, 255).astype(np.uint8)
    target = np.zeros((n, n, 5), np.uint8)
    target[..., 0] = 255
    target[..., 4] = tmap
    return target, np.ascontiguousarray(corpus)


def config3_request(k, base_seed=3000):
    """Request k of the uneven-request batch: an RGB image with sides U[128,1536] from the
    grating+noise generator (seed base_seed+k), a mask of 1-6 seeded ellipses covering about
    3-25% of the image, and patchSize drawn from {10, 30, 50}.  Returns (image, mask, patch)."""
    rng = np.random.RandomState(base_seed + k)
    w, h = rng.randint(128, 1537, size=2)
    img = grating_noise(int(h), int(w), 3, base_seed + k)
    count = rng.randint(1, 7)
    frac = rng.uniform(0.03, 0.25)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    m = np.zeros((h, w), bool)
    for _ in range(count):
        area = frac * w * h / count
        aspect = rng.uniform(0.5, 2.0)
        a = np.sqrt(area * aspect / np.pi)
        b = area / (np.pi * a)
        cx, cy = rng.uniform(0.15 * w, 0.85 * w), rng.uniform(0.15 * h, 0.85 * h)
        t = rng.uniform(0, np.pi)
        dx, dy = x - cx, y - cy
        u = dx * np.cos(t) + dy * np.sin(t)
        v = -dx * np.sin(t) + dy * np.cos(t)
        m |= (u / a) ** 2 + (v / b) ** 2 <= 1.0
    patch = int(rng.choice([10, 30, 50]))
    return img, (m * 255).astype(np.uint8), patch


def config3(n=96, base_seed=3000):
    """The batch of n uneven heal requests: lists of images, masks and patch sizes."""
    reqs = [config3_request(k, base_seed) for k in range(n)]
    return [r[0] for r in reqs], [r[1] for r in reqs], [r[2] for r in reqs]


def config5(seed=5, size=4096, hole=None):
    """Learned-proposal workload: size^2 RGB value noise + grating (config 2's generator,
    3 channels) with a centred square hole of side size/4 (1024 at 4096).  Returns
    (image, mask)."""
    img = value_noise(size, size, 3, seed)
    h = hole if hole is not None else size // 4
    mask = np.zeros((size, size), np.uint8)
    o = (size - h) // 2
    mask[o:o + h, o:o + h] = 255
    return img, mask


def training_textures(count=200, size=512, base_seed=10000):
    """The training set of the proposal network: config 2's generator (RGB) with a seeded
    12-ellipse hole, one per seed base_seed .. base_seed+count-1.  Yields (image, mask)."""
    for k in range(count):
        img = value_noise(size, size, 3, base_seed + k)
        s = size / 2048.0
        m = ellipse_mask(size, size, 12, base_seed + 7919 * (k + 1), axes=(60 * s, 260 * s))
        yield img, (m * 255).astype(np.uint8)
