"""Deterministic synthetic mosaic workloads (numpy PCG64, fixed seeds).

Workloads by name (sizes as the benchmark configurations state them):
  small   640x480 ramp target (R=x ramp, G=y ramp, B=128, noise s=10);
          200 materials 64x64, flat U{0..255}^3 + noise s=12; grid 40x30,
          usage 3, weight 0.25, colour 0.4, 500 SA iterations.
  medium  1920x1080 diagonal gradient + 8 Gaussian blobs (s=150 px), noise
          s=8; 1,500 materials 128x128 flat uniform + noise s=12; grid 80x60,
          usage 4, weight 0.25, colour 0.4, 2,000 SA iterations.
  large   4000x2400 as medium with 32 blobs; 10,000 materials 64x64 whose
          flat colours come from a 3-component Gaussian mixture (means
          U[0,255]^3, s=12); grid 200x120, usage 5, weight 0.25, 20,000 SA.
  batch   64 jobs: grid U{20..160} x U{15..120}, 16 px cells, library
          U{256..4096} materials 64x64 (uniform flat + noise 12), usage
          U{3..5}, 1,000 SA iterations each.
  stream  5,000 jobs arriving as a Poisson process of rate lambda; cells
          log-uniform in [300, 24000], aspect 4:3 or 16:9, 16 px cells;
          each job uses one of 4 resident libraries (1,000 / 2,500 / 5,000 /
          10,000 materials 64x64, mixture colours as in large); usage
          U{3..5}, weight 0.25, colour 0.4, 1,000 SA iterations.  Targets are
          crops of 4 canvases of 4000x2400 built like the large target.
"""
import numpy as np

WORKLOADS = {
    "small": dict(width=640, height=480, n=200, msize=64, grid_w=40, grid_h=30, max_usage=3,
                  weight=0.25, color=0.4, iters=500),
    "medium": dict(width=1920, height=1080, n=1500, msize=128, grid_w=80, grid_h=60, max_usage=4,
                   weight=0.25, color=0.4, iters=2000),
    "large": dict(width=4000, height=2400, n=10000, msize=64, grid_w=200, grid_h=120,
                  max_usage=5, weight=0.25, color=0.4, iters=20000),
}


def _u8(a):
    return np.clip(np.rint(a), 0, 255).astype(np.uint8)


_NOISE_TABLES = {}


def _noise_table(sigma):
    """rint(N(0, sigma)) at the 65,536 mid-quantiles: indexing it with uniform
    16-bit draws samples the rounded normal (inverse-CDF sampling)."""
    t = _NOISE_TABLES.get(sigma)
    if t is None:
        from scipy.special import ndtri
        u = (np.arange(65536, dtype=np.float64) + 0.5) / 65536.0
        t = np.rint(ndtri(u) * sigma).astype(np.int16)
        _NOISE_TABLES[sigma] = t
    return t


def ramp_target(width, height, rng):
    x = np.arange(width, dtype=np.float32)
    y = np.arange(height, dtype=np.float32)
    img = np.empty((height, width, 3), np.float32)
    img[..., 0] = np.rint(255.0 * x / (width - 1))[None, :]
    img[..., 1] = np.rint(255.0 * y / (height - 1))[:, None]
    img[..., 2] = 128.0
    img += rng.standard_normal(img.shape, dtype=np.float32) * 10.0
    return _u8(img)


def blob_target(width, height, n_blobs, rng, sigma=150.0, noise=8.0):
    x = np.arange(width, dtype=np.float32) / width
    y = np.arange(height, dtype=np.float32) / height
    base = 255.0 * (x[None, :] + y[:, None]) / 2.0
    img = np.repeat(base[..., None], 3, axis=2).astype(np.float32)
    cx = rng.uniform(0, width, n_blobs)
    cy = rng.uniform(0, height, n_blobs)
    col = rng.uniform(0, 255, (n_blobs, 3)).astype(np.float32)
    r = int(4 * sigma)
    for b in range(n_blobs):
        x0, x1 = max(int(cx[b]) - r, 0), min(int(cx[b]) + r, width)
        y0, y1 = max(int(cy[b]) - r, 0), min(int(cy[b]) + r, height)
        if x0 >= x1 or y0 >= y1:
            continue
        xs = np.arange(x0, x1, dtype=np.float32) - cx[b]
        ys = np.arange(y0, y1, dtype=np.float32) - cy[b]
        g = np.exp(-(xs[None, :] ** 2 + ys[:, None] ** 2) / (2 * sigma * sigma)).astype(np.float32)
        amp = col[b][None, None, :] - img[y0:y1, x0:x1, :]
        img[y0:y1, x0:x1, :] += g[..., None] * amp
    img += rng.standard_normal(img.shape, dtype=np.float32) * noise
    return _u8(img)


def flat_materials(n, size, rng, colours=None, noise=12.0):
    """n flat tiles: colour (rounded to integers) + rint(N(0, noise)) per
    channel and pixel, clipped to [0, 255].  Since the colour is integral,
    this equals rint(colour + N(0, noise)) in distribution."""
    if colours is None:
        colours = rng.integers(0, 256, (n, 3)).astype(np.float32)
    col = np.rint(np.asarray(colours, np.float64)).astype(np.int16)
    table = _noise_table(float(noise))
    out = np.empty((n, size, size, 3), np.uint8)
    chunk = max(1, (1 << 24) // (size * size * 3))
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        idx = rng.integers(0, 65536, (e - s, size, size, 3), dtype=np.uint16)
        v = table[idx]
        v += col[s:e, None, None, :]
        np.clip(v, 0, 255, out=v)
        out[s:e] = v
    return out


def mixture_colours(n, rng, k=3, sigma=12.0):
    means = rng.uniform(0, 255, (k, 3))
    comp = rng.integers(0, k, n)
    return (means[comp] + rng.standard_normal((n, 3)) * sigma).astype(np.float32)


def make(name, seed=0):
    """Returns (params, target u8 HxWx3, materials u8 NxSxSx3)."""
    p = dict(WORKLOADS[name])
    rng = np.random.default_rng(seed)
    if name == "small":
        target = ramp_target(p["width"], p["height"], rng)
        mats = flat_materials(p["n"], p["msize"], rng)
    elif name == "medium":
        target = blob_target(p["width"], p["height"], 8, rng)
        mats = flat_materials(p["n"], p["msize"], rng)
    elif name == "large":
        target = blob_target(p["width"], p["height"], 32, rng)
        mats = flat_materials(p["n"], p["msize"], rng, colours=mixture_colours(p["n"], rng))
    else:
        raise KeyError(name)
    return p, target, mats


def _parallel(fn, args):
    """fn(*a) for each a on a thread pool (numpy releases the GIL); results in
    order, so the output does not depend on the thread count."""
    import concurrent.futures
    import os
    workers = max(1, min(16, os.cpu_count() or 1, len(args)))
    with concurrent.futures.ThreadPoolExecutor(workers) as ex:
        return list(ex.map(lambda a: fn(*a), args))


def make_batch(n_jobs=64, seed=0, cell=16, msize=64):
    """Independent jobs of the batched form: list of (params, target, materials).
    Job j's pixels come from its own child stream of SeedSequence(seed)."""
    rng = np.random.default_rng(seed)
    children = np.random.SeedSequence(seed).spawn(n_jobs)
    specs = []
    for _ in range(n_jobs):
        gw = int(rng.integers(20, 161))
        gh = int(rng.integers(15, 121))
        n = int(rng.integers(256, 4097))
        usage = int(rng.integers(3, 6))
        specs.append(dict(width=gw * cell, height=gh * cell, n=n, msize=msize, grid_w=gw, grid_h=gh,
                          max_usage=usage, weight=0.25, color=0.4, iters=1000))

    def one(p, ss):
        r = np.random.default_rng(ss)
        target = blob_target(p["width"], p["height"], 8, r, sigma=max(8.0, p["width"] / 12.0))
        return p, target, flat_materials(p["n"], msize, r)

    return _parallel(one, list(zip(specs, children)))


def make_stream(n_jobs=5000, rate=2000.0, seed=0, libraries=(1000, 2500, 5000, 10000),
                cells=(300, 24000), canvas=(4000, 2400), n_canvases=4, cell=16, msize=64,
                iters=1000):
    """The served job stream.  Returns a dict:
      libs      list of material arrays (n x msize x msize x 3 u8)
      canvases  list of canvas targets (H x W x 3 u8)
      jobs      list of dicts: arrival (s), lib, canvas, ox, oy (crop origin in
                pixels), grid_w, grid_h, width, height, max_usage, weight, color,
                iters, seed
    Arrival times are the partial sums of exponential gaps of mean 1/rate."""
    rng = np.random.default_rng(seed)
    kids = np.random.SeedSequence(seed).spawn(len(libraries) + n_canvases)
    cw, ch = canvas

    def lib(n, ss):
        r = np.random.default_rng(ss)
        return flat_materials(n, msize, r, colours=mixture_colours(n, r))

    libs = _parallel(lib, list(zip(libraries, kids[:len(libraries)])))
    canvases = _parallel(lambda ss: blob_target(cw, ch, 32, np.random.default_rng(ss)),
                         [(k,) for k in kids[len(libraries):]])
    lo, hi = np.log(cells[0]), np.log(cells[1])
    t = 0.0
    jobs = []
    for j in range(n_jobs):
        t += float(rng.exponential(1.0 / rate))
        c = float(np.exp(rng.uniform(lo, hi)))
        aspect = 16.0 / 9.0 if rng.integers(0, 2) else 4.0 / 3.0
        gw = max(1, int(round(np.sqrt(c * aspect))))
        gh = max(1, int(round(c / gw)))
        gw = min(gw, cw // cell)
        gh = min(gh, ch // cell)
        w, h = gw * cell, gh * cell
        jobs.append(dict(arrival=t, lib=int(rng.integers(0, len(libraries))),
                         canvas=int(rng.integers(0, n_canvases)),
                         ox=int(rng.integers(0, cw - w + 1)), oy=int(rng.integers(0, ch - h + 1)),
                         grid_w=gw, grid_h=gh, width=w, height=h,
                         max_usage=int(rng.integers(3, 6)), weight=0.25, color=0.4, iters=iters,
                         seed=j))
    return dict(libs=libs, canvases=canvases, jobs=jobs, canvas_size=canvas)
