"""mh_generate_device_noalloc, the allocation-free form of mh_generate_device:
its working memory is the library's static per-GPU scratch array.  Its
outputs must equal mh_generate_device's for the same job bit for bit (grid,
composited image, placement counters, annealing costs and counts), on two
calls in a row where the second job is smaller (the array keeps the first
call's contents), and it refuses jobs beyond the array or without cached
tiles before enqueuing anything."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _job(ctx, mh, synth, name, seed):
    p, target, mats = synth.make(name)
    lib = ctx.library_from_pixels(mats)
    d_t = ctx.to_device(target)
    prm = mh.JobParams(p["grid_w"], p["grid_h"], p["max_usage"], p["weight"], p["color"], 1,
                       p["iters"], 100.0, 0.99995, seed)
    return p, lib, d_t, prm


def _run(ctx, fn, lib, d_t, p, prm, own_grid=True):
    cells = p["grid_w"] * p["grid_h"]
    tw, th = p["width"] // p["grid_w"], p["height"] // p["grid_h"]
    shape = (p["grid_h"] * th, p["grid_w"] * tw, 3)
    d_g = ctx.device_alloc(cells * 4) if own_grid else None
    d_i = ctx.device_alloc(int(np.prod(shape)))
    fn(lib, d_t, p["width"], p["height"], prm, d_g, d_i)
    ctx.synchronize()
    res = ctx.generate_device_result()
    grid = ctx.to_host(d_g, cells, np.int32) if own_grid else None
    img = ctx.to_host(d_i, shape, np.uint8)
    for ptr in (d_g, d_i):
        if ptr:
            ctx.device_free(ptr)
    return grid, img, res


def _same(a, b):
    if a[0] is not None and b[0] is not None:
        np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    assert a[2] == b[2]  # placement counters, annealing costs/counts, output size


def test_noalloc_equals_generate_device_on_two_calls(mh, synth):
    ctx = mh.Context(0)
    big = _job(ctx, mh, synth, "medium", 3)
    small = _job(ctx, mh, synth, "small", 4)
    # the existing entry point first: its results, and the libraries' resized
    # tiles for these cell sizes (the allocation-free form only reads them)
    want = [_run(ctx, ctx.generate_device, lib, d_t, p, prm) for p, lib, d_t, prm in (big, small)]
    print("noalloc: reference jobs done", flush=True)
    got = []
    for p, lib, d_t, prm in (big, small):  # two calls in a row, the second smaller
        print(f"noalloc: {p['grid_w']}x{p['grid_h']} cells, {p['n']} materials", flush=True)
        got.append(_run(ctx, ctx.generate_device_noalloc, lib, d_t, p, prm))
    for w, g in zip(want, got):
        _same(w, g)
        assert g[2]["place"]["placed"] > 0 and g[2]["anneal"]["iterations"] > 0
    # the grid kept in the scratch array too (no grid buffer passed), after
    # the smaller job left its contents behind
    p, lib, d_t, prm = big
    _same(want[0], _run(ctx, ctx.generate_device_noalloc, lib, d_t, p, prm, own_grid=False))
    for _, lib, d_t, _ in (big, small):
        ctx.device_free(d_t)
        lib.close()
    ctx.close()


def test_noalloc_refuses_what_it_cannot_hold(mh, synth):
    ctx = mh.Context(0)
    p, lib, d_t, prm = _job(ctx, mh, synth, "small", 0)
    # 300 x 300 cells need far more than the scratch array holds
    big = mh.JobParams(300, 300, 3, 0.25, 0.4, 1, 500, 100.0, 0.99995, 0)
    with pytest.raises(mh.MhError) as e:
        ctx.generate_device_noalloc(lib, d_t, 600, 450, big, None, None)
    assert e.value.code == mh.MH_ERR_UNSUPPORTED and "allocation-free" in str(e.value)
    # no resized tiles for this cell size yet: refused before any work
    d_i = ctx.device_alloc(p["width"] * p["height"] * 3)
    with pytest.raises(mh.MhError) as e:
        ctx.generate_device_noalloc(lib, d_t, p["width"], p["height"], prm, None, d_i)
    assert e.value.code == mh.MH_ERR_UNSUPPORTED and "resized tiles" in str(e.value)
    # placement and annealing alone need no tiles
    d_g = ctx.device_alloc(p["grid_w"] * p["grid_h"] * 4)
    ctx.generate_device_noalloc(lib, d_t, p["width"], p["height"], prm, d_g, None)
    ctx.synchronize()
    got = ctx.to_host(d_g, p["grid_w"] * p["grid_h"], np.int32)
    ctx.generate_device(lib, d_t, p["width"], p["height"], prm, d_g, None)
    ctx.synchronize()
    np.testing.assert_array_equal(got, ctx.to_host(d_g, p["grid_w"] * p["grid_h"], np.int32))
    for ptr in (d_i, d_g, d_t):
        ctx.device_free(ptr)
    lib.close()
    ctx.close()
