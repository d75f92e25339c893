"""wpbr_film_loss_noalloc (the film loss without device allocation, its
partials in the library's file-scope scratch) against wpbr_film_loss through
the C ABI: two calls in a row, the second on a smaller film; the loss equal
bit for bit, the image and the loss gradient equal bit for bit, and the loss
alone (no gradient or image outputs) as well."""
import ctypes as C

import numpy as np
import pytest

from woodenpbrengine_amd import lib, scenes

pytestmark = pytest.mark.gpu


def _call(ctx, fn, r, d_target, with_outputs):
    n = r.W * r.H
    d_dl = ctx.empty((n, 3), np.float32) if with_outputs else None
    d_rgb = ctx.empty((n, 3), np.float32) if with_outputs else None
    out = C.c_double(-1.0)
    ctx.check(fn(ctx.h, r.h, r.d_film.ptr, d_target.ptr, d_dl.ptr if d_dl else None,
                 d_rgb.ptr if d_rgb else None, C.byref(out)))
    res = (out.value, d_dl.numpy() if d_dl else None, d_rgb.numpy() if d_rgb else None)
    for a in (d_dl, d_rgb):
        if a is not None:
            a.free()
    return res


@pytest.mark.parametrize("with_outputs", [True, False])
def test_film_loss_noalloc_matches_film_loss(ctx, with_outputs):
    so = lib.load()
    g = np.random.default_rng(77)
    for w, h, spp in ((256, 192, 2), (64, 64, 1)):  # the second call's film is smaller
        s = scenes.c0_scene(width=w, height=h, spp=spp)
        r = lib.DiffRenderer(ctx, s)
        r.forward()
        target = (r.film_rgb() * (1 + 0.2 * g.uniform(-1, 1, (h, w, 3)))).astype(np.float32)
        d_target = ctx.array(target)
        ref = _call(ctx, so.wpbr_film_loss, r, d_target, with_outputs)
        got = _call(ctx, so.wpbr_film_loss_noalloc, r, d_target, with_outputs)
        print("film %dx%d: loss %.9g / %.9g" % (w, h, ref[0], got[0]), flush=True)
        assert ref[0] > 0.0
        assert np.float64(got[0]).view(np.uint64) == np.float64(ref[0]).view(np.uint64)
        if with_outputs:
            np.testing.assert_array_equal(got[1].view(np.uint32), ref[1].view(np.uint32))
            np.testing.assert_array_equal(got[2].view(np.uint32), ref[2].view(np.uint32))
        d_target.free()
        r.destroy()


def test_film_loss_noalloc_rejects_null_arguments(ctx):
    so = lib.load()
    out = C.c_double(0.0)
    assert so.wpbr_film_loss_noalloc(ctx.h, None, None, None, None, None, C.byref(out)) == 1
