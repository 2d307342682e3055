"""bsm_area_intake_filtered_noalloc (staging in a static device array, no
allocation) against bsm_area_intake_filtered from a host frame, through the C
ABI, bit-exact: two calls in a row on one area (the second with a smaller
frame), a frame of exactly the static maximum, and the refusal of a larger one.
Both entry points are also compared with the oracle's retina."""
import ctypes as C

import numpy as np
import pytest

from bismit_amd import _lib as L
from bismit_amd import recipes
from bismit_amd.cortex import CorticalArea
from oracle import pyoracle as P

pytestmark = pytest.mark.gpu

MAX = 4 << 20   # BSM_FILTER_STATIC_MAX


def frame(rng, n):
    return np.where(rng.random(n) < 0.2, rng.integers(1, 256, n), 0).astype(np.uint8)


def noalloc(area, slc, depth, f):
    return area.lib.bsm_area_intake_filtered_noalloc(area.h, L.FILTER_RETINA, slc, depth, L.vptr(f))


def staged(area, slc, depth, f):
    return area.lib.bsm_area_intake_filtered(area.h, L.FILTER_RETINA, slc, depth, L.vptr(f), 0)


def test_noalloc_filtered_intake_matches_allocating_form(ctx):
    scheme = recipes.cortical_area("na", 64, input_side=24, pyr_depth=3)
    a = CorticalArea(scheme, ctx, init_layout=False)
    b = CorticalArea(scheme, ctx, init_layout=False)
    o = P.OracleArea(scheme)
    rng = np.random.default_rng(90)
    # three pyramidal slices of 64 x 64, then the 24 x 24 input slice
    for slc, depth in ((2, 3), (0, 1)):
        sl = scheme.slices[slc]
        f = frame(rng, depth * sl.v_size * sl.u_size)
        print("call", slc, depth, f.size, flush=True)
        L.check(noalloc(a, slc, depth, f))
        L.check(staged(b, slc, depth, f))
        o.retina(slc, depth, f)
        got = a.read("AXN_STATES")
        assert np.array_equal(got, b.read("AXN_STATES")), (slc, depth)
        assert np.array_equal(got, o.read("AXN_STATES")), (slc, depth)
    assert np.count_nonzero(a.read("AXN_STATES")) > 0
    a.close()
    b.close()
    o.close()


def test_noalloc_filtered_intake_at_and_past_the_maximum(ctx):
    # a 2048 x 2048 input slice is exactly the static maximum; 2100 x 2100 is past it
    for side, ok in ((2048, True), (2100, False)):
        scheme = recipes.cortical_area("big", 16, pyr_depth=1, ssc_spd=4, pyr_spd=4, distal=False)
        scheme.slices[0] = recipes.SliceScheme(side, side)
        a = CorticalArea(scheme, ctx, init_layout=False)
        f = frame(np.random.default_rng(side), side * side)
        print("side", side, f.size, f.size <= MAX, flush=True)
        rc = noalloc(a, 0, 1, f)
        if ok:
            L.check(rc)
            b = CorticalArea(scheme, ctx, init_layout=False)
            L.check(staged(b, 0, 1, f))
            assert np.array_equal(a.read("AXN_STATES"), b.read("AXN_STATES"))
            b.close()
        else:
            assert rc != 0
            msg = L.load().bsm_last_error().decode()
            assert "exceeds the static staging maximum" in msg, msg
        a.close()
