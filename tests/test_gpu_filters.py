"""Sensory filter (retina) and activity smoother against the oracle, bit-exact:
the kernels alone on random inputs, the filtered intake from host and device
memory, and whole Config 1 cycles with both attached (the smoother raises
energies until the spiny stellate soma's restless/dark rules fire)."""
import dataclasses

import numpy as np
import pytest

from bismit_amd import inputs, recipes
from bismit_amd.cortex import Batch, CorticalArea
from oracle import pyoracle as P

from helpers import compare
from helpers import make_pair

pytestmark = pytest.mark.gpu


def frame_like(rng, n, kind):
    if kind == "glyph":
        return np.where(rng.random(n) < 0.15, 255, 0).astype(np.uint8)
    if kind == "dense":
        return rng.integers(0, 256, n).astype(np.uint8)
    return (rng.random(n) < 0.3).astype(np.uint8) * rng.integers(0, 256, n).astype(np.uint8)


@pytest.mark.parametrize("kind", ["glyph", "dense", "sparse"])
def test_retina_kernel_matches_oracle(ctx, kind):
    scheme = recipes.config1_area()                   # 24x24 input slice 0, 64x64 slices 1..3
    g, o = make_pair(ctx, scheme)
    rng = np.random.default_rng({"glyph": 1, "dense": 2, "sparse": 3}[kind])
    for slc, depth in [(0, 1), (2, 2)]:               # the input slice; two pyramidal slices
        sl = scheme.slices[slc]
        f = frame_like(rng, depth * sl.v_size * sl.u_size, kind)
        d = ctx.alloc(f.nbytes)
        ctx.upload(d, f)
        g.k_retina(slc, depth, d)
        g.sync()
        o.retina(slc, depth, f)
        ctx.free(d)
        assert np.array_equal(g.read("AXN_STATES"), o.read("AXN_STATES")), (kind, slc)
    assert np.count_nonzero(g.read("AXN_STATES")) > 0
    g.close()
    o.close()


def test_filtered_intake_host_and_device(ctx):
    scheme = dataclasses.replace(recipes.config1_area(), input_filter="retina")
    g, o = make_pair(ctx, scheme)
    fr = inputs.GlyphStream(side=24, p=0.15, hold=1, seed=8).frames(2)
    g.intake(fr[0])
    o.intake(fr[0])
    assert np.array_equal(g.read("AXN_STATES"), o.read("AXN_STATES"))
    d = ctx.alloc(fr[1].nbytes)
    ctx.upload(d, fr[1])
    g.intake_device(d, fr[1].size)
    o.intake(fr[1])
    g.sync()
    assert np.array_equal(g.read("AXN_STATES"), o.read("AXN_STATES"))
    ctx.free(d)
    with pytest.raises(ValueError):
        Batch([g], ctx)
    g.close()
    o.close()


@pytest.mark.parametrize("r,ssc_depth,side", [(1, 1, 16), (2, 2, 20), (4, 1, 64), (6, 2, 33)])
def test_smoother_kernel_matches_oracle(ctx, r, ssc_depth, side):
    scheme = recipes.cortical_area("sm", side, ssc_depth=ssc_depth, ssc_dens=2, ssc_spd=8,
                                   pyr_dens=2, pyr_spd=8)
    g, o = make_pair(ctx, scheme)
    g.set_smoother(0, r)
    o.set_smoother(0, r)
    rng = np.random.default_rng(r * 10 + ssc_depth)
    n = scheme.cells(0)
    for rep in range(3):
        # few distinct activity values (ties everywhere) and energies at both bounds
        act = rng.choice(np.array([0, 3, 3, 40, 200, 255], np.uint8), n)
        en = rng.choice(np.array([0, 1, 128, 254, 255], np.uint8), n)
        for a in (g, o):
            a.write("SOMA_ACTIVITIES", 0, act)
            if rep == 0:
                a.write("SOMA_ENERGIES", 0, en)
        g.k_smooth(0)
        g.sync()
        o.k("smooth", 0)
        assert np.array_equal(g.read("SOMA_ENERGIES", 0), o.read("SOMA_ENERGIES", 0)), rep
    assert np.count_nonzero(g.read("SOMA_ENERGIES", 0) != en) > 0
    g.close()
    o.close()


def test_cycle_config1_with_smoother_and_retina(ctx):
    """Config 1 with a radius-4 smoother on the spiny stellate layer and the
    retina on the glyph input, 300 cycles, every buffer after every cycle."""
    base = recipes.config1_area()
    layers = list(base.layers)
    layers[0] = dataclasses.replace(layers[0], smoother_radius=4)
    scheme = dataclasses.replace(base, layers=layers, input_filter="retina")
    g, o = make_pair(ctx, scheme, layout_seed=3)
    assert any(st["kind"] == 4 for st in g.stage_plan())
    n = 300
    frames = inputs.GlyphStream(side=24, p=0.15, hold=10, seed=4).frames(n)
    rnd = g.draw_rnd(n)
    for c in range(n):
        g.intake(frames[c])
        o.intake(frames[c])
        r = rnd[c * g.n_slots:(c + 1) * g.n_slots]
        g.cycle(1, r)
        o.cycle(r, 1)
        compare(g, o, scheme, f"smoother+retina cycle {c}")
    # the smoother drove energies past the restless threshold (191) somewhere
    assert o.read("SOMA_ENERGIES", 0).max() > 191
    with pytest.raises(Exception):
        Batch([g], ctx)
    g.close()
    o.close()
