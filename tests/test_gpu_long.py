"""The workloads at their full sizes and run lengths.  The oracle is too slow
for these, so they use properties that hold at any size: two different kernel
schedules and tilings of the same area, and the batched form against the plain
form, must end bit-exact after the workload's full number of cycles.  The plain
form itself is pinned to the oracle cycle by cycle in test_gpu_parity."""
import numpy as np
import pytest

from bismit_amd import inputs, recipes
from bismit_amd.cortex import Batch, CorticalArea

from helpers import layer_bufs

pytestmark = pytest.mark.gpu


def assert_same(a, b, scheme, tag):
    assert np.array_equal(a.read("AXN_STATES"), b.read("AXN_STATES")), f"{tag} axons"
    for li in range(len(scheme.layers)):
        for name in layer_bufs(scheme, li):
            x, y = a.read(name, li), b.read(name, li)
            if not np.array_equal(x, y):
                bad = np.flatnonzero(x != y)
                raise AssertionError(f"{tag} layer {li} {name}: {bad.size} mismatches at "
                                     f"{bad[:6].tolist()}")


def test_config2_10k_cycles_schedules_agree(ctx):
    """Config 2 (256x256, 142.6M synapses), 10,000 cycles of 2% SDR input with
    50% overlap, learning on: the two-stream schedule with 2x16 tiles and 16
    synapses per lane equals the one-stream schedule with 8x8 tiles and 8 per
    lane (different kernel variants, work lists and launch order)."""
    scheme = recipes.config2_area()
    n, nf = 10000, 256
    fr = inputs.SdrStream(scheme.input_range()[1], 0.02, 0.5, seed=21).frames(nf)
    dev = ctx.alloc(fr.nbytes)
    ctx.upload(dev, fr)
    a = CorticalArea(scheme, ctx, rnd_seed=5, layout_seed=6)
    b = CorticalArea(scheme, ctx, rnd_seed=5, layout_seed=6)
    a.set_tuning(tile_v=2, tile_u=16, chunk=16, halo=6, overlap=1)
    b.set_tuning(tile_v=8, tile_u=8, chunk=8, halo=8, overlap=2, inhib_tile=32)
    assert a.stage_plan() != b.stage_plan()
    for area in (a, b):
        done = 0
        while done < n:   # the frame ring repeats every nf cycles
            k = min(nf, n - done)
            area.run(dev, fr.shape[1], fr.shape[1], k)
            done += k
        area.sync()
    assert np.count_nonzero(a.read("SYN_STRENGTHS", 1)) > 0
    assert_same(a, b, scheme, "c2 10k")
    a.close()
    b.close()
    ctx.free(dev)


def test_config3_5k_cycles_batched_equals_plain(ctx):
    """Config 3: the 64-area mix, 5,000 graph-captured batched cycles (two
    streams, autotuner-sized tiles) against each area cycled alone."""
    schemes = recipes.config3_areas(64, seed=0)
    n, nf = 5000, 64
    batched = [CorticalArea(s, ctx, rnd_seed=3000 + i, layout_seed=4000 + i)
               for i, s in enumerate(schemes)]
    plain = [CorticalArea(s, ctx, rnd_seed=3000 + i, layout_seed=4000 + i)
             for i, s in enumerate(schemes)]
    devs, lens = [], []
    for i, s in enumerate(schemes):
        f = inputs.SdrStream(s.input_range()[1], 0.02, 0.5, seed=500 + i).frames(nf)
        d = ctx.alloc(f.nbytes)
        ctx.upload(d, f)
        devs.append(d)
        lens.append(f.shape[1])
    b = Batch(batched, ctx)
    b.set_tuning(tile_v=4, tile_u=16, chunk=16, halo=6)
    b.set_inputs(devs, lens, lens, [s.input_range()[0] for s in schemes], nf)
    b.cycle(n, graph=True)
    b.sync()
    b.close()
    for i, a in enumerate(plain):
        done = 0
        while done < n:
            k = min(nf, n - done)
            a.run(devs[i], lens[i], lens[i], k)
            done += k
        a.sync()
    for i, s in enumerate(schemes):
        assert_same(batched[i], plain[i], s, f"c3 5k area {i}")
    for a in batched + plain:
        a.close()
    for d in devs:
        ctx.free(d)


def test_config4_20k_cycles_batched_equals_plain(ctx):
    """Config 4: three 128x128 areas with feed-forward and feedback tracts and
    the Zipf/Markov symbol stream, 20,000 cycles, batched + graph-captured
    against the plain form."""
    from bismit_amd.hier import Hierarchy
    hp = Hierarchy(ctx, seed=9, batched=False)
    hb = Hierarchy(ctx, seed=9, batched=True)
    frames = hp.frames(20000)
    hp.run(frames)
    hb.run(frames, graph=True)
    for i, s in enumerate(hp.schemes):
        assert_same(hb.areas[i], hp.areas[i], s, f"hier 20k area {i}")
    hb.batch.close()
    hb.batch = None
    hp.close()
    hb.close()
