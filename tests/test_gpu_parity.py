"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle,
bit-exact on every buffer.  Integer path: the bar is exact equality."""
import ctypes as C
import os

import numpy as np
import pytest

from bismit_amd import _lib as L
from bismit_amd import inputs, recipes
from bismit_amd.cortex import CorticalArea
from oracle import pyoracle as P

from helpers import compare, layer_bufs, make_pair, run_cycles

pytestmark = pytest.mark.gpu

P.set_threads(min(16, os.cpu_count() or 1))


def randomize(g, o, scheme, rng, strengths=True):
    """Random contents for every mutable buffer, identical on both sides."""
    def put(name, li, arr):
        g.write(name, li, arr)
        o.write(name, li, arr)

    axn = rng.choice(np.array([0, 0, 0, 0, 17, 128, 255], dtype=np.uint8), g.axn_len)
    put("AXN_STATES", 0, axn)
    for li, lay in enumerate(scheme.layers):
        for name in layer_bufs(scheme, li, synapses=False):
            if name == "SOMA_STATES" and lay.kind == L.LAYER_SSC:
                continue  # alias of DEN_STATES
            n = g.buf_len(name, li)
            vals = rng.integers(0, 256, n).astype(np.uint8)
            if name in ("TFT_BEST_DEN_IDS", "TFT_PREV_BEST_DEN_IDS"):
                dmin = min(t.dens_per_tft for t in lay.tufts)
                vals = rng.integers(0, dmin, n).astype(np.uint8)
            if name in ("SOMA_ENERGIES",):
                vals = rng.choice(np.array([0, 0, 100, 192, 255], dtype=np.uint8), n)
            put(name, li, vals)
        n = g.buf_len("SYN_STATES", li)
        put("SYN_STATES", li, rng.choice(np.array([0, 0, 0, 128, 200, 255], dtype=np.uint8), n))
        put("SYN_FLAG_SETS", li, rng.integers(0, 256, n).astype(np.uint8))
        if strengths:
            put("SYN_STRENGTHS", li, rng.integers(-127, 128, n).astype(np.int8))


SCHEMES = {
    "small": lambda: recipes.cortical_area("s", 16, input_side=12, pyr_depth=2, ssc_dens=2,
                                           ssc_spd=8, pyr_dens=3, pyr_spd=12, inhib_radius=3,
                                           thresholds=(200, 0, 300)),
    "c1": recipes.config1_area,
    "apical": lambda: _with_apical(),
}


def _with_apical():
    s = recipes.cortical_area("a", 24, pyr_depth=3, ssc_dens=4, ssc_spd=16, pyr_dens=2,
                              pyr_spd=20, thresholds=(300, 0, 100))
    s.layers[1].tufts.append(recipes.TuftScheme.apical_distal(3, 8, 50, [1, 2]))
    return s


@pytest.mark.parametrize("which", ["small", "c1", "apical"])
def test_kernels_one_by_one(ctx, which):
    """Each bsm_k_* entry point against its oracle restatement on random state."""
    scheme = SCHEMES[which]()
    print("case", which, flush=True)
    g, o = make_pair(ctx, scheme)
    rng = np.random.default_rng(11)
    randomize(g, o, scheme, rng)
    compare(g, o, scheme, "init")
    for li, lay in enumerate(scheme.layers):
        steps = [("syn_flags", (li,))]
        for t in range(len(lay.tufts)):
            steps.append(("syn_cycle", (li, t)))
            steps.append(("den_cycle", (li, t, int(rng.integers(-2**31, 2**31)))))
        if lay.kind == L.LAYER_SSC:
            steps += [("ssc_cycle", (li,)), ("inhib_simple", (li, int(rng.integers(-2**31, 2**31)))),
                      ("ssc_mtp", (li, int(rng.integers(0, 2**32)))),
                      ("inhib_passthrough", (li, int(rng.integers(-2**31, 2**31))))]
        else:
            for t in range(len(lay.tufts)):
                steps.append(("tft_cycle", (li, t)))
            steps += [("pyr_cycle", (li,)), ("inhib_intra_column", (li,)), ("pyr_output", (li,))]
            for t in range(len(lay.tufts)):
                if lay.tufts[t].den_kind == L.DEN_DISTAL:
                    steps.append(("tft_dst_mtp", (li, t, int(rng.integers(-2**31, 2**31)))))
            steps.append(("inhib_intra_column", (li,)))
            for t in range(len(lay.tufts)):
                if lay.tufts[t].den_kind == L.DEN_DISTAL:
                    steps.append(("tft_dst_mtp", (li, t, int(rng.integers(-2**31, 2**31)))))
        for name, args in steps:
            getattr(g, "k_" + name)(*args)
            o.k(name, *args)
            compare(g, o, scheme, f"{which} after {name}{args}")
    g.close()


def test_cycle_config1_1000_cycles(ctx):
    """Config 1: 64x64, 24x24 glyph stream (p=0.15, new glyph every 10 cycles),
    learning on, 1000 cycles, every buffer compared after every cycle."""
    scheme = recipes.config1_area()
    g, o = make_pair(ctx, scheme, layout_seed=0)
    # host-written layouts within 6 columns: every tuft launch check-free; the
    # distal tuft (own layer's slices) runs on the side stream
    plan = g.stage_plan()
    assert [(p["kind"], p["stream"]) for p in plan] == [(1, 0), (2, 0), (1, 1), (1, 0), (3, 0)]
    assert all(p["fast"] == 1 and p["packed"] == 1 for p in plan if p["kind"] == 1)
    frames = inputs.GlyphStream(side=24, p=0.15, hold=10, seed=0).frames(1000)
    run_cycles(g, o, scheme, frames, 1000, tag="c1")
    # the run did learn and fire
    assert np.count_nonzero(g.read("SYN_STRENGTHS", 0)) > 0
    assert np.count_nonzero(g.read("SYN_STRENGTHS", 1)) > 0
    assert np.count_nonzero(g.axon_states()[576:]) > 0
    g.close()


def test_fused_intake_runs_match_oracle(ctx):
    """The Config 2 recipe at 48x48 columns through multi-cycle calls from
    device frames: each frame is exactly the spiny stellate tuft's one staged
    slice, which no pyramidal tuft reads, so the spiny stellate F1 launch takes
    it (reads its windows from the frame, writes it into the axon slice) in
    place of an intake copy.  Every buffer equals the oracle's after each call
    (calls of 1 to 17 cycles), and the input slice holds the call's last frame."""
    scheme = recipes.cortical_area("c2s", 48, pyr_dens=8, pyr_spd=64)
    g, o = make_pair(ctx, scheme, layout_seed=3)
    twin = CorticalArea(scheme, ctx, rnd_seed=0, init_layout=False)
    off, ln = scheme.input_range()
    assert ln == 48 * 48
    frames = inputs.SdrStream(ln, 0.02, 0.5, seed=4).frames(40)
    dev = ctx.alloc(frames.nbytes)
    ctx.upload(dev, frames)
    done = 0
    for k in (1, 2, 5, 17, 15):
        print("run", done, k, flush=True)
        rnd = twin.draw_rnd(k)
        g.run(C.c_void_p(dev.value + done * ln), ln, ln, k, offset=off)
        for c in range(k):
            o.intake(frames[done + c])
            o.cycle(rnd[c * g.n_slots:(c + 1) * g.n_slots], 1)
        done += k
        g.sync()
        assert np.array_equal(g.axon_states()[off:off + ln], frames[done - 1])
        compare(g, o, scheme, f"run to cycle {done}")
    assert np.count_nonzero(g.read("SYN_STRENGTHS", 0)) > 0
    g.close()
    twin.close()
    ctx.free(dev)


def test_cycle_config1_multi_cycle_runs(ctx):
    """Config 1 through multi-cycle calls (bsm_area_run: intake + cycle per
    frame from device memory, rnd from the area's own stream; then
    bsm_area_cycle of 40 cycles with caller rnd): every buffer equals the
    oracle's run after each call, for calls of 2 to 150 cycles."""
    scheme = recipes.config1_area()
    g, o = make_pair(ctx, scheme, layout_seed=0)
    # the twin area draws the same rnd stream the run calls consume
    twin = CorticalArea(scheme, ctx, rnd_seed=0, init_layout=False)
    off, ln = scheme.input_range()
    frames = inputs.GlyphStream(side=24, p=0.15, hold=10, seed=0).frames(350)
    dev = ctx.alloc(frames.nbytes)
    ctx.upload(dev, frames)
    done = 0
    for k in (2, 3, 7, 50, 91, 150):
        print("run", done, k, flush=True)
        rnd = twin.draw_rnd(k)
        g.run(C.c_void_p(dev.value + done * ln), ln, ln, k, offset=off)
        for c in range(k):
            o.intake(frames[done + c])
            o.cycle(rnd[c * g.n_slots:(c + 1) * g.n_slots], 1)
        done += k
        g.sync()
        compare(g, o, scheme, f"run to cycle {done}")
    assert np.count_nonzero(g.read("SYN_STRENGTHS", 1)) > 0
    rnd = twin.draw_rnd(40)
    g.cycle(40, rnd)
    for c in range(40):
        o.cycle(rnd[c * g.n_slots:(c + 1) * g.n_slots], 1)
    compare(g, o, scheme, "cycle(40)")
    g.close()
    twin.close()
    ctx.free(dev)


@pytest.mark.parametrize("tuning", [
    dict(tile_v=4, tile_u=4, halo=1, chunk=4, overlap=2),
    dict(tile_v=16, tile_u=8, halo=8, chunk=8, inhib_tile=32),
    dict(tile_v=8, tile_u=16, halo=3, chunk=16),
    dict(tile_v=3, tile_u=5, halo=2, chunk=1, inhib_tile=8),
    dict(tile_v=2, tile_u=16, halo=7, chunk=32),
    dict(tile_v=4, tile_u=32, halo=7, chunk=32, overlap=2),
])
def test_cycle_tunings(ctx, tuning):
    """Any tile/halo/chunk/inhibition-tile choice gives identical results
    (chunk 1 forces the checked gather, whose small halos take the out-of-window
    global path; 32 x 32 inhibition tiles put four cells on each thread)."""
    print("tuning", tuning, flush=True)
    scheme = recipes.cortical_area("t", 40, input_side=32, pyr_depth=3, ssc_dens=4, ssc_spd=16,
                                   pyr_dens=4, pyr_spd=24, inhib_radius=4)
    g, o = make_pair(ctx, scheme, reach=7)
    g.set_tuning(**tuning)
    rng = np.random.default_rng(3)
    randomize(g, o, scheme, rng)
    frames = inputs.SdrStream(32 * 32, 0.05, 0.5, seed=1).frames(40)
    run_cycles(g, o, scheme, frames, 40, tag=str(tuning))
    g.close()


def test_cycle_scalar_quirk_nonspatial(ctx):
    """syns_per_tft % 4 != 0 takes the scalar gather (char-wrapped mids) on a
    non-spatial source slice with a large mid (cl/bismit.cl:331-332)."""
    s = recipes.cortical_area("q", 24, pyr_depth=2, ssc_dens=3, ssc_spd=5, pyr_dens=3,
                              pyr_spd=7, thresholds=(100, 0, 100))
    s.slices[0] = recipes.SliceScheme.nonspatial(230, 230)
    g, o = make_pair(ctx, s, reach=60)
    rng = np.random.default_rng(8)
    randomize(g, o, s, rng)
    frames = (rng.random((30, 230 * 230)) < 0.1).astype(np.uint8) * 255
    run_cycles(g, o, s, frames, 30, tag="quirk")
    g.close()


@pytest.mark.parametrize("shape", [
    # rectangular area, the widest dendrites (256 synapses, the limit of the
    # packed 16-bit sums) beside 4-synapse ones, two and three slices
    dict(side=24, u_side=72, ssc_depth=2, pyr_depth=3, ssc_dens=2, ssc_spd=4, pyr_dens=8,
         pyr_spd=256),
    # an area of one 8 x 8 tile, three spatial slices
    dict(side=8, ssc_depth=3, pyr_depth=1, ssc_dens=3, ssc_spd=8, pyr_dens=2, pyr_spd=12),
])
def test_cycle_extreme_shapes(ctx, shape):
    """Planner edge cases against the oracle: per-job tiles where one tuft's
    item holds a fraction of a chunk per lane (its tile grows, bounded by the
    area) and another's would exceed the LDS or the chunk bound (its tile
    shrinks), on rectangular and minimal areas; 12 cycles, every buffer after
    every cycle."""
    print("shape", shape, flush=True)
    scheme = recipes.cortical_area("x", **shape)
    g, o = make_pair(ctx, scheme)
    plan = g.stage_plan()
    print(plan, flush=True)
    assert all(p["lds_bytes"] <= 64 * 1024 for p in plan)
    frames = inputs.SdrStream(scheme.input_range()[1], 0.08, 0.5, seed=3).frames(12)
    run_cycles(g, o, scheme, frames, 12, tag=str(shape))
    assert np.count_nonzero(g.read("SYN_STRENGTHS", 0)) > 0
    g.close()
    o.close()


def test_cycle_settings(ctx):
    """bypass_inhib (passthrough with rnd 0) and disable_learning."""
    for st in (L.SET_BYPASS_INHIB, L.SET_DISABLE_LEARNING, L.SET_DISABLE_PYRS):
        print("settings", st, flush=True)
        s = recipes.cortical_area("x", 32, pyr_depth=2, ssc_dens=4, ssc_spd=16, pyr_dens=4,
                                  pyr_spd=16, settings=st)
        g, o = make_pair(ctx, s)
        frames = inputs.SdrStream(32 * 32, 0.05, 0.5, seed=2).frames(25)
        run_cycles(g, o, s, frames, 25, tag=f"settings {st}")
        g.close()


def test_cycle_config2_full_size(ctx):
    """Config 2 at its full size (256x256, 142.6M synapses): 3 cycles of 2%
    SDR input compared buffer for buffer with the (multithreaded) oracle."""
    scheme = recipes.config2_area()
    g, o = make_pair(ctx, scheme)
    frames = inputs.SdrStream(256 * 256, 0.02, 0.5, seed=0).frames(3)
    run_cycles(g, o, scheme, frames, 3, tag="c2")
    # the path measured at this size: both tuft launches check-free, 32 synapses per lane
    plan = g.stage_plan()
    assert [(p["kind"], p["stream"]) for p in plan] == [(1, 0), (2, 0), (1, 1), (1, 0), (3, 0)]
    for p in plan:
        if p["kind"] == 1:
            assert p["chunk"] == 32 and p["fast"] == 1 and p["packed"] == 1, p
    # bytes of the pyramidal tuft launches, one tuft each: per synapse the packed
    # offsets (1 B) and the strength's sign bit, plus a 2-bit staged-slice code
    # for the distal tuft (two staged slices), no state written; 4 B/dendrite,
    # 12 B/cell-tuft; the snapshot of the staged slices read and written
    cells = 2 * 256 * 256
    assert [p["alg_bytes"] for p in plan if p["kind"] == 1][1:] == \
        [cells * 512 * e // 8 + cells * 8 * 4 + cells * 12 + 2 * n * 256 * 256
         for e, n in ((11, 2), (9, 1))]
    # the spiny stellate launch: packed offsets, sign bit and the stored fired
    # bit per synapse (10 eighths), 4 B/dendrite, 1 B/cell (energy), the
    # snapshot of its one staged slice read and written
    assert plan[0]["alg_bytes"] == 256 * 256 * (128 * 10 // 8 + 4 * 4 + 1) + 2 * 256 * 256
    g.close()
    o.close()


def test_packed_offsets_follow_writes(ctx):
    """The gather's packed copy of the offsets follows every write: a partial
    write across a tuft boundary, a write past |offset| 7 (that tuft falls back
    to the unpacked reads), and a fill; each followed by cycles against the
    oracle."""
    s = recipes.cortical_area("pk", 32, pyr_depth=2, ssc_dens=4, ssc_spd=16, pyr_dens=4,
                              pyr_spd=16)
    g, o = make_pair(ctx, s)
    rng = np.random.default_rng(12)
    frames = inputs.SdrStream(32 * 32, 0.05, 0.5, seed=5).frames(40)
    k = [0]

    def cycles(tag, n=8):
        run_cycles(g, o, s, frames[k[0]:k[0] + n], n, tag=tag)
        k[0] += n

    def put(name, li, arr, offset=0):
        g.write(name, li, arr, offset)
        o.write(name, li, arr, offset)

    def packed():
        return [(p["layer"], p["packed"]) for p in g.stage_plan() if p["kind"] == 1]

    assert all(pk == 1 for _, pk in packed())
    cycles("initial")
    cells = s.cells(1)
    t0 = cells * s.layers[1].tufts[0].syns_per_tft
    put("SYN_SRC_COL_U_OFFS", 1, rng.integers(-7, 8, 300).astype(np.int8), t0 - 100)
    assert all(pk == 1 for _, pk in packed())
    cycles("partial write across tufts")
    put("SYN_SRC_COL_V_OFFS", 1, rng.integers(-9, 10, 64).astype(np.int8), 200)
    assert (1, 0) in packed() and (1, 1) in packed()   # tuft 0 unpacked, tuft 1 still packed
    cycles("offsets past 7")
    for name in ("SYN_SRC_COL_V_OFFS", "SYN_SRC_COL_U_OFFS"):
        g.fill(name, 0, 3)
        o.buf(name, 0)[:] = 3
    cycles("fill")
    g.close()
    o.close()


def test_slice_codes_follow_writes(ctx):
    """The gather's 2-bit staged-slice codes follow writes of the slice ids: a
    write at an odd offset and length across a tuft boundary (partial groups of
    four synapses at both ends), then a fill; each followed by cycles against
    the oracle."""
    s = recipes.cortical_area("sc", 32, pyr_depth=2, ssc_dens=4, ssc_spd=16, pyr_dens=4,
                              pyr_spd=16)
    g, o = make_pair(ctx, s)
    rng = np.random.default_rng(21)
    frames = inputs.SdrStream(32 * 32, 0.05, 0.5, seed=6).frames(24)
    tufts = s.layers[1].tufts
    assert any(len(set(t.src_slcs)) > 1 for t in tufts)
    run_cycles(g, o, s, frames[:8], 8, tag="initial")
    cells = s.cells(1)
    t0 = cells * tufts[0].syns_per_tft
    lo, n = t0 - 101, 203
    ids = np.array([rng.choice(tufts[0 if i < t0 else 1].src_slcs) for i in range(lo, lo + n)],
                   dtype=np.uint8)
    g.write("SYN_SRC_SLC_IDS", 1, ids, lo)
    o.write("SYN_SRC_SLC_IDS", 1, ids, lo)
    assert all(p["packed"] == 1 for p in g.stage_plan() if p["kind"] == 1)
    run_cycles(g, o, s, frames[8:16], 8, tag="partial write across tufts")
    val = int(tufts[1].src_slcs[-1])
    g.fill("SYN_SRC_SLC_IDS", 1, val)
    o.buf("SYN_SRC_SLC_IDS", 1)[:] = val
    run_cycles(g, o, s, frames[16:24], 8, tag="fill")
    g.close()
    o.close()


def test_virtual_states_across_writes(ctx):
    """Check-free tufts write no states (they are recomputed from a snapshot of
    the staged slices) and read strength-sign bits instead of strengths.  Both
    follow every route that changes them outside the cycle: partial writes of
    strengths (odd offset and length), of states and of flags (the previous
    state), a strength fill, and cycles run with no buffer read in between (so
    that the learning stages take previous states from the snapshots).  Every
    buffer is compared with the oracle."""
    s = recipes.cortical_area("vs", 32, pyr_depth=2, ssc_dens=4, ssc_spd=16, pyr_dens=4,
                              pyr_spd=16)
    g, o = make_pair(ctx, s)
    rng = np.random.default_rng(33)
    frames = inputs.SdrStream(32 * 32, 0.08, 0.5, seed=9).frames(64)
    k = [0]

    def cycles(tag, n=8, every=1):
        run_cycles(g, o, s, frames[k[0]:k[0] + n], n, every=every, tag=tag)
        k[0] += n

    def put(name, li, arr, offset=0):
        g.write(name, li, arr, offset)
        o.write(name, li, arr, offset)

    assert all(p["fast"] == 1 for p in g.stage_plan() if p["kind"] == 1)
    cycles("initial", 10, every=5)
    for li in range(2):
        n = g.buf_len("SYN_STRENGTHS", li)
        put("SYN_STRENGTHS", li, rng.integers(-60, 60, 999).astype(np.int8), n // 3 + 5)
    cycles("strength write", 8, every=4)
    n = g.buf_len("SYN_STATES", 1)
    put("SYN_STATES", 1, rng.choice(np.array([0, 128, 200], dtype=np.uint8), 777), n // 2 + 3)
    put("SYN_FLAG_SETS", 1, rng.integers(0, 256, 555).astype(np.uint8), n // 4 + 1)
    cycles("state and flag writes", 8)
    g.fill("SYN_STRENGTHS", 1, 0xF0)
    o.buf("SYN_STRENGTHS", 1)[:] = -16
    cycles("strength fill", 12, every=6)
    g.close()
    o.close()


def test_buffer_io_roundtrip(ctx):
    s = recipes.cortical_area("io", 16, ssc_spd=8, ssc_dens=2, pyr_dens=2, pyr_spd=8)
    g = CorticalArea(s, ctx)
    rng = np.random.default_rng(1)
    for name in ("SYN_FLAG_SETS", "SYN_STATES", "SYN_STRENGTHS", "DEN_ACTIVITIES"):
        n = g.buf_len(name, 1)
        v = rng.integers(0, 256, n).astype(np.uint8)
        if name in L.SIGNED_BUFS:
            v = v.view(np.int8)
        g.write(name, 1, v)
        assert np.array_equal(g.read(name, 1), v)
        part = g.read(name, 1, offset=5, length=17)
        assert np.array_equal(part, v[5:22])
    g.fill("SYN_FLAG_SETS", 1, 0x13)
    assert np.all(g.read("SYN_FLAG_SETS", 1) == 0x13)
    with pytest.raises(L.BsmError):
        g.read("TFT_STATES", 0)   # SSC layers have no tufts
    with pytest.raises(L.BsmError):
        g.read("SYN_STATES", 1, offset=10**9, length=1)
    g.close()


def test_cycle_generic_gather_paths(ctx):
    """Tufts the host cannot prove covered by the staged LDS window take the
    checked gather: a synapse sourcing a slice outside its tuft's sources, and
    offsets beyond the check-free reach."""
    scheme = recipes.cortical_area("g", 48, pyr_depth=2, ssc_dens=4, ssc_spd=16, pyr_dens=4,
                                   pyr_spd=16, inhib_radius=3, thresholds=(200, 0, 300))
    g, o = make_pair(ctx, scheme, reach=6)
    # SSC synapse 5 sources a pyramidal slice; pyramidal distal offsets reach 45
    for side in (g, o):
        side.write("SYN_SRC_SLC_IDS", 0, np.array([3], dtype=np.uint8), offset=5)
    n1 = scheme.cells(1) * 64
    far = np.random.default_rng(2).integers(-45, 46, n1).astype(np.int8)
    for side in (g, o):
        side.write("SYN_SRC_COL_V_OFFS", 1, far, offset=n1)
    frames = inputs.SdrStream(48 * 48, 0.04, 0.5, seed=5).frames(20)
    run_cycles(g, o, scheme, frames, 20, tag="generic")
    g.close()


@pytest.mark.parametrize("which", ["small", "c1", "apical", "nonspatial"])
def test_regrow_matches_oracle(ctx, which):
    """bsm_area_regrow against the oracle's restatement on random state: every
    synapse with strength <= -25 gets the same new source and strength."""
    if which == "nonspatial":
        scheme = recipes.cortical_area("n", 24, pyr_depth=2, ssc_dens=3, ssc_spd=8, pyr_dens=3,
                                       pyr_spd=8)
        scheme.slices[0] = recipes.SliceScheme.nonspatial(230, 230)
    else:
        scheme = SCHEMES[which]()
    print("case", which, flush=True)
    g, o = make_pair(ctx, scheme)
    rng = np.random.default_rng(17)
    randomize(g, o, scheme, rng)
    before = g.read("SYN_SRC_COL_V_OFFS", 0).copy()
    for epoch in range(3):
        g.regrow(seed=12345, epoch=epoch, reach=6)
        o.regrow(seed=12345, epoch=epoch, reach=6)
        compare(g, o, scheme, f"{which} regrow epoch {epoch}")
    assert not np.array_equal(before, g.read("SYN_SRC_COL_V_OFFS", 0))
    # the regrown area still cycles identically (the planner re-derives the gather path)
    frames = inputs.SdrStream(scheme.input_range()[1], 0.05, 0.5, seed=3).frames(5)
    run_cycles(g, o, scheme, frames, 5, tag=f"{which} after regrow")
    g.close()


def test_cycle_config1_with_regrowth(ctx):
    """Config 1 with regrowth at the end of cycles 201, 402 and 603, as the
    reference's counter schedules it (src/cortex/cortical_area.rs:792-806);
    strengths start in [-40, 10] so that regrowth has work; every buffer
    compared after every cycle."""
    scheme = recipes.config1_area()
    g, o = make_pair(ctx, scheme, layout_seed=4)
    rng = np.random.default_rng(23)
    for li in range(len(scheme.layers)):
        st = rng.integers(-40, 11, g.buf_len("SYN_STRENGTHS", li)).astype(np.int8)
        g.write("SYN_STRENGTHS", li, st)
        o.write("SYN_STRENGTHS", li, st)
    frames = inputs.GlyphStream(side=24, p=0.15, hold=10, seed=2).frames(603)
    rnd = g.draw_rnd(603)
    changed = 0
    for c in range(603):
        g.intake(frames[c])
        o.intake(frames[c])
        r = rnd[c * g.n_slots:(c + 1) * g.n_slots]
        g.cycle(1, r)
        o.cycle(r, 1)
        if (c + 1) % 201 == 0:
            pre = g.read("SYN_SRC_COL_V_OFFS", 0)
            g.regrow(seed=77, epoch=(c + 1) // 201 - 1, reach=6)
            o.regrow(seed=77, epoch=(c + 1) // 201 - 1, reach=6)
            changed += int(np.count_nonzero(pre != g.read("SYN_SRC_COL_V_OFFS", 0)))
        compare(g, o, scheme, f"c1 regrow cycle {c}")
    assert changed > 0
    g.close()
