"""Batched and serving forms: every area stepped inside a batch (static or
dynamic/masked, with or without graph capture) must end bit-exact with the same
area stepped alone through the plain path, which test_gpu_parity pins against
the oracle."""
import numpy as np
import pytest

from bismit_amd import inputs, recipes
from bismit_amd.cortex import Batch, CorticalArea
from bismit_amd.serve import Server, SessionSpec

from helpers import layer_bufs

pytestmark = pytest.mark.gpu


def uneven_schemes():
    return [
        recipes.cortical_area("b0", 24, input_side=16, ssc_depth=1, pyr_depth=2, ssc_spd=16,
                              pyr_spd=20),
        recipes.cortical_area("b1", 40, ssc_depth=2, pyr_depth=1, ssc_spd=32, pyr_spd=64),
        recipes.cortical_area("b2", 56, ssc_depth=3, pyr_depth=3, ssc_spd=20, pyr_spd=16,
                              pyr_dens=3),
        recipes.cortical_area("b3", 32, ssc_depth=1, pyr_depth=2, ssc_spd=28, pyr_spd=36,
                              settings=8),
        recipes.cortical_area("b4", 72, ssc_depth=1, pyr_depth=1, ssc_spd=64, pyr_spd=44),
    ]


def make_areas(ctx, schemes, base_seed=100):
    return [CorticalArea(s, ctx, rnd_seed=base_seed + i, layout_seed=base_seed + 50 + i)
            for i, s in enumerate(schemes)]


def assert_same(a, b, scheme, tag):
    x, y = a.read("AXN_STATES"), b.read("AXN_STATES")
    assert np.array_equal(x, y), f"{tag} axons"
    for li in range(len(scheme.layers)):
        for name in layer_bufs(scheme, li):
            x, y = a.read(name, li), b.read(name, li)
            if not np.array_equal(x, y):
                bad = np.flatnonzero(x != y)
                raise AssertionError(f"{tag} layer {li} {name}: {bad.size} mismatches at "
                                     f"{bad[:6].tolist()}")


def frames_for(scheme, n, seed):
    return inputs.SdrStream(scheme.input_range()[1], 0.03, 0.5, seed=seed).frames(n)


@pytest.mark.parametrize("graph", [True, False])
def test_static_batch_matches_plain(ctx, graph):
    schemes = uneven_schemes()
    n = 30
    batched = make_areas(ctx, schemes)
    plain = make_areas(ctx, schemes)
    frames = [frames_for(s, n, 7 + i) for i, s in enumerate(schemes)]
    devs = []
    for f in frames:
        d = ctx.alloc(f.nbytes)
        ctx.upload(d, f)
        devs.append(d)
    b = Batch(batched, ctx)
    if not graph:
        b.set_tuning(overlap=2)   # one stream; the graphed run forks a side stream
    b.set_inputs(devs, [f.shape[1] for f in frames], [f.shape[1] for f in frames],
                 [s.input_range()[0] for s in schemes], n)
    b.cycle(n // 2, graph=graph)
    b.cycle(n - n // 2, graph=graph)
    b.sync()
    for i, (a, s) in enumerate(zip(plain, schemes)):
        for c in range(n):
            a.intake(frames[i][c])
            a.cycle(1)
    b.close()
    for i, s in enumerate(schemes):
        assert_same(batched[i], plain[i], s, f"area {i}")
    # areas continue plainly after leaving the batch
    for i, s in enumerate(schemes):
        for a in (batched[i], plain[i]):
            a.intake(frames[i][0])
            a.cycle(2)
        assert_same(batched[i], plain[i], s, f"area {i} after batch")
    for d in devs:
        ctx.free(d)


def test_batch_profile_stages_advances_like_cycles(ctx):
    """Profiled (event-bracketed, ungraphed) cycles are ordinary batch cycles."""
    schemes = uneven_schemes()[:3]
    n = 12
    batched = make_areas(ctx, schemes, 700)
    plain = make_areas(ctx, schemes, 700)
    frames = [frames_for(s, n, 70 + i) for i, s in enumerate(schemes)]
    devs = []
    for f in frames:
        d = ctx.alloc(f.nbytes)
        ctx.upload(d, f)
        devs.append(d)
    b = Batch(batched, ctx)
    b.set_inputs(devs, [f.shape[1] for f in frames], [f.shape[1] for f in frames],
                 [s.input_range()[0] for s in schemes], n)
    b.cycle(5, graph=True)
    prof = b.profile_stages(4)
    assert [k for k, _ in prof] == [1, 2, 1, 1, 3]
    assert all(t > 0 for _, t in prof)
    b.cycle(3, graph=True)
    b.sync()
    b.close()
    for i, a in enumerate(plain):
        for c in range(n):
            a.intake(frames[i][c])
            a.cycle(1)
    for i, s in enumerate(schemes):
        assert_same(batched[i], plain[i], s, f"profiled area {i}")
    for a in batched + plain:
        a.close()
    for d in devs:
        ctx.free(d)


def test_cortex_cycle_matches_per_area_calls(ctx):
    """bsm_cortex_cycle (host-driven cycle of several plain areas) equals each
    area's own intake + cycle calls."""
    from bismit_amd.cortex import cycle_areas
    schemes = uneven_schemes()[:3]
    n = 7
    looped = make_areas(ctx, schemes, 900)
    single = make_areas(ctx, schemes, 900)
    frames = [frames_for(s, 4, 90 + i) for i, s in enumerate(schemes)]
    devs = []
    for f in frames:
        d = ctx.alloc(f.nbytes)
        ctx.upload(d, f)
        devs.append(d)
    lens = [f.shape[1] for f in frames]
    cycle_areas(looped, devs, lens, 4, n)
    for i, a in enumerate(single):
        for c in range(n):
            a.intake(frames[i][c % 4])
            a.cycle(1)
    for i, s in enumerate(schemes):
        assert_same(looped[i], single[i], s, f"cortex loop area {i}")
    for a in looped + single:
        a.close()
    for d in devs:
        ctx.free(d)


def test_config3_plan_uses_fast_wide_variant(ctx):
    """Every C3 tuft has synapses per dendrite a multiple of 4 and offsets within
    6 columns: both persistent tuft launches must take the check-free gather at
    16 synapses per lane (a launch takes the narrowest chunk its tufts allow, and
    some C3 tufts' synapses per tuft are not multiples of 32)."""
    schemes = recipes.config3_areas(64, seed=0)
    areas = make_areas(ctx, schemes, 1000)
    b = Batch(areas, ctx)
    plan = b.stage_plan()
    print(plan, flush=True)
    f1 = [p for p in plan if p["kind"] == 1]
    assert [p["stream"] for p in f1] == [0, 1, 0]   # distal tufts beside the SSC stages
    for p in f1:
        assert p["chunk"] == 16 and p["fast"] == 1 and p["packed"] == 1, p

    def launch_bytes(p, pick):
        # per synapse, in eighth bytes: a 2-bit staged-slice code unless the tuft
        # stages one slice, packed offsets, the strength's sign bit; states are
        # virtual (the snapshot of the staged slices read and written instead),
        # and a spiny stellate layer's tuft 0 stores each synapse's fired bit
        def syn_e(lay, t, ti):
            return (0 if len(set(t.src_slcs)) <= 1 else 2) + 8 + 1 + (
                1 if lay.kind == 1 and ti == 0 else 0)

        def snap(s, lay, t):
            sl = s.slices[min(t.src_slcs)]
            return len(set(t.src_slcs)) * sl.v_size * sl.u_size
        return sum(s.cells(li) * t.syns_per_tft * syn_e(lay, t, ti) // 8 +
                   s.cells(li) * (t.dens_per_tft * 4 + (12 if lay.kind == 2 else 0)) +
                   (s.cells(li) if lay.kind == 1 else 0) + 2 * snap(s, lay, t)
                   for s in schemes for li, lay in enumerate(s.layers)
                   for ti, t in enumerate(lay.tufts) if pick(lay, ti))

    # SSC tufts, distal tufts (side stream), proximal tufts
    assert [p["alg_bytes"] for p in f1] == [
        launch_bytes(f1[0], lambda lay, ti: lay.kind == 1),
        launch_bytes(f1[1], lambda lay, ti: lay.kind == 2 and ti == 1),
        launch_bytes(f1[2], lambda lay, ti: lay.kind == 2 and ti == 0)]
    b.close()
    for a in areas:
        a.close()


def test_config3_batch_matches_plain(ctx):
    """The 64-area mix: 12 graph-captured batched cycles vs each area alone."""
    schemes = recipes.config3_areas(64, seed=0)
    n = 12
    batched = make_areas(ctx, schemes, 1000)
    plain = make_areas(ctx, schemes, 1000)
    frames = [frames_for(s, n, 300 + i) for i, s in enumerate(schemes)]
    devs = []
    for f in frames:
        d = ctx.alloc(f.nbytes)
        ctx.upload(d, f)
        devs.append(d)
    b = Batch(batched, ctx)
    b.set_inputs(devs, [f.shape[1] for f in frames], [f.shape[1] for f in frames],
                 [s.input_range()[0] for s in schemes], n)
    b.cycle(n, graph=True)
    b.sync()
    b.close()
    for i, a in enumerate(plain):
        for c in range(n):
            a.intake(frames[i][c])
            a.cycle(1)
    for i, s in enumerate(schemes):
        assert_same(batched[i], plain[i], s, f"c3 area {i}")
    for a in batched + plain:
        a.close()
    for d in devs:
        ctx.free(d)


@pytest.mark.parametrize("graph", [True, False])
def test_dynamic_batch_masks(ctx, graph):
    """Random per-cycle active masks: each area's result equals a plain run of the
    frames it actually received, in order."""
    schemes = uneven_schemes()
    batched = make_areas(ctx, schemes, 500)
    plain = make_areas(ctx, schemes, 500)
    b = Batch(batched, ctx, dynamic=True)
    rng = np.random.default_rng(4)
    streams = [inputs.SdrStream(s.input_range()[1], 0.03, 0.5, seed=i) for i, s in
               enumerate(schemes)]
    fed = [[] for _ in schemes]
    for c in range(25):
        mask = (rng.random(len(schemes)) < 0.6).astype(np.uint8)
        ofs = np.full(len(schemes), -1, dtype=np.int64)
        chunks, pos = [], 0
        for i in np.flatnonzero(mask):
            f = streams[i].next_frame()
            fed[i].append(f)
            ofs[i] = pos
            chunks.append(f)
            pos += f.size
        frames = np.concatenate(chunks) if chunks else np.zeros(0, np.uint8)
        b.cycle_masked(mask, frames, ofs, graph=graph)
    b.sync()
    b.close()
    for i, a in enumerate(plain):
        for f in fed[i]:
            a.intake(f)
            a.cycle(1)
        assert len(fed[i]) > 3
    for i, s in enumerate(schemes):
        assert_same(batched[i], plain[i], s, f"dyn area {i}")


def test_static_batch_regrowth_matches_plain(ctx):
    """Regrowth of batched areas (bsm_batch_regrow, a masked call then a full
    one) between graph-captured batched cycles equals each plain area's own
    regrowth at the same cycles; strengths start in [-40, 10] so regrowth has
    work, and the batch re-plans when a regrown source summary widens."""
    schemes = uneven_schemes()
    n = 50
    batched = make_areas(ctx, schemes, base_seed=300)
    plain = make_areas(ctx, schemes, base_seed=300)
    rng = np.random.default_rng(5)
    for s, x, y in zip(schemes, batched, plain):
        for li in range(len(s.layers)):
            st = rng.integers(-40, 11, x.buf_len("SYN_STRENGTHS", li)).astype(np.int8)
            x.write("SYN_STRENGTHS", li, st)
            y.write("SYN_STRENGTHS", li, st)
    frames = [frames_for(s, n, 70 + i) for i, s in enumerate(schemes)]
    devs = []
    for f in frames:
        d = ctx.alloc(f.nbytes)
        ctx.upload(d, f)
        devs.append(d)
    seeds = [900 + 13 * i for i in range(len(schemes))]
    which = np.array([1, 0, 1, 1, 0], np.uint8)
    b = Batch(batched, ctx)
    b.set_inputs(devs, [f.shape[1] for f in frames], [f.shape[1] for f in frames],
                 [s.input_range()[0] for s in schemes], n)
    b.cycle(20, graph=True)
    before = batched[0].read("SYN_SRC_COL_V_OFFS", 0).copy()
    b.regrow(seeds, [0] * len(schemes), reach=6, which=which)
    assert not np.array_equal(before, batched[0].read("SYN_SRC_COL_V_OFFS", 0))
    b.cycle(20, graph=True)
    b.regrow(seeds, [1] * len(schemes), reach=6)
    b.cycle(10, graph=True)
    b.sync()
    b.close()
    for i, a in enumerate(plain):
        for c in range(n):
            a.intake(frames[i][c])
            a.cycle(1)
            if c == 19 and which[i]:
                a.regrow(seeds[i], 0, reach=6)
            if c == 39:
                a.regrow(seeds[i], 1, reach=6)
    for i, s in enumerate(schemes):
        assert_same(batched[i], plain[i], s, f"regrow area {i}")
    for a in batched + plain:
        a.close()
    for d in devs:
        ctx.free(d)


def test_serving_sessions_match_solo_replay(ctx):
    """Continuous batching: sessions arrive while others run, are admitted as
    capacity frees up, and each session's final state equals its solo replay.
    Regrowth every 25 of a session's cycles (batched, masked to the due
    sessions) matches the solo area's own regrowth at the same counts."""
    specs = [SessionSpec(i, 0.02 * i, [16 + 8 * (i % 3)] * (1 + i % 2), 200.0 + 50 * i,
                         0.15 + 0.05 * (i % 3), 1000 + i) for i in range(8)]
    finals = {}

    def on_retire(sess):
        finals[sess.spec.sid] = (sess.spec, sess.done, [
            {(name, li): a.read(name, li) for li in range(len(sc.layers))
             for name in layer_bufs(sc, li)} | {("AXN_STATES", 0): a.read("AXN_STATES")}
            for a, sc in zip(sess.areas, sess.schemes)], sess.last_outputs)

    n_out = [0]

    def on_output(sess, frames):
        n_out[0] += 1
        # delivered once the cycle is counted: the frame is of the session's latest cycle
        assert all(f.cycle == sess.done for f in frames)

    every = 25
    srv = Server(ctx, hbm_budget_bytes=1 << 30, max_sessions=4, epoch_s=0.01, max_wait_s=None,
                 on_retire=on_retire, regrow_every=every, on_output=on_output)
    st = srv.run(specs, wall_limit_s=60.0)
    summ = st.summary()
    print(summ, flush=True)
    assert summ["sessions_completed"] == len(specs)
    assert summ["max_resident_sessions"] <= 4
    assert summ["membership_changes"] > 2
    assert n_out[0] == summ["frames"]      # one delivery per session-cycle
    from bismit_amd.serve import area_scheme, output_layer
    for sid, (spec, done, bufs, outs) in finals.items():
        assert done == spec.n_frames
        for i, side in enumerate(spec.sides):
            sc = area_scheme(side, f"r{sid}{i}")
            solo = CorticalArea(sc, ctx, rnd_seed=spec.seed + 7919 * i,
                                layout_seed=spec.seed + 104729 * i, reach=6)
            stream = inputs.SdrStream(sc.input_range()[1], 0.02, 0.5, seed=spec.seed + i)
            for c in range(1, done + 1):
                solo.intake(stream.next_frame())
                solo.cycle(1)
                if c % every == 0:
                    solo.regrow(spec.seed + 15485863 * i + 1, c // every - 1, reach=6)
            for (name, li), arr in bufs[i].items():
                assert np.array_equal(solo.read(name, li), arr), (sid, i, name, li)
            # the last output frame the session received is the final output slices
            off, n = sc.layer_axn_range(output_layer(sc))
            assert outs[i].cycle == done
            assert np.array_equal(outs[i].data, solo.read("AXN_STATES")[off:off + n]), (sid, i)
            solo.close()


def test_hierarchy_batched_matches_plain(ctx):
    """Config 4 (three 128x128 areas, feed-forward and feedback tracts, symbol
    stream): batched + graph-captured run equals the plain run, every area and
    buffer, after 40 cycles."""
    from bismit_amd.hier import Hierarchy
    hp = Hierarchy(ctx, seed=3, batched=False)
    hb = Hierarchy(ctx, seed=3, batched=True)
    frames = hp.frames(40)
    hp.run(frames)
    hb.run(frames, graph=True)
    for i, s in enumerate(hp.schemes):
        assert_same(hb.areas[i], hp.areas[i], s, f"hier area {i}")
    # the feedback and feed-forward slices actually carry activity
    b_axn = hp.areas[1].axon_states()
    idz, _ = hp.schemes[1].slice_idzs()
    assert np.count_nonzero(b_axn[idz[0]:idz[1]]) > 0
    hb.batch.close()
    hb.batch = None
    hp.close()
    hb.close()
