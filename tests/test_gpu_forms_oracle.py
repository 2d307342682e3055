"""Control and settings branches of the fused cycle, and the batched, tract and
serving forms, each compared directly with the oracle on every buffer (the
other batch tests compare the batched form with the plain one).

Branches (src/cortex/cortical_area.rs:253-338,716-789):
  pyr_output      a pyramidal layer whose column output is pyr_output
                  (cl/control.cl:298-314, src/cortex/pyr_outputter.rs:30-44)
  pyr_none        a pyramidal layer with no control layer: nothing writes its
                  axons; distal learning reads them (cl/bismit.cl:1195-1268)
  ssc_none        a spiny stellate layer with no control layer: no inhibition,
                  SSC learning reads the layer's axons as they are
  disable_sscs    CorticalAreaSettings::disable_sscs: neither the spatial cycle
                  nor its learning runs
For the layers whose axons no kernel writes, both sides get the same random
axon states before every cycle, so that their learning paths see activity.
"""
import ctypes as C
import os

import numpy as np
import pytest

from bismit_amd import _lib as L
from bismit_amd import hier, inputs, recipes
from bismit_amd.cortex import Batch, CorticalArea
from oracle import pyoracle as P

from helpers import area_rnd, compare, host_layout, load_layout, make_pair

pytestmark = pytest.mark.gpu

P.set_threads(min(16, os.cpu_count() or 1))


def control_cases():
    def base(name, **kw):
        return recipes.cortical_area(name, 24, input_side=16, pyr_depth=2, ssc_dens=4, ssc_spd=16,
                                     pyr_dens=4, pyr_spd=20, thresholds=(150, 0, 300), **kw)
    po = base("po")
    po.layers[1].control = L.CTRL_PYR_OUTPUT
    pn = base("pn")
    pn.layers[1].control = L.CTRL_NONE
    sn = base("sn")
    sn.layers[0].control = L.CTRL_NONE
    ds = base("ds", settings=L.SET_DISABLE_SSCS)
    # (scheme, layers whose axon slices are fed random states before each cycle)
    return {"pyr_output": (po, []), "pyr_none": (pn, [1]), "ssc_none": (sn, [0]),
            "disable_sscs": (ds, [0])}


def feed(sides, scheme, layers, rng):
    for li in layers:
        off, n = scheme.layer_axn_range(li)
        vals = rng.choice(np.array([0, 0, 0, 0, 0, 0, 90, 200, 255], np.uint8), n)
        for x in sides:
            x.write("AXN_STATES", 0, vals, off)


def lit(scheme, g):
    """The run fired and learned somewhere (the branch was exercised)."""
    axn = g.read("AXN_STATES")
    n0 = scheme.input_range()[1]
    return np.count_nonzero(axn[n0:]) > 0 and any(
        np.count_nonzero(g.read("SYN_STRENGTHS", li)) > 0 for li in range(len(scheme.layers)))


@pytest.mark.parametrize("case", ["pyr_output", "pyr_none", "ssc_none", "disable_sscs"])
def test_control_branch_plain_matches_oracle(ctx, case):
    """20 plain cycles of each branch, every buffer compared after every cycle."""
    scheme, fed = control_cases()[case]
    print("case", case, flush=True)
    g, o = make_pair(ctx, scheme, layout_seed=3)
    rng = np.random.default_rng(50)
    frames = inputs.SdrStream(scheme.input_range()[1], 0.08, 0.5, seed=4).frames(20)
    rnd = g.draw_rnd(20)
    for c in range(20):
        feed((g, o), scheme, fed, rng)
        g.intake(frames[c])
        o.intake(frames[c])
        r = rnd[c * g.n_slots:(c + 1) * g.n_slots]
        g.cycle(1, r)
        o.cycle(r, 1)
        compare(g, o, scheme, f"{case} cycle {c}")
    assert lit(scheme, g), case
    g.close()
    o.close()


@pytest.mark.parametrize("dynamic", [False, True])
def test_control_branches_batched_match_oracle(ctx, dynamic):
    """The four branches as areas of one batch (static and graph-captured, or
    dynamic under random masks); each area against its own oracle area after
    every cycle.  The dynamic run reads every buffer and rewrites SYN_FLAG_SETS
    of a layer right after masked cycles with no explicit batch sync (buffer
    I/O settles the batch and its device-side state parities first)."""
    cases = control_cases()
    names = list(cases)
    schemes = [cases[k][0] for k in names]
    seeds = [200 + i for i in range(len(schemes))]
    gs, os_ = [], []
    for i, s in enumerate(schemes):
        g = CorticalArea(s, ctx, rnd_seed=seeds[i], init_layout=False)
        o = P.OracleArea(s)
        lay = host_layout(s, 10 + i)
        load_layout(g, s, lay)
        load_layout(o, s, lay)
        gs.append(g)
        os_.append(o)
    n = 16
    rnds = [area_rnd(seeds[i], gs[i].n_slots, n) for i in range(len(schemes))]
    frames = [inputs.SdrStream(s.input_range()[1], 0.08, 0.5, seed=60 + i).frames(n)
              for i, s in enumerate(schemes)]
    rng = np.random.default_rng(51)
    b = Batch(gs, ctx, dynamic=dynamic)
    devs = []
    if not dynamic:
        for f in frames:
            d = ctx.alloc(f.nbytes)
            ctx.upload(d, f)
            devs.append(d)
        b.set_inputs(devs, [f.shape[1] for f in frames], [f.shape[1] for f in frames],
                     [s.input_range()[0] for s in schemes], n)
    done = [0] * len(schemes)
    for c in range(n):
        for i, s in enumerate(schemes):
            feed((gs[i], os_[i]), s, cases[names[i]][1], rng)
        if dynamic:
            mask = (rng.random(len(schemes)) < 0.7).astype(np.uint8)
            ofs = np.full(len(schemes), -1, np.int64)
            parts, pos = [], 0
            for i in np.flatnonzero(mask):
                ofs[i] = pos
                parts.append(frames[i][done[i]])
                pos += frames[i].shape[1]
            b.cycle_masked(mask, np.concatenate(parts) if parts else np.zeros(0, np.uint8), ofs,
                           graph=bool(c % 2))
        else:
            mask = np.ones(len(schemes), np.uint8)
            b.cycle(1, graph=True)
        for i, s in enumerate(schemes):
            if mask[i]:
                os_[i].intake(frames[i][done[i]] if dynamic else frames[i][c])
                os_[i].cycle(rnds[i][done[i]], 1)
                done[i] += 1
            compare(gs[i], os_[i], s, f"{'dyn' if dynamic else 'static'} {names[i]} cycle {c}")
        if dynamic and c == n // 2:
            # flag writes between masked cycles land in the current buffers
            for i, s in enumerate(schemes):
                fl = rng.integers(0, 256, gs[i].buf_len("SYN_FLAG_SETS", 1)).astype(np.uint8)
                gs[i].write("SYN_FLAG_SETS", 1, fl)
                os_[i].write("SYN_FLAG_SETS", 1, fl)
                compare(gs[i], os_[i], s, f"dyn {names[i]} after flag write")
    assert all(d >= 6 for d in done), done
    b.close()
    for g in gs:
        g.close()
    for d in devs:
        ctx.free(d)


def test_config3_recipe_batched_matches_oracle(ctx):
    """Six areas of the Config 3 recipe (uneven sides 32-64, 1-3 slices per
    layer, 16-64 synapses per dendrite, own 2% SDR input each) in one static,
    graph-captured batch on the two-stream schedule; 20 cycles, each area
    against an oracle area after every cycle."""
    schemes = recipes.config3_areas(6, seed=5, sides=(32, 64))
    n = 20
    seeds = [300 + i for i in range(len(schemes))]
    gs, os_ = [], []
    for i, s in enumerate(schemes):
        g = CorticalArea(s, ctx, rnd_seed=seeds[i], init_layout=False)
        o = P.OracleArea(s)
        lay = host_layout(s, 40 + i)
        load_layout(g, s, lay)
        load_layout(o, s, lay)
        gs.append(g)
        os_.append(o)
    frames = [inputs.SdrStream(s.input_range()[1], 0.02, 0.5, seed=70 + i).frames(n)
              for i, s in enumerate(schemes)]
    rnds = [area_rnd(seeds[i], gs[i].n_slots, n) for i in range(len(schemes))]
    devs = []
    for f in frames:
        d = ctx.alloc(f.nbytes)
        ctx.upload(d, f)
        devs.append(d)
    b = Batch(gs, ctx)
    plan = b.stage_plan()
    assert any(p["stream"] == 1 for p in plan), plan
    assert all(p["fast"] == 1 for p in plan if p["kind"] == 1), plan
    b.set_inputs(devs, [f.shape[1] for f in frames], [f.shape[1] for f in frames],
                 [s.input_range()[0] for s in schemes], n)
    for c in range(n):
        b.cycle(1, graph=True)
        for i, s in enumerate(schemes):
            os_[i].intake(frames[i][c])
            os_[i].cycle(rnds[i][c], 1)
            compare(gs[i], os_[i], s, f"c3 area {i} cycle {c}")
    assert all(np.count_nonzero(g.read("SYN_STRENGTHS", 0)) > 0 for g in gs)
    b.close()
    for g in gs:
        g.close()
    for d in devs:
        ctx.free(d)


def hier_oracles(ctx, hs):
    """Oracle twins of a Hierarchy's areas, with the device-grown layouts."""
    out = []
    for a, s in zip(hs.areas, hs.schemes):
        o = P.OracleArea(s)
        for li in range(len(s.layers)):
            for name in ("SYN_SRC_SLC_IDS", "SYN_SRC_COL_V_OFFS", "SYN_SRC_COL_U_OFFS"):
                o.write(name, li, a.read(name, li))
        out.append(o)
    return out


@pytest.mark.parametrize("batched", [False, True])
def test_config4_hierarchy_tracts_match_oracle(ctx, batched):
    """The Config 4 hierarchy at side 32 (three areas, feed-forward and
    feedback tracts, Zipf/Markov symbol input): the oracle replays the same
    tract copies -- every tract reads its source as the previous cycle left it,
    before any area cycles (src/subcortex/thalamus/thalamus.rs:71-285,
    src/cortex/axon_space.rs:509-629) -- then area A's intake and each area's
    cycle; 30 cycles, every buffer of every area after every cycle, through
    bsm_area_copy_axons (plain) or the batch's tract launch (graph-captured)."""
    seed = 9
    hs = hier.Hierarchy(ctx, seed=seed, batched=False, side=32)
    orcs = hier_oracles(ctx, hs)
    n = 30
    frames = hs.frames(n)
    rnds = [area_rnd(seed + 11 * i, a.n_slots, n) for i, a in enumerate(hs.areas)]
    if batched:
        b = Batch(hs.areas, ctx)
        arr = (L.Tract * len(hs.ranges))(*[L.Tract(*r) for r in hs.ranges])
        L.check(ctx.lib.bsm_batch_set_tracts(b.h, arr, len(hs.ranges)))
        dev = ctx.alloc(frames.nbytes)
        ctx.upload(dev, frames)
        k = frames.shape[1]
        b.set_inputs([dev, None, None], [k, 0, 0], [k, 0, 0],
                     [hs.schemes[0].input_range()[0], 0, 0], n)
    for c in range(n):
        if batched:
            b.cycle(1, graph=True)
        else:
            for s_i, s_ofs, d_i, d_ofs, ln in hs.ranges:
                L.check(ctx.lib.bsm_area_copy_axons(hs.areas[d_i].h, d_ofs, hs.areas[s_i].h,
                                                    s_ofs, ln))
            hs.areas[0].intake(frames[c])
            for i, a in enumerate(hs.areas):
                a.cycle(1, rnds[i][c])
        prev = [o.read("AXN_STATES") for o in orcs]
        for s_i, s_ofs, d_i, d_ofs, ln in hs.ranges:
            orcs[d_i].buf("AXN_STATES")[d_ofs:d_ofs + ln] = prev[s_i][s_ofs:s_ofs + ln]
        orcs[0].intake(frames[c])
        for i, o in enumerate(orcs):
            o.cycle(rnds[i][c], 1)
        for i, s in enumerate(hs.schemes):
            compare(hs.areas[i], orcs[i], s, f"hier {'batched' if batched else 'plain'} "
                                              f"area {i} cycle {c}")
    # the tracts carried activity both ways
    idz, _ = hs.schemes[1].slice_idzs()
    axn_b = orcs[1].read("AXN_STATES")
    assert np.count_nonzero(axn_b[idz[0]:idz[1]]) > 0       # feed-forward into B
    assert np.count_nonzero(axn_b[idz[1]:idz[2]]) > 0       # feedback into B
    if batched:
        b.close()
        ctx.free(dev)
    hs.close()


def test_device_layout_equals_host_generator(ctx):
    """bsm_area_init_layout (on the device) lays out every tuft exactly as the
    host generator bsm_gen_tuft_layout does, with zero strengths."""
    sc = recipes.cortical_area("lay", 64, ssc_depth=2, pyr_depth=3, ssc_spd=24, pyr_spd=36)
    g = CorticalArea(sc, ctx, layout_seed=77, reach=6)
    for li, (slc, v, u) in enumerate(host_layout(sc, 77, 6)):
        assert np.array_equal(g.read("SYN_SRC_SLC_IDS", li), slc)
        assert np.array_equal(g.read("SYN_SRC_COL_V_OFFS", li), v)
        assert np.array_equal(g.read("SYN_SRC_COL_U_OFFS", li), u)
        assert not np.any(g.read("SYN_STRENGTHS", li))
    g.close()


def test_serving_sessions_at_config5_sizes_match_oracle(ctx):
    """Config 5 sizes: sessions of one to three areas with sides 64-256
    (one 256 x 256 area of 42M synapses), arriving while others run, admitted
    as slots free up (at most three resident), regrowing every 20 of their own
    cycles.  Each session's final state equals the oracle's replay of that
    session alone: its frames, its areas' rnd streams, layouts and regrowth
    epochs."""
    from bismit_amd.serve import Server, SessionSpec, area_scheme
    specs = [SessionSpec(0, 0.00, [256], 50.0, 0.9, 11),
             SessionSpec(1, 0.05, [64, 136], 60.0, 0.7, 12),
             SessionSpec(2, 0.10, [200], 40.0, 0.8, 13),
             SessionSpec(3, 0.15, [64, 96, 128], 60.0, 0.6, 14)]
    finals = {}

    def on_retire(sess):
        finals[sess.spec.sid] = (sess.spec, sess.done, [
            {(name, li): a.read(name, li) for li in range(len(sc.layers))
             for name in helpers_bufs(sc, li)} | {("AXN_STATES", 0): a.read("AXN_STATES")}
            for a, sc in zip(sess.areas, sess.schemes)])

    every = 20
    srv = Server(ctx, hbm_budget_bytes=8 << 30, max_sessions=3, epoch_s=0.01, max_wait_s=None,
                 on_retire=on_retire, regrow_every=every)
    st = srv.run(specs, wall_limit_s=60.0)
    summ = st.summary()
    print(summ, flush=True)
    assert summ["sessions_completed"] == len(specs)
    assert summ["max_resident_sessions"] <= 3
    for sid, (spec, done, bufs) in sorted(finals.items()):
        assert done == spec.n_frames
        for i, side in enumerate(spec.sides):
            sc = area_scheme(side, f"o{sid}{i}")
            print("oracle replay: session", sid, "area", i, "side", side, "cycles", done,
                  flush=True)
            o = P.OracleArea(sc)
            load_layout(o, sc, host_layout(sc, spec.seed + 104729 * i, 6))
            stream = inputs.SdrStream(sc.input_range()[1], 0.02, 0.5, seed=spec.seed + i)
            rnd = area_rnd(spec.seed + 7919 * i, sc.rnd_slots(), done)
            for c in range(1, done + 1):
                o.intake(stream.next_frame())
                o.cycle(rnd[c - 1], 1)
                if c % every == 0:
                    o.regrow(spec.seed + 15485863 * i + 1, c // every - 1, reach=6)
            for (name, li), arr in bufs[i].items():
                want = o.read(name, li)
                if not np.array_equal(want, arr):
                    bad = np.flatnonzero(want != arr)
                    raise AssertionError(f"session {sid} area {i} layer {li} {name}: "
                                         f"{bad.size} mismatches at {bad[:6].tolist()}")
            o.close()


def helpers_bufs(scheme, li):
    from helpers import layer_bufs
    return layer_bufs(scheme, li)


def test_dynamic_batch_membership_and_sdr_intake_match_oracle(ctx):
    """Continuous batching through bsm_batch_add_areas / bsm_batch_remove_areas:
    a dynamic batch created empty; areas join at cycles 0 and 6, one leaves at
    cycle 12, runs two plain cycles on its own (its rnd stream continues where
    the batch left it) and joins again at cycle 18.  Frames alternate between
    dense entries and SDR entries (count + active indices, scattered on the
    device).  Each area against its own oracle area after every cycle."""
    schemes = [recipes.cortical_area(f"m{i}", side, ssc_depth=sd, pyr_depth=pd, ssc_spd=spd,
                                     pyr_spd=pspd)
               for i, (side, sd, pd, spd, pspd) in enumerate(
                   [(24, 1, 2, 16, 20), (40, 2, 1, 32, 24), (32, 1, 3, 20, 16),
                    (48, 3, 2, 24, 36), (16, 1, 1, 8, 12)])]
    seeds = [400 + i for i in range(len(schemes))]
    gs, os_, streams = [], [], []
    for i, s in enumerate(schemes):
        g = CorticalArea(s, ctx, rnd_seed=seeds[i], init_layout=False)
        o = P.OracleArea(s)
        lay = host_layout(s, 80 + i)
        load_layout(g, s, lay)
        load_layout(o, s, lay)
        gs.append(g)
        os_.append(o)
        streams.append(inputs.SdrStream(s.input_range()[1], 0.03, 0.5, seed=90 + i))
    n = 24
    rnds = [area_rnd(seeds[i], gs[i].n_slots, 64) for i in range(len(schemes))]
    done = [0] * len(schemes)
    rng = np.random.default_rng(52)
    b = Batch([], ctx, dynamic=True)
    members = []

    def step_oracle(i, dense):
        os_[i].intake(dense)
        os_[i].cycle(rnds[i][done[i]], 1)
        done[i] += 1

    for c in range(n):
        if c == 0:
            b.add([gs[0], gs[1], gs[2]])
            members = [0, 1, 2]
        elif c == 6:
            b.add([gs[3], gs[4]])
            members = [0, 1, 2, 3, 4]
        elif c == 12:
            b.remove([gs[1]])
            members = [0, 2, 3, 4]
            for _ in range(2):
                idx = streams[1].next_indices()
                dense = np.zeros(streams[1].n, np.uint8)
                dense[idx] = 255
                gs[1].intake(dense)
                gs[1].cycle(1)
                step_oracle(1, dense)
            compare(gs[1], os_[1], schemes[1], "area 1 plain while out of the batch")
        elif c == 18:
            b.add([gs[1]])
            members = [0, 2, 3, 4, 1]
        assert [id(a) for a in b.areas] == [id(gs[i]) for i in members]
        mask = (rng.random(len(members)) < 0.75).astype(np.uint8)
        ofs = np.full(len(members), np.uint64(2**64 - 1), np.uint64)
        buf = bytearray()
        dense_of = {}
        for j in np.flatnonzero(mask):
            i = members[j]
            idx = streams[i].next_indices()
            dense = np.zeros(streams[i].n, np.uint8)
            dense[idx] = 255
            dense_of[i] = dense
            if (c + j) % 2:
                pad = (-len(buf)) % 4
                buf += bytes(pad)
                ofs[j] = np.uint64(len(buf)) | np.uint64(L.FRAME_SDR)
                buf += np.concatenate([[idx.size], idx]).astype(np.uint32).tobytes()
            else:
                ofs[j] = len(buf)
                buf += dense.tobytes()
        b.cycle_masked(mask, np.frombuffer(bytes(buf), np.uint8), ofs, graph=bool(c % 3 == 0))
        for j, i in enumerate(members):
            if mask[j]:
                step_oracle(i, dense_of[i])
        for i in range(len(schemes)):
            compare(gs[i], os_[i], schemes[i], f"membership area {i} cycle {c}")
    assert min(done) >= 8, done
    # an SDR index beyond the input range is refused on the host, before any launch
    bad = np.array([1, schemes[0].input_range()[1]], np.uint32).tobytes()
    ofs = np.full(len(members), np.uint64(2**64 - 1), np.uint64)
    ofs[0] = np.uint64(L.FRAME_SDR)
    mask = np.zeros(len(members), np.uint8)
    mask[0] = 1
    with pytest.raises(L.BsmError, match="beyond its input range"):
        b.cycle_masked(mask, np.frombuffer(bad, np.uint8), ofs, graph=False)
    compare(gs[0], os_[0], schemes[0], "after the refused cycle")
    b.close()
    for g in gs:
        g.close()
