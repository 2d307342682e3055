"""Shared test helpers: build matching oracle/GPU areas and compare buffers."""
import numpy as np

from bismit_amd import _lib as L

# buffers compared per layer kind
SYN_BUFS = ["SYN_STATES", "SYN_STRENGTHS", "SYN_FLAG_SETS", "SYN_SRC_SLC_IDS",
            "SYN_SRC_COL_V_OFFS", "SYN_SRC_COL_U_OFFS"]
DEN_BUFS = ["DEN_STATES", "DEN_STATES_RAW", "DEN_ACTIVITIES", "DEN_ENERGIES"]
CEL_BUFS = ["SOMA_STATES", "SOMA_ENERGIES", "SOMA_ACTIVITIES"]
PYR_BUFS = ["SOMA_FLAG_SETS", "TFT_STATES", "TFT_BEST_DEN_IDS", "TFT_BEST_DEN_STATES_RAW",
            "TFT_BEST_DEN_STATES", "TFT_PREV_STATES", "TFT_PREV_BEST_DEN_IDS",
            "TFT_PREV_BEST_DEN_STATES_RAW", "TFT_PREV_BEST_DEN_STATES"]


def layer_bufs(scheme, li, synapses=True):
    out = list(DEN_BUFS) + list(CEL_BUFS)
    if scheme.layers[li].kind == L.LAYER_PYR:
        out += PYR_BUFS
    if synapses:
        out += SYN_BUFS
    return out


def host_layout(scheme, seed=0, reach=6):
    """Per layer: (slc, v, u) arrays from the engine's host layout generator."""
    out = []
    for li, lay in enumerate(scheme.layers):
        cells = scheme.cells(li)
        parts = [L.gen_tuft_layout(seed, li, t, cells * ts.syns_per_tft, ts.src_slcs, reach)
                 for t, ts in enumerate(lay.tufts)]
        out.append(tuple(np.concatenate([p[k] for p in parts]) for k in range(3)))
    return out


def load_layout(area_like, scheme, layout):
    for li, (s, v, u) in enumerate(layout):
        area_like.write("SYN_SRC_SLC_IDS", li, s)
        area_like.write("SYN_SRC_COL_V_OFFS", li, v)
        area_like.write("SYN_SRC_COL_U_OFFS", li, u)


def make_pair(ctx, scheme, layout_seed=0, reach=6, rnd_seed=0):
    """A device area and an oracle area of the same scheme and synapse layout."""
    from bismit_amd.cortex import CorticalArea
    from oracle import pyoracle as P
    g = CorticalArea(scheme, ctx, rnd_seed=rnd_seed, init_layout=False)
    o = P.OracleArea(scheme)
    lay = host_layout(scheme, layout_seed, reach)
    load_layout(g, scheme, lay)
    load_layout(o, scheme, lay)
    return g, o


def run_cycles(g, o, scheme, frames, n, rng_seed=0, every=1, tag=""):
    """n plain cycles on both sides with the device area's own rnd draws;
    every buffer compared after every `every` cycles and after the last."""
    rnd = g.draw_rnd(n)
    for c in range(n):
        if frames is not None:
            g.intake(frames[c])
            o.intake(frames[c])
        r = rnd[c * g.n_slots:(c + 1) * g.n_slots]
        g.cycle(1, r)
        o.cycle(r, 1)
        if (c + 1) % every == 0 or c == n - 1:
            compare(g, o, scheme, f"{tag} cycle {c}")


def area_rnd(seed, n_slots, n_cycles):
    """The draws an area created with rnd_seed=seed consumes over n_cycles
    (its xorshift128 stream), cycle-major, as int32."""
    return L.xorshift128(seed, n_slots * n_cycles).view(np.int32).reshape(n_cycles, n_slots)


def compare(gpu, orc, scheme, tag="", synapses=True):
    """Assert every buffer of the area is identical; return the list compared."""
    names = []
    a = gpu.read("AXN_STATES")
    b = orc.read("AXN_STATES")
    _eq(a, b, f"{tag} AXN_STATES")
    for li in range(len(scheme.layers)):
        for name in layer_bufs(scheme, li, synapses):
            _eq(gpu.read(name, li), orc.read(name, li), f"{tag} layer {li} {name}")
            names.append(name)
    return names


def _eq(a, b, what):
    if a.shape != b.shape:
        raise AssertionError(f"{what}: shape {a.shape} != {b.shape}")
    if not np.array_equal(a, b):
        idx = np.flatnonzero(a != b)
        raise AssertionError(f"{what}: {idx.size} mismatches, first at {idx[:8].tolist()}: "
                             f"gpu {a[idx[:8]].tolist()} vs oracle {b[idx[:8]].tolist()}")
