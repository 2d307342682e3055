"""Pins of the CPU oracle (oracle/bsm_oracle.c) on the CPU.

The reference ships no golden vectors or known-answer tests for this path and
its OpenCL/Rust build cannot run here, so each oracle function is pinned by
what the mathematics fixes: published known answers of the textbook generators
it builds on, exact counting identities, closed forms of hand-built inputs,
brute force against independent formulations on tiny grids, and the invariants
the reference's own tests assert (src/tests/dens_tfts.rs:196-363,
src/tests/kernels.rs:21-91).
"""
import numpy as np
import pytest

from bismit_amd import _lib as L
from bismit_amd import recipes
from oracle import pyoracle as P

M32 = 0xFFFFFFFF


def s32(x):
    x &= M32
    return x - (1 << 32) if x & 0x80000000 else x


def xorshift32(y):
    """Marsaglia (2003) xorshift32, triple (13, 17, 5), logical shifts."""
    y ^= (y << 13) & M32
    y ^= y >> 17
    y ^= (y << 5) & M32
    return y


def inv_lshift13(x):
    """Inverse of x -> x ^ (x << 13) on 32 bits."""
    return (x ^ (x << 13) ^ (x << 26)) & M32


@pytest.fixture(scope="module")
def lib():
    return P.load()


# ---------------------------------------------------------------- rnd_mix


def test_xorshift32_known_answer():
    # Marsaglia, "Xorshift RNGs", J. Stat. Softw. 8(14), 2003: y=2463534242 -> 723471715
    assert xorshift32(2463534242) == 723471715


def test_rnd_mix_is_xorshift32_with_arithmetic_shift(lib):
    """rnd_mix(0, s) equals the textbook xorshift32 step except that the >> 17 is
    arithmetic: when s ^ (s << 13) is negative the result differs by exactly
    0x000F8000 (sign bits 15..31 from the shift, folded by the << 5)."""
    rng = np.random.default_rng(1)
    seeds = list(rng.integers(0, 1 << 32, 4000, dtype=np.uint64)) + [0, 1, 0x7FFFFFFF, 0x80000000,
                                                                       0xFFFFFFFF, 2463534242]
    n_neg = 0
    for s in seeds:
        s = int(s)
        x1 = (s ^ (s << 13)) & M32
        want = xorshift32(s)
        if x1 & 0x80000000:
            want ^= 0x000F8000
            n_neg += 1
        assert lib.orc_rnd_mix(0, s32(s)) & M32 == want, hex(s)
    assert 100 < n_neg < len(seeds) - 100


def test_rnd_mix_rnd_enters_before_the_first_shift(lib):
    """(seed ^ rnd) << 13: rnd_mix(rnd, s) == rnd_mix(0, s') where s' solves
    s' ^ (s' << 13) == s ^ (s << 13) ^ (rnd << 13)."""
    rng = np.random.default_rng(2)
    for _ in range(3000):
        s, r = (int(v) for v in rng.integers(0, 1 << 32, 2, dtype=np.uint64))
        x1 = (s ^ (s << 13) ^ (r << 13)) & M32
        s2 = inv_lshift13(x1)
        assert ((s2 ^ (s2 << 13)) & M32) == x1
        assert lib.orc_rnd_mix(s32(r), s32(s)) == lib.orc_rnd_mix(0, s32(s2))


def test_lshft_masks(lib):
    # learning-rate masks of the three LTP paths (cl/bismit.cl:572-577, 616-623)
    assert lib.orc_lshft_mask(0) == 0x7F
    assert lib.orc_lshft_mask(2) == 0x1FF
    assert lib.orc_lshft_mask(5) == 0xFFF


@pytest.mark.parametrize("mask_l2i", [0, 2, 5])
def test_inc_dec_exact_counts(lib, mask_l2i):
    """Over an aligned block of 2^bits seeds the low `bits` of rnd_mix take every
    value once (seed bits 0..12 enter the low bits bijectively), so the number
    of increments is exactly 127-|v| and of decrements 127-|v|+(v==127)."""
    mask = lib.orc_lshft_mask(mask_l2i)
    bits = 7 + mask_l2i
    rng = np.random.default_rng(3 + mask_l2i)
    for v in [-128, -127, -100, -1, 0, 1, 5, 64, 126, 127]:
        for _ in range(2):
            r = s32(int(rng.integers(0, 1 << 32)))
            base = int(rng.integers(0, 1 << (32 - bits))) << bits
            inc = sum(lib.orc_rnd_inc(r, s32(base + k), v, mask) for k in range(1 << bits))
            dec = sum(lib.orc_rnd_dec(r, s32(base + k), v, mask) for k in range(1 << bits))
            av = abs(v)
            assert inc == max(0, 127 - av), (v, mask)
            assert dec == max(0, 127 - av + (v == 127)), (v, mask)


def test_activity_rating_composition(lib):
    """update_activity_rating: +1 when (mix(rnd, seed) & 0xFF) > a and active,
    then -1 when (mix(rnd, seed << 1) & 0xFFFF) < decay and the *updated* a > 0."""
    rng = np.random.default_rng(4)
    for _ in range(4000):
        a = int(rng.integers(0, 256))
        act = int(rng.integers(0, 2))
        r, s = (s32(int(v)) for v in rng.integers(0, 1 << 32, 2))
        decay = int(rng.choice([768, 1536, 40000]))
        a1 = a + (((lib.orc_rnd_mix(r, s) & 0xFF) > a) and act)
        s2 = s32((s << 1) & M32)
        a2 = a1 - (((lib.orc_rnd_mix(r, s2) & 0xFFFF) < decay) and a1 > 0)
        assert lib.orc_update_activity_rating(a, act, r, s, decay) == a2


def test_activity_rating_bounds_and_rates(lib):
    # rating never wraps: 0 stays >= 0, 255 cannot increment
    for s in range(2000):
        assert lib.orc_update_activity_rating(0, 0, 12345, s, 65535) == 0
        assert lib.orc_update_activity_rating(255, 1, 777, s, 0) == 255
    # the decay draw fires at rate decay/65536 (16-bit mask): 2^16 trials at
    # 768/65536 expect 768 decrements, sigma ~28
    dec = sum(1 - lib.orc_update_activity_rating(1, 0, 99, s, 768) for s in range(1 << 16))
    assert 768 - 150 < dec < 768 + 150


def test_syn_fire_closed_form(lib):
    vals = [lib.orc_syn_fire(a) for a in range(256)]
    assert vals[0] == 0 and vals[1] == 128 and vals[2] == 129 and vals[255] == 255
    assert all(128 <= v <= 255 for v in vals[1:])
    assert all(vals[i] <= vals[i + 1] for i in range(255))


# ---------------------------------------------------------------- kernels


def tiny_area(side=16, input_side=None, pyr_depth=2, ssc_spd=8, pyr_spd=8, dens=2, radius=2,
              thresholds=(0, 0, 0), **kw):
    return recipes.cortical_area("t", side, input_side=input_side, pyr_depth=pyr_depth,
                                 ssc_dens=dens, ssc_spd=ssc_spd, pyr_dens=dens, pyr_spd=pyr_spd,
                                 inhib_radius=radius, thresholds=thresholds, **kw)


def test_gather_all_synapses_on_one_axon():
    """Reference invariant (src/tests/dens_tfts.rs:196-363): all synapses of a
    cell-tuft sourced from one active axon -> every dendrite of it is active."""
    sch = tiny_area(thresholds=(0, 0, 0))
    o = P.OracleArea(sch)
    rng = np.random.default_rng(5)
    idz, _ = sch.slice_idzs()
    for trial in range(40):
        o.buf("AXN_STATES")[:] = 0
        li = int(rng.integers(0, 2))
        lay = sch.layers[li]
        t = int(rng.integers(0, len(lay.tufts)))
        ts = lay.tufts[t]
        src = ts.src_slcs[0]
        v, u = (int(x) for x in rng.integers(0, 16, 2))
        slc_l = int(rng.integers(0, lay.depth))
        vo, uo = (int(x) for x in rng.integers(-3, 4, 2))
        tv, tu = v + vo, u + uo
        if not (0 <= tv < 16 and 0 <= tu < 16):
            continue
        o.buf("AXN_STATES")[idz[src] + tv * 16 + tu] = 200
        cells = sch.cells(li)
        tz = sum(cells * x.syns_per_tft for x in lay.tufts[:t])
        cel = slc_l * 256 + v * 16 + u
        z = tz + cel * ts.syns_per_tft
        for name, val in (("SYN_SRC_SLC_IDS", src), ("SYN_SRC_COL_V_OFFS", vo),
                          ("SYN_SRC_COL_U_OFFS", uo)):
            o.buf(name, li)[z:z + ts.syns_per_tft] = np.array(val).astype(o.buf(name, li).dtype)
        o.k("syn_cycle", li, t)
        o.k("den_cycle", li, t, 0)
        st = o.buf("SYN_STATES", li)[z:z + ts.syns_per_tft]
        assert np.all(st == 128 + 100)
        dz = sum(cells * x.dens_per_tft for x in lay.tufts[:t]) + cel * ts.dens_per_tft
        dens = o.buf("DEN_STATES", li)[dz:dz + ts.dens_per_tft]
        # 8 synapses x 228 = 1824 >> 4 = 114 per dendrite
        assert np.all(dens == min(255, (ts.syns_per_den * 228) >> 4)), dens


def test_gather_brute_force_geometry():
    """Every synapse state equals the textbook hex-grid lookup: scaled cell
    position plus offset, 0 off the slice, on a slice mapped from a smaller
    source area (scale calc_scale(24, 64) = 6) and on local slices."""
    sch = recipes.cortical_area("g", 64, input_side=24, ssc_dens=1, ssc_spd=4, pyr_dens=1,
                                pyr_spd=4, pyr_depth=1, distal=False)
    assert sch.slices[0].v_scale == 6 and sch.slices[1].v_scale == 16
    o = P.OracleArea(sch)
    rng = np.random.default_rng(6)
    axn = o.buf("AXN_STATES")
    axn[:] = rng.integers(0, 256, axn.size).astype(np.uint8)
    idz, _ = sch.slice_idzs()
    for li in range(2):
        n = o.buf("SYN_STATES", li).size
        srcs = sch.layers[li].tufts[0].src_slcs + [2]
        o.buf("SYN_SRC_SLC_IDS", li)[:] = rng.choice(srcs, n)
        o.buf("SYN_SRC_COL_V_OFFS", li)[:] = rng.integers(-30, 31, n)
        o.buf("SYN_SRC_COL_U_OFFS", li)[:] = rng.integers(-30, 31, n)
        o.k("syn_cycle", li, 0)
        sl = o.buf("SYN_SRC_SLC_IDS", li)
        vo = o.buf("SYN_SRC_COL_V_OFFS", li)
        uo = o.buf("SYN_SRC_COL_U_OFFS", li)
        st = o.buf("SYN_STATES", li)
        spt = 4
        for i in rng.choice(n, 3000, replace=False):
            cel = i // spt
            v, u = (cel % 4096) // 64, (cel % 4096) % 64
            s = sch.slices[int(sl[i])]
            # independent formulation: nearest source row/col by integer scaling
            sv = v * s.v_scale // 16 + int(vo[i])
            su = u * s.u_scale // 16 + int(uo[i])
            a = 0
            if 0 <= sv < s.v_size and 0 <= su < s.u_size:
                a = int(axn[idz[int(sl[i])] + sv * s.u_size + su])
            want = 0 if a == 0 else 128 + a // 2
            assert st[i] == want, (li, i)


def test_den_cycle_closed_forms():
    sch = tiny_area(ssc_spd=32, dens=4, thresholds=(5000, 0, 0))
    o = P.OracleArea(sch)
    st = o.buf("SYN_STATES", 0)
    sw = o.buf("SYN_STRENGTHS", 0)
    st[:] = 128
    sw[:] = 0
    sw[32:64:2] = -1  # dendrite 1: half its synapses have negative strength
    st[64:96] = 0     # dendrite 2: silent
    o.k("den_cycle", 0, 0, 0)
    raw = o.buf("DEN_STATES_RAW", 0)
    ds = o.buf("DEN_STATES", 0)
    # 32 x 128 = 4096 -> raw 256 clamps to 255; the sum never exceeds 5000 -> state 0
    assert raw[0] == 255 and ds[0] == 0
    assert raw[1] == 255 and ds[1] == 0
    assert raw[2] == 0 and ds[2] == 0
    sch2 = tiny_area(ssc_spd=32, dens=4, thresholds=(1000, 0, 0))
    o2 = P.OracleArea(sch2)
    o2.buf("SYN_STATES", 0)[:] = 128
    o2.buf("SYN_STRENGTHS", 0)[32:64:2] = -1
    o2.k("den_cycle", 0, 0, 0)
    ds2 = o2.buf("DEN_STATES", 0)
    assert ds2[0] == 255           # 4096 > 1000 -> min(256, 255)
    assert ds2[1] == 2048 >> 4     # only the 16 non-negative synapses count


def test_tft_cycle_first_max_and_cap():
    sch = tiny_area(dens=4, pyr_spd=4)
    sch.layers[1].tufts[0].max_active_dens_l2 = 0
    o = P.OracleArea(sch)
    raw = o.buf("DEN_STATES_RAW", 1)
    ds = o.buf("DEN_STATES", 1)
    raw[0:4] = [5, 9, 9, 3]
    ds[0:4] = [1, 0, 7, 2]
    raw[4:8] = [0, 0, 0, 0]
    ds[4:8] = [100, 100, 50, 10]   # sum 260 >= 256 -> 255
    o.k("tft_cycle", 1, 0)
    ids = o.buf("TFT_BEST_DEN_IDS", 1)
    braw = o.buf("TFT_BEST_DEN_STATES_RAW", 1)
    best = o.buf("TFT_BEST_DEN_STATES", 1)
    tst = o.buf("TFT_STATES", 1)
    assert ids[0] == 1 and braw[0] == 9 and best[0] == 7 and tst[0] == 10
    assert ids[1] == 0 and braw[1] == 0 and best[1] == 100 and tst[1] == 255
    o.k("tft_cycle", 1, 0)  # previous values shift into prev_*
    assert o.buf("TFT_PREV_BEST_DEN_IDS", 1)[0] == 1
    assert o.buf("TFT_PREV_STATES", 1)[1] == 255


def hex_dist(dv, du):
    return (abs(dv) + abs(du) + abs(dv + du)) // 2


@pytest.mark.parametrize("radius", [1, 2, 4])
def test_inhib_simple_brute_force(radius):
    """Winner-take-all against an independent formulation: for each cell, every
    grid cell within hex distance r (cube-coordinate distance, all grid cells
    enumerated) must satisfy s*H^2 >= n*(H^2 - (d-1)^2), H = r + 3."""
    side = 12
    sch = tiny_area(side=side, radius=radius)
    o = P.OracleArea(sch)
    rng = np.random.default_rng(10 + radius)
    soma = o.buf("SOMA_STATES", 0)
    soma[:side * side] = rng.choice([0, 0, 0, 10, 50, 120, 200, 255], side * side)
    o.k("inhib_simple", 0, 0)
    idz, _ = sch.slice_idzs()
    axn = o.buf("AXN_STATES")[idz[1]:idz[1] + side * side]
    H2 = (radius + 3) ** 2
    n_offsets = sum(1 for dv in range(-radius, radius + 1) for du in range(-radius, radius + 1)
                    if hex_dist(dv, du) <= radius)
    assert n_offsets == 3 * radius * radius + 3 * radius + 1
    s = soma[:side * side].reshape(side, side).astype(int)
    for v in range(side):
        for u in range(side):
            ok = True
            for nv in range(side):
                for nu in range(side):
                    d = hex_dist(nv - v, nu - u)
                    if d <= radius and s[v, u] * H2 < s[nv, nu] * (H2 - (d - 1) ** 2):
                        ok = False
            assert axn[v * side + u] == (s[v, u] if ok else 0), (v, u)


def test_intra_column_reset_quirk():
    sch = tiny_area(pyr_depth=3)
    o = P.OracleArea(sch)
    soma = o.buf("SOMA_STATES", 1)
    cols = 256
    cases = [([5, 3, 7], [0, 0, 7]), ([7, 3, 5], [7, 0, 5]), ([4, 4, 4], [4, 4, 4]),
             ([9, 2, 1], [9, 2, 1]), ([2, 9, 1], [2, 9, 1]), ([9, 2, 0], [9, 2, 0]),
             ([1, 5, 5], [0, 5, 5]), ([0, 0, 0], [0, 0, 0])]
    for c, (states, _) in enumerate(cases):
        for s in range(3):
            soma[s * cols + c] = states[s]
    o.k("inhib_intra_column", 1)
    idz, _ = sch.slice_idzs()
    axn = o.buf("AXN_STATES")
    base = idz[sch.layers[1].base_slc]
    for c, (_, want) in enumerate(cases):
        got = [int(axn[base + s * cols + c]) for s in range(3)]
        assert got == want, (c, got)


def test_pyr_cycle_closed_forms():
    sch = tiny_area()
    o = P.OracleArea(sch)
    C_ = sch.cells(1)
    tst = o.buf("TFT_STATES", 1)
    braw = o.buf("TFT_BEST_DEN_STATES_RAW", 1)
    cases = [(0, 200, 0), (3, 0, 1), (3, 100, 51), (200, 100, 100), (255, 255, 63 + 127),
             (1, 1, 1), (4, 9, 1 + 4)]
    for c, (p, d, _) in enumerate(cases):
        tst[0 * C_ + c] = p          # proximal tuft 0: tuft state
        braw[1 * C_ + c] = d         # distal tuft 1: best raw dendrite state
    o.k("pyr_cycle", 1)
    soma = o.buf("SOMA_STATES", 1)
    for c, (_, _, want) in enumerate(cases):
        assert soma[c] == want, c


def test_ssc_mtp_seeds_are_twice_the_cell_base(lib):
    """ssc_mtp_simple seeds synapse k of cell c with syn_idz + i = 2*c*spt + k."""
    sch = tiny_area(ssc_spd=8, dens=2)
    o = P.OracleArea(sch)
    spt = 16
    idz, _ = sch.slice_idzs()
    axn = o.buf("AXN_STATES")
    active = [3, 77, 200]
    for c in active:
        axn[idz[1] + c] = 9
    st = o.buf("SYN_STATES", 0)
    st[:] = 0
    for c in active:
        st[c * spt:c * spt + 8] = 200  # first half active, second half inactive
    rnd = 0x1234567
    o.k("ssc_mtp", 0, rnd)
    sw = o.buf("SYN_STRENGTHS", 0)
    for c in range(256):
        for k in range(spt):
            i = c * spt + k
            if c not in active:
                assert sw[i] == 0
                continue
            seed = 2 * c * spt + k
            mix = lib.orc_rnd_mix(rnd, seed)
            if k < 8:
                assert sw[i] == ((mix & 0x7F) > 0)
            else:
                assert sw[i] == -((mix & 0x1FF) > 0x180)


def test_tft_dst_mtp_newly_active_and_shared_flags(lib):
    """Learning on the previous best dendrite of newly active cells only; the
    cell flag is shared by all distal tufts so a second distal tuft never
    learns (src/cortex/tufts.rs:205)."""
    sch = tiny_area(pyr_depth=1, dens=2, pyr_spd=4)
    sch.layers[1].tufts.append(recipes.TuftScheme.apical_distal(2, 4, 0, [2]))
    o = P.OracleArea(sch)
    C_ = sch.cells(1)
    idz, _ = sch.slice_idzs()
    axn = o.buf("AXN_STATES")
    base = idz[sch.layers[1].base_slc]
    flags = o.buf("SOMA_FLAG_SETS", 1)
    pbid = o.buf("TFT_PREV_BEST_DEN_IDS", 1)
    pbraw = o.buf("TFT_PREV_BEST_DEN_STATES_RAW", 1)
    sflags = o.buf("SYN_FLAG_SETS", 1)
    # cell 5: newly active, tuft prev active, best den 1; cell 6: was active
    axn[base + 5] = 40
    axn[base + 6] = 40
    flags[6] = 0x08
    for t in (1, 2):
        pbid[t * C_ + 5] = 1
        pbraw[t * C_ + 5] = 3
        pbraw[t * C_ + 6] = 3
    spt = 8
    tz1 = C_ * spt
    z = tz1 + 5 * spt + 1 * 4
    sflags[z:z + 2] = 0x10
    rnd = -99
    o.k("tft_dst_mtp", 1, 1, rnd)
    o.k("tft_dst_mtp", 1, 2, rnd)
    sw = o.buf("SYN_STRENGTHS", 1)
    changed = np.flatnonzero(sw)
    assert set(changed.tolist()) <= set(range(z, z + 4))
    for k in range(4):
        i = z + k
        mix = lib.orc_rnd_mix(rnd, z + i)
        want = ((mix & 0x7F) > 0) if k < 2 else -((mix & 0xFFF) > 0xF80)
        assert sw[i] == want
    assert flags[5] == 0x08 and flags[6] == 0x08 and flags[7] == 0


def test_ssc_cycle_energy_rules():
    sch = tiny_area()
    o = P.OracleArea(sch)
    e = o.buf("SOMA_ENERGIES", 0)
    s = o.buf("SOMA_STATES", 0)
    cases = [(0, 10, 0, 10), (192, 10, 128, 255), (192, 0, 192, 0), (255, 0, 191, 255),
             (255, 7, 191, 255), (191, 50, 191, 50)]
    for c, (en, st, _, _) in enumerate(cases):
        e[c], s[c] = en, st
    o.k("ssc_cycle", 0)
    for c, (_, _, en2, st2) in enumerate(cases):
        assert (e[c], s[c]) == (en2, st2), c


# ------------------------------------------------------------------ regrowth

def _regrow_area(scheme, seed=1, weak_frac=None):
    """Oracle area whose synapses all start on (slice 0, 0, 0) with random
    strengths (all <= -25 when weak_frac is None)."""
    o = P.OracleArea(scheme)
    rng = np.random.default_rng(seed)
    for li in range(len(scheme.layers)):
        n = o.buf("SYN_STRENGTHS", li).size
        if weak_frac is None:
            st = np.full(n, -100, dtype=np.int8)
        else:
            st = rng.integers(-127, 128, n).astype(np.int8)
        o.write("SYN_STRENGTHS", li, st)
        # every synapse starts on a valid source of its tuft
        lay = scheme.layers[li]
        cells = scheme.cells(li)
        parts = [np.full(cells * t.syns_per_tft, t.src_slcs[0], dtype=np.uint8) for t in lay.tufts]
        o.write("SYN_SRC_SLC_IDS", li, np.concatenate(parts))
    return o


def _tuft_view(o, scheme, li, t):
    """(slc, v, u, strength) of tuft t of layer li, shaped (dendrites, spd)."""
    lay = scheme.layers[li]
    cells = scheme.cells(li)
    z = sum(cells * x.syns_per_tft for x in lay.tufts[:t])
    n = cells * lay.tufts[t].syns_per_tft
    spd = lay.tufts[t].syns_per_den
    out = []
    for name in ("SYN_SRC_SLC_IDS", "SYN_SRC_COL_V_OFFS", "SYN_SRC_COL_U_OFFS", "SYN_STRENGTHS"):
        out.append(o.read(name, li)[z:z + n].astype(np.int32).reshape(-1, spd))
    return out


def test_regrow_spatial_tile_is_the_hex_ball_and_unique_per_dendrite():
    r = 6
    s = recipes.cortical_area("rg", 16, input_side=16, ssc_dens=4, ssc_spd=32, pyr_dens=4,
                              pyr_spd=32)
    o = _regrow_area(s)
    o.regrow(seed=7, epoch=0, reach=r)
    slc, v, u, st = _tuft_view(o, s, 0, 0)           # SSC tuft: source slice 0, scale 16
    assert np.all(slc == 0)
    ball = {(a, b) for a in range(-r, r + 1) for b in range(-r, r + 1)
            if max(abs(a), abs(b), abs(a + b)) <= r}
    assert len(ball) == 3 * r * (r + 1) + 1
    seen = set(zip(v.ravel().tolist(), u.ravel().tolist()))
    assert seen == ball                               # inside the tile, and all of it drawn
    # 32 synapses per dendrite over a 127-position tile: sources unique per dendrite
    keys = slc * 65536 + (v + 128) * 256 + (u + 128)
    for row in keys:
        assert np.unique(row).size == row.size
    # strength = ((r - |v|) + (r - |u|)) >> 3 times a draw from -3..3
    inten = ((r - np.abs(v)) + (r - np.abs(u))) >> 3
    assert np.all(st[inten == 0] == 0)
    k = st[inten > 0] // inten[inten > 0]
    assert np.all(st[inten > 0] % inten[inten > 0] == 0)
    assert set(np.unique(k).tolist()) == set(range(-3, 4))
    o.close()


def test_regrow_scaled_and_nonspatial_sources():
    # input 24x24 on a 64x64 area: scale 6, reach 6 -> scaled reach (6 * 6) >> 4 = 2
    s = recipes.cortical_area("rs", 64, input_side=24, ssc_dens=2, ssc_spd=8, pyr_dens=2,
                              pyr_spd=8)
    o = _regrow_area(s)
    o.regrow(seed=3, epoch=5, reach=6)
    slc, v, u, st = _tuft_view(o, s, 0, 0)
    ball2 = {(a, b) for a in range(-2, 3) for b in range(-2, 3) if max(abs(a), abs(b), abs(a + b)) <= 2}
    assert set(zip(v.ravel().tolist(), u.ravel().tolist())) == ball2
    assert np.all(st == 0)                            # ((2 - |v|) + (2 - |u|)) >> 3 == 0
    o.close()
    # a non-spatial source slice (scale 0, size 230): v, u in [-115, 115], strength 0
    s = recipes.cortical_area("rn", 24, ssc_dens=3, ssc_spd=8, pyr_dens=3, pyr_spd=8)
    s.slices[0] = recipes.SliceScheme.nonspatial(230, 230)
    o = _regrow_area(s)
    o.regrow(seed=3, epoch=5, reach=6)
    slc, v, u, st = _tuft_view(o, s, 0, 0)
    assert v.min() >= -115 and v.max() <= 115 and u.min() >= -115 and u.max() <= 115
    assert v.min() < -100 and v.max() > 100           # the whole range is used
    assert np.all(st == 0)
    o.close()


def test_regrow_touches_only_weak_synapses_and_is_deterministic():
    s = recipes.cortical_area("rw", 16, input_side=16, ssc_dens=4, ssc_spd=16, pyr_dens=4,
                              pyr_spd=16)
    a = _regrow_area(s, seed=11, weak_frac=True)
    b = _regrow_area(s, seed=11, weak_frac=True)
    c = _regrow_area(s, seed=11, weak_frac=True)
    before = {(li, n): a.read(n, li) for li in range(2)
              for n in ("SYN_SRC_SLC_IDS", "SYN_SRC_COL_V_OFFS", "SYN_SRC_COL_U_OFFS", "SYN_STRENGTHS")}
    a.regrow(seed=9, epoch=2, reach=6)
    b.regrow(seed=9, epoch=2, reach=6)
    c.regrow(seed=9, epoch=3, reach=6)
    changed_any = False
    for li in range(2):
        weak = before[(li, "SYN_STRENGTHS")] <= -25
        for n in ("SYN_SRC_SLC_IDS", "SYN_SRC_COL_V_OFFS", "SYN_SRC_COL_U_OFFS", "SYN_STRENGTHS"):
            x = a.read(n, li)
            assert np.array_equal(x[~weak], before[(li, n)][~weak]), (li, n)
            assert np.array_equal(x, b.read(n, li))
            changed_any |= not np.array_equal(x, c.read(n, li))
        # regrown strengths are small: |s| <= 3 * intensity <= 3
        assert np.all(np.abs(a.read("SYN_STRENGTHS", li)[weak]) <= 3)
    assert changed_any                                # another epoch draws other sources
    for o in (a, b, c):
        o.close()


# ------------------------------------------------------------------ smoother
@pytest.mark.parametrize("r,V,U", [(4, 64, 64), (2, 17, 23), (6, 40, 31), (1, 9, 12)])
def test_smoother_groups_tile_every_cell_once(r, V, U):
    """The reference's smoother_overlap invariant (src/tests/hex_tile_fields.rs:
    140-180): with every field cell counted, each cell is hit exactly once per
    group, i.e. six times per cycle.  Brute force over the centers' hex fields."""
    v, u, counts = P.smoother_centers(r, V, U)
    assert len(counts) == 6 and sum(counts) == v.size
    hits = np.zeros((6, V, U), np.int32)
    c0 = 0
    for g in range(6):
        for k in range(c0, c0 + counts[g]):
            for dv in range(-r, r + 1):
                for du in range(max(-r, -dv - r), min(r, -dv + r) + 1):
                    vv, uu = v[k] + dv, u[k] + du
                    if 0 <= vv < V and 0 <= uu < U:
                        hits[g, vv, uu] += 1
        c0 += counts[g]
    assert np.all(hits == 1)
    assert np.all(hits.sum(axis=0) == 6)
    # every center lies inside the reference's bounds [-r, dims + r)
    assert v.min() >= -r and u.min() >= -r and v.max() < V + r and u.max() < U + r
    # the engine's generator gives the same centers
    ve, ue, ce = L.smoother_centers(r, V, U)
    assert np.array_equal(v, ve) and np.array_equal(u, ue) and ce == counts


def _smooth_area(depth=1, side=16, r=3):
    s = recipes.cortical_area("sm", side, ssc_depth=depth, ssc_dens=2, ssc_spd=4, pyr_dens=2,
                              pyr_spd=4)
    o = P.OracleArea(s)
    o.set_smoother(0, r)
    return s, o


def test_smoother_equal_activities_change_nothing():
    """All activities equal: least and most are the same (last) cell of every
    field, so no energy moves (cl/control.cl:280-291)."""
    s, o = _smooth_area()
    o.write("SOMA_ACTIVITIES", 0, np.full(s.cells(0), 77, np.uint8))
    e = np.random.default_rng(0).integers(0, 256, s.cells(0)).astype(np.uint8)
    o.write("SOMA_ENERGIES", 0, e)
    o.k("smooth", 0)
    assert np.array_equal(o.read("SOMA_ENERGIES", 0), e)
    o.close()


def test_smoother_moves_one_unit_per_group_and_conserves():
    """Distinct activities: each center adds one to its field's least active cell
    and takes one from its most active (energies away from 0 and 255), so the
    total is conserved and a cell changes by at most six."""
    s, o = _smooth_area(depth=2, side=20, r=2)
    n = s.cells(0)
    rng = np.random.default_rng(1)
    act = rng.permutation(np.arange(n) % 250).astype(np.uint8) + 1   # 1..250: never 0 or 255
    o.write("SOMA_ACTIVITIES", 0, act)
    o.write("SOMA_ENERGIES", 0, np.full(n, 100, np.uint8))
    o.k("smooth", 0)
    e = o.read("SOMA_ENERGIES", 0).astype(np.int32)
    assert e.sum() == 100 * n
    assert np.abs(e - 100).max() <= 6 and np.count_nonzero(e != 100) > 0
    # the global extremes are always their fields' pick: they move in all six groups
    lo = np.flatnonzero(act == act.min())[-1]
    hi = np.flatnonzero(act == act.max())[-1]
    assert e[lo] == 106 and e[hi] == 94
    o.close()


def test_smoother_ties_go_to_the_later_cell_and_saturate():
    s, o = _smooth_area(side=12, r=1)
    V = U = 12
    n = s.cells(0)
    act = np.full(n, 100, np.uint8)
    # one field of group 0 gets two minima and two maxima
    v, u, counts = P.smoother_centers(1, V, U)
    k = next(i for i in range(counts[0]) if 2 <= v[i] < V - 2 and 2 <= u[i] < U - 2)
    cv, cu = int(v[k]), int(u[k])
    field = [(cv + dv) * U + (cu + du) for dv in (-1, 0, 1)
             for du in range(max(-1, -dv - 1), min(1, -dv + 1) + 1)]
    act[field[0]] = act[field[3]] = 5        # two least active, field order 0 < 3
    act[field[1]] = act[field[5]] = 200      # two most active, 1 < 5
    o.write("SOMA_ACTIVITIES", 0, act)
    en = np.full(n, 50, np.uint8)
    o.write("SOMA_ENERGIES", 0, en)
    o.k("smooth", 0)
    e = o.read("SOMA_ENERGIES", 0)
    assert e[field[3]] > 50 and e[field[5]] < 50          # later ones picked
    assert e[field[0]] <= 50 + 5 and e[field[1]] >= 50 - 5
    # saturation: at 255 the least active cell cannot rise, at 0 the most cannot fall
    en = np.full(n, 50, np.uint8)
    en[field[3]] = 255
    en[field[5]] = 0
    o.write("SOMA_ACTIVITIES", 0, act)
    o.write("SOMA_ENERGIES", 0, en)
    o.k("smooth", 0)
    e = o.read("SOMA_ENERGIES", 0)
    assert e[field[3]] == 255 and e[field[5]] == 0
    o.close()


# ------------------------------------------------------------------ retina
def _retina_area(side=16, depth=1):
    s = recipes.cortical_area("rt", 32, input_side=side, ssc_dens=2, ssc_spd=4, pyr_dens=2,
                              pyr_spd=4)
    return s, P.OracleArea(s)


def _retina(o, s, frame):
    s0 = min(s.input_slcs)
    o.retina(s0, 1, frame)
    off, n = s.input_range()
    return o.read("AXN_STATES")[off:off + n]


def test_retina_hand_computed_cases():
    """cl/filters.cl:34-100 by hand: cell index bits pick the cell type."""
    side = 16
    s, o = _retina_area(side)
    # impulse at on-center small (idx & 3 == 0): 255 - (255 // 37 + 48) = 201
    for idx, want in [(5 * side + 4, 201),       # idx & 3 == 0: small on-center
                      (5 * side + 6, 205),       # idx & 3 == 2: large on-center, 255 // 91 = 2
                      (5 * side + 5, 0)]:        # odd: off-center, a bright center is never on
        f = np.zeros(side * side, np.uint8)
        f[idx] = 255
        out = _retina(o, s, f)
        assert out[idx] == want, idx
        assert np.count_nonzero(out) == (1 if want else 0)
    # dark off-center cells in a bright surround
    for c, want in [(6 * side + 7, 200),          # idx & 3 == 3: small, (36 * 255) // 37 - 48
                    (6 * side + 5, 204)]:         # idx & 3 == 1: large, (90 * 255) // 91 - 48
        f = np.full(side * side, 255, np.uint8)
        f[c] = 0
        assert _retina(o, s, f)[c] == want, c
    # uniform field: interior cells see mean == center, nothing passes the threshold
    f = np.full(side * side, 180, np.uint8)
    out = _retina(o, s, f).reshape(side, side)
    assert np.all(out[5:-5, 5:-5] == 0)
    # corner of a uniform 255 field: 10 of the 37 field cells are on the grid,
    # off-grid ones count as zeros: 255 - ((10 * 255) // 37 + 48) = 139
    f = np.full(side * side, 255, np.uint8)
    assert _retina(o, s, f)[0] == 139
    o.close()


def test_retina_against_a_convolution():
    """Hex-field sums by 2-D convolution (scipy), an independent formulation of
    the neighbourhood; cell types from the index bits."""
    from scipy.signal import convolve2d
    side = 24
    s, o = _retina_area(side)
    rng = np.random.default_rng(4)
    f = (rng.random(side * side) < 0.3).astype(np.uint8) * rng.integers(0, 256, side * side,
                                                                         dtype=np.uint8)
    out = _retina(o, s, f)
    img = f.reshape(side, side).astype(np.int64)

    def field_sum(r):
        m = np.zeros((2 * r + 1, 2 * r + 1), np.int64)
        for dv in range(-r, r + 1):
            for du in range(-r, r + 1):
                if max(abs(dv), abs(du), abs(dv + du)) <= r:
                    m[dv + r, du + r] = 1
        return convolve2d(img, m[::-1, ::-1], mode="same"), int(m.sum())

    s3, n3 = field_sum(3)
    s5, n5 = field_sum(5)
    idx = np.arange(side * side).reshape(side, side)
    off = (idx & 1) == 1
    large = (((idx & 3) >> 1) ^ (idx & 1)) == 1
    avg = np.where(large, s5 // n5, s3 // n3)
    want = np.where(off, np.maximum(avg - (img + 48), 0), np.maximum(img - (avg + 48), 0))
    assert np.array_equal(out, want.ravel().astype(np.uint8))
    o.close()


# ------------------------------------------ axon index forms (scalar vs vec4)


def test_axn_state_scalar_equals_vec4_on_spatial_slices():
    """Reference invariant (src/tests/kernels.rs:21-91 with the test kernels of
    cl/tests.cl:15-85): the scalar and vec4 axon index functions agree on every
    spatial slice (mid 0), for every cell and every offset in [-8, 8] -- on a
    local slice (scale 16) and on a slice mapped from a smaller area (scale 6)."""
    sch = recipes.cortical_area("q", 16, input_side=6, ssc_dens=1, ssc_spd=4, pyr_dens=1,
                                pyr_spd=4, pyr_depth=1, distal=False)
    assert [s.v_scale for s in sch.slices] == [6, 16, 16]
    o = P.OracleArea(sch)
    rng = np.random.default_rng(30)
    axn = o.buf("AXN_STATES")
    axn[:] = rng.integers(1, 256, axn.size).astype(np.uint8)
    n_safe = n_unsafe = 0
    for slc in range(3):
        for v in range(16):
            for u in range(16):
                for vo in range(-8, 9):
                    for uo in range(-8, 9):
                        a = o.axn_state(slc, v, vo, u, uo, 0)
                        b = o.axn_state(slc, v, vo, u, uo, 1)
                        assert a == b, (slc, v, u, vo, uo, a, b)
                        n_safe += a[2]
                        n_unsafe += 1 - a[2]
    assert n_safe > 0 and n_unsafe > 0


def test_axn_state_char_wrapped_mid_hand_cases():
    """Non-spatial slice of 230 x 230 (scale 0, mid 115): the scalar index adds
    the mid into the char offset (cl/bismit.cl:331-332), which wraps past 127,
    while vec4 adds in int (:373-374).  Expected indices worked by hand:
    row * 230 + col, e.g. 127 * 230 + 110 = 29320; 115 + 13 = 128 wraps to
    -128 in char (off the slice)."""
    sch = recipes.cortical_area("w", 16, ssc_dens=1, ssc_spd=4, pyr_dens=1, pyr_spd=4,
                                pyr_depth=1, distal=False)
    sch.slices[0] = recipes.SliceScheme.nonspatial(230, 230)
    o = P.OracleArea(sch)
    axn = o.buf("AXN_STATES")
    axn[:230 * 230] = (np.arange(230 * 230) % 255 + 1).astype(np.uint8)
    # (v_ofs, u_ofs) -> (int-form index or None when off the slice,
    #                    scalar-form index or None)
    cases = {
        (0, 0): (26565, 26565),           # (115, 115)
        (12, -5): (29320, 29320),         # (127, 110): no wrap yet
        (13, 0): (29555, None),           # int (128, 115); char -128
        (20, 0): (31165, None),           # int (135, 115); char -121
        (114, 0): (52785, None),          # int (229, 115); char -27
        (0, 13): (26578, None),           # u wraps the same way
        (-115, -115): (0, 0),             # (0, 0)
        (-116, 0): (None, None),          # row -1 in both forms
        (-128, 0): (None, None),          # -13 in both forms
        (127, 0): (None, None),           # int 242 >= 230; char -14
    }
    for (vo, uo), want in cases.items():
        for quirk in (0, 1):
            for v_id, u_id in ((0, 0), (5, 9), (15, 15)):   # scale 0: the cell position drops out
                st, idx, safe = o.axn_state(0, v_id, vo, u_id, uo, quirk)
                w = want[quirk]
                if w is None:
                    assert safe == 0 and st == 0, (vo, uo, quirk)
                else:
                    assert safe == 1 and idx == w and st == w % 255 + 1, (vo, uo, quirk, idx)


# ---------------------------------------------------------- tft_set_syn_flags


def test_syn_flags_every_byte_and_state():
    """tft_set_syn_flags (cl/syns.cl:23-33): for every flag byte, with the
    synapse's previous state 0 and not 0, only bit 4 (SYN_PREV_ACTIVE, 0x10)
    changes, and it ends up set exactly when the state is not 0; states are
    untouched."""
    sch = tiny_area(pyr_spd=8, dens=2)          # layer 1: 512 cells x 2 tufts x 16 synapses
    o = P.OracleArea(sch)
    fl = o.buf("SYN_FLAG_SETS", 1)
    st = o.buf("SYN_STATES", 1)
    assert fl.size >= 1024
    fl[:] = 0
    st[:] = 0
    fl[:512] = np.tile(np.arange(256), 2)
    st[256:512] = np.array([1, 128, 255, 200] * 64, dtype=np.uint8)
    st_before = st.copy()
    o.k("syn_flags", 1)
    for i in range(512):
        byte, active = i % 256, i >= 256
        want = (byte | 0x10) if active else (byte & 0xEF)
        assert fl[i] == want, (i, fl[i], want)
    assert np.array_equal(st, st_before)
    assert np.all(fl[512:] == 0)


# ------------------------------------------- inhib_passthrough and pyr_output


def test_inhib_passthrough_is_inhib_simple_of_radius_zero():
    """inhib_passthrough (cl/control.cl:194-217) writes each cell's soma to its
    own axon, unmasked, and rates activity with seed soma | cell and decay 768.
    inhib_simple at radius 0 (pinned by brute force above) tests a cell only
    against itself (distance 0: s*9 >= s*8), so it must give identical axons
    and activities for the same rnd."""
    side = 16
    rng = np.random.default_rng(40)
    sch_p = tiny_area(side=side, radius=4)
    sch_s = tiny_area(side=side, radius=0)
    op, os_ = P.OracleArea(sch_p), P.OracleArea(sch_s)
    soma = rng.choice(np.array([0, 0, 0, 1, 50, 200, 255], np.uint8), side * side)
    act = rng.integers(0, 256, side * side).astype(np.uint8)
    axn0 = rng.integers(0, 256, op.buf("AXN_STATES").size).astype(np.uint8)
    for o in (op, os_):
        o.buf("AXN_STATES")[:] = axn0
        o.buf("DEN_STATES", 0)[:side * side] = soma     # the SSC soma aliases dendrite states
        o.buf("SOMA_ACTIVITIES", 0)[:] = act
    for rnd in (0, -12345, 987654321):
        op.k("inhib_passthrough", 0, rnd)
        os_.k("inhib_simple", 0, rnd)
        assert np.array_equal(op.read("AXN_STATES"), os_.read("AXN_STATES"))
        assert np.array_equal(op.read("SOMA_ACTIVITIES", 0), os_.read("SOMA_ACTIVITIES", 0))
    # closed form: the layer's axon slice is the soma, every other axon untouched
    idz, _ = sch_p.slice_idzs()
    got = op.read("AXN_STATES")
    assert np.array_equal(got[idz[1]:idz[1] + side * side], soma)
    keep = np.ones(got.size, bool)
    keep[idz[1]:idz[1] + side * side] = False
    assert np.array_equal(got[keep], axn0[keep])
    # activities move by at most one step per call and only up where the soma fires
    o1 = P.OracleArea(sch_p)
    o1.buf("DEN_STATES", 0)[:side * side] = soma
    o1.buf("SOMA_ACTIVITIES", 0)[:] = act
    o1.k("inhib_passthrough", 0, 777)
    d = o1.read("SOMA_ACTIVITIES", 0).astype(int) - act.astype(int)
    assert np.all(np.abs(d) <= 1)
    assert np.all(d[soma == 0] <= 0)


def test_pyr_output_writes_the_soma_unmasked():
    """pyr_output (cl/control.cl:298-314): every slice of the layer's axons
    equals its soma, including zeros, and nothing else changes.  With one slice
    it equals inhib_intra_column (pinned by the reset-quirk cases above), whose
    only slice is always the column's best."""
    rng = np.random.default_rng(41)
    for depth in (1, 2, 3):
        sch = tiny_area(pyr_depth=depth)
        o = P.OracleArea(sch)
        axn0 = rng.integers(0, 256, o.buf("AXN_STATES").size).astype(np.uint8)
        o.buf("AXN_STATES")[:] = axn0
        soma = rng.choice(np.array([0, 0, 3, 90, 190], np.uint8), sch.cells(1))
        o.buf("SOMA_STATES", 1)[:] = soma
        o.k("pyr_output", 1)
        off, n = sch.layer_axn_range(1)
        got = o.read("AXN_STATES")
        assert np.array_equal(got[off:off + n], soma)
        keep = np.ones(got.size, bool)
        keep[off:off + n] = False
        assert np.array_equal(got[keep], axn0[keep])
        if depth == 1:
            o2 = P.OracleArea(sch)
            o2.buf("AXN_STATES")[:] = axn0
            o2.buf("SOMA_STATES", 1)[:] = soma
            o2.k("inhib_intra_column", 1)
            assert np.array_equal(o2.read("AXN_STATES"), got)


# ------------------------------------------- activity rating: decrement rule


def test_activity_rating_decrement_tests_the_updated_rating(lib):
    """update_activity_rating (cl/bismit.cl:531-540) decrements on the rating
    as just incremented.  From a = 0 with the cell active and decay 40000, the
    result is 1 with probability (255/256) * (1 - 40000/65536) = 0.3886 when
    the decrement tests the updated rating, and 255/256 = 0.9961 if it tested
    the old one (0 never decrements).  2^16 seeds: sigma = 0.0019."""
    n = 1 << 16
    ones = sum(lib.orc_update_activity_rating(0, 1, 4242, s, 40000) for s in range(n))
    p = ones / n
    assert abs(p - (255 / 256) * (1 - 40000 / 65536)) < 0.012, p
    # and from a = 1 inactive the rate is decay/65536 whatever a was before
    # (no increment can happen): 2^16 trials at 40000 expect 40000 decrements
    dec = sum(1 - lib.orc_update_activity_rating(1, 0, 4242, s, 40000) for s in range(n))
    assert abs(dec - 40000) < 600, dec


# ------------------------------------------------- area cycle: enqueue order


def _three_layer_scheme(settings):
    """Input slice, a spiny stellate layer, a pyramidal layer with intra-column
    inhibition (proximal + basal distal tuft) and a pyramidal layer whose
    control is pyr_output (proximal + apical distal tuft)."""
    S = recipes.SliceScheme
    T = recipes.TuftScheme
    side = 16
    slices = [S.spatial_input(16, 16, side, side), S(side, side), S(side, side), S(side, side),
              S(side, side)]
    layers = [
        recipes.LayerScheme("iv", L.LAYER_SSC, 1, 1, [T.basal_proximal(2, 8, 40, [0])],
                            L.CTRL_INHIB_SIMPLE, 2),
        recipes.LayerScheme("iii", L.LAYER_PYR, 2, 2,
                            [T.basal_proximal(2, 8, 0, [1]), T.basal_distal(2, 8, 30, [2, 3])],
                            L.CTRL_INTRA_COLUMN),
        recipes.LayerScheme("v", L.LAYER_PYR, 1, 4,
                            [T.basal_proximal(2, 8, 0, [2, 3]), T.apical_distal(2, 8, 30, [4])],
                            L.CTRL_PYR_OUTPUT),
    ]
    return recipes.AreaScheme("order", side, side, slices, layers, settings, input_slcs=[0])


def _cycle_by_kernels(o, sch, r):
    """One CorticalArea::cycle driven kernel by kernel in the reference's
    enqueue order (src/cortex/cortical_area.rs:716-789): spatial layers cycle
    (spiny_stellates.rs:220-252: flags, gather and dendrites per tuft
    (dendrites.rs:191-209, synapses.rs:363-387), soma, control post), then
    spatial layers learn (:256-268), then every other layer cycles
    (pyramidals.rs:297-335: dendrites, tft_cycle per tuft, pyr_cycle, control
    post, tft_dst_mtp per distal tuft, tufts.rs:304-358).  The draws are
    taken per layer in descriptor order as den[t], control, learning[t]."""
    s = sch.settings
    learn = not s & L.SET_DISABLE_LEARNING
    slot0 = np.cumsum([0] + [2 * len(l.tufts) + 1 for l in sch.layers])

    def dens(li, lay):
        o.k("syn_flags", li)
        for t in range(len(lay.tufts)):
            o.k("syn_cycle", li, t)
        for t in range(len(lay.tufts)):
            o.k("den_cycle", li, t, int(r[slot0[li] + t]))

    def control(li, lay):
        rc = int(r[slot0[li] + len(lay.tufts)])
        if lay.control == L.CTRL_INHIB_SIMPLE:
            if s & L.SET_BYPASS_INHIB:
                o.k("inhib_passthrough", li, 0)
            else:
                o.k("inhib_simple", li, rc)
        elif lay.control == L.CTRL_INTRA_COLUMN:
            o.k("inhib_intra_column", li)
        elif lay.control == L.CTRL_PYR_OUTPUT:
            o.k("pyr_output", li)

    spatial = [(li, l) for li, l in enumerate(sch.layers) if l.spatial]
    other = [(li, l) for li, l in enumerate(sch.layers) if not l.spatial]
    if not s & L.SET_DISABLE_SSCS:
        for li, lay in spatial:
            dens(li, lay)
            o.k("ssc_cycle", li)
            control(li, lay)
        if learn:
            for li, lay in spatial:
                o.k("ssc_mtp", li, int(r[slot0[li] + len(lay.tufts) + 1]) & M32)
    if not s & L.SET_DISABLE_PYRS:
        for li, lay in other:
            dens(li, lay)
            for t in range(len(lay.tufts)):
                o.k("tft_cycle", li, t)
            o.k("pyr_cycle", li)
            control(li, lay)
            if learn:
                for t, ts in enumerate(lay.tufts):
                    if ts.den_kind == L.DEN_DISTAL:
                        o.k("tft_dst_mtp", li, t, int(r[slot0[li] + len(lay.tufts) + 1 + t]))


@pytest.mark.parametrize("settings", [0, L.SET_DISABLE_LEARNING, L.SET_BYPASS_INHIB,
                                      L.SET_DISABLE_SSCS, L.SET_DISABLE_PYRS])
def test_area_cycle_is_the_reference_enqueue_order(settings):
    """orc_area_cycle equals the single-kernel entry points (each pinned above)
    run in the reference's enqueue order, over 12 cycles with learning on real
    activity: a swapped stage (learning before inhibition, tft_dst_mtp before
    pyr_cycle, spatial learning inside the spatial cycle loop) or a draw taken
    from the wrong slot changes some buffer."""
    import sys, os
    sys.path.insert(0, os.path.dirname(__file__))
    from helpers import host_layout, load_layout, layer_bufs
    sch = _three_layer_scheme(settings)
    a, b = P.OracleArea(sch), P.OracleArea(sch)
    lay = host_layout(sch, seed=3, reach=3)
    load_layout(a, sch, lay)
    load_layout(b, sch, lay)
    rng = np.random.default_rng(50)
    st0 = []
    for li in range(len(sch.layers)):
        st = rng.integers(-20, 30, sch.layer_synapses(li)).astype(np.int8)
        a.write("SYN_STRENGTHS", li, st)
        b.write("SYN_STRENGTHS", li, st)
        st0.append(st)
    assert a.n_slots == sch.rnd_slots() == 3 + 5 + 5
    n_cyc = 12
    rnd = rng.integers(-2**31, 2**31, (n_cyc, a.n_slots)).astype(np.int32)
    for c in range(n_cyc):
        frame = ((rng.random(256) < 0.3) * rng.integers(1, 256, 256)).astype(np.uint8)
        a.intake(frame)
        b.intake(frame)
        a.cycle(rnd[c], 1)
        _cycle_by_kernels(b, sch, rnd[c])
        assert np.array_equal(a.read("AXN_STATES"), b.read("AXN_STATES")), c
        for li in range(len(sch.layers)):
            for name in layer_bufs(sch, li):
                assert np.array_equal(a.read(name, li), b.read(name, li)), (c, li, name)
    # not vacuous: layers that cycle on input fired, and with learning on
    # their strengths moved; a disabled kind stays silent and unlearned (with
    # SSCs off the pyramidal layers get no input, so only the flags change)
    for li, lay in enumerate(sch.layers):
        off = settings & (L.SET_DISABLE_SSCS if lay.spatial else L.SET_DISABLE_PYRS)
        fired = a.read("SOMA_STATES", li).any()
        moved = not np.array_equal(a.read("SYN_STRENGTHS", li), st0[li])
        if off or settings & L.SET_DISABLE_SSCS:
            assert not fired, (li, settings)
        else:
            assert fired, (li, settings)
            assert moved == (not settings & L.SET_DISABLE_LEARNING), (li, settings)


def test_slice_offsets_and_rnd_slots_by_hand():
    """orc_slice_idz and orc_rnd_slots on a scheme worked by hand: slices of
    24x24 (scaled input), then 16x16 for the spiny stellate slice and each of
    the two pyramidal slices, then 230x230 (non-spatial) start at 0, 576, 832,
    1088, 1344; slots are 2 * tufts + 1 per layer (den[t], control, learning[t])."""
    sch = recipes.cortical_area("h", 16, input_side=24, pyr_depth=2, ssc_dens=1, ssc_spd=4,
                                pyr_dens=1, pyr_spd=4)
    sch.slices.append(recipes.SliceScheme.nonspatial(230, 230))
    o = P.OracleArea(sch)
    lib = o.lib
    import ctypes as C
    got = []
    for s in range(5):
        z = C.c_uint32()
        assert lib.orc_slice_idz(o.h, s, C.byref(z)) == 0
        got.append(z.value)
    assert got == [0, 576, 832, 1088, 1344]
    z = C.c_uint32()
    assert lib.orc_slice_idz(o.h, 5, C.byref(z)) != 0
    assert o.read("AXN_STATES").size == 1344 + 230 * 230
    assert o.n_slots == (2 * 1 + 1) + (2 * 2 + 1) == 8
