"""CPU tests of the host side: the C-ABI library loads and exports every symbol
include/bismit_hip.h declares, the ctypes structs match the C layout, the rnd
schedule and layout generators, recipes and input streams.  No GPU is touched
(no device compute calls)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from bismit_amd import _lib as L
from bismit_amd import inputs, recipes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bismit_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bsm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = L.load()
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    bound = {n for n, _, _ in L.SIGNATURES}
    assert set(names) <= bound, sorted(set(names) - bound)


def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "bismit_hip.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n",'
                   'sizeof(bsm_slice_desc), sizeof(bsm_tuft_desc), sizeof(bsm_layer_desc),'
                   'sizeof(bsm_area_desc), offsetof(bsm_layer_desc, tufts),'
                   'offsetof(bsm_area_desc, layers), offsetof(bsm_tuft_desc, threshold),'
                   'sizeof(bsm_tuning), sizeof(bsm_stage_info),'
                   'offsetof(bsm_stage_info, alg_bytes), offsetof(bsm_stage_info, packed));'
                   'return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    want = [C.sizeof(L.SliceDesc), C.sizeof(L.TuftDesc), C.sizeof(L.LayerDesc),
            C.sizeof(L.AreaDesc), L.LayerDesc.tufts.offset, L.AreaDesc.layers.offset,
            L.TuftDesc.threshold.offset, C.sizeof(L.Tuning), C.sizeof(L.StageInfo),
            L.StageInfo.alg_bytes.offset, L.StageInfo.packed.offset]
    assert got == want


def test_splitmix64_known_answers():
    # Steele, Lea, Flood (2014) reference generator from state 0
    st = C.c_uint64(0)
    lib = L.load()
    got = [lib.bsm_splitmix64(C.byref(st)) for _ in range(3)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def xor128(state, n):
    """Marsaglia (2003) xor128 on a 4-word state."""
    x, y, z, w = state
    M = 0xFFFFFFFF
    out = []
    for _ in range(n):
        t = (x ^ (x << 11)) & M
        x, y, z = y, z, w
        w = (w ^ (w >> 19)) ^ (t ^ (t >> 8))
        out.append(w)
    return out


def test_xorshift128_stream():
    assert xor128((123456789, 362436069, 521288629, 88675123), 1) == [3701687786]
    # the engine seeds the 4 words with two splitmix64 draws (low word first)
    st = C.c_uint64(77)
    lib = L.load()
    a = lib.bsm_splitmix64(C.byref(st))
    b = lib.bsm_splitmix64(C.byref(st))
    state = (a & 0xFFFFFFFF, a >> 32, b & 0xFFFFFFFF, b >> 32)
    assert L.xorshift128(77, 50).tolist() == xor128(state, 50)


def test_calc_scale():
    # src/cmn/slice_dims.rs:200-223
    assert recipes.calc_scale(64, 64) == 16
    assert recipes.calc_scale(24, 64) == 6
    assert recipes.calc_scale(32, 128) == 4
    assert recipes.calc_scale(128, 64) == 32
    with pytest.raises(ValueError):
        recipes.calc_scale(64 * 17, 64)


def test_layout_generator():
    src = [1, 3, 4]
    slc, v, u = L.gen_tuft_layout(9, 1, 0, 200000, src, 6)
    assert set(np.unique(slc).tolist()) == set(src)
    assert v.min() == -6 and v.max() == 6 and u.min() == -6 and u.max() == 6
    hist = np.bincount(v.astype(int) + 6, minlength=13) / v.size
    assert np.all(np.abs(hist - 1 / 13) < 0.01)
    slc2, v2, u2 = L.gen_tuft_layout(9, 1, 0, 1000, src, 6)
    assert np.array_equal(slc2, slc[:1000]) and np.array_equal(v2, v[:1000])
    slc3, _, _ = L.gen_tuft_layout(9, 1, 1, 1000, src, 6)
    assert not np.array_equal(slc3, slc[:1000])


def footprint_rc(scheme):
    desc, keep = scheme.to_c()
    n = C.c_size_t()
    rc = L.load().bsm_area_footprint(C.byref(desc), C.byref(n))
    return rc, n.value, L.load().bsm_last_error().decode()


def test_descriptor_validation():
    ok = recipes.config1_area()
    rc, n, _ = footprint_rc(ok)
    assert rc == 0 and n >= 7 * ok.n_synapses()
    bad = recipes.config1_area()
    bad.layers[0].tufts.append(recipes.TuftScheme.basal_distal(2, 8, 0, [0]))
    rc, _, msg = footprint_rc(bad)
    assert rc != 0 and "exactly 1 tuft" in msg
    bad = recipes.config1_area()
    bad.layers[1].tufts[0].syns_per_den = 300
    rc, _, msg = footprint_rc(bad)
    assert rc != 0 and "syns_per_den" in msg
    bad = recipes.config1_area()
    bad.layers[1].base_slc = 0
    rc, _, msg = footprint_rc(bad)
    assert rc != 0 and "local spatial" in msg


def test_footprint_config2():
    s = recipes.config2_area()
    assert s.n_synapses() == 142606336
    rc, n, _ = footprint_rc(s)
    assert rc == 0
    # 8 B per synapse: two state buffers, static flags, slice id, v and u offsets,
    # their packed copy and the strength; dendrites and cells add the rest
    assert 8 * s.n_synapses() <= n < 9 * s.n_synapses()


def test_config_sizes():
    c1 = recipes.config1_area()
    assert c1.n_synapses() == 2621440
    assert c1.axn_len() == 576 + 4096 + 8192
    areas = recipes.config3_areas()
    assert len(areas) == 64
    sides = [a.v_size for a in areas]
    assert min(sides) >= 32 and max(sides) <= 192 and all(s % 8 == 0 for s in sides)
    assert len(set(sides)) > 10


def test_sdr_stream():
    st = inputs.SdrStream(65536, 0.02, 0.5, seed=3)
    fr = st.frames(20)
    act = (fr != 0).sum(axis=1)
    assert np.all(act == 1311)
    for i in range(19):
        ov = np.logical_and(fr[i] != 0, fr[i + 1] != 0).sum()
        assert ov == 656
    assert set(np.unique(fr).tolist()) == {0, 255}


def test_glyph_stream():
    g = inputs.GlyphStream(seed=1).frames(1000)
    assert g.shape == (1000, 576)
    assert abs(inputs.uniform_glyph_check(g) - 0.15) < 0.01
    assert all(np.array_equal(g[10 * k], g[10 * k + 9]) for k in range(100))
    assert not np.array_equal(g[0], g[10])


def test_symbol_stream():
    s = inputs.SymbolStream(seed=2)
    f = s.frames(500)
    assert np.all((f != 0).sum(axis=1) == 40)
    syms = inputs.SymbolStream(seed=2).symbols(20000)
    counts = np.bincount(syms, minlength=96)
    assert counts.argmax() < 20  # Zipf head dominates


def test_sdr_packed_many_streams_equal_each_stream_alone():
    """next_packed over many streams (split over threads in the library) gives
    every stream the frame it gives alone, in both the dense and the SDR form."""
    from bismit_amd.inputs import SdrStream, next_packed
    rng = np.random.default_rng(3)
    sides = [8 * int(rng.integers(8, 33)) for _ in range(130)]
    a = [SdrStream(s * s, 0.02, 0.5, seed=i) for i, s in enumerate(sides)]
    b = [SdrStream(s * s, 0.02, 0.5, seed=i) for i, s in enumerate(sides)]
    for sdr in (True, False, True):
        sizes = [4 * (st.k + 1) if sdr else st.n for st in a]
        ofs = np.zeros(len(sizes), np.uint64)
        ofs[1:] = np.cumsum(sizes[:-1], dtype=np.uint64)
        buf = np.zeros(int(ofs[-1]) + sizes[-1], np.uint8)
        next_packed(a, buf, ofs, sdr=sdr)
        for i, st in enumerate(b):
            got = buf[int(ofs[i]):int(ofs[i]) + sizes[i]]
            if sdr:
                e = got.view(np.uint32)
                assert e[0] == st.k and np.array_equal(e[1:], st.next_indices())
            else:
                assert np.array_equal(got, st.next_frame())
    for st in a + b:
        st.close()
