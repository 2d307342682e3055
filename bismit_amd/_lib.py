"""ctypes binding of the engine's C ABI (include/bismit_hip.h).

The product path is libbismit_hip.so built in-tree by bismit_amd/build.py; there
is no fallback: if the library is missing, loading raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbismit_hip.so")

MAX_SLICES = 255
MAX_LAYERS = 8
MAX_TUFTS = 3
MAX_TUFT_SRCS = 8

LAYER_SSC, LAYER_PYR = 1, 2
DEN_BASAL, DEN_APICAL = 0, 1
DEN_PROXIMAL, DEN_DISTAL = 0, 1
CTRL_NONE, CTRL_INHIB_SIMPLE, CTRL_INTRA_COLUMN, CTRL_PYR_OUTPUT = 0, 1, 2, 3
SET_BYPASS_INHIB, SET_DISABLE_PYRS, SET_DISABLE_SSCS, SET_DISABLE_LEARNING = 1, 2, 4, 8
RND_DEN, RND_CTRL, RND_MTP = 0, 1, 2
BATCH_DYNAMIC = 1

# buffers, mirroring SamplerKind (src/cortex/cortical_area.rs:61-90)
BUF_NAMES = [
    "AXN_STATES", "SOMA_STATES", "SOMA_ENERGIES", "SOMA_ACTIVITIES", "SOMA_FLAG_SETS",
    "TFT_STATES", "TFT_BEST_DEN_IDS", "TFT_BEST_DEN_STATES_RAW", "TFT_BEST_DEN_STATES",
    "TFT_PREV_STATES", "TFT_PREV_BEST_DEN_IDS", "TFT_PREV_BEST_DEN_STATES_RAW",
    "TFT_PREV_BEST_DEN_STATES", "DEN_STATES", "DEN_STATES_RAW", "DEN_ENERGIES",
    "DEN_ACTIVITIES", "DEN_THRESHOLDS", "SYN_STATES", "SYN_STRENGTHS", "SYN_SRC_SLC_IDS",
    "SYN_SRC_COL_V_OFFS", "SYN_SRC_COL_U_OFFS", "SYN_FLAG_SETS",
]
BUF = {n: i for i, n in enumerate(BUF_NAMES)}
FILTER_RETINA = 1
FRAME_SDR = 1 << 63          # bsm_batch_cycle_masked: the offset names an SDR entry
SIGNED_BUFS = {"SYN_STRENGTHS", "SYN_SRC_COL_V_OFFS", "SYN_SRC_COL_U_OFFS"}


class SliceDesc(C.Structure):
    _fields_ = [("v_size", C.c_uint32), ("u_size", C.c_uint32), ("v_scale", C.c_uint8),
                ("u_scale", C.c_uint8), ("v_mid", C.c_uint8), ("u_mid", C.c_uint8)]


class TuftDesc(C.Structure):
    _fields_ = [("den_class", C.c_uint8), ("den_kind", C.c_uint8),
                ("max_active_dens_l2", C.c_uint8), ("n_src_slcs", C.c_uint8),
                ("src_slc_ids", C.c_uint8 * MAX_TUFT_SRCS), ("dens_per_tft", C.c_uint32),
                ("syns_per_den", C.c_uint32), ("threshold", C.c_uint32)]


class LayerDesc(C.Structure):
    _fields_ = [("kind", C.c_uint8), ("spatial", C.c_uint8), ("depth", C.c_uint8),
                ("base_axn_slc", C.c_uint8), ("n_tufts", C.c_uint8), ("control", C.c_uint8),
                ("_pad", C.c_uint8 * 2), ("inhib_radius", C.c_int32),
                ("tufts", TuftDesc * MAX_TUFTS)]


class AreaDesc(C.Structure):
    _fields_ = [("v_size", C.c_uint32), ("u_size", C.c_uint32), ("n_slices", C.c_uint32),
                ("n_layers", C.c_uint32), ("settings", C.c_uint32), ("_pad", C.c_uint32),
                ("slices", C.POINTER(SliceDesc)), ("layers", C.POINTER(LayerDesc))]


class Tract(C.Structure):
    _fields_ = [("src_area", C.c_uint32), ("src_offset", C.c_uint32), ("dst_area", C.c_uint32),
                ("dst_offset", C.c_uint32), ("len", C.c_uint32)]


class Tuning(C.Structure):
    _fields_ = [("tile_v", C.c_uint32), ("tile_u", C.c_uint32), ("halo", C.c_uint32),
                ("chunk", C.c_uint32), ("inhib_tile", C.c_uint32), ("overlap", C.c_uint32)]


class StageInfo(C.Structure):
    _fields_ = [("kind", C.c_int32), ("layer", C.c_uint32), ("chunk", C.c_uint32),
                ("staged", C.c_uint32), ("fast", C.c_uint32), ("tile_v", C.c_uint32),
                ("tile_u", C.c_uint32), ("n_items", C.c_uint32), ("lds_bytes", C.c_uint32),
                ("grid", C.c_uint32), ("stream", C.c_uint32), ("alg_bytes", C.c_uint64),
                ("packed", C.c_uint32), ("_reserved", C.c_uint32)]


def stage_dicts(arr, n):
    return [{f: getattr(arr[i], f) for f, _ in StageInfo._fields_} for i in range(n)]


_P = C.c_void_p
_U8P = C.POINTER(C.c_uint8)
_I32P = C.POINTER(C.c_int32)

# (name, restype, argtypes)
SIGNATURES = [
    ("bsm_last_error", C.c_char_p, []),
    ("bsm_ctx_create", C.c_int, [C.c_int, C.POINTER(_P)]),
    ("bsm_ctx_destroy", C.c_int, [_P]),
    ("bsm_sync", C.c_int, [_P]),
    ("bsm_device_info", C.c_int, [_P, C.c_char_p, C.c_size_t, C.POINTER(C.c_int),
                                  C.POINTER(C.c_size_t)]),
    ("bsm_area_create", C.c_int, [_P, C.POINTER(AreaDesc), C.c_uint64, C.POINTER(_P)]),
    ("bsm_area_destroy", C.c_int, [_P]),
    ("bsm_area_sync", C.c_int, [_P]),
    ("bsm_area_footprint", C.c_int, [C.POINTER(AreaDesc), C.POINTER(C.c_size_t)]),
    ("bsm_area_axn_len", C.c_int, [_P, C.POINTER(C.c_uint32)]),
    ("bsm_area_syn_count", C.c_int, [_P, C.POINTER(C.c_uint64)]),
    ("bsm_area_slice_idz", C.c_int, [_P, C.c_uint32, C.POINTER(C.c_uint32)]),
    ("bsm_area_buf_len", C.c_int, [_P, C.c_int, C.c_uint32, C.POINTER(C.c_size_t)]),
    ("bsm_area_read", C.c_int, [_P, C.c_int, C.c_uint32, C.c_size_t, C.c_size_t, _P]),
    ("bsm_sdr_create", C.c_int, [C.c_uint32, C.c_double, C.c_double, C.c_uint64, C.POINTER(_P)]),
    ("bsm_sdr_destroy", C.c_int, [_P]),
    ("bsm_sdr_info", C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("bsm_sdr_next", C.c_int, [_P, _P]),
    ("bsm_sdr_frames", C.c_int, [_P, C.c_uint32, _P]),
    ("bsm_sdr_next_packed", C.c_int, [_P, C.c_uint32, _P, _P]),
    ("bsm_sdr_next_indices", C.c_int, [_P, _P]),
    ("bsm_sdr_next_packed_sdr", C.c_int, [_P, C.c_uint32, _P, _P]),
    ("bsm_area_intake_filtered", C.c_int, [_P, C.c_int, C.c_uint32, C.c_uint32, _P, C.c_int]),
    ("bsm_area_intake_filtered_noalloc", C.c_int, [_P, C.c_int, C.c_uint32, C.c_uint32, _P]),
    ("bsm_k_retina", C.c_int, [_P, C.c_uint32, C.c_uint32, _P]),
    ("bsm_area_set_smoother", C.c_int, [_P, C.c_uint32, C.c_int32]),
    ("bsm_k_smooth", C.c_int, [_P, C.c_uint32]),
    ("bsm_smoother_centers", C.c_int, [C.c_int32, C.c_uint32, C.c_uint32, C.POINTER(C.c_int32),
                                       C.POINTER(C.c_int32), C.POINTER(C.c_uint32), C.c_uint32,
                                       C.POINTER(C.c_uint32)]),
    ("bsm_sampler_create", C.c_int, [_P, C.c_int, C.c_uint32, C.c_uint64, C.c_uint64, C.c_int,
                                     C.POINTER(_P)]),
    ("bsm_sampler_recv", C.c_int, [_P, C.c_int, _P, C.c_uint64, C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_int)]),
    ("bsm_sampler_recv_many", C.c_int, [C.POINTER(_P), C.c_uint32, C.c_int, C.POINTER(_P),
                                        C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_int)]),
    ("bsm_sampler_set_backpressure", C.c_int, [_P, C.c_int, C.POINTER(C.c_int)]),
    ("bsm_sampler_set_buffers", C.c_int, [_P, C.c_uint32]),
    ("bsm_sampler_info", C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_int),
                                   C.POINTER(C.c_int)]),
    ("bsm_sampler_destroy", C.c_int, [_P]),
    ("bsm_area_write", C.c_int, [_P, C.c_int, C.c_uint32, C.c_size_t, C.c_size_t, _P]),
    ("bsm_area_fill", C.c_int, [_P, C.c_int, C.c_uint32, C.c_uint8]),
    ("bsm_area_intake", C.c_int, [_P, C.c_uint32, _P, C.c_uint32, C.c_int]),
    ("bsm_area_output", C.c_int, [_P, C.c_uint32, _P, C.c_uint32]),
    ("bsm_k_syn_flags", C.c_int, [_P, C.c_uint32]),
    ("bsm_k_syn_cycle", C.c_int, [_P, C.c_uint32, C.c_uint32]),
    ("bsm_k_den_cycle", C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_int32]),
    ("bsm_k_tft_cycle", C.c_int, [_P, C.c_uint32, C.c_uint32]),
    ("bsm_k_ssc_cycle", C.c_int, [_P, C.c_uint32]),
    ("bsm_k_inhib_simple", C.c_int, [_P, C.c_uint32, C.c_int32]),
    ("bsm_k_inhib_passthrough", C.c_int, [_P, C.c_uint32, C.c_int32]),
    ("bsm_k_ssc_mtp", C.c_int, [_P, C.c_uint32, C.c_uint32]),
    ("bsm_k_pyr_cycle", C.c_int, [_P, C.c_uint32]),
    ("bsm_k_inhib_intra_column", C.c_int, [_P, C.c_uint32]),
    ("bsm_k_pyr_output", C.c_int, [_P, C.c_uint32]),
    ("bsm_k_tft_dst_mtp", C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_int32]),
    ("bsm_area_rnd_slots", C.c_int, [_P, C.POINTER(C.c_uint32)]),
    ("bsm_area_rnd_slot_info", C.c_int, [_P, C.c_uint32, C.POINTER(C.c_int),
                                         C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("bsm_area_cycle", C.c_int, [_P, _I32P, C.c_uint32]),
    ("bsm_area_run", C.c_int, [_P, C.c_uint32, _P, C.c_uint64, C.c_uint32, C.c_uint32]),
    ("bsm_area_draw_rnd", C.c_int, [_P, _I32P, C.c_uint32]),
    ("bsm_area_set_tuning", C.c_int, [_P, C.POINTER(Tuning)]),
    ("bsm_area_get_tuning", C.c_int, [_P, C.POINTER(Tuning)]),
    ("bsm_dev_alloc", C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    ("bsm_dev_free", C.c_int, [_P, _P]),
    ("bsm_dev_upload", C.c_int, [_P, _P, _P, C.c_size_t]),
    ("bsm_area_time_cycles", C.c_int, [_P, C.c_uint32, _P, C.c_uint64, C.c_uint32, C.c_uint32,
                                       C.POINTER(C.c_float)]),
    ("bsm_area_profile_stages", C.c_int, [_P, C.c_uint32, C.c_uint32, C.POINTER(C.c_float),
                                          C.POINTER(C.c_int), C.POINTER(C.c_uint32),
                                          C.POINTER(C.c_uint32)]),
    ("bsm_calib_copy", C.c_int, [_P, _P, _P, C.c_size_t, C.POINTER(C.c_float)]),
    ("bsm_calib_stream", C.c_int, [_P, _P, _P, C.c_size_t, C.c_uint32, C.POINTER(C.c_float)]),
    ("bsm_gen_tuft_layout", C.c_int, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64,
                                      C.c_uint32, _U8P, C.c_int32, _P, _P, _P]),
    ("bsm_area_init_layout", C.c_int, [_P, C.c_uint64, C.c_int32]),
    ("bsm_splitmix64", C.c_uint64, [C.POINTER(C.c_uint64)]),
    ("bsm_xorshift128_fill", C.c_int, [C.c_uint64, C.POINTER(C.c_uint32), C.c_uint64]),
    ("bsm_batch_create", C.c_int, [_P, C.POINTER(_P), C.c_uint32, C.POINTER(_P)]),
    ("bsm_batch_create_ex", C.c_int, [_P, C.POINTER(_P), C.c_uint32, C.c_uint32, C.POINTER(_P)]),
    ("bsm_batch_add_areas", C.c_int, [_P, C.POINTER(_P), C.c_uint32, C.POINTER(C.c_uint32),
                                      C.POINTER(C.c_uint32)]),
    ("bsm_batch_remove_areas", C.c_int, [_P, C.POINTER(_P), C.c_uint32]),
    ("bsm_batch_busy_ms", C.c_int, [_P, C.POINTER(C.c_double)]),
    ("bsm_batch_set_input_ranges", C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("bsm_batch_cycle_masked", C.c_int, [_P, _U8P, _P, C.POINTER(C.c_uint64), C.c_uint64,
                                         C.c_int]),
    ("bsm_batch_query", C.c_int, [_P, C.POINTER(C.c_int)]),
    ("bsm_batch_set_tracts", C.c_int, [_P, C.POINTER(Tract), C.c_uint32]),
    ("bsm_area_copy_axons", C.c_int, [_P, C.c_uint32, _P, C.c_uint32, C.c_uint32]),
    ("bsm_batch_destroy", C.c_int, [_P]),
    ("bsm_batch_cycle", C.c_int, [_P, C.c_uint32, C.c_int]),
    ("bsm_batch_set_inputs", C.c_int, [_P, C.POINTER(_P), C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_uint32]),
    ("bsm_batch_sync", C.c_int, [_P]),
    ("bsm_area_regrow", C.c_int, [_P, C.c_uint64, C.c_uint64, C.c_int32]),
    ("bsm_batch_regrow", C.c_int, [_P, C.POINTER(C.c_uint8), C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_uint64), C.c_int32]),
    ("bsm_cortex_cycle", C.c_int, [_P, C.c_uint32, _P, _P, _P, _P, C.c_uint32, C.c_uint32]),
    ("bsm_batch_time_cycles", C.c_int, [_P, C.c_uint32, C.c_int, C.POINTER(C.c_float)]),
    ("bsm_area_stage_plan", C.c_int, [_P, C.c_uint32, C.POINTER(StageInfo),
                                      C.POINTER(C.c_uint32)]),
    ("bsm_batch_stage_plan", C.c_int, [_P, C.c_uint32, C.POINTER(StageInfo),
                                       C.POINTER(C.c_uint32)]),
    ("bsm_batch_profile_stages", C.c_int, [_P, C.c_uint32, C.c_uint32, C.POINTER(C.c_float),
                                           C.POINTER(C.c_int), C.POINTER(C.c_uint32)]),
    ("bsm_batch_set_tuning", C.c_int, [_P, C.POINTER(Tuning)]),
]

_lib = None


def load():
    """Load libbismit_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m bismit_amd.build`")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class BsmError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        msg = load().bsm_last_error().decode(errors="replace")
        raise BsmError(f"bismit_hip error {rc}: {msg}")
    return rc


def u8ptr(a):
    return a.ctypes.data_as(_U8P)


def vptr(a):
    return C.c_void_p(a.ctypes.data)


def i32ptr(a):
    return a.ctypes.data_as(_I32P)


def xorshift128(seed, n):
    """n draws of the engine's xorshift128 stream seeded by splitmix64(seed)."""
    out = np.empty(int(n), dtype=np.uint32)
    check(load().bsm_xorshift128_fill(C.c_uint64(seed), out.ctypes.data_as(C.POINTER(C.c_uint32)),
                                      C.c_uint64(int(n))))
    return out


def gen_tuft_layout(seed, layer, tuft, n_syns, src_ids, reach):
    """Host-side synapse layout generator (src slice, v offset, u offset)."""
    src = np.asarray(src_ids, dtype=np.uint8)
    slc = np.empty(n_syns, dtype=np.uint8)
    v = np.empty(n_syns, dtype=np.int8)
    u = np.empty(n_syns, dtype=np.int8)
    check(load().bsm_gen_tuft_layout(C.c_uint64(seed), layer, tuft, C.c_uint64(n_syns), len(src),
                                     u8ptr(src), reach, vptr(slc), vptr(v), vptr(u)))
    return slc, v, u


def smoother_centers(radius, v_size, u_size):
    """(v, u, group_counts) of the engine's smoother center groups (host-only)."""
    lib = load()
    counts = (C.c_uint32 * 6)()
    n = C.c_uint32()
    check(lib.bsm_smoother_centers(radius, v_size, u_size, None, None, counts, 0, C.byref(n)))
    v = np.empty(n.value, np.int32)
    u = np.empty(n.value, np.int32)
    check(lib.bsm_smoother_centers(radius, v_size, u_size, v.ctypes.data_as(C.POINTER(C.c_int32)),
                                   u.ctypes.data_as(C.POINTER(C.c_int32)), counts, n.value,
                                   C.byref(n)))
    return v, u, list(counts)
