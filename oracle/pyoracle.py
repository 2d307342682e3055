"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY (the checker).

OracleArea holds host copies of every buffer in the reference layouts and runs
the plain-C restatement of each reference kernel (oracle/bsm_oracle.c).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

# same buffer numbering as include/bismit_hip.h (enum bsm_buf)
BUF_NAMES = [
    "AXN_STATES", "SOMA_STATES", "SOMA_ENERGIES", "SOMA_ACTIVITIES", "SOMA_FLAG_SETS",
    "TFT_STATES", "TFT_BEST_DEN_IDS", "TFT_BEST_DEN_STATES_RAW", "TFT_BEST_DEN_STATES",
    "TFT_PREV_STATES", "TFT_PREV_BEST_DEN_IDS", "TFT_PREV_BEST_DEN_STATES_RAW",
    "TFT_PREV_BEST_DEN_STATES", "DEN_STATES", "DEN_STATES_RAW", "DEN_ENERGIES",
    "DEN_ACTIVITIES", "DEN_THRESHOLDS", "SYN_STATES", "SYN_STRENGTHS", "SYN_SRC_SLC_IDS",
    "SYN_SRC_COL_V_OFFS", "SYN_SRC_COL_U_OFFS", "SYN_FLAG_SETS",
]
BUF = {n: i for i, n in enumerate(BUF_NAMES)}
SIGNED = {"SYN_STRENGTHS", "SYN_SRC_COL_V_OFFS", "SYN_SRC_COL_U_OFFS"}

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from . import build
            build.build()
        lib = C.CDLL(LIB)
        P = C.c_void_p
        lib.orc_rnd_mix.restype = C.c_int32
        lib.orc_rnd_mix.argtypes = [C.c_int32, C.c_int32]
        lib.orc_rnd_inc.restype = C.c_int32
        lib.orc_rnd_inc.argtypes = [C.c_int32, C.c_int32, C.c_int8, C.c_int32]
        lib.orc_rnd_dec.restype = C.c_int32
        lib.orc_rnd_dec.argtypes = [C.c_int32, C.c_int32, C.c_int8, C.c_int32]
        lib.orc_update_activity_rating.restype = C.c_uint8
        lib.orc_update_activity_rating.argtypes = [C.c_uint8, C.c_int32, C.c_int32, C.c_int32,
                                                   C.c_uint16]
        lib.orc_syn_fire.restype = C.c_uint8
        lib.orc_syn_fire.argtypes = [C.c_uint8]
        lib.orc_lshft_mask.restype = C.c_int32
        lib.orc_lshft_mask.argtypes = [C.c_int32]
        lib.orc_area_create.argtypes = [P, C.POINTER(P)]
        lib.orc_area_destroy.argtypes = [P]
        lib.orc_area_destroy.restype = None
        lib.orc_buf_ptr.argtypes = [P, C.c_int, C.c_uint32, C.POINTER(P), C.POINTER(C.c_size_t)]
        lib.orc_slice_idz.argtypes = [P, C.c_uint32, C.POINTER(C.c_uint32)]
        lib.orc_set_threads.argtypes = [C.c_int]
        lib.orc_set_threads.restype = None
        lib.orc_rnd_slots.argtypes = [P, C.POINTER(C.c_uint32)]
        for n in ["syn_flags", "ssc_cycle", "pyr_cycle", "inhib_intra_column", "pyr_output"]:
            getattr(lib, "orc_k_" + n).argtypes = [P, C.c_uint32]
        for n in ["syn_cycle", "tft_cycle"]:
            getattr(lib, "orc_k_" + n).argtypes = [P, C.c_uint32, C.c_uint32]
        for n in ["den_cycle", "tft_dst_mtp"]:
            getattr(lib, "orc_k_" + n).argtypes = [P, C.c_uint32, C.c_uint32, C.c_int32]
        for n in ["inhib_simple", "inhib_passthrough"]:
            getattr(lib, "orc_k_" + n).argtypes = [P, C.c_uint32, C.c_int32]
        lib.orc_k_ssc_mtp.argtypes = [P, C.c_uint32, C.c_uint32]
        lib.orc_axn_state.restype = C.c_uint8
        lib.orc_axn_state.argtypes = [P, C.c_uint32, C.c_uint32, C.c_int8, C.c_uint32, C.c_int8,
                                      C.c_int32, C.POINTER(C.c_uint32), C.POINTER(C.c_int32)]
        lib.orc_area_cycle.argtypes = [P, C.POINTER(C.c_int32), C.c_uint32]
        lib.orc_area_regrow.argtypes = [P, C.c_uint64, C.c_uint64, C.c_int32]
        lib.orc_retina.argtypes = [P, C.c_uint32, C.c_uint32, P]
        lib.orc_smoother_centers.argtypes = [C.c_int32, C.c_int32, C.c_int32,
                                             C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                             C.POINTER(C.c_uint32), C.c_uint32,
                                             C.POINTER(C.c_uint32)]
        lib.orc_set_smoother.argtypes = [P, C.c_uint32, C.c_int32]
        lib.orc_k_smooth.argtypes = [P, C.c_uint32]
        _lib = lib
    return _lib


def smoother_centers(radius, v_size, u_size):
    """(v, u, group_counts) of the smoother's six center groups."""
    lib = load()
    counts = (C.c_uint32 * 6)()
    n = C.c_uint32()
    if lib.orc_smoother_centers(radius, v_size, u_size, None, None, counts, 0, C.byref(n)):
        raise RuntimeError("orc_smoother_centers failed")
    v = np.empty(n.value, np.int32)
    u = np.empty(n.value, np.int32)
    lib.orc_smoother_centers(radius, v_size, u_size, v.ctypes.data_as(C.POINTER(C.c_int32)),
                             u.ctypes.data_as(C.POINTER(C.c_int32)), counts, n.value, C.byref(n))
    return v, u, list(counts)


def set_threads(n):
    load().orc_set_threads(int(n))


class OracleArea:
    def __init__(self, scheme):
        self.lib = load()
        self.scheme = scheme
        desc, self._keep = scheme.to_c()
        self.h = C.c_void_p()
        if self.lib.orc_area_create(C.cast(C.pointer(desc), C.c_void_p), C.byref(self.h)) != 0:
            raise RuntimeError("orc_area_create failed")
        n = C.c_uint32()
        self.lib.orc_rnd_slots(self.h, C.byref(n))
        self.n_slots = n.value
        self.layer_names = [l.name for l in scheme.layers]
        for li, lay in enumerate(scheme.layers):
            if getattr(lay, "smoother_radius", 0):
                self.set_smoother(li, lay.smoother_radius)

    def set_smoother(self, layer, radius):
        if self.lib.orc_set_smoother(self.h, self.li(layer), radius):
            raise RuntimeError("orc_set_smoother failed")

    def retina(self, slc, depth, frame):
        f = np.ascontiguousarray(frame, dtype=np.uint8)
        if self.lib.orc_retina(self.h, slc, depth, f.ctypes.data):
            raise RuntimeError("orc_retina failed")

    def li(self, layer):
        return self.layer_names.index(layer) if isinstance(layer, str) else int(layer)

    def buf(self, name, layer=0):
        """Writable numpy view of an oracle buffer."""
        p = C.c_void_p()
        n = C.c_size_t()
        if self.lib.orc_buf_ptr(self.h, BUF[name], self.li(layer), C.byref(p), C.byref(n)) != 0:
            raise KeyError(name)
        if n.value == 0:
            return np.zeros(0, dtype=np.int8 if name in SIGNED else np.uint8)
        arr = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(n.value,))
        return arr.view(np.int8) if name in SIGNED else arr

    def read(self, name, layer=0):
        return self.buf(name, layer).copy()

    def write(self, name, layer, arr, offset=0):
        b = self.buf(name, layer)
        a = np.asarray(arr)
        b[offset:offset + a.size] = a.view(b.dtype) if a.dtype.itemsize == 1 else a
        return b

    def intake(self, frame, offset=None):
        if offset is None and getattr(self.scheme, "input_filter", None) == "retina":
            s0 = min(self.scheme.input_slcs)
            self.retina(s0, len(self.scheme.input_slcs), frame)
            return
        if offset is None:
            offset = self.scheme.input_range()[0]
        f = np.asarray(frame, dtype=np.uint8)
        self.buf("AXN_STATES")[offset:offset + f.size] = f

    def cycle(self, rnd, n=1):
        rnd = np.ascontiguousarray(rnd, dtype=np.int32)
        if rnd.size != n * self.n_slots:
            raise ValueError("rnd must hold n * n_slots draws")
        rc = self.lib.orc_area_cycle(self.h, rnd.ctypes.data_as(C.POINTER(C.c_int32)), n)
        if rc:
            raise RuntimeError("orc_area_cycle failed")

    def regrow(self, seed, epoch, reach=6):
        """Deterministic regrowth of every synapse with strength <= -25."""
        if self.lib.orc_area_regrow(self.h, seed, epoch, reach):
            raise RuntimeError("orc_area_regrow failed")

    def axn_state(self, slc, v_id, v_ofs, u_id, u_ofs, scalar_quirk):
        """(state, axon index, safe) the gather reads for one synapse source."""
        idx = C.c_uint32()
        safe = C.c_int32()
        st = self.lib.orc_axn_state(self.h, slc, v_id, v_ofs, u_id, u_ofs, scalar_quirk,
                                    C.byref(idx), C.byref(safe))
        return int(st), idx.value, safe.value

    def k(self, name, *args):
        rc = getattr(self.lib, "orc_k_" + name)(self.h, *args)
        if rc:
            raise RuntimeError("orc_k_" + name + " failed")

    def close(self):
        if self.h:
            self.lib.orc_area_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
