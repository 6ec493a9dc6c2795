"""ctypes binding of libfis.so (the C ABI declared in include/fis.h).

The library is built in-tree by :func:`build` (hipcc, gfx950) and loaded
with ``ctypes`` -- no JIT, no torch extension.  There is no fallback: if the
library is missing or does not match this binding, :class:`KernelUnavailable`
is raised, the same loud-failure contract as ``EventKernelUnavailable``
(lab/flybrain/evspmv.py:173-190).
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent.parent  # repo root
CSRC = PKG.parent / "csrc"
SOURCES = [CSRC / "lif.hip"]
HEADER = ROOT / "include" / "fis.h"
LIB_PATH = PKG / "libfis.so"
ABI_VERSION = 15

MAX_WINDOW = 20
SCRATCH_BYTES = 64 << 20  # FIS_SCRATCH_BYTES: fis_window_noalloc's staging
RING_LEN = 40  # int32 ring slots (two delays)
DELAY = 20     # steps from a spike to the input it makes


def write_slot(u):
    """Ring slot that the spikes of step ``u`` are summed into."""
    return (u + DELAY + 1) % RING_LEN


def read_slot(t):
    """Ring slot that step ``t`` reads (the spikes of step t - DELAY)."""
    return (t + 1) % RING_LEN


FIRE_OFF, OUT_OFF, IN_OFF, HOLD, NOREFRAC = 0x01, 0x02, 0x04, 0x08, 0x10
BLOCK_WAVES, LINE_WORDS, LINE_INTS, EVENT_ALIGN = 4, 16, 32, 8
REC_SLOTS, PULL_TILE, PULL_MAX = 40, 256, 1024
PULL_TILES = (128, 256, 512)  # pull-road workgroup shapes compiled
ST_SPIKE, ST_RMAX = 0x80, 127

HIPCC_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-fPIC", "-shared", "-std=c++17",
    # one IEEE rounding per source operation, denormals kept (the reference
    # runs IEEE float32 with gradual underflow)
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    # no packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


class KernelUnavailable(RuntimeError):
    """libfis.so is missing, stale or unloadable.  Never silently replaced."""


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise KernelUnavailable("hipcc not found (looked at $HIPCC and "
                            "/opt/rocm/bin/hipcc)")


def build(verbose=False):
    """Compile libfis.so for gfx950 next to this file (always recompiles)."""
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", str(ROOT / "include"), "-o", str(tmp),
           *map(str, SOURCES)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise KernelUnavailable("hipcc failed (%d):\n%s" % (res.returncode,
                                                            res.stderr))
    if verbose and res.stderr:
        print(res.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


class FisSim(ctypes.Structure):
    P = ctypes.c_void_p
    _fields_ = [
        ("n", ctypes.c_int32), ("batch", ctypes.c_int32),
        ("bstride", ctypes.c_int32), ("ev_cap", ctypes.c_int32),
        ("offsets", P), ("cols", P), ("vals", P),
        ("v", P), ("g", P), ("st", P), ("fl", P), ("vhold", P),
        ("ring", P), ("sbits", P), ("events", P), ("ev_count", P),
        ("clock", P), ("seeds", P), ("trial_ids", P),
        ("w_scale", ctypes.c_float), ("a_syn", ctypes.c_float),
        ("a_mem", ctypes.c_float), ("v_rest", ctypes.c_float),
        ("v_reset", ctypes.c_float), ("v_th", ctypes.c_float),
        ("gain", ctypes.c_float), ("prob_scale", ctypes.c_float),
        ("refrac_steps", ctypes.c_int32),
        ("spikes_out", P), ("counts", P), ("pop_counts", P),
        ("raster", P), ("raster_len", ctypes.c_int32),
        ("in_offsets", P), ("in_cols", P), ("in_vals", P),
        ("rec", P), ("act", P), ("order", P),
        ("t_origin", P), ("t_end", P), ("word_end", P),
        ("drive_group", P), ("drive_rate", P),
        ("drive_n_groups", ctypes.c_int32), ("drive_n_seg", ctypes.c_int32),
        ("drive_seg_steps", ctypes.c_int32),
        ("pull_tile", ctypes.c_int32),
        ("clock_ahead", ctypes.c_int32), ("defer_tick", ctypes.c_int32),
        ("drive_dense", P), ("due", P), ("calm", P), ("rec_ids", P),
        ("work", P),
    ]


class FisDriveEntry(ctypes.Structure):
    _fields_ = [("idx", ctypes.c_uint32), ("group", ctypes.c_uint32)]


EVENT_BYTES = 16
DRIVE_ENTRY_BYTES = 8

_lib = None


def lib():
    """Load libfis.so once; raise KernelUnavailable if it cannot be used."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise KernelUnavailable(
            "%s is not built.  Run __graft_entry__.build() (hipcc, gfx950)."
            % LIB_PATH)
    try:
        L = ctypes.CDLL(str(LIB_PATH))
    except OSError as e:
        raise KernelUnavailable("cannot load %s: %s" % (LIB_PATH, e)) from e
    P, I, I64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
    sims = ctypes.POINTER(FisSim)
    sig = {
        "fis_abi_version": ([], I),
        "ev_spmv_abi": ([], I),
        "fis_struct_layout": ([I], I64),
        "fis_n_waves": ([ctypes.c_int32, ctypes.c_int32], I64),
        "fis_n_blocks": ([ctypes.c_int32, ctypes.c_int32], I64),
        "ev_spmv_i32": ([P] * 6 + [I, I, I, P], I),
        "ev_spmv_i64": ([P] * 6 + [I, I, I, P], I),
        "fis_drive_rates": ([sims, P, I, P], I),
        "fis_drive_table": ([sims, P, I, P, I, I, I, I, P], I),
        "fis_drive_tape": ([sims, P, P, I64, I, P], I),
        "fis_update": ([sims, I, P], I),
        "fis_scatter": ([sims, I, P], I),
        "fis_window_noalloc": ([sims, I, P], I),
        "fis_update_scatter": ([sims, I, P], I),
        "fis_tick": ([sims, I, P], I),
        "fis_seed_ring": ([sims, I, P], I),
        "fis_resend": ([sims, P, I, I, P], I),
        "fis_pull_window": ([sims, I, I, P], I),
        "fis_pull_resend": ([sims, P, I, P], I),
        "fis_pull_to_ring": ([sims, P, P], I),
        "fis_reset_trials": ([sims, P, I, ctypes.c_float, P], I),
        "fis_flag_pairs": ([sims, P, I, ctypes.c_uint32, P], I),
    }
    for name, (args, res) in sig.items():
        try:
            fn = getattr(L, name)
        except AttributeError as e:
            raise KernelUnavailable("%s lacks %s: rebuild it" % (LIB_PATH, name)) from e
        fn.argtypes = args
        fn.restype = res
    if L.fis_abi_version() != ABI_VERSION:
        raise KernelUnavailable("libfis ABI %d, binding expects %d: rebuild"
                                % (L.fis_abi_version(), ABI_VERSION))
    want = {0: ctypes.sizeof(FisSim), 1: FisSim.offsets.offset,
            2: FisSim.w_scale.offset, 3: FisSim.refrac_steps.offset,
            4: FisSim.raster_len.offset, 5: EVENT_BYTES,
            6: ctypes.sizeof(FisDriveEntry), 7: FisSim.in_offsets.offset,
            8: FisSim.t_origin.offset, 9: FisSim.pull_tile.offset,
            10: FisSim.clock_ahead.offset, 11: FisSim.drive_dense.offset,
            12: FisSim.due.offset, 13: FisSim.calm.offset,
            14: FisSim.rec_ids.offset, 15: FisSim.work.offset}
    for k, v in want.items():
        got = L.fis_struct_layout(k)
        if got != v:
            raise KernelUnavailable("struct layout %d: library %d, binding %d"
                                    % (k, got, v))
    _lib = L
    return L


def check(rc, what):
    """Turn a nonzero hipError_t into an exception naming the call."""
    if rc != 0:
        raise RuntimeError("%s failed: hipError %d" % (what, rc))


def exported_symbols():
    """Every entry point include/fis.h declares (parsed from the header)."""
    import re

    text = HEADER.read_text()
    return sorted(set(re.findall(
        r"^\s*(?:int|int64_t)\s+((?:fis|ev_spmv)_\w+)\s*\(", text, re.M)))
