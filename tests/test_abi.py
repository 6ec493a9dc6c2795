"""The C-ABI library: loads, exports every declared symbol, and its struct
layout agrees with the ctypes mirror.  No kernel is launched (CPU)."""

from __future__ import annotations

import ctypes
import subprocess

import pytest

from flybrain_amd import _lib as L


@pytest.fixture(scope="module")
def lib():
    if not L.LIB_PATH.exists():
        L.build()
    return L.lib()


def test_exports_every_declared_symbol(lib):
    declared = L.exported_symbols()
    assert len(declared) >= 12
    # the reference's event-SpMV entry points, under their own names
    # (lab/flybrain/ev_spmv.cu:114-137, bound by name at evspmv.py:886-903)
    assert {"ev_spmv_i32", "ev_spmv_i64", "ev_spmv_abi"} <= set(declared)
    raw = ctypes.CDLL(str(L.LIB_PATH))
    for name in declared:
        assert hasattr(raw, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(L.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(declared) <= exported


def test_versions_and_layout(lib):
    assert lib.fis_abi_version() == L.ABI_VERSION
    assert lib.ev_spmv_abi() == 3  # the reference's ev_spmv_abi()
    assert lib.fis_struct_layout(0) == ctypes.sizeof(L.FisSim)
    assert lib.fis_struct_layout(5) == 16
    assert lib.fis_struct_layout(6) == ctypes.sizeof(L.FisDriveEntry)
    for what, field in ((1, "offsets"), (2, "w_scale"), (3, "refrac_steps"),
                        (4, "raster_len"), (7, "in_offsets"), (8, "t_origin"),
                        (9, "pull_tile"), (10, "clock_ahead"),
                        (11, "drive_dense"), (12, "due"), (13, "calm"),
                        (14, "rec_ids"), (15, "work")):
        assert lib.fis_struct_layout(what) == getattr(L.FisSim, field).offset


def test_wave_count(lib):
    assert lib.fis_n_waves(138_639, 1) == (138_639 + 63) // 64
    assert lib.fis_n_waves(138_639, 256) == 138_639 * 4
    assert lib.fis_n_waves(10, 4) == 1


def test_bad_arguments_rejected_before_launch(lib):
    """Argument checks run on the host and return an error code."""
    s = L.FisSim()
    s.n, s.batch, s.bstride = 10, 3, 3  # 3 is not a valid stride
    assert lib.fis_update(ctypes.byref(s), 1, None) != 0
    assert lib.fis_scatter(ctypes.byref(s), 1, None) != 0
    assert lib.ev_spmv_i32(None, None, None, None, None, None, 0, 1, 128,
                           None) != 0


def test_stride_rule():
    from flybrain_amd.device import stride_for

    assert [stride_for(b) for b in (1, 2, 3, 5, 63, 64, 65, 256)] == \
        [1, 2, 4, 8, 64, 64, 128, 256]


def test_header_constants_match_binding():
    text = L.HEADER.read_text()
    for name, val in (("FIS_MAX_WINDOW", L.MAX_WINDOW),
                      ("FIS_RING_LEN", L.RING_LEN),
                      ("FIS_FIRE_OFF", L.FIRE_OFF), ("FIS_OUT_OFF", L.OUT_OFF),
                      ("FIS_IN_OFF", L.IN_OFF), ("FIS_HOLD", L.HOLD),
                      ("FIS_NOREFRAC", L.NOREFRAC),
                      ("FIS_ST_SPIKE", L.ST_SPIKE), ("FIS_ST_RMAX", L.ST_RMAX)):
        line = [ln for ln in text.splitlines()
                if ln.startswith("#define %s " % name)][0]
        assert int(line.split()[2].rstrip("u"), 0) == val, name
