"""The C-ABI library loads and exports every entry point include/rbmpp.h
declares, and the Fortran binding module declares the same set (no compute
calls: these run without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "rbmpp.h")
LIB = os.path.join(ROOT, "rbm_pp_amd", "librbmpp.so")


def _declared():
    src = open(HDR).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rbmpp_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = _declared()
    for must in ("rbmpp_create", "rbmpp_make_sample", "rbmpp_measure", "rbmpp_sr_step",
                 "rbmpp_measure_energy", "rbmpp_eval", "rbmpp_ratio_exchange", "rbmpp_local"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_config_struct_matches_header():
    from rbm_pp_amd import api
    # rbmpp_config: 14 int32 + 2 double + uint64 + 2 double, no padding
    assert ctypes.sizeof(api.Config) == 14 * 4 + 2 * 8 + 8 + 3 * 8
    assert ctypes.sizeof(api.SrConfig) == 4 * 8 + 2 * 4
    assert ctypes.sizeof(api.SrReport) == 7 * 8 + 2 * 4
