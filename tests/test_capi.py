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
    assert ctypes.sizeof(api.SrReport) == 8 * 8 + 2 * 4


FMOD = os.path.join(ROOT, "rbm_pp_amd", "fortran", "rbmpp_mod.f90")


def test_fortran_binding_declares_every_entry_point():
    src = open(FMOD).read()
    bound = sorted(set(re.findall(r'bind\(C,\s*name="(rbmpp_[a-z_0-9]+)"\)', src)))
    assert bound == _declared()


_PROBE = """
program probe
  use iso_c_binding
  use rbmpp_mod
  implicit none
  type(rbmpp_config), target :: c
  type(rbmpp_sr_config), target :: s
  type(rbmpp_sr_report), target :: r
  integer(c_intptr_t) :: b
  b = transfer(c_loc(c), b)
  print '(5I8)', c_sizeof(c), c_sizeof(s), c_sizeof(r), &
    transfer(c_loc(c%seed), b) - b, transfer(c_loc(c%cond_guard), b) - b
end program
"""


def test_fortran_struct_layout_matches_c(tmp_path):
    import shutil
    import subprocess
    from rbm_pp_amd import api
    fc = shutil.which("amdflang") or shutil.which("flang")
    if fc is None:
        pytest.skip("no Fortran compiler")
    if not os.path.exists(LIB):
        pytest.skip("library not built")
    (tmp_path / "probe.f90").write_text(_PROBE)
    subprocess.check_call([fc, "-J", str(tmp_path), "-c", FMOD, "-o", str(tmp_path / "m.o")])
    libdir = os.path.dirname(LIB)
    subprocess.check_call([fc, "-J", str(tmp_path), str(tmp_path / "probe.f90"), str(tmp_path / "m.o"),
                           "-L" + libdir, "-lrbmpp", "-Wl,-rpath," + libdir, "-o", str(tmp_path / "probe")])
    out = subprocess.check_output([str(tmp_path / "probe")]).split()
    sizes = [int(t) for t in out]
    assert sizes == [ctypes.sizeof(api.Config), ctypes.sizeof(api.SrConfig), ctypes.sizeof(api.SrReport),
                     api.Config.seed.offset, api.Config.cond_guard.offset]
