"""Builds the gfx950 shared library librbmpp.so (the C ABI of include/rbmpp.h)
in-tree with hipcc."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librbmpp.so")
SOURCES = ["rbmpp_capi.hip", "rbmpp_sampler.hip", "rbmpp_local.hip", "rbmpp_meas.hip", "rbmpp_meas_blk.hip",
           "rbmpp_meas_blk_big.hip", "rbmpp_corr.hip", "rbmpp_corr_blk.hip", "rbmpp_sr.hip", "rbmpp_syrk.hip"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         # packed FP32 instructions are kept out of every kernel
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# per-source additions.  The sampler is built without interprocedural
# register allocation: with it the cold out-of-line helpers (recompute,
# amplitudes) shape the register assignment of the proposal loop that calls
# them (97 spilled registers instead of 36, 1.3x the sweep time); the other
# kernels are slightly faster with it.
EXTRA_FLAGS = {"rbmpp_sampler.hip": ["-mllvm", "-enable-ipra=0"]}


def flags_for(src: str) -> list:
    return FLAGS + EXTRA_FLAGS.get(src, [])


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = ["hipcc", *flags_for(src), "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((src, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    objs = []
    for src, obj, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l and l.strip()
                 and "nodiscard" not in l]
        if p.returncode != 0:
            sys.stderr.write(text)
            raise RuntimeError(f"hipcc failed on {src}")
        if verbose and lines:
            sys.stderr.write("\n".join(lines) + "\n")
        objs.append(obj)
    tmp = LIB + ".tmp"
    subprocess.check_call(["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    build_fortran()
    return LIB


FDIR = os.path.join(HERE, "fortran")
DRIVER = os.path.join(FDIR, "rbmpp_vmc")


def build_fortran() -> str | None:
    """The Fortran binding module and the main.f90-shaped driver, linked
    against librbmpp.so (amdflang; skipped when the image has no Fortran
    compiler)."""
    fc = shutil.which("amdflang") or shutil.which("flang")
    if fc is None:
        return None
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    mod_o = os.path.join(objdir, "rbmpp_mod.o")
    drv_o = os.path.join(objdir, "rbmpp_vmc.o")
    subprocess.check_call([fc, "-O2", "-J", objdir, "-c", os.path.join(FDIR, "rbmpp_mod.f90"), "-o", mod_o])
    subprocess.check_call([fc, "-O2", "-J", objdir, "-c", os.path.join(FDIR, "rbmpp_vmc.f90"), "-o", drv_o])
    subprocess.check_call([fc, "-o", DRIVER, drv_o, mod_o, "-L" + HERE, "-lrbmpp",
                           "-Wl,-rpath,$ORIGIN/.."])
    return DRIVER


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
