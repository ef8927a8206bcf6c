"""Machine-code checks of the gfx950 kernels (CPU only: hipcc emits the
device assembly, nothing runs).

A vector-memory store of more than 64 bits reads its data registers after
issue; a VALU instruction that rewrites one of them must wait one cycle.
When the wait is missing the store writes the new register contents.  That
happened once in the sampler's update loop (128-bit buffer stores followed
directly by index arithmetic into their data registers) and corrupted X^-1
at rounding-invisible but trajectory-visible levels.  This test keeps every
kernel of the library free of the pattern, and also rejects packed FP32
arithmetic, which the build turns off."""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

from rbm_pp_amd import build

WIDE_STORE = re.compile(r"^(buffer|global|flat|scratch)_store_(dwordx3|dwordx4|b96|b128)\s+(.*)$")


def _vregs(op: str) -> set:
    m = re.match(r"v\[(\d+):(\d+)\]", op)
    if m:
        return set(range(int(m.group(1)), int(m.group(2)) + 1))
    m = re.match(r"v(\d+)$", op)
    return {int(m.group(1))} if m else set()


def wide_store_hazards(asm: str, window: int = 2):
    """(line, store, instruction) for each VALU write of a wide store's data
    registers within `window` issue slots after the store (s_nop n counts as
    n+1 slots)."""
    lines = asm.splitlines()
    found = []
    for i, line in enumerate(lines):
        m = WIDE_STORE.match(line.strip())
        if not m:
            continue
        ops = [o.strip() for o in m.group(3).split(",")]
        data = _vregs(ops[1] if m.group(1) in ("global", "flat") else ops[0])
        j, slots = i + 1, 0
        while j < len(lines) and slots < window:
            t = lines[j].strip()
            j += 1
            if not t or t.startswith((";", ".")) or t.endswith(":"):
                continue
            if t.startswith("s_nop"):
                slots += int(t.split()[1], 0) + 1
                continue
            if t.startswith("v_") and len(t.split()) > 1 and _vregs(t.split()[1].rstrip(",")) & data:
                found.append((i, line.strip(), t))
                break
            slots += 1
    return found


def test_scanner_flags_the_pattern():
    bad = "  buffer_store_dwordx4 v[32:35], v106, s[16:19], s0 offen\n  v_add_u32_e32 v32, 2, v68\n"
    ok = "  buffer_store_dwordx4 v[32:35], v106, s[16:19], s0 offen\n  s_nop 1\n  v_add_u32_e32 v32, 2, v68\n"
    narrow = "  buffer_store_dwordx2 v[32:33], v106, s[16:19], s0 offen\n  v_add_u32_e32 v32, 2, v68\n"
    glob = "  global_store_dwordx4 v[0:1], v[8:11], off\n  v_mov_b32_e32 v9, 0\n"
    assert len(wide_store_hazards(bad)) == 1
    assert wide_store_hazards(ok) == []
    assert wide_store_hazards(narrow) == []
    assert len(wide_store_hazards(glob)) == 1


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="hipcc not available")
def test_kernels_have_no_wide_store_data_hazard_and_no_packed_fp32():
    with tempfile.TemporaryDirectory() as tmp:
        procs = []
        for src in build.SOURCES:
            flags = [f for f in build.flags_for(src) if f != "-fPIC"]
            out = os.path.join(tmp, src.replace(".hip", ".s"))
            cmd = ["hipcc", *flags, "--cuda-device-only", "-S", "-o", out, os.path.join(build.CSRC, src)]
            procs.append((src, out, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        problems = []
        for src, out, p in procs:
            text, _ = p.communicate()
            assert p.returncode == 0, text.decode(errors="replace")[-2000:]
            asm = open(out).read()
            for i, st, nxt in wide_store_hazards(asm):
                problems.append(f"{src}:{i}: {st} -> {nxt}")
            for m in re.finditer(r"^\s*(v_pk_(add|mul|fma)_f32)\b.*$", asm, re.M):
                problems.append(f"{src}: packed FP32 instruction: {m.group(0).strip()}")
    assert not problems, "\n".join(problems[:20])
