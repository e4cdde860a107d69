"""The library's gfx950 code, disassembled from this build's own objects:
no range-checked buffer load carries an immediate offset.

The projection kernels read image columns and projection-data runs through
raw buffers whose range check turns a plane outside the run into a 0. That
check does not wrap the sum of a load's register offset and its immediate
offset, so if the compiler splits a lane offset into slot q's register offset
plus a constant (256 bytes for the next z slot), slot q + 1 reads 0 wherever
slot q's offset is negative, although its own plane is inside the run. The
kernels keep such offsets whole (whole_offset in pmrt_kernels.hpp); this test
fails if any `buffer_load ... offen offset:N` appears in the device code.
"""
import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "_build", "csrc")
LLVM = "/opt/rocm/llvm/bin"
TOOLS = [os.path.join(LLVM, t) for t in ("llvm-objcopy", "clang-offload-bundler", "llvm-objdump")]
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"

FOLDED = re.compile(r"\bbuffer_load\w*\s.*\boffen\b.*\boffset:\d+")
LABEL = re.compile(r"^[0-9a-f]+ <([^>]+)>:")


def _objects():
    return sorted(glob.glob(os.path.join(BUILD, "*.o")))


def _device_disassembly(obj, tmp):
    """Disassembly of the gfx950 code object in obj's HIP fat binary, or None
    for an object without device code."""
    fat = os.path.join(tmp, "fat.bin")
    co = os.path.join(tmp, "gfx950.co")
    r = subprocess.run([TOOLS[0], "--dump-section", f".hip_fatbin={fat}", obj, os.devnull],
                       capture_output=True, text=True)
    if r.returncode != 0 or not os.path.exists(fat) or os.path.getsize(fat) == 0:
        return None
    subprocess.run([TOOLS[1], "--unbundle", "--type=o", f"--targets={TARGET}", f"--input={fat}", f"--output={co}"],
                   check=True, capture_output=True)
    out = subprocess.run([TOOLS[2], "-d", "--mcpu=gfx950", co], check=True, capture_output=True, text=True).stdout
    os.remove(fat)
    os.remove(co)
    return out


def test_no_folded_offsets_on_range_checked_loads(tmp_path):
    if not all(os.path.exists(t) for t in TOOLS):
        pytest.skip("ROCm LLVM tools not found")
    objs = _objects()
    if not objs:
        pytest.skip("no objects under _build/csrc (run __graft_entry__.build() first)")
    checked, found = 0, []
    for obj in objs:
        dis = _device_disassembly(obj, str(tmp_path))
        if dis is None:
            continue
        checked += 1
        func = "?"
        for line in dis.splitlines():
            m = LABEL.match(line)
            if m:
                func = m.group(1)
            elif FOLDED.search(line):
                found.append(f"{os.path.basename(obj)}: {func}: {line.strip()[:90]}")
    assert checked >= 5, f"only {checked} objects with gfx950 code"
    assert not found, "buffer loads with an immediate offset:\n" + "\n".join(found[:20])


def test_pattern_catches_a_folded_load():
    """The check itself: the folded form matches, the plain form does not."""
    assert FOLDED.search("buffer_load_dword v6, v14, s[16:19], 0 offen offset:256")
    assert not FOLDED.search("buffer_load_dword v160, v161, s[52:55], 0 offen            // 0000000638F4")
