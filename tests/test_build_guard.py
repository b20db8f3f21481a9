"""Build guard: the gfx950 code objects hold no packed-FP32 instructions.

The Makefile compiles every HIP source with
`-Xclang -target-feature -Xclang -packed-fp32-ops`, which turns off the
packed-FP32 feature in the gfx950 device pass (the host pass warns that the
feature is unknown to x86-64 and ignores it).  This test disassembles the
device code of every object in chaoslibrary_amd/csrc/_build (AMD assembly of
code built here) and checks that no v_pk_add_f32 / v_pk_mul_f32 /
v_pk_fma_f32 appears, so a build that lost the flag, or a source that
writes them by hand, fails on the CPU.
"""
import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "chaoslibrary_amd", "csrc", "_build")
LLVM = "/opt/rocm/llvm/bin"
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"
PACKED = re.compile(r"\bv_pk_(add|mul|fma)_f32\b")


def _device_asm(obj, tmp):
    fb = os.path.join(tmp, os.path.basename(obj) + ".fb")
    co = os.path.join(tmp, os.path.basename(obj) + ".co")
    r = subprocess.run([f"{LLVM}/llvm-objcopy", f"--dump-section=.hip_fatbin={fb}", obj],
                       capture_output=True)
    if r.returncode != 0:
        return None  # host-only object (no device code)
    subprocess.run([f"{LLVM}/clang-offload-bundler", "--unbundle", "--type=o", f"--input={fb}",
                    f"--targets={TARGET}", f"--output={co}"], check=True)
    return subprocess.run([f"{LLVM}/llvm-objdump", "-d", "--mcpu=gfx950", co], check=True,
                          capture_output=True, text=True).stdout


def test_no_packed_fp32_in_device_code(tmp_path):
    objs = sorted(glob.glob(os.path.join(BUILD, "*.o")))
    if not objs or not os.path.exists(f"{LLVM}/llvm-objdump"):
        pytest.skip("no built objects or no ROCm LLVM tools")
    seen = 0
    for o in objs:
        asm = _device_asm(o, str(tmp_path))
        if asm is None:
            continue
        n_insn = sum(1 for line in asm.splitlines() if line.strip().startswith("v_"))
        if n_insn == 0:
            continue
        seen += 1
        bad = [line.strip() for line in asm.splitlines() if PACKED.search(line)]
        assert not bad, f"{os.path.basename(o)}: {len(bad)} packed-FP32 instructions, e.g. {bad[:3]}"
    assert seen >= 5, "expected device code in the kernel objects"


def test_makefile_passes_the_feature_flag():
    mk = open(os.path.join(ROOT, "chaoslibrary_amd", "csrc", "Makefile")).read()
    assert "-Xclang -target-feature -Xclang -packed-fp32-ops" in mk
