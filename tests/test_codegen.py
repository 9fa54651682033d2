"""Checks on the gfx950 machine code of the library's own kernels, read from
the build objects (csrc/_build/*.o, written by build()): the device code is
unbundled with clang-offload-bundler and disassembled with llvm-objdump.

* No kernel calls a device function (no s_swappc_b64).  A call needs a return
  address in s[30:31]; a called function large enough for long branches
  (s_getpc / s_add / s_setpc over a scavenged SGPR pair) has had that pair be
  s[30:31] itself, which turns the function's return into a jump back to its
  own exit block: a wave that never ends.  Inlined code has no return.
* The FP32 BP kernel (ldpc_bp.hip) uses no scratch and contains no long
  branch: its whole body is within short-branch range.
"""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

BUILD = os.path.join(ROOT, "deeplearning_5g_ldpc_decoder_amd", "csrc", "_build")
LLVM = "/opt/rocm/llvm/bin"
OBJECTS = ["ldpc_kernels", "ldpc_fixed", "ldpc_bp", "ldpc_unrolled", "ldpc_synth"]


def _device_code(tmp_path, name):
    obj = os.path.join(BUILD, name + ".o")
    if not os.path.exists(obj):
        pytest.skip(f"{obj} not built")
    if not (shutil.which("objcopy") and os.path.exists(os.path.join(LLVM, "llvm-objdump"))):
        pytest.skip("objcopy / ROCm llvm tools not available")
    fb = tmp_path / (name + ".fb")
    co = tmp_path / (name + ".co")
    subprocess.run(["objcopy", "--dump-section", f".hip_fatbin={fb}", obj], check=True)
    subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--unbundle", "--type=o",
                    f"--input={fb}", "--targets=hipv4-amdgcn-amd-amdhsa--gfx950",
                    f"--output={co}"], check=True)
    dis = subprocess.run([os.path.join(LLVM, "llvm-objdump"), "-d", str(co)], check=True,
                         capture_output=True, text=True).stdout
    notes = subprocess.run([os.path.join(LLVM, "llvm-readelf"), "--notes", str(co)],
                           check=True, capture_output=True, text=True).stdout
    return dis, notes


@pytest.mark.parametrize("name", OBJECTS)
def test_no_device_function_calls(tmp_path, name):
    dis, _ = _device_code(tmp_path, name)
    assert "<" in dis and "s_endpgm" in dis, "empty disassembly"
    assert "s_swappc_b64" not in dis, f"{name}: a kernel calls a device function"


def test_bp_kernel_no_scratch_no_long_branch(tmp_path):
    dis, notes = _device_code(tmp_path, "ldpc_bp")
    assert "ldpc_bp_kernel" in dis
    assert "s_setpc_b64" not in dis
    sizes = [int(line.split(":")[1]) for line in notes.splitlines()
             if ".private_segment_fixed_size:" in line]
    assert sizes and all(s == 0 for s in sizes), sizes


def test_specialized_kernels_no_scratch(tmp_path):
    """The lifting-size specialized kernels (the headline decode) keep their
    state in VGPRs and LDS: a spill to scratch inside the row loop costs a
    memory round trip per access, so any private segment fails the build's
    tests."""
    _, notes = _device_code(tmp_path, "ldpc_fixed")
    sizes = [int(line.split(":")[1]) for line in notes.splitlines()
             if ".private_segment_fixed_size:" in line]
    assert sizes and all(s == 0 for s in sizes), sizes
