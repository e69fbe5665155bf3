"""Builds the host emulation of the fused search kernel (test infrastructure).

The device code of second-candidate_amd/csrc/sc_search.hip (its anonymous
namespace up to the host-side plan code) is compiled as C++ against shim.h and
linked with the oracle library and driver.inc.  Usage:
    python make_emu.py OUT_BINARY SANITIZER [encode|trav3|travw|scan]   (SANITIZER: address | thread)
With `encode`, the MFMA encode kernel of sc_encode_mfma.hip instead; with
`trav3` / `travw`, the three-candidate / one-workgroup-per-query traversal
kernels, and with `scan`, the per-query scan kernel, with their own drivers.
"""
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(ROOT, "second-candidate_amd", "csrc", "sc_search.hip")
ORACLE = os.path.join(ROOT, "oracle")


def generate(path, driver="driver.inc"):
    src = open(SRC).read()
    # launch helpers use the HIP runtime; the emulation calls the kernel directly
    src = re.sub(r"using FusedKernel = .*?(?=__global__ void k_zero_ints)", "", src, flags=re.S)
    a, b = src.index("namespace {"), src.index("struct Plan {")
    body = src[a:b]
    body = body.replace("extern __shared__ __align__(16) unsigned char smem[];",
                        "unsigned char *smem = g_smem;")
    body = body.replace('asm volatile("" : "+v"(v));', "")
    out = ['#include "shim.h"', "#include <vector>", "extern unsigned char *g_smem;", body,
           "} // namespace",
           "void emu_run_fused(const FusedArgs &a, bool mixed) {",
           "    if (mixed) k_search_fused<3, true>(a); else k_search_fused<3, false>(a);",
           "}", open(os.path.join(HERE, driver)).read()]
    open(path, "w").write("\n".join(out))


ENC_SRC = os.path.join(ROOT, "second-candidate_amd", "csrc", "sc_encode_mfma.hip")


def generate_encode(path):
    """The device code of sc_encode_mfma.hip (its anonymous namespace up to
    the host-side codebook preparation) with the MFMA builtins emulated."""
    src = open(ENC_SRC).read()
    a, b = src.index("namespace {"), src.index("int prep_codebook(")
    body = src[a:b]
    body = body.replace("extern __shared__ __align__(16) unsigned char smem[];",
                        "unsigned char *smem = g_smem;")
    body = body.replace("__builtin_amdgcn_mfma_f32_16x16x16bf16_1k", "emu_mfma_bf16")
    body = body.replace("__builtin_amdgcn_mfma_f32_16x16x32_bf16", "emu_mfma32_bf16")
    out = ['#include "shim.h"', "#include <vector>", "extern unsigned char *g_smem;", body,
           "} // namespace", open(os.path.join(HERE, "driver_encode.inc")).read()]
    open(path, "w").write("\n".join(out))


def build(out, sanitizer, cxx=None, encode=False, trav3=False, travw=False, scan=False):
    cxx = cxx or os.environ.get("EMU_CXX", "/opt/rocm/llvm/bin/clang++")
    cpp = out + ".cpp"
    if encode:
        generate_encode(cpp)
    elif trav3:
        generate(cpp, "driver_trav3.inc")
    elif travw:
        generate(cpp, "driver_travw.inc")
    elif scan:
        generate(cpp, "driver_scan.inc")
    else:
        generate(cpp)
    cmd = [cxx, "-std=c++17", "-O1", "-g", "-fsanitize=" + sanitizer, "-fno-omit-frame-pointer",
           "-I" + HERE, "-I" + ORACLE, cpp, "-o", out, "-L" + ORACLE, "-loracle",
           "-Wl,-rpath," + ORACLE, "-lpthread"]
    subprocess.run(cmd, check=True)
    return out


if __name__ == "__main__":
    mode = sys.argv[3] if len(sys.argv) > 3 else ""
    build(sys.argv[1], sys.argv[2], encode=mode == "encode", trav3=mode == "trav3",
          travw=mode == "travw", scan=mode == "scan")
