"""Build recipe for the gfx950 library and CLI (hipcc, no CMake needed).

    python -m cuspfft_amd.build        # or cuspfft_amd.build.build()

Outputs (git-ignored, shipped in-tree):
    cuspfft_amd/_lib/libspfft_hip.so   C ABI of include/spfft_c.h + C++ sparse_fft.h
    cuspfft_amd/_lib/spfft_cli         benchmark driver with the reference's flags
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIB_DIR, "libspfft_hip.so")
CLI = os.path.join(LIB_DIR, "spfft_cli")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed FP32 ops (v_pk_*_f32) are disabled for all device code: some of their forms have
# given wrong results on this hardware now and then, with no error
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall",
          "-Wno-unused-result", "-I", os.path.join(ROOT, "include"), "-I", CSRC]

LIB_SOURCES = ["csc_build.hip", "csc_noalloc.hip", "rowpass.hip", "rowpass_conv_a.hip", "rowpass_conv_b.hip",
               "rowpass_conv_c.hip", "rowpass_conv_d.hip", "rowpass_conv_e.hip",
               "rowpass_fft1_a.hip", "rowpass_fft1_b.hip", "rowpass_fft1_c.hip", "rowpass_mfma.hip",
               "rowpass_fused_a.hip", "rowpass_fused_b.hip", "rowpass_fused_c.hip",
               "rowpass_fused_d.hip", "rowpass_fused_e.hip", "rowpass_fused_f.hip",
               "rowpass_fused_g.hip", "rowpass_fused_h.hip", "rowpass_rader.hip", "rowpass_raderruns.hip",
               "rowpass_rader_gen_a.hip", "rowpass_rader_gen_b.hip", "rowpass_rader_gen_c.hip",
               "rowpass_rader_gen_d.hip",
               "rowpass_split.hip", "rowpass_split_a.hip", "rowpass_split_b.hip", "rowpass_split_c.hip",
               "spfft_plan.cpp", "spfft_batch.cpp", "mlp.hip",
               "spfft_gen.cpp", "sparse_fft_shim.cpp"]
CLI_SOURCES = ["cli_main.cpp", "cpu_fft.cpp"]


# The row-FFT files prune their codelets to the inputs and outputs that can be
# non-zero / are needed: with signed zeros ignored, x + 0 folds to x.  The Rader
# kernel shares their stage code and is built the same way (its results differ from
# a build without the flag at the rounding level only: the compiler picks other FMA
# and negation forms).
NO_SIGNED_ZEROS = {"rowpass_fused_a.hip", "rowpass_fused_b.hip", "rowpass_fused_c.hip",
                   "rowpass_fused_d.hip", "rowpass_fused_e.hip", "rowpass_fused_f.hip",
                   "rowpass_fused_g.hip", "rowpass_fused_h.hip", "rowpass_rader.hip", "rowpass_raderruns.hip",
                   "rowpass_rader_gen_a.hip", "rowpass_rader_gen_b.hip", "rowpass_rader_gen_c.hip",
                   "rowpass_rader_gen_d.hip",
                   "rowpass_fft1_a.hip", "rowpass_fft1_b.hip", "rowpass_fft1_c.hip"}


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, *DEVICE_FLAGS]
    if src in NO_SIGNED_ZEROS:
        cmd += ["-fno-signed-zeros"]
    if src.endswith(".cpp"):
        cmd += ["-x", "hip"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, LIB_SOURCES + CLI_SOURCES))
    lib_objs = objs[:len(LIB_SOURCES)]
    cli_objs = objs[len(LIB_SOURCES):]
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *lib_objs],
                   check=True)
    os.replace(tmp, LIB)
    tmp = CLI + ".tmp"
    subprocess.run([HIPCC, f"--offload-arch={ARCH}", "-o", tmp, *cli_objs, "-L", LIB_DIR,
                    "-lspfft_hip", "-Wl,-rpath,$ORIGIN"], check=True)
    os.replace(tmp, CLI)
    if verbose:
        print(f"built {LIB} and {CLI}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
