"""Build libgcompress.so (gfx950 code object embedded) and the CPU oracle.

    python gpu-compressor_amd/build.py            # library + oracle
    python gpu-compressor_amd/build.py --lib-only

Steps:
  1. hipcc --genco --offload-arch=gfx950 csrc/kernels.hip -> _build/kernels.hsaco and
     _build/kernels_parse.hsaco (the parse kernels, built with PARSE_FLAGS)
  2. g++ the host sources (the code object is embedded with .incbin and
     loaded at run time with hipModuleLoadData) -> libgcompress.so
  3. gcc oracle/gc_oracle.c -> oracle/liboracle.so  (test infrastructure)

Outputs stay in the tree (git-ignored) so that they travel with the source.
"""
import argparse
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
LIB = os.path.join(PKG, "libgcompress.so")
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")

HOST_SOURCES = [
    "gc_runtime.cpp",
    "gc_matchfind.cpp",
    "gc_batch.cpp",
    "gc_corpus.cpp",
    "gc_codec.cpp",
    "gc_serve.cpp",
    "gc_lzma_enc.cpp",
    "gc_stream.cpp",
]
KERNEL_SOURCES = ["kernels.hip"]
# HIP sources compiled for host and device into the library itself (device
# variables registered with the runtime, looked up with hipGetSymbolAddress)
HOST_HIP_SOURCES = ["gc_noalloc.hip"]

# No packed-FP32 instructions anywhere (none are needed: the work is integer).
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


# The parse kernels' code object: uniform branches are left unstructurized
# (LLVM's StructurizeCFG otherwise rebuilds the walk's uniform control flow
# with extra flow blocks, whose phis become register copies on every path).
PARSE_FLAGS = ["-mllvm", "-structurizecfg-skip-uniform-regions"]


def build_code_objects(flags=(), defines=(), out_dir=BUILD):
    """(main, parse) code objects from kernels.hip."""
    os.makedirs(out_dir, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in KERNEL_SOURCES]
    defs = [f"-D{d}" for d in defines]
    main_co = os.path.join(out_dir, "kernels.hsaco")
    parse_co = os.path.join(out_dir, "kernels_parse.hsaco")
    run([HIPCC, "--genco", *HIP_FLAGS, *flags, *defs, "-DGC_BUILD_PARSE=0", "-I", CSRC, *srcs, "-o", main_co])
    run([HIPCC, "--genco", *HIP_FLAGS, *PARSE_FLAGS, *flags, *defs, "-DGC_BUILD_MAIN=0", "-I", CSRC, *srcs,
         "-o", parse_co])
    return main_co, parse_co


def build_library(cos):
    objs = []
    for s in HOST_SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(s)[0] + ".o")
        run([
            "g++", "-c", "-O2", "-g", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter",
            "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", "-I", CSRC,
            f'-DGC_CO_PATH="{cos[0]}"', f'-DGC_PARSE_CO_PATH="{cos[1]}"',
            os.path.join(CSRC, s), "-o", obj,
        ])
        objs.append(obj)
    for s in HOST_HIP_SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(s)[0] + ".o")
        run([HIPCC, "-c", "-fPIC", *HIP_FLAGS, "-I", CSRC, os.path.join(CSRC, s), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    run(["g++", "-shared", "-o", tmp, *objs, f"-L{ROCM}/lib", f"-Wl,-rpath,{ROCM}/lib",
         "-lamdhip64", "-ldl", "-lpthread"])
    os.replace(tmp, LIB)


def build_oracle():
    tmp = ORACLE_LIB + ".tmp"
    run(["gcc", "-O2", "-g", "-std=c11", "-fPIC", "-shared", "-Wall", "-Wextra",
         os.path.join(ROOT, "oracle", "gc_oracle.c"), "-o", tmp, "-lpthread"])
    os.replace(tmp, ORACLE_LIB)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib-only", action="store_true")
    ap.add_argument("--oracle-only", action="store_true")
    a = ap.parse_args(argv)
    if not a.oracle_only:
        build_library(build_code_objects())
    if not a.lib_only:
        build_oracle()


if __name__ == "__main__":
    main()
