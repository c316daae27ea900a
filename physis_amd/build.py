"""In-tree build of libphysis_rt_hip and of the translated Physis programs.

Everything is compiled for gfx950 with hipcc and placed in physis_amd/lib/,
next to the Python host mirror that loads it.  Called by
__graft_entry__.build() and usable as `python -m physis_amd.build`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "lib")
CSRC = os.path.join(PKG, "csrc")
PROGRAMS = os.path.join(PKG, "programs")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off keeps a*b+c as two roundings, like the REFERENCE target
# compiled by gcc for x86-64 (no FMA), which is what bit-exact parity needs.
# Packed FP32 instructions (v_pk_*_f32) are disabled: some of their forms have
# given silently wrong results on this hardware.
HIP_FLAGS = [
    "-x", "hip", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
]

RUNTIME_SOURCES = [
    os.path.join(CSRC, "runtime", "libphysis_rt_hip.cc"),
    os.path.join(CSRC, "runtime", "reduce_hip.hip"),
    os.path.join(CSRC, "runtime", "buffer_hip.hip"),
    os.path.join(CSRC, "runtime", "grid_user_hip.hip"),
]

# Poisson solver extensions (precision tier, plain FP32 solve) over the
# runtime's grids: libphysis_poisson_hip.
POISSON_SOURCES = [os.path.join(CSRC, "poisson", "mixed_hip.hip")]
# Learned warm start (U-Net forward/backward/Adam): libphysis_warmstart_hip.
WARMSTART_SOURCES = [os.path.join(CSRC, "warmstart", "unet_hip.hip")]

# Streaming-bandwidth reference kernels (copy, 13-stream sum) used by the
# measurement scripts: libphysis_ceiling_hip.
CEILING_SOURCES = [os.path.join(CSRC, "ceiling", "stream_hip.hip")]

TRANSLATOR_SOURCES = [os.path.join(CSRC, "translator", f) for f in
                      ("lexer.cc", "parser.cc", "translator.cc", "physisc.cc")]
PHYSISC = os.path.join(PKG, "bin", "physisc-hip")

# Physis programs translated by physisc-hip and built into their own shared
# libraries: the benchmark programs, and the test programs (own versions of
# the reference's system tests) under tests/programs.
PROGRAM_SOURCES = [os.path.join(PROGRAMS, "diffusion7.c"),
                   os.path.join(PROGRAMS, "himeno.c"),
                   os.path.join(PROGRAMS, "redblack7.c"),
                   os.path.join(PROGRAMS, "udt7.c"),
                   os.path.join(PROGRAMS, "stencil2d.c"),
                   os.path.join(PROGRAMS, "mixed7.c"),
                   os.path.join(ROOT, "tests", "programs", "systests.c")]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True)
    out = "\n".join(line for line in proc.stdout.splitlines()
                    if "packed-fp32-ops" not in line
                    and "argument unused during compilation" not in line)
    if proc.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out)


def _compile_object(src: str, obj: str) -> None:
    _run([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])


def build_runtime(objdir: str) -> str:
    objs = []
    jobs = []
    for src in RUNTIME_SOURCES:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        jobs.append((src, obj))
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda j: _compile_object(*j), jobs))
    out = os.path.join(LIB, "libphysis_rt_hip.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs,
          "-o", out])
    return out


def _build_extension(sources: list[str], name: str, objdir: str) -> str:
    """A shared library over libphysis_rt_hip from HIP sources."""
    objs = []
    for src in sources:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        _compile_object(src, obj)
        objs.append(obj)
    out = os.path.join(LIB, f"lib{name}.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o",
          out, f"-L{LIB}", "-lphysis_rt_hip", "-Wl,-rpath,$ORIGIN"])
    return out


def build_poisson(objdir: str) -> str:
    return _build_extension(POISSON_SOURCES, "physis_poisson_hip", objdir)


def build_warmstart(objdir: str) -> str:
    return _build_extension(WARMSTART_SOURCES, "physis_warmstart_hip", objdir)


def build_ceiling(objdir: str) -> str:
    return _build_extension(CEILING_SOURCES, "physis_ceiling_hip", objdir)


def build_translator() -> str:
    os.makedirs(os.path.dirname(PHYSISC), exist_ok=True)
    _run(["g++", "-std=c++17", "-O2", "-Wall", *TRANSLATOR_SOURCES, "-o",
          PHYSISC])
    return PHYSISC


def translate(src: str) -> str:
    """physisc-hip src.c -> src.hip.cpp next to it."""
    out = src[:-2] + ".hip.cpp"
    _run([PHYSISC, src, "-o", out])
    return out


def build_program(src: str) -> str:
    name = os.path.basename(src)[:-2]
    gen = translate(src)
    out = os.path.join(LIB, f"libphysis_prog_{name}.so")
    _run([HIPCC, *HIP_FLAGS, "-Wno-unused-variable", "-shared", gen, "-o", out,
          f"-L{LIB}", "-lphysis_rt_hip", "-Wl,-rpath,$ORIGIN"])
    return out


def build_all(verbose: bool = True) -> list[str]:
    os.makedirs(LIB, exist_ok=True)
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    built = [build_translator(), build_runtime(objdir)]
    with ThreadPoolExecutor(max_workers=4) as ex:
        ext = [ex.submit(build_poisson, objdir),
               ex.submit(build_warmstart, objdir),
               ex.submit(build_ceiling, objdir)]
        built += list(ex.map(build_program, PROGRAM_SOURCES))
        built += [e.result() for e in ext]
    shutil.rmtree(objdir, ignore_errors=True)
    if verbose:
        for b in built:
            print("built", os.path.relpath(b, ROOT))
    return built


if __name__ == "__main__":
    build_all()
    sys.exit(0)
