"""Build recipe for the native libraries, run in-tree.

libfastkpm_hip.so  HIP kernels + C-ABI (include/fastkpm_hip.h), hipcc for gfx950.
libfastkpm.so      C++ fkpm:: host API (include/fastkpm.h) over the C-ABI, g++.

Both land in the package directory (fastkpm_amd/) so that they travel with the
source tree; objects go to fastkpm_amd/_build/.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = CSRC / "host"
INCLUDE = ROOT / "include"
BUILD = PKG / "_build"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["kpm_capi.hip", "kpm_inst_dc.hip", "kpm_inst_dr.hip", "kpm_inst_fc.hip",
               "kpm_inst_fr.hip", "kpm_batch.hip", "kpm_noalloc.hip"]
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             # packed FP32 instructions (v_pk_*_f32) stay off: every kernel is built
             # from plain scalar FP32 code
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I", str(INCLUDE), "-I", str(CSRC)]

DEVICE_LIB = PKG / "libfastkpm_hip.so"
HOST_LIB = PKG / "libfastkpm.so"


def _run_parallel(cmds: list[list[str]]) -> None:
    procs = [(c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
             for c in cmds]
    failed = []
    for c, p in procs:
        out, _ = p.communicate()
        # hipcc's host pass does not know the packed-fp32 feature; drop that note
        lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln]
        if p.returncode != 0:
            failed.append((c, "\n".join(lines)))
        elif any("error" in ln or "warning" in ln for ln in lines):
            sys.stderr.write("\n".join(lines) + "\n")
    if failed:
        msg = "\n\n".join(f"$ {' '.join(c)}\n{o}" for c, o in failed)
        raise RuntimeError(f"native build failed:\n{msg}")


def build_device_lib() -> Path:
    BUILD.mkdir(exist_ok=True)
    cmds = []
    objs = []
    for src in HIP_SOURCES:
        obj = BUILD / (Path(src).stem + ".o")
        objs.append(str(obj))
        cmds.append([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    _run_parallel(cmds)
    tmp = DEVICE_LIB.with_suffix(".so.tmp")
    _run_parallel([[HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(tmp)]])
    os.replace(tmp, DEVICE_LIB)
    return DEVICE_LIB


def build_host_lib() -> Path:
    srcs = sorted(str(p) for p in HOST.glob("*.cpp"))
    if not srcs:
        return HOST_LIB
    tmp = HOST_LIB.with_suffix(".so.tmp")
    _run_parallel([["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra",
                    "-Wno-unused-parameter", "-I", str(INCLUDE), *srcs,
                    "-L", str(PKG), "-lfastkpm_hip", "-Wl,-rpath,$ORIGIN",
                    "-o", str(tmp)]])
    os.replace(tmp, HOST_LIB)
    return HOST_LIB


CPP_TESTS = ["engine_flow", "ref_adapter"]


def build_cpp_test() -> list[Path]:
    """The C++ programs of tests/cpp: the application-style flow over the
    fkpm:: API (engine_flow) and the reference-side Engine<T> that forwards
    to the C-ABI (ref_adapter)."""
    cmds, outs = [], []
    for name in CPP_TESTS:
        src = ROOT / "tests" / "cpp" / f"{name}.cpp"
        if not src.exists():
            continue
        out = ROOT / "tests" / "cpp" / name
        cmds.append(["g++", "-O2", "-std=c++17", "-Wall", "-I", str(INCLUDE), str(src),
                     "-L", str(PKG), "-lfastkpm", "-lfastkpm_hip",
                     "-Wl,-rpath,$ORIGIN/../../fastkpm_amd", "-o", str(out.with_suffix(".tmp"))])
        outs.append(out)
    _run_parallel(cmds)
    for out in outs:
        os.replace(out.with_suffix(".tmp"), out)
    return outs


def build_all() -> None:
    build_device_lib()
    build_host_lib()
    build_cpp_test()


if __name__ == "__main__":
    build_all()
    print("built", DEVICE_LIB, HOST_LIB)
