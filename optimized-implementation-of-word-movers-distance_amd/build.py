"""Build libwmd.so (HIP, gfx950) in-tree with hipcc.

    python -m <package>.build        (or __graft_entry__.build())

The library is the product: the C ABI of include/wmd.h over the kernels in
csrc/.  Every translation unit compiles to an object in parallel (the
resident-solver shapes are spread over eight units), then one link makes the
shared library.  Packed-FP32 instruction selection is disabled for all
device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libwmd.so")
UNITS = ("wmd_kernels.hip", *(f"wmd_resident_{p}.hip" for p in range(8)), "wmd_mlp.hip",
         "wmd_topk.hip", "wmd_capi.hip", "wmd_io.cpp")
SOURCES = [os.path.join(PKG_DIR, "csrc", f) for f in UNITS]
# The reference program's command line (reads its data files, writes scores).
CLI_SOURCE = os.path.join(PKG_DIR, "csrc", "wmd_b.cpp")
CLI_PATH = os.path.join(PKG_DIR, "wmd_b")
ARCH = os.environ.get("WMD_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


# Units whose matrix-core kernels keep their accumulators in VGPRs (the
# ReLU and output work reads them directly, with no AGPR copies).
MFMA_VGPR_UNITS = ("wmd_mlp.hip",)


def compile_command(src: str, obj: str) -> list[str]:
    extra = (["-mllvm", "-amdgpu-mfma-vgpr-form"]
             if os.path.basename(src) in MFMA_VGPR_UNITS else [])
    return [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", *extra,
            "-I", os.path.join(REPO, "include"), "-c", src, "-o", obj]


def link_command(objs: list[str], out: str) -> list[str]:
    return [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", out]


def cli_command(out: str) -> list[str]:
    """Host-only program over the C ABI: g++ against libwmd.so, found next to
    the executable at run time."""
    return [shutil.which("g++") or "g++", "-O2", "-std=c++17", "-I", os.path.join(REPO, "include"),
            CLI_SOURCE, "-L", PKG_DIR, "-lwmd", "-Wl,-rpath,$ORIGIN", "-o", out]


def _run(cmd: list[str]) -> subprocess.CompletedProcess:
    return subprocess.run(cmd, capture_output=True, text=True)


def build(verbose: bool = False, jobs: int | None = None) -> str:
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 1, 8)
    with tempfile.TemporaryDirectory(prefix="wmd_build_") as tmpdir:
        objs = [os.path.join(tmpdir, os.path.basename(s) + ".o") for s in SOURCES]
        with ThreadPoolExecutor(max_workers=jobs) as pool:
            results = list(pool.map(_run, [compile_command(s, o) for s, o in zip(SOURCES, objs)]))
        failed = [(s, r) for s, r in zip(SOURCES, results) if r.returncode != 0]
        for s, r in failed:
            sys.stderr.write(f"--- {os.path.basename(s)}\n{r.stdout}{r.stderr}")
        if failed:
            raise RuntimeError("hipcc failed building libwmd.so")
        if verbose:
            for r in results:
                sys.stderr.write(r.stderr)
        tmp = LIB_PATH + ".tmp"
        res = _run(link_command(objs, tmp))
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("hipcc failed linking libwmd.so")
    os.replace(tmp, LIB_PATH)
    res = _run(cli_command(CLI_PATH + ".tmp"))
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed building wmd_b")
    os.replace(CLI_PATH + ".tmp", CLI_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
