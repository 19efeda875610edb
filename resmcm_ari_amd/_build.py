"""Build recipes for the native libraries (in-tree, no JIT cache).

* ``lib/libsmcm_hin.so`` - host C++ HIN utilities (generators, text IO, meta-path parser).
* ``lib/libsmcm.so``     - the drop-in C ABI: hand-written HIP kernels for gfx950 plus the host
  C++ chain executor.  Built with ``hipcc --offload-arch=gfx950``.
* ``lib/smcm_resmcm``    - the drop-in command line program (same argv and output line as the
  reference's ``efficient/resmcm``).
* ``lib/smcm_res``       - the drop-in estimate-only program (``accuracy/res``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
OBJDIR = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIN_SOURCES = ["hin.cpp"]
HIP_SOURCES = [
    "device_graph.hip",
    "estimator.hip",
    "spgemm.hip",
    "dense.hip",
    "scan.hip",
    "batch.hip",
    "providers.hip",
    "noalloc.hip",
    "smcm_api.cpp",
    "batch_api.cpp",
    "server.cpp",
]
CLI_SOURCES = ["cli_resmcm.cpp"]

# -ffp-contract=off keeps a*b+c from being fused: the estimator must round every product
# and sum as the reference does.  The packed-fp32 feature is disabled for all device code.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-command-line-argument",
    "-I" + INCLUDE, "-I" + CSRC,
]


def _run(cmd: list[str]) -> None:
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + p.stdout)


def build_hin() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libsmcm_hin.so")
    srcs = [os.path.join(CSRC, s) for s in HIN_SOURCES]
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I" + INCLUDE, *srcs, "-o", out])
    return out


def _compile_hip(src: str) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    _run([HIPCC, *HIP_FLAGS, *lang, "-c", src, "-o", obj])
    return obj


def build_smcm(jobs: int = 8) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile_hip, srcs))
    out = os.path.join(LIBDIR, "libsmcm.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_cli() -> str:
    out = os.path.join(LIBDIR, "smcm_resmcm")
    srcs = [os.path.join(CSRC, s) for s in CLI_SOURCES]
    _run(["g++", "-O2", "-std=c++17", "-I" + INCLUDE, "-I" + CSRC, *srcs, "-o", out,
          "-L" + LIBDIR, "-lsmcm", "-lsmcm_hin", "-Wl,-rpath,$ORIGIN"])
    # the estimate-only program (accuracy/res)
    _run(["g++", "-O2", "-std=c++17", "-I" + INCLUDE, "-I" + CSRC, os.path.join(CSRC, "cli_res.cpp"), "-o",
          os.path.join(LIBDIR, "smcm_res"), "-L" + LIBDIR, "-lsmcm", "-lsmcm_hin", "-Wl,-rpath,$ORIGIN"])
    # meta-path enumeration (host only)
    _run(["g++", "-O2", "-std=c++17", "-I" + INCLUDE, os.path.join(CSRC, "cli_metapaths.cpp"), "-o",
          os.path.join(LIBDIR, "smcm_metapaths"), "-L" + LIBDIR, "-lsmcm_hin", "-Wl,-rpath,$ORIGIN"])
    return out


def build_vendor() -> str:
    """rocSPARSE baseline library (timing comparisons only; the product path never loads it)."""
    out = os.path.join(LIBDIR, "libsmcm_vendor.so")
    _run([HIPCC, "-O2", "-std=c++17", "-fPIC", "-shared", os.path.join(CSRC, "vendor_spgemm.cpp"), "-o", out,
          "-lrocsparse"])
    return out


def build_oracle(with_reference: bool | None = None) -> None:
    """Compile the oracle's C restatement; also the reference's own programs from their
    sources when the reference tree is present (only in the build container)."""
    oracle_dir = os.path.join(ROOT, "oracle")
    if shutil.which("make") is None:
        raise RuntimeError("make not found")
    target = ["all"] if with_reference in (None, True) else ["oracle"]
    _run(["make", "-s", "-C", oracle_dir, *target])


def build_all() -> None:
    build_hin()
    build_smcm()
    build_cli()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", LIBDIR)
