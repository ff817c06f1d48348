"""Builds the native parts of dqsim_bosonic_amd in-tree.

* ``libdqsim_gpu.so``: the HIP engine (csrc/dq_engine.hip) for gfx950, exporting
  the C ABI of include/dqsim_gpu.h.
* ``_core*.so``: the C++ host (pybind11) that mirrors the reference's Python API
  and calls the engine through that C ABI.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ARCH = os.environ.get("DQ_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off keeps every complex product/sum rounded as in the
# reference (no FMA) in the exact objects; packed fp32 ops are disabled for
# every HIP object.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, flush=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def engine_path() -> str:
    return os.path.join(HERE, "libdqsim_gpu.so")


def core_path() -> str:
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(HERE, "_core" + suffix)


# (source, contraction): the exact kernels keep every product and sum rounded
# separately; the contracted ones may fuse them into FMAs.
ENGINE_SOURCES = (("dq_engine.hip", "off"), ("dq_tile_f64.hip", "off"), ("dq_tile_f64_fma.hip", "fast"),
                  ("dq_tile_f32.hip", "fast"), ("dq_tile_f32_mfma.hip", "fast"), ("dq_batch.hip", "off"),
                  ("dq_shots.hip", "off"), ("dq_mps.hip", "fast"),
                  ("dq_adjoint.hip", "fast"))


def build_engine() -> str:
    """Compiles the HIP translation units in parallel, then links them."""
    from concurrent.futures import ThreadPoolExecutor

    out = engine_path()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    jobs = []
    for src, contract in ENGINE_SOURCES:
        flags = [f if f != "-ffp-contract=off" else f"-ffp-contract={contract}" for f in HIP_FLAGS if f != "-shared"]
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        jobs.append(([HIPCC, *flags, "-c", os.path.join(HERE, "csrc", src), "-o", obj], obj))
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        list(ex.map(lambda j: _run(j[0]), jobs))
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *[o for _, o in jobs], "-o", out])
    return out


def build_core() -> str:
    import pybind11

    out = core_path()
    inc = [f"-I{sysconfig.get_paths()['include']}", f"-I{pybind11.get_include()}"]
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fvisibility=hidden", *inc,
          os.path.join(HERE, "csrc", "host", "core.cpp"), "-o", out,
          f"-L{HERE}", "-ldqsim_gpu", "-Wl,-rpath,$ORIGIN"])
    return out


def build_all() -> None:
    build_engine()
    build_core()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"engine": build_engine, "core": build_core, "all": build_all}[w]()
