"""Build recipes for the in-tree native code (gfx950 HIP library and the
Fortran host driver)."""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
FC = os.path.join(ROCM, "bin", "amdflang")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions (v_pk_*_f32) stay disabled: some of their
    # forms have given wrong results on these GPUs without any error
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, cwd=ROOT):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "not a recognized feature" not in l
                    and "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


HIP_SOURCES = ["nemo_gpu.hip", "stage_m3.hip", "stage_m4.hip", "stage_m5.hip", "batch.hip"]


def build_hip(verbose: bool = False) -> str:
    """Compile the HIP translation units in parallel and link libnemo_gpu.so."""
    from concurrent.futures import ThreadPoolExecutor
    out = os.path.join(HERE, "libnemo_gpu.so")
    bdir = os.path.join(HERE, "csrc", "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = [os.path.join(bdir, s.replace(".hip", ".o")) for s in HIP_SOURCES]

    def compile_one(k):
        return _run([HIPCC, *HIP_FLAGS, "-c", "-o", objs[k], os.path.join(HERE, "csrc", HIP_SOURCES[k])])

    with ThreadPoolExecutor(max_workers=len(HIP_SOURCES)) as ex:
        logs = list(ex.map(compile_one, range(len(HIP_SOURCES))))
    tmp = out + ".tmp"
    logs.append(_run([HIPCC, "--offload-arch=gfx950", "-shared", "-o", tmp, *objs]))
    os.replace(tmp, out)
    if verbose:
        print("\n".join(l for l in logs if l))
    return out


def build_fortran_driver() -> str:
    """Fortran host driver (nemo_amd/fortran) linked against libnemo_gpu.so."""
    fdir = os.path.join(HERE, "fortran")
    bdir = os.path.join(fdir, "_build")
    os.makedirs(bdir, exist_ok=True)
    exe = os.path.join(HERE, "nemo_gpu_driver")
    srcs = ["nemo_gpu_mod.f90", "nemo_setup_mod.f90", "nemo_driver.f90"]
    for s in srcs:
        _run([FC, "-O2", "-J", bdir, "-c", os.path.join(fdir, s), "-o",
              os.path.join(bdir, s.replace(".f90", ".o"))])
    objs = [os.path.join(bdir, s.replace(".f90", ".o")) for s in srcs]
    _run([FC, "-o", exe, *objs, "-L" + HERE, "-lnemo_gpu", "-Wl,-rpath,$ORIGIN",
          "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    return exe


HDF5_PREFIX = os.environ.get("NEMO_HDF5_PREFIX", "/opt/conda")


def build_flash() -> str | None:
    """Host FLASH checkpoint writer libnemo_flash.so (needs the HDF5 C
    library; skipped with a message where it is absent)."""
    inc = os.path.join(HDF5_PREFIX, "include")
    libdir = os.path.join(HDF5_PREFIX, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        print(f"HDF5 headers not found under {HDF5_PREFIX}: FLASH checkpoint writer not built")
        return None
    out = os.path.join(HERE, "libnemo_flash.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I" + inc, "-o", tmp,
          os.path.join(HERE, "csrc", "flash_checkpoint.cpp"), "-L" + libdir, "-Wl,-rpath," + libdir, "-lhdf5"])
    os.replace(tmp, out)
    return out


def build_oracle() -> None:
    _run(["make", "-C", os.path.join(ROOT, "oracle")])
    if os.path.isdir("/root/reference/source"):
        _run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])
