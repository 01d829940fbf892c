"""In-tree build of the native parts of hastycompute_amd.

- ``libhasty_hip.so``: the gfx950 HIP kernels and the C ABI of include/hasty_hip.h,
  compiled with hipcc (one object per source, in parallel) into hastycompute_amd/lib/.
- ``_hasty_torch.so``: the C++ operator layer (hasty::fft / mri / op / svt with the
  reference's class names) and its TorchScript registrations, linked against
  libhasty_hip.so; built with torch.utils.cpp_extension into hastycompute_amd/lib/.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_DIR = os.path.join(PKG, "lib")
HIP_SRC = os.path.join(PKG, "csrc", "hip")
HOST_SRC = os.path.join(PKG, "csrc", "host")
BUILD_DIR = os.path.join(ROOT, "_build")
HIP_LIB = os.path.join(LIB_DIR, "libhasty_hip.so")
TORCH_LIB = os.path.join(LIB_DIR, "_hasty_torch.so")

ARCH = os.environ.get("HASTY_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-munsafe-fp-atomics",
    # packed FP32 VALU ops are kept out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + out)
    return out


def hip_sources():
    return sorted(os.path.join(HIP_SRC, f) for f in os.listdir(HIP_SRC) if f.endswith(".hip"))


def build_hip(jobs: int = 8, verbose: bool = False) -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    obj_dir = os.path.join(BUILD_DIR, "hip")
    os.makedirs(obj_dir, exist_ok=True)
    srcs = hip_sources()
    objs = [os.path.join(obj_dir, os.path.basename(s)[:-4] + ".o") for s in srcs]

    def comp(pair):
        src, obj = pair
        return _run([HIPCC, *HIP_FLAGS, "-I", HIP_SRC, "-c", src, "-o", obj])

    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for out in ex.map(comp, zip(srcs, objs)):
            if verbose and out.strip():
                print(out)
    tmp = HIP_LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-soname,libhasty_hip.so", "-o", tmp, *objs])
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


def build_torch(verbose: bool = False) -> str:
    import ctypes

    import torch
    from torch.utils.cpp_extension import load

    # the loader opens the fresh build from its build folder, where $ORIGIN does not
    # reach libhasty_hip.so; a global preload (matched by soname) satisfies it
    ctypes.CDLL(HIP_LIB, mode=ctypes.RTLD_GLOBAL)
    os.makedirs(LIB_DIR, exist_ok=True)
    bdir = os.path.join(BUILD_DIR, "torch_ext")
    os.makedirs(bdir, exist_ok=True)
    srcs = sorted(os.path.join(HOST_SRC, f) for f in os.listdir(HOST_SRC) if f.endswith(".cpp"))
    torch_lib = os.path.join(os.path.dirname(torch.__file__), "lib")
    load(
        name="_hasty_torch",
        sources=srcs,
        extra_cflags=["-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1", "-I/opt/rocm/include",
                      "-I" + os.path.join(ROOT, "include"), "-I" + HOST_SRC],
        # "$$" is ninja's escape and "\$" the shell's: the linker sees -rpath,$ORIGIN
        extra_ldflags=["-L" + LIB_DIR, "-lhasty_hip", "-Wl,-rpath,\\$$ORIGIN", "-L" + torch_lib, "-lc10_hip",
                       "-ltorch_hip"],
        build_directory=bdir,
        is_python_module=False,
        verbose=verbose,
    )
    built = os.path.join(bdir, "_hasty_torch.so")
    tmp = TORCH_LIB + ".tmp"
    shutil.copyfile(built, tmp)
    os.replace(tmp, TORCH_LIB)
    return TORCH_LIB


def build_all(jobs: int = 8, verbose: bool = False):
    build_hip(jobs, verbose)
    if os.path.isdir(HOST_SRC) and any(f.endswith(".cpp") for f in os.listdir(HOST_SRC)):
        build_torch(verbose)


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    if what == "hip":
        print(build_hip(verbose=True))
    elif what == "torch":
        print(build_torch(verbose=True))
    else:
        build_all(verbose=True)
