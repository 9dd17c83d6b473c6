"""Build recipe for libcubvh_amd.so (the C-ABI BVH query library, gfx950).

Compiles every source under cubvh_amd/csrc with hipcc into cubvh_amd/libcubvh_amd.so,
in-tree, and the `_cubvh` extension module (csrc/bindings.cpp) against torch and that
library.  Run ``python -m cubvh_amd.build`` or call :func:`build_library` and
:func:`build_binding`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
REPO_DIR = PKG_DIR.parent
CSRC = PKG_DIR / "csrc"
LIB_PATH = PKG_DIR / "libcubvh_amd.so"

SOURCES = ["capi.cpp", "build.cpp", "query.hip", "order.hip", "forest.hip", "floodfill.hip", "grad.hip", "gpubuild.hip", "hashtable.hip", "smc.hip", "probe.hip"]

ARCH = os.environ.get("CUBVH_AMD_ARCH", "gfx950")

# -ffp-contract=off: the only fused multiply-adds are the explicit ones in the sources,
# which makes results reproducible bit for bit.  Packed fp32 instructions are disabled.
COMMON_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Wall",
    "-Wno-unused-variable",
    "-Wno-unused-function",
]
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libcubvh_amd.so")


def build_library(verbose: bool = False, jobs: int = 4) -> Path:
    """Compile all sources and link libcubvh_amd.so.  Always rebuilds."""
    obj_dir = PKG_DIR / "_build"
    obj_dir.mkdir(exist_ok=True)
    cc = hipcc()
    inc = ["-I", str(REPO_DIR / "include"), "-I", str(CSRC)]
    cmds = []
    objs = []
    for src in SOURCES:
        obj = obj_dir / (src.replace(".", "_") + ".o")
        objs.append(obj)
        if src.endswith(".hip"):
            cmd = [cc, "-x", "hip", *HIP_FLAGS, *COMMON_FLAGS, *inc, "-c", str(CSRC / src), "-o", str(obj)]
        else:
            cmd = [cc, *COMMON_FLAGS, "-D__HIP_PLATFORM_AMD__", *inc, "-c", str(CSRC / src), "-o", str(obj)]
        cmds.append(cmd)
    procs = []
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
        if len([p for p in procs if p.poll() is None]) >= jobs:
            procs[-jobs].wait()
    failed = []
    for cmd, p in zip(cmds, procs):
        out, _ = p.communicate()
        out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l)
        if p.returncode != 0:
            failed.append((cmd, out))
        elif verbose and out.strip():
            print(out)
    if failed:
        msg = "\n\n".join(" ".join(c) + "\n" + o for c, o in failed)
        raise RuntimeError("libcubvh_amd build failed:\n" + msg)
    tmp = LIB_PATH.with_suffix(".so.tmp")
    link = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print(" ".join(link), flush=True)
    r = subprocess.run(link, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("libcubvh_amd link failed:\n" + r.stdout)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


BINDING_SRC = "bindings.cpp"


def binding_path() -> Path:
    """In-tree path of the `_cubvh` extension module (importable as cubvh_amd._cubvh,
    and as top-level `_cubvh` with cubvh_amd/ on sys.path)."""
    import sysconfig

    return PKG_DIR / ("_cubvh" + sysconfig.get_config_var("EXT_SUFFIX"))


def build_binding(verbose: bool = False) -> Path:
    """Compile the `_cubvh` PyTorch-ROCm extension (csrc/bindings.cpp: the reference
    package's native module names over the C ABI) against torch's headers and libraries
    and libcubvh_amd.so (found next to it through $ORIGIN).  Always rebuilds."""
    import torch
    from torch.utils import cpp_extension

    cc = hipcc()
    out = binding_path()
    tmp = out.with_name(out.name + ".tmp")
    torch_lib = cpp_extension.library_paths()[0]
    inc = [f"-I{p}" for p in cpp_extension.include_paths()]
    import sysconfig

    inc.append(f"-I{sysconfig.get_paths()['include']}")
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cmd = [cc, "-O2", "-std=c++17", "-fPIC", "-shared", "-D__HIP_PLATFORM_AMD__", "-DUSE_ROCM=1",
           "-DTORCH_EXTENSION_NAME=_cubvh", "-DTORCH_API_INCLUDE_EXTENSION_H", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
           "-Wno-unused-result", "-Wno-deprecated-declarations", *inc, "-I", str(REPO_DIR / "include"),
           str(CSRC / BINDING_SRC), "-o", str(tmp),
           f"-L{torch_lib}", "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-ltorch_hip", "-ltorch_python",
           f"-L{PKG_DIR}", "-lcubvh_amd", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{torch_lib}"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("_cubvh extension build failed:\n" + r.stdout)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    path = build_library(verbose="-v" in sys.argv)
    print(path)
    print(build_binding(verbose="-v" in sys.argv))
