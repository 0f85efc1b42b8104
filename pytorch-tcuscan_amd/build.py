"""Builds the tcuscan MI355X libraries in-tree.

1. every csrc/*.hip source -> object with hipcc for gfx950, linked into
   libtcuscan_hip.so (the C ABI declared in include/tcuscan.h);
2. every csrc/*.cpp source (tcuscan_ops.cpp, serve.cpp) -> the `tcuscan_ops`
   Python extension module, a C++ layer over that C ABI, linked against
   torch and libtcuscan_hip.so.

Both land next to this file, so putting this directory on sys.path makes
`import tcuscan_ops` and `import tcuscan` work.  Usage: python build.py [-j N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # Packed FP32 VALU forms are kept out of every kernel on this target.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{os.path.join(REPO, 'include')}",
]

LIB_NAME = "libtcuscan_hip.so"


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "unknown target feature" not in l
                    and "packed-fp32-ops" not in l and l.strip() != "1 warning generated." )
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def hip_sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def ext_sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".cpp"))


def _ext_flags():
    import torch
    from torch.utils import cpp_extension as ce

    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    inc = [f"-I{p}" for p in ce.include_paths()] + [
        f"-I{sysconfig.get_paths()['include']}", f"-I{os.path.join(ROCM, 'include')}",
        f"-I{os.path.join(REPO, 'include')}",
    ]
    cflags = ["-O2", "-fPIC", "-std=c++17", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
              "-DTORCH_EXTENSION_NAME=tcuscan_ops", "-DTORCH_API_INCLUDE_EXTENSION_H",
              "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1", "-Wno-deprecated-declarations", *inc]
    libdirs = ce.library_paths()
    ldflags = [*[f"-L{d}" for d in libdirs], f"-L{HERE}",
               "-lc10", "-ltorch", "-ltorch_cpu", "-ltorch_python", "-lc10_hip", "-ltorch_hip",
               "-ltcuscan_hip", "-Wl,-rpath,$ORIGIN", *[f"-Wl,-rpath,{d}" for d in libdirs]]
    return cflags, ldflags


def build(jobs: int | None = None) -> None:
    """Compiles every HIP and C++ object in one parallel pool, then links
    libtcuscan_hip.so and the extension (which links against it)."""
    jobs = jobs or min(8, os.cpu_count() or 4)
    os.makedirs(BUILD, exist_ok=True)
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    cflags, ldflags = _ext_flags()
    hip_objs = [os.path.join(BUILD, os.path.basename(s) + ".o") for s in hip_sources()]
    ext_objs = [os.path.join(BUILD, os.path.basename(s) + ".o") for s in ext_sources()]
    cmds = [[hipcc, *HIPCC_FLAGS, "-c", s, "-o", o] for s, o in zip(hip_sources(), hip_objs)]
    cmds += [["g++", *cflags, "-c", s, "-o", o] for s, o in zip(ext_sources(), ext_objs)]
    # Longest first: scan sources and the torch-heavy C++ files dominate.
    cmds.sort(key=lambda c: -os.path.getsize(c[c.index("-c") + 1]))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    lib = os.path.join(HERE, LIB_NAME)
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", lib + ".tmp", *hip_objs,
          f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}"])
    os.replace(lib + ".tmp", lib)
    ext = os.path.join(HERE, "tcuscan_ops" + sysconfig.get_config_var("EXT_SUFFIX"))
    _run(["g++", "-shared", "-o", ext + ".tmp", *ext_objs, *ldflags])
    os.replace(ext + ".tmp", ext)


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--clean", action="store_true")
    a = ap.parse_args()
    if a.clean:
        clean()
    build(a.j)
    print("built", LIB_NAME, "and tcuscan_ops")
    sys.exit(0)
