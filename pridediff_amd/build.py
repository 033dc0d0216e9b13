"""In-tree build of the native parts of pridediff_amd.

* ``libdd.so``      — the HIP kernels and the C ABI (include/dd.h, dd3d.h),
                      compiled by hipcc for gfx950 only;
* ``radontorch.so`` — the torch-op boundary (csrc/radontorch.cpp), a plain
                      C++ extension built with g++ against torch's headers and
                      linked to libdd.so.  No .cu sources and no hipify step.

Usage: ``python -m pridediff_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

HIP_SOURCES = ["dd2d.hip", "dd3d.hip", "dd_profile.hip", "ddbatch.hip"]
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=fast",
    # keep the compiler from emitting packed FP32 (v_pk_*_f32) instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if out:
        print(out, flush=True)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}")


def build_libdd() -> str:
    out = os.path.join(PKG, "libdd.so")
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    _run([os.path.join(ROCM, "bin", "hipcc"), *HIPCC_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}",
          "-o", out + ".tmp", *srcs])
    os.replace(out + ".tmp", out)
    return out


def build_radontorch() -> str:
    import torch
    from torch.utils import cpp_extension as ce

    out = os.path.join(PKG, "radontorch.so")
    torch_lib = os.path.join(os.path.dirname(torch.__file__), "lib")
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    incs = [*ce.include_paths(), sysconfig.get_paths()["include"], os.path.join(ROCM, "include"),
            INCLUDE, CSRC]
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-w",
           f"-D_GLIBCXX_USE_CXX11_ABI={abi}", "-DTORCH_EXTENSION_NAME=radontorch",
           "-DTORCH_API_INCLUDE_EXTENSION_H", "-DUSE_ROCM=1", "-D__HIP_PLATFORM_AMD__=1",
           *[f"-I{i}" for i in incs],
           os.path.join(CSRC, "radontorch.cpp"),
           f"-L{PKG}", "-ldd", f"-L{torch_lib}", "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu",
           "-ltorch_hip", "-ltorch_python",
           "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{torch_lib}",
           "-o", out + ".tmp"]
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_all() -> None:
    build_libdd()
    build_radontorch()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "all" in what:
        build_all()
    else:
        if "libdd" in what:
            build_libdd()
        if "radontorch" in what:
            build_radontorch()
