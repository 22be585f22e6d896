"""Build the native FDTD library for gfx950 with hipcc (no JIT, no torch extension).

Produces ``witwin_maxwell_amd/_lib/libmaxwell_fdtd.so`` exporting the C ABI of
``include/maxwell_fdtd.h``.  The library links the HIP runtime that torch ships
(rpath to torch/lib) so that streams and device pointers handed over from torch
belong to the same runtime instance.
"""

from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "_lib")
LIBNAME = "libmaxwell_fdtd.so"
SOURCES = ["maxwell_fdtd.hip"]
HEADERS = ["yee.h", "terms.h", "fused.h", "shell.h", "zedge.h", "adjoint.h", "batch.h"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=fast",
    # keep packed-FP32 (v_pk_*_f32) instructions out of the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _torch_libdir():
    try:
        import torch
        return os.path.join(os.path.dirname(torch.__file__), "lib")
    except Exception:  # pragma: no cover - torch is part of the image
        return None


def library_path():
    return os.path.join(LIBDIR, LIBNAME)


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    out = library_path()
    tmp = out + ".tmp"
    cmd = ["hipcc"] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    tl = _torch_libdir()
    if tl and os.path.exists(os.path.join(tl, "libamdhip64.so")):
        cmd += ["-L", tl, f"-Wl,-rpath,{tl}"]
    cmd += [os.path.join(CSRC, s) for s in SOURCES] + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed building {LIBNAME}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
