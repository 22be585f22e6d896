"""The reference-named per-op interface (fdtd/native_module.py) on the CPU
emulation of the HIP library, under AddressSanitizer: the same cases as
tests/test_gpu_native_module.py, run on host tensors and compared with the
float64 oracle.  Run by tests/test_emulated_kernels.py in a child process.

usage: python run_shim.py LIB"""
import ctypes
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import torch  # noqa: E402

from witwin_maxwell_amd.fdtd import native as nat  # noqa: E402

lib = nat.load_library(sys.argv[1])
nat.load_library = lambda path=None: lib
nat.ptr = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
nat.current_stream = lambda: ctypes.c_void_p(None)
torch.cuda.synchronize = lambda *a, **k: None

from witwin_maxwell_amd.fdtd import native_module as nm  # noqa: E402


def _host_f32(t, name):
    if not isinstance(t, torch.Tensor) or t.dtype != torch.float32:
        raise TypeError(f"{name} must be a float32 tensor")
    return t.contiguous()


nm._f32 = _host_f32
nm.get_native_fdtd_module = lambda: nm._NATIVE_MODULE

import test_gpu_native_module as t  # noqa: E402

t.DEVICE = "cpu"
for mode in ("standard", "compressed", "dense"):
    t.test_curl_updates_match_oracle(mode)
    print(f"curl updates {mode}: ok", flush=True)
t.test_pole_and_polarisation_ops_match_oracle()
print("pole currents: ok", flush=True)
t.test_source_monitor_and_line_ops_match_reference_sums()
print("sources, monitors, line: ok", flush=True)
t.test_unknown_op_is_refused()
print("OK")
