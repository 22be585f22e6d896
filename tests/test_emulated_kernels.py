"""The HIP kernels compiled for the host against a CPU emulation of the HIP subset
they use (tests/emu: wavefronts in lockstep, AddressSanitizer on), run on small
grids and compared with the float64 oracle: every index the kernels form is
checked and the values match the oracle to 1e-5, without a GPU."""

import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EMU = os.path.join(HERE, "emu")
CSRC = os.path.join(ROOT, "witwin_maxwell_amd", "csrc")


def _asan_runtime():
    try:
        out = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True, check=True)
    except (OSError, subprocess.CalledProcessError):
        return None
    path = out.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.fixture(scope="module")
def emulated_library(tmp_path_factory):
    asan = _asan_runtime()
    if asan is None:
        pytest.skip("g++ with AddressSanitizer is not available")
    out = tmp_path_factory.mktemp("emu")
    src = out / "maxwell_fdtd_emu.cpp"
    src.write_text(open(os.path.join(CSRC, "maxwell_fdtd.hip")).read())
    lib = out / "libmaxwell_fdtd_emu.so"
    cmd = ["g++", "-std=c++20", "-O1", "-g", "-fPIC", "-shared", "-pthread", "-fsanitize=address", "-w",
           "-I", EMU, "-I", os.path.join(ROOT, "include"), "-I", CSRC, str(src),
           os.path.join(EMU, "emu_globals.cpp"), "-o", str(lib)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr[-4000:]
    return str(lib), asan


def test_single_pass_step_matches_oracle_under_emulation(emulated_library):
    """The case groups of run_case.py, each in its own process, side by side."""
    lib, asan = emulated_library
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0", OMP_NUM_THREADS="1")
    groups = ("boxes", "staged", "dispersive", "periodic", "adjoint")
    procs = {g: subprocess.Popen([sys.executable, os.path.join(EMU, "run_case.py"), lib, g], stdout=subprocess.PIPE,
                                 stderr=subprocess.PIPE, text=True, env=env) for g in groups}
    failed = []
    for g, proc in procs.items():
        out, err = proc.communicate(timeout=900)
        print(out)
        if proc.returncode != 0 or not out.strip().endswith("OK"):
            failed.append((g, out[-3000:], err[-3000:]))
    assert not failed, failed


def test_reference_named_ops_under_emulation(emulated_library):
    """Every op of the NativeFDTDModule interface through the emulated library:
    argument structs built from the reference's keywords index only inside their
    arrays, and each op matches the oracle."""
    lib, asan = emulated_library
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0")
    proc = subprocess.run([sys.executable, os.path.join(EMU, "run_shim.py"), lib], capture_output=True, text=True,
                          env=env, timeout=900)
    print(proc.stdout)
    assert proc.returncode == 0, (proc.stdout[-3000:], proc.stderr[-3000:])
    assert proc.stdout.strip().endswith("OK")
