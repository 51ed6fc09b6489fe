"""The reference's unit-test suite restated in C++ against the GPU-backed facade classes
(molecular-dynamics_amd/facade/tests/test_facade.cpp), run as one process on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "molecular-dynamics_amd", "facade", "md_facade_tests")


@pytest.mark.gpu
def test_facade_suite_passes():
    assert os.path.exists(BIN), "facade tests not built: run __graft_entry__.build()"
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=900)
    print(p.stdout[-6000:])
    assert p.returncode == 0, p.stdout[-4000:] + p.stderr[-2000:]
    assert "37 tests, 0 failed" in p.stdout or ", 0 failed" in p.stdout


def test_facade_library_loads_without_a_device():
    import ctypes

    lib = os.path.join(ROOT, "molecular-dynamics_amd", "facade", "libmdfacade.so")
    assert os.path.exists(lib)
    ctypes.CDLL(lib)


def test_xyz_formatter_matches_printf():
    """The XYZ line formatter of the trajectory writers against printf's "%.8f" / "%zu" on 18 M values (ties,
    carries, signed zeros, subnormals, random bit patterns) and whole lines (facade/tests/test_format.cpp)."""
    exe = os.path.join(ROOT, "molecular-dynamics_amd", "facade", "md_format_tests")
    assert os.path.exists(exe), "md_format_tests not built: run __graft_entry__.build()"
    p = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-3000:]
    assert " 0 mismatches" in p.stdout
