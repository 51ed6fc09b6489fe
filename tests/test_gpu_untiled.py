"""The per-particle (untiled) list and force kernels, the path for systems past the tiled layout's limit and the
one MDGPU_UNTILED=1 selects. The switch is read once per process, so the checks run in one child process
(tests/untiled_child.py: list and FP64 forces bit-exact against the oracle, mixed forces within 1e-6, the golden
deterministic trajectory hash, and a device-driven mixed md_run against FP64)."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_untiled_path_matches_oracle():
    env = dict(os.environ, MDGPU_UNTILED="1")
    out = subprocess.run([sys.executable, os.path.join(HERE, "untiled_child.py")], env=env, capture_output=True,
                         text=True, timeout=240)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-4000:]
    assert out.stdout.strip().endswith("OK")
