"""GPU test of the library loader: torch must keep working in a process that
called the library before importing torch. The PyTorch ROCm wheel ships its
own HIP and HSA runtimes, and the HSA runtime shares its soname with the one
librwig_hip.so links, so the first loaded serves both; rwig_amd._lib imports
torch (where installed) before loading the library. A fresh child process
is needed, since this test process has imported torch already."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import numpy as np
import rwig_amd
a = np.full(40, 1 / 40); b = np.full(30, 1 / 30)
x = np.linspace(0, 1, 40)[:, None]; y = np.linspace(0, 1, 30)[None, :]
r = rwig_amd.sinkhorn(a, b, (x - y) ** 2, {"reg": 0.1})
assert r["iter"] > 0
import torch
t = torch.arange(8, dtype=torch.float64, device="cuda")
assert float(t.sum()) == 28.0
print("library first, then torch: ok", flush=True)
"""


def test_torch_works_after_the_library_was_used_first():
    print("child: library call, then torch on the GPU", flush=True)
    out = subprocess.run([sys.executable, "-c", CHILD], cwd=ROOT, capture_output=True,
                         text=True, timeout=240)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "ok" in out.stdout
