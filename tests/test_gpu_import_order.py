"""A context created before any PyTorch call works: the package loads
PyTorch's HIP runtime before its own library, so the process holds a single
runtime (regression test: a runtime initialized by the library first left
torch.cuda reporting no devices).  Runs in a fresh interpreter, since this
test process has long since initialized PyTorch."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = """
import fdk_cuda_impl_amd as yk
from fdk_cuda_impl_amd import phantoms
g = yk.ScanGeometry(16, 16, 150.0, 300.0, 32, 24, 0.8, 0.8, volume=(16, 16, 8), voxel_mm=(0.6, 0.6, 0.6))
ctx = yk.Context(0)
vol = ctx.phantom(phantoms.shepp_logan(4.0), g.grid())
assert float(vol.abs().max()) > 0
ctx.close()
print("context before torch: ok")
"""


def test_context_before_any_torch_call():
    print("fresh interpreter: Context(0) as the first GPU call", flush=True)
    r = subprocess.run([sys.executable, "-c", SCRIPT], cwd=REPO, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "context before torch: ok" in r.stdout
