"""The Session facade (include/yk_session.hpp) on the GPU: the reference's
DLL scenarios ported in tests/session_check.cpp (forward projection with each
projector and with host / device buffers, execution validation, FDK offline
vs 13 + 19-view streaming, the streaming state machine and completion-time
host output, reset / release / reuse, the forward-projection pose, the
DLL Shepp-Logan FP -> FDK quality gates with 47-view streaming) and the
iterative pipelines (SIRT smoke and reset / reuse, the SART / SIRT / OSSART /
CGLS mapping onto the C-ABI solvers, buffer and request rules), the
flat-frame tail (external calibrated frames: tail checks, invariance of
SART / SIRT / OSSART / CGLS / FDK under nominal frames, the forward
projection of calibrated frames, an older SSystemSpec) and per-iteration
progress (example/multispectrum_sim/tests/recon/backend/
test_reconstruction_progress.cpp)."""
import subprocess

import pytest

from fdk_cuda_impl_amd.build import SESSION_CHECK

pytestmark = pytest.mark.gpu


def test_session_scenarios():
    res = subprocess.run([SESSION_CHECK], capture_output=True, text=True, timeout=240)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "FAIL" not in res.stdout
    assert res.stdout.count("PASS") == 26
