"""The C++ consumer of the boundary headers (tests/csrc/integration_example.cpp,
built by build()): LegitEngine's frame loop over a BaseRenderer-shaped
SSVGIRenderer and the reference-signature builders, bit-identical to the C
ABI renderer on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "csrc", "integration_example")


def test_integration_example_bit_identical():
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is missing: run __graft_entry__.build()")
    print("running the C++ integration example", flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    assert "bit-identical to lgx_renderer_render" in r.stdout
    assert "match the renderer's 4 pyramid levels" in r.stdout
