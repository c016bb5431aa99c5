"""Host-side logic that needs no GPU: the bench's equal-error interpolation,
the U-Net FLOP model, the learning-rate schedule and the uneven-batch
resolution draw."""
import math
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

torch = pytest.importorskip("torch")


def test_equal_error_interpolation():
    import bench
    # a power law ms = c * err^-2 is reproduced exactly by log-log interpolation
    pts = [(e, 3.0 * e ** -2.0) for e in (0.2, 0.1, 0.05, 0.025)]
    for e in (0.15, 0.07, 0.03):
        assert math.isclose(bench.equal_error_ms(e, pts), 3.0 * e ** -2.0, rel_tol=1e-12)
    assert bench.equal_error_ms(0.1, pts) == pytest.approx(300.0)
    # outside the sweep: no extrapolation
    assert bench.equal_error_ms(0.3, pts) is None
    assert bench.equal_error_ms(0.01, pts) is None


def test_unet_flop_model():
    from legitengine_amd import denoiser as dn
    # 3 levels of 32/64/128 channels, two 3x3 convolutions per level, concatenated skips,
    # 1x1 head: 113,280 multiply-adds per full-resolution pixel
    assert dn.unet_flops_per_pixel() == 2 * 113280
    assert sum(ci * co * 9 + co for ci, co, _ in dn.UNET_CONVS) + 32 * 3 + 3 == 473347


def test_lr_schedule():
    from legitengine_amd import denoiser as dn
    assert dn.lr_at(0, 100, 1e-3, "cosine") == pytest.approx(1e-3)
    assert dn.lr_at(99, 100, 1e-3, "cosine") == pytest.approx(5e-5)
    assert dn.lr_at(37, 100, 1e-3, "const") == 1e-3
    vals = [dn.lr_at(i, 100, 1e-3, "cosine") for i in range(100)]
    assert all(a >= b for a, b in zip(vals, vals[1:]))


def test_batch_resolutions_are_seeded():
    import bench
    sizes = [bench.batch_resolution(s) for s in bench.BATCH_SEEDS]
    assert sizes == [bench.batch_resolution(s) for s in bench.BATCH_SEEDS]
    assert set(sizes) <= set(bench.BATCH_SIZES) and len(set(sizes)) == 4
    assert all(w % 4 == 0 and h % 4 == 0 for w, h in sizes)  # the U-Net's two 2x poolings


def test_integration_doc_is_the_compiled_example():
    """INTEGRATION.md section 1's C++ is the code between the snippet markers
    of tests/csrc/integration_example.cpp, which build() compiles."""
    import os
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = open(os.path.join(root, "tests", "csrc", "integration_example.cpp")).read()
    snippet = src.split("// [integration-snippet-begin]\n", 1)[1].split("    // [integration-snippet-end]", 1)[0]
    doc = open(os.path.join(root, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", doc, flags=re.S)
    dedent = "\n".join(line[4:] if line.startswith("    ") else line for line in snippet.splitlines()) + "\n"
    assert any(b == dedent for b in blocks), "INTEGRATION.md does not hold the compiled snippet"
