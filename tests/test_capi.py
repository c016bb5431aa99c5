"""The C-ABI library loads and exports every entry point include/lgx.h
declares; host-only entry points behave (no GPU needed).

Also pins the gather's exact division by the viewport extent: it must equal
the IEEE quotient for every float that can select a texel, for each extent
the configurations and tests use.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from legitengine_amd import lgx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "lgx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(lgx_[a-z0-9_]+)\s*\(", text)
    return sorted(set(n for n in names if not n.endswith("_t")))


def test_library_exports_every_declared_symbol():
    L = lgx.lib()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    # the ctypes binding covers the header
    assert set(names) <= set(lgx.SIGNATURES), set(names) - set(lgx.SIGNATURES)


def test_strerror_and_version():
    L = lgx.lib()
    assert L.lgx_version() >= 1
    for code in (lgx.OK, lgx.EINVAL, lgx.EFORMAT, lgx.ESIZE, lgx.EUNSUPPORTED, lgx.EWORKSPACE, lgx.ENODEVICE):
        assert L.lgx_strerror(code)
    assert L.lgx_strerror(-99) == b"unknown lgx error"


def test_scene_is_seeded_and_valid():
    a, b = lgx.make_scene(100, 32, 8), lgx.make_scene(100, 32, 8)
    assert bytes(a) == bytes(b)
    assert bytes(lgx.make_scene(101, 32, 8)) != bytes(a)
    for i in range(32):
        for k in range(3):
            ext = a.box_max[i][k] - a.box_min[i][k]
            assert 0.05 <= ext <= 0.3 + 1e-6 and 0 <= a.box_min[i][k] and a.box_max[i][k] <= 1 + 1e-6
    for i in range(8):
        assert 0.05 <= a.sph_radius[i] <= 0.15
    gb = lgx.render_gbuffer_cpu(a, 160, 90, "f32")
    d = gb["depth"]
    assert (d > 0).all() and (d <= np.sqrt(3) + 1e-5).all()  # inside the unit room, no background
    from oracle import binding as orc
    n = np.stack([np.array(orc_decode(e)) for e in gb["normal"].reshape(-1, 2)[::97]])
    assert np.allclose(np.linalg.norm(n, axis=1), 1, atol=1e-5)
    # FP16 storage is the round-to-nearest-even cast of the FP32 render
    gm = lgx.render_gbuffer_cpu(a, 160, 90, "mixed")
    assert np.array_equal(gm["radiance"], gb["radiance"].astype(np.float16))
    assert np.array_equal(gm["normal"], gb["normal"].astype(np.float16))
    assert np.array_equal(gm["depth"], gb["depth"])


def orc_decode(e):
    import ctypes as C
    from oracle import binding as orc
    out = (C.c_float * 3)()
    orc.lib().orc_oct_decode(float(e[0]), float(e[1]), out)
    return list(out)


def test_gather_workspace_and_argument_checks():
    import ctypes as C
    L = lgx.lib()
    p = lgx.GatherParams()
    p.width, p.height, p.dirs, p.max_steps, p.grid, p.sampler = 1920, 1080, 32, 64, 4, 1
    n32 = L.lgx_gather_workspace_bytes(C.byref(p))
    # 16 slices x (32 dirs x 4 + S steps x 2) floats + 11 level starts, with
    # S <= 64 the steps the reference rule can take on a 1080p diagonal
    rest = n32 - 16 * 32 * 4 * 4 - 16 * 11 * 4
    assert rest % (16 * 2 * 4) == 0 and 30 <= rest // (16 * 2 * 4) <= 64
    p.precision = lgx.PREC_F64
    assert L.lgx_gather_workspace_bytes(C.byref(p)) > n32
    p.dirs = lgx.LGX_MAX_LEVELS * 0 + 257
    assert L.lgx_gather_workspace_bytes(C.byref(p)) == 0
    assert L.lgx_gather(None, None, None, None, None, None) == lgx.EINVAL


def test_batch_argument_checks():
    """Host-side checks of the batch entry points: they fail before touching
    a device."""
    import ctypes as C
    L = lgx.lib()
    h = C.c_void_p()
    assert L.lgx_gather_batch_create(None, None, None, None, None, 0, C.byref(h)) == lgx.EINVAL
    assert L.lgx_batch_create(None, None, 0, None, C.byref(h)) == lgx.EINVAL
    assert L.lgx_gather_batch_launch(None, 0, None) == lgx.EINVAL
    assert L.lgx_batch_run(None, None) == lgx.EINVAL and L.lgx_batch_replay(None, None) == lgx.EINVAL
    assert L.lgx_gather_batch_blocks(None) == -1
    # descriptors that pass lgx_gather's checks (never dereferenced on the host)
    w, hh = 64, 32
    fake = 0x100000
    p = lgx.GatherParams()
    p.width, p.height, p.dirs, p.max_steps, p.grid, p.sampler = w, hh, 4, 16, 4, lgx.SAMPLER_HIZ_POINT
    g = lgx.GBuffer()
    g.depth = lgx.image(fake, w, hh, lgx.R32F)
    g.normal = lgx.image(fake, w, hh, lgx.RG16F)
    pyr = lgx.Pyramid()
    pyr.n_levels = 1
    pyr.depth[0] = lgx.image(fake, w, hh, lgx.R32F)
    pyr.radiance[0] = lgx.image(fake, w, hh, lgx.RGBA16F)
    out = lgx.image(fake, w, hh, lgx.RGBA32F)
    ws = (C.c_void_p * 1)(fake)
    swz = swizzling_tiles()
    for field, value, code in (("precision", lgx.PREC_F64, lgx.EFORMAT), ("tile", 99, lgx.EUNSUPPORTED),
                               ("dirs", 4, lgx.EWORKSPACE)):
        q = lgx.GatherParams()
        C.memmove(C.byref(q), C.byref(p), C.sizeof(q))
        setattr(q, field, value)
        rc = L.lgx_gather_batch_create(C.byref(q), ws, C.byref(g), C.byref(pyr), C.byref(out), 1, C.byref(h))
        assert rc == code and not h.value, (field, rc)
    # a workspace lgx_gather_prepare never filled is rejected before any launch
    assert L.lgx_gather(C.byref(p), C.c_void_p(fake), C.byref(g), C.byref(pyr), C.byref(out), None) == lgx.EWORKSPACE
    # the allocation-free gather keeps only tables in its array
    q = lgx.GatherParams()
    C.memmove(C.byref(q), C.byref(p), C.sizeof(q))
    q.tile = swz[0]
    assert L.lgx_gather_noalloc(C.byref(q), C.byref(g), C.byref(pyr), C.byref(out), None) == lgx.EUNSUPPORTED


def tile_info(t):
    import ctypes as C
    info = (C.c_int32 * 5)()
    assert lgx.lib().lgx_gather_tile_info(t, info) == lgx.OK
    return list(info)


def swizzling_tiles():
    return [t for t in range(lgx.lib().lgx_gather_tiles()) if tile_info(t)[4]]


def test_gather_tile_info():
    """Every launch shape is described; each swizzling shape has a twin with
    the same tile, registers and step loop that reads the levels as built."""
    import ctypes as C
    L = lgx.lib()
    n = L.lgx_gather_tiles()
    infos = [tile_info(t) for t in range(n)]
    assert infos[0] == [16, 16, 1, 0, 0]  # the default shape
    for w, h, occ, pipe, swz in infos:
        assert w * h == 256 and occ in (1, 8) and pipe in (0, 1) and swz in (0, 1)
    plain = [tuple(i[:4]) for i in infos if not i[4]]
    twins = [tuple(i[:4]) for i in infos if i[4]]
    assert len(set(plain)) == len(plain) and sorted(plain) == sorted(twins) and len(twins) > 0
    info = (C.c_int32 * 5)()
    assert L.lgx_gather_tile_info(n, info) == lgx.EINVAL and L.lgx_gather_tile_info(-1, info) == lgx.EINVAL
    assert L.lgx_gather_tile_info(0, None) == lgx.EINVAL


def test_swizzled_workspace_size():
    """A launch shape that swizzles adds room for copies of the levels the
    pattern's columns share (pattern 8: levels 0-2, 4 + 16 bytes a texel);
    others, and samplers or precisions that never swizzle, do not."""
    import ctypes as C
    L = lgx.lib()
    p = lgx.GatherParams()
    p.width, p.height, p.dirs, p.max_steps, p.grid, p.footprint_dirs = 1920, 1080, 4, 0, 8, 256
    p.sampler = lgx.SAMPLER_HIZ_POINT
    base = L.lgx_gather_workspace_bytes(C.byref(p))
    texels = sum((1920 >> l) * (1080 >> l) for l in range(3))
    swz = swizzling_tiles()
    for t in swz:
        p.tile = t
        extra = L.lgx_gather_workspace_bytes(C.byref(p)) - base
        assert 20 * texels <= extra <= 20 * texels + 7 * 256
    for field, value in (("sampler", lgx.SAMPLER_TRILINEAR), ("precision", lgx.PREC_F32_ACC64), ("grid", 1)):
        q = lgx.GatherParams()
        C.memmove(C.byref(q), C.byref(p), C.sizeof(q))
        setattr(q, field, value)
        b0 = L.lgx_gather_workspace_bytes(C.byref(q))
        q.tile = 0
        assert b0 == L.lgx_gather_workspace_bytes(C.byref(q)), field


@pytest.fixture(scope="module")
def const_div_checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("cdiv") / "check_const_div")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe,
                    os.path.join(ROOT, "tests", "csrc", "check_const_div.c"), "-lm"], check=True)
    return exe


@pytest.mark.parametrize("extents", [("1920", "1080"), ("256", "320", "192", "96", "64", "48", "128", "640", "360"),
                                     ("512", "1280", "720", "2560", "1440")])
def test_exact_extent_division(const_div_checker, extents):
    r = subprocess.run([const_div_checker, *extents], capture_output=True, text=True, timeout=300)
    assert r.stdout.strip() == "ok", r.stdout


def test_shared_denominator_division(tmp_path):
    """The gather's normalisations and unprojection divide three numerators by
    one denominator with one IEEE reciprocal and a Markstein correction per
    quotient; every significand of the denominator (with fixed and random
    numerators, signs and signed zeros) must give the IEEE quotient."""
    exe = str(tmp_path / "check_common_div")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe,
                    os.path.join(ROOT, "tests", "csrc", "check_common_div.c"), "-lm"], check=True)
    r = subprocess.run([exe, "32"], capture_output=True, text=True, timeout=300)
    assert r.stdout.startswith("ok "), r.stdout
    assert int(r.stdout.split()[1]) > 3 * 10 ** 8


@pytest.mark.skipif(os.environ.get("LGX_EXPECT_GPU") == "1", reason="checks behaviour without a device")
def test_renderer_create_fails_cleanly_without_device():
    """Failure handling: no HIP device -> an error code, no crash, no handle."""
    try:
        import torch
        if torch.cuda.device_count() > 0:
            pytest.skip("a device is present")
    except ImportError:
        pass
    import ctypes as C
    c = lgx.RendererConfig()
    c.width, c.height, c.dirs, c.max_steps, c.levels, c.grid, c.sampler = 64, 64, 16, 32, 5, 4, 1
    c.normal_format, c.albedo_format, c.radiance_format = lgx.RG16F, lgx.RGBA8, lgx.RGBA16F
    h = C.c_void_p()
    rc = lgx.lib().lgx_renderer_create(C.byref(c), C.byref(h))
    assert rc != lgx.OK and not h.value
