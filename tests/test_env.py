"""The .env front-end (SDK/sutil/Environment.cpp:593-792) on the reference's
shipped configuration files and on malformed input (CPU only)."""
import ctypes
import os

import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _read(path):
    from pmdas_amd import _lib
    lib = _lib.load()
    env = _lib.MdasEnv()
    rc = lib.mdas_env_read(path.encode(), ctypes.byref(env))
    return rc, env, lib.mdas_env_last_error().decode()


def test_pool_env():
    rc, e, _ = _read(os.path.join(GOLDEN, "pool.env"))
    assert rc == 0
    assert e.mdas == 1 and e.samples == 4.0
    assert e.scale_factor == pytest.approx(0.0625) and e.alpha == 0.25
    assert e.bits_per_dim == 1 and e.extra_img_bits == 8 and e.candidates_num == 4
    assert (e.width, e.height) == (1024, 1024)


def test_chess_env():
    rc, e, _ = _read(os.path.join(GOLDEN, "chess.env"))
    assert rc == 0
    assert e.samples == 8.0 and e.scale_factor == 1.0
    assert (e.width, e.height) == (768, 1024)
    assert e.scene_kind == -1 and e.dim == 3


def test_cornell_env_and_config():
    from pmdas_amd import _lib
    rc, e, _ = _read(os.path.join(GOLDEN, "cornell-box.env"))
    assert rc == 0
    assert e.alpha == 0.0625 and e.bits_per_dim == 1 and e.extra_img_bits == 7
    cfg = _lib.MdasConfig()
    assert _lib.load().mdas_env_config(ctypes.byref(e), 30000000, ctypes.byref(cfg)) == 0
    assert (cfg.scale_x, cfg.scale_y) == (1024.0, 1024.0)
    assert cfg.max_samples == 30000000 and cfg.dim == 3


def test_defaults_when_absent(tmp_path):
    p = tmp_path / "min.env"
    p.write_text("Sampler {\n  mdas true\n}\n")
    rc, e, _ = _read(str(p))
    assert rc == 0
    # optixMotionBlur.cpp:118-125 registered defaults
    assert (e.samples, e.scale_factor, e.alpha) == (1.0, 1.0, 0.25)
    assert (e.bits_per_dim, e.extra_img_bits, e.candidates_num) == (0, 8, 4)
    assert (e.width, e.height) == (1024, 768)


def test_first_value_wins_and_scene_block(tmp_path):
    p = tmp_path / "s.env"
    p.write_text("Mdas { alpha 0.5 alpha 0.125 } # comment\nScene { kind mbdof seed 7 }\n")
    rc, e, _ = _read(str(p))
    assert rc == 0 and e.alpha == 0.5
    assert e.scene_kind == 3 and e.scene_seed == 7 and e.dim == 5


@pytest.mark.parametrize("text,needle", [
    ("Mdas {\n  alphaa 0.5\n}\n", "line 2"),
    ("}\n", "unpaired"),
    ("Mdas { bitsPerDim x }\n", "bad value"),
    ("Sampler { mdas maybe }\n", "bad value"),
    ("Mdas { alpha }\n", "Mdas.alpha"),
    ("Mdas { alpha", "missing value"),
    ("Mdas { alpha 0.5\n", "unterminated"),
    ("Scene { kind teapot }\n", "Scene.kind"),
])
def test_malformed(tmp_path, text, needle):
    p = tmp_path / "bad.env"
    p.write_text(text)
    rc, _, msg = _read(str(p))
    assert rc == -1
    assert needle in msg


def test_missing_file():
    rc, _, msg = _read("/nonexistent/x.env")
    assert rc == -1 and "cannot open" in msg
