"""The integrand's sphere groups (scenes.hpp: build_accel, group_may_hit and
the pruned loops of trace / occluded) give the same values as the plain loops
over every sphere, bit for bit, CPU only.

The pruning is product code (the GPU integrand kernels evaluate every scene
through it); compiled into the CPU library it is held here against the plain
loops, which the geometric pins in test_oracle_scene.py fix.  Checked: the
three sphere scenes over seeds and both image sizes, in float and double,
coordinates on the shutter and lens ends, and hand-built scenes (one sphere,
a sphere of radius 1e4, a ray origin far out on the plane).  A bound that
shrinks the group spheres by 0.05 fails this test (checked once by hand);
the rounding margins below that are fixed by the bound stated in scenes.hpp,
not by sampling.
"""
import numpy as np
import pytest

from test_oracle_scene import F_EYE, F_PLANE, Scene


def _same(a, b):
    u = np.uint64 if a.dtype == np.float64 else np.uint32
    bad = np.argwhere(a.view(u) != b.view(u))
    assert len(bad) == 0, (len(bad), bad[:3].tolist())


def _coords(rng, n, dim, W, H):
    c = rng.random((n, dim), dtype=np.float32)
    c[:, 0] *= np.float32(W)
    c[:, 1] *= np.float32(H)
    ends = np.array([0.0, 0.5, 0.99999994, 1.0], np.float32)
    c[: n // 20, 2:] = rng.choice(ends, (n // 20, dim - 2))
    return c


@pytest.mark.parametrize("kind,dim", [(2, 4), (3, 5), (4, 6)])
@pytest.mark.parametrize("seed", [0, 1, 7])
def test_grouped_equals_plain(orc, kind, dim, seed):
    W, H = (1920.0, 1080.0) if seed != 1 else (3840.0, 2160.0)
    params = orc.make_scene(kind, seed, dim, W, H)
    rng = np.random.default_rng(100 * kind + seed)
    c = _coords(rng, 200000, dim, W, H)
    _same(orc.eval_scene(params, c, dim), orc.eval_scene_grouped(params, c, dim))
    c64 = c[:20000].astype(np.float64)
    _same(orc.eval_scene_f64(params, c64, dim), orc.eval_scene_grouped(params, c64, dim, fp64=True))


def test_grouped_hand_built_scenes(orc):
    rng = np.random.default_rng(5)
    # one sphere on the axis, moving
    s = Scene(orc, 3, 5)
    s.objects([[0.0, 0.0, 6.0, 1.0, 0.3, -0.2, 0.1, 0.2, 0.4, 0.6]])
    c = _coords(rng, 50000, 5, 1.0, 1.0)
    _same(s.eval32(c), orc.eval_scene_grouped(bytes(s.buf), c, 5))
    # a sphere of radius 1e4 as a ceiling over the AO plane, and a few small ones
    s = Scene(orc, 4, 6)
    plane = float(s.f[F_PLANE])
    s.f[F_EYE:F_EYE + 3] = (0.0, plane + 0.2, 0.0)
    rows = [[0.0, plane + 0.5 + 1e4, 0.0, 1e4, 0, 0, 0, 1, 1, 1]]
    rows += [[float(x), plane + 0.3, float(z), 0.2, 0, 0, 0, 1, 1, 1]
             for x, z in rng.uniform(-3, 3, (9, 2))]
    s.objects(rows)
    c = _coords(rng, 50000, 6, 1.0, 1.0)
    _same(s.eval32(c), orc.eval_scene_grouped(bytes(s.buf), c, 6))
    # camera low over the plane: shadow and AO rays start far out on it
    s = Scene(orc, 3, 5, seed=3)
    s.f[F_EYE:F_EYE + 3] = (0.0, plane + 0.01, -50.0)
    c = _coords(rng, 50000, 5, 1.0, 1.0)
    c[:, 1] *= np.float32(0.5)  # lower half of the image: the plane, far away
    _same(s.eval32(c), orc.eval_scene_grouped(bytes(s.buf), c, 5))
