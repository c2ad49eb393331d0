"""Pins of the analytic integrands (the stand-ins for the reference's OptiX
radiance programs, SDK/cuda/whitted.cu:164-293 and SDK/cuda/ao.cu:112-152)
by hand-built scenes whose answers geometry fixes, CPU only.

- parameter block: the field offsets, read back from a generated scene;
- disks: a moving disk covers a point over the t-interval where
  |p - c0 - v t| <= r (a quadratic), the highest-index disk is on top;
- spheres: a centre ray hits a sphere on the axis at distance z - r with
  normal -z, so the Lambertian value is albedo (ambient + n.l light); a
  downward ray shades the plane with n = +y; a sphere on the light ray
  leaves only the ambient term; a miss gives the sky gradient at d.y = 0;
- thin lens: every lens sample of a pixel passes through the point of the
  chief ray at z = focal distance, and the lens disk has radius lens_radius
  with uniform area density (a sphere halfway to focus is hit by the lens
  samples with |l| < 2 rho);
- motion: the (x, y, t, u, v) scene moves spheres by v t with t = x[2];
- ambient occlusion: a plane point under a parallel ceiling at height h
  sees, with cosine-weighted directions and occlusion radius R, the
  unoccluded fraction (h / R)^2; with nothing above it sees 1.
"""
import numpy as np
import pytest

# float32 indices into the parameter block
F_SX, F_SY, F_EYE, F_TANX, F_TANY, F_LENS, F_FOCAL = 4, 5, 6, 9, 10, 11, 12
F_LDIR, F_LCOL, F_AMB, F_PLANE, F_ALB, F_AO, F_CONST, F_OBJ = 13, 16, 19, 20, 21, 24, 25, 33
STRIDE = 12


class Scene:
    """A writable copy of a generated scene's parameter block."""

    def __init__(self, orc, kind, dim, sx=1.0, sy=1.0, seed=0):
        self.buf = bytearray(orc.make_scene(kind, seed, dim, sx, sy))
        self.f = np.frombuffer(self.buf, np.float32)
        self.i = np.frombuffer(self.buf, np.int32)
        self.orc = orc
        self.dim = dim

    def objects(self, rows):
        self.i[1] = len(rows)
        self.f[F_OBJ:F_OBJ + 256 * STRIDE] = 0
        for k, r in enumerate(rows):
            self.f[F_OBJ + k * STRIDE:F_OBJ + k * STRIDE + len(r)] = r

    def eval64(self, x):
        x = np.asarray(x, np.float64).reshape(-1, self.dim)
        return self.orc.eval_scene_f64(bytes(self.buf), x, self.dim)

    def eval32(self, x):
        x = np.asarray(x, np.float32).reshape(-1, self.dim)
        return self.orc.eval_scene(bytes(self.buf), x, self.dim)


def test_params_layout(orc):
    s = Scene(orc, 2, 4, 16.0, 8.0)
    assert len(s.buf) == 33 * 4 + 256 * STRIDE * 4
    assert (s.i[0], s.i[1], s.i[2]) == (2, 64, 4)
    assert (s.f[F_SX], s.f[F_SY]) == (16.0, 8.0)
    assert s.f[F_TANY] == np.float32(0.41421356237)
    assert s.f[F_TANX] == s.f[F_TANY] * np.float32(2.0)
    assert (s.f[F_LENS], s.f[F_FOCAL], s.f[F_AMB], s.f[F_PLANE], s.f[F_AO]) == \
        (np.float32(0.05), 5.0, np.float32(0.1), -2.0, 1.0)
    assert abs(np.linalg.norm(s.f[F_LDIR:F_LDIR + 3].astype(np.float64)) - 1) < 1e-7  # unit
    assert list(s.f[F_CONST:F_CONST + 3]) == [0.25, 0.5, 0.75]
    obj = s.f[F_OBJ:F_OBJ + 64 * STRIDE].reshape(64, STRIDE)
    assert (obj[:, 3] >= 0.2).all() and (obj[:, 3] <= 0.8).all()  # radii
    assert (obj[:, 2] >= 3).all() and (obj[:, 2] <= 12).all()  # depth
    assert (obj[:, 4:7] == 0).all()  # the DOF scene does not move
    d = Scene(orc, 1, 3)
    disks = d.f[F_OBJ:F_OBJ + 16 * STRIDE].reshape(16, STRIDE)
    assert (disks[:, 2] >= 0.02).all() and (disks[:, 2] <= 0.1).all()
    assert (np.abs(disks[:, 3:5]) <= 0.1).all()
    assert (disks[:, 5:8] >= 0).all() and (disks[:, 5:8] < 1).all()


def _cover_interval(p, c0, v, r):
    """t in [0, 1] with |p - c0 - v t| <= r: the roots of a quadratic."""
    q = np.subtract(p, c0)
    a, b, c = np.dot(v, v), -2 * np.dot(q, v), np.dot(q, q) - r * r
    disc = b * b - 4 * a * c
    if disc < 0:
        return 0.0
    t1, t2 = (-b - np.sqrt(disc)) / (2 * a), (-b + np.sqrt(disc)) / (2 * a)
    return max(0.0, min(t2, 1.0) - max(t1, 0.0))


@pytest.mark.parametrize("p", [(0.45, 0.5), (0.5, 0.47), (0.62, 0.44), (0.3, 0.5), (0.9, 0.9)])
def test_disk_coverage_over_shutter(orc, p):
    s = Scene(orc, 1, 3)
    c0, v, r, col = (0.4, 0.5), (0.2, -0.06), 0.1, (0.3, 0.6, 0.9)
    s.objects([[c0[0], c0[1], r, v[0], v[1], *col]])
    n = 4096
    t = (np.arange(n) + 0.5) / n
    x = np.stack([np.full(n, p[0]), np.full(n, p[1]), t], 1)
    val = s.eval64(x)
    hit = val[:, 0] != 0
    assert np.allclose(val[hit], np.float32(col)) and (val[~hit] == 0).all()
    want = _cover_interval(p, np.float32(c0), np.float32(v), np.float32(r))
    assert abs(hit.mean() - want) <= 2.0 / n, (hit.mean(), want)


def test_disks_highest_index_on_top(orc):
    s = Scene(orc, 1, 3, 2.0, 4.0)  # screen coordinates are x / scale
    s.objects([[0.4, 0.5, 0.1, 0, 0, 1, 0, 0], [0.5, 0.5, 0.1, 0, 0, 0, 1, 0]])
    pts = np.array([[0.45, 0.5, 0.3], [0.35, 0.5, 0.3], [0.58, 0.5, 0.3], [0.2, 0.2, 0.3]])
    pts[:, 0] *= 2.0
    pts[:, 1] *= 4.0
    v = s.eval64(pts)
    assert v.tolist() == [[0, 1, 0], [1, 0, 0], [0, 1, 0], [0, 0, 0]]
    assert np.array_equal(s.eval32(pts), v.astype(np.float32))


def _sphere_scene(orc, kind, dim, rows, sx=1.0, sy=1.0):
    s = Scene(orc, kind, dim, sx, sy)
    s.objects(rows)
    return s


def _lambert(s, albedo, n, lit=True):
    l = s.f[F_LDIR:F_LDIR + 3].astype(np.float64)
    lc = s.f[F_LCOL:F_LCOL + 3].astype(np.float64)
    ndl = max(float(np.dot(n, l)), 0.0) if lit else 0.0
    return np.asarray(albedo, np.float64) * (float(s.f[F_AMB]) + ndl * lc)


def test_centre_ray_hits_sphere_on_axis(orc):
    alb = (0.2, 0.4, 0.6)
    s = _sphere_scene(orc, 2, 4, [[0, 0, 5, 1, 0, 0, 0, *alb]])
    # pixel centre, lens centre: the ray (0, 0, 1) from the eye hits z = 4,
    # normal (0, 0, -1)
    v = s.eval64([[0.5, 0.5, 0.5, 0.5]])[0]
    want = _lambert(s, np.float32(alb), (0, 0, -1))
    assert np.allclose(v, want, rtol=0, atol=1e-7), (v, want)


def test_sky_plane_and_shadow(orc):
    s = _sphere_scene(orc, 2, 4, [])
    sky = s.eval64([[0.5, 0.5, 0.5, 0.5]])[0]
    assert np.allclose(sky, [0.75, 0.85, 1.0], atol=1e-12)  # a = 0.5 at d.y = 0
    # bottom edge of the image: d = (0, -tan_y, 1) / |.| reaches y = -2
    bottom = [[0.5, 0.0, 0.5, 0.5]]
    alb = s.f[F_ALB:F_ALB + 3].astype(np.float64)
    assert np.allclose(s.eval64(bottom)[0], _lambert(s, alb, (0, 1, 0)), atol=1e-7)
    ty = float(s.f[F_TANY])
    d = np.array([0.0, -ty, 1.0]) / np.hypot(ty, 1.0)
    p = d * (-2.0 / d[1])
    l = s.f[F_LDIR:F_LDIR + 3].astype(np.float64)
    c = p + 1.0 * l
    rho = 0.3
    # the blocker sits on the light ray but off the camera ray
    assert np.linalg.norm(np.cross(c, d)) > 2 * rho
    s.objects([[*c, rho, 0, 0, 0, 0.9, 0.9, 0.9]])
    assert np.allclose(s.eval64(bottom)[0], _lambert(s, alb, (0, 1, 0), lit=False), atol=1e-7)


def _lens_grid(n):
    g = (np.arange(n) + 0.5) / n
    return np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2)


def test_thin_lens_focus_and_aperture(orc):
    n = 64
    lens = _lens_grid(n)
    # a small black sphere at the focus point of pixel (dx, dy) = (0.3, -0.2)
    s = _sphere_scene(orc, 2, 4, [], 40.0, 20.0)
    dx, dy = 0.3, -0.2
    dc = np.array([dx * float(s.f[F_TANX]), dy * float(s.f[F_TANY]), 1.0])
    dc /= np.linalg.norm(dc)
    focus = dc * (float(s.f[F_FOCAL]) / dc[2])
    s.objects([[*focus, 0.002, 0, 0, 0, 0, 0, 0]])
    pix = [(dx + 1) / 2 * 40.0, (dy + 1) / 2 * 20.0]
    x = np.concatenate([np.tile(pix, (len(lens), 1)), lens], 1)
    assert (s.eval64(x) == 0).all()  # every lens sample reaches the focus point
    # halfway to focus on the axis the lens rays are |l| / 2 off the chief ray
    rho = 0.01
    s.objects([[0, 0, float(s.f[F_FOCAL]) / 2, rho, 0, 0, 0, 0, 0, 0]])
    x = np.concatenate([np.tile([20.0, 10.0], (len(lens), 1)), lens], 1)
    frac = (s.eval64(x)[:, 2] == 0).mean()
    want = (2 * rho / float(s.f[F_LENS])) ** 2  # uniform lens disk
    assert abs(frac - want) < 0.01, (frac, want)


def test_motion_time_is_third_coordinate(orc):
    row = [-1, 0, 5, 0.5, 2, 0, 0, 0, 0, 0]  # centre (-1 + 2t, 0, 5)
    s = _sphere_scene(orc, 3, 5, [row])
    n = 1000
    t = (np.arange(n) + 0.5) / n
    x = np.stack([np.full(n, 0.5), np.full(n, 0.5), t, np.full(n, 0.5), np.full(n, 0.5)], 1)
    hit = s.eval64(x)[:, 2] == 0
    assert np.array_equal(hit, np.abs(-1 + 2 * t) < 0.5)
    # the depth-of-field scene ignores the velocity (t = 0: centre at -1)
    s2 = _sphere_scene(orc, 2, 4, [row])
    assert s2.eval64([[0.5, 0.5, 0.5, 0.5]])[0, 2] != 0


def test_ambient_occlusion_under_ceiling(orc):
    n = 512  # the concentric map's radius runs in steps of 2 / n on this grid
    hemi = _lens_grid(n)
    s = Scene(orc, 4, 6)
    plane = float(s.f[F_PLANE])
    s.f[F_EYE:F_EYE + 3] = (0.0, plane + 0.2, 0.0)
    s.objects([])
    # the bottom edge of the image looks down onto the plane; lens centre
    x = np.concatenate([np.tile([0.5, 0.0, 0.5, 0.5], (len(hemi), 1)), hemi], 1)
    assert (s.eval64(x) == 1).all()  # nothing above the plane
    # a ceiling 0.5 above the plane: a sphere of radius 1e4 resting on it
    big, h = 1e4, 0.5
    s.objects([[0, plane + h + big, 0, big, 0, 0, 0, 1, 1, 1]])
    vis = s.eval64(x)[:, 0]
    assert set(np.unique(vis)) <= {0.0, 1.0}
    eps = 1e-3  # the occlusion ray starts this far above the surface
    R = float(s.f[F_AO])
    want = ((h - eps) / R) ** 2  # cos^2 of a cosine-weighted direction is uniform
    assert abs(vis.mean() - want) < 2.0 / n, (vis.mean(), want)
    # a camera ray into the sky is unoccluded
    up = np.concatenate([np.tile([0.5, 1.0, 0.5, 0.5], (4, 1)), hemi[:4]], 1)
    s.objects([])
    assert (s.eval64(up) == 1).all()
