"""ctypes wrapper of the CPU oracle (liboracle_mdas.so).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, to check results and as the CPU baseline -- never
the product.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_mdas.so")

_lib = None
_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_LL = ctypes.c_longlong
_U = ctypes.c_uint32

_SIGS = {
    "orc_create": (_P, [_I, _LL, _I, _I, _I, _F, _F, _F]),
    "orc_destroy": (None, [_P]),
    "orc_set_threads": (None, [_I]),
    "orc_initial_sampling": (_I, [_P]),
    "orc_compute_errors": (_I, [_P]),
    "orc_adaptive_sampling": (_I, [_P]),
    "orc_integrate": (_I, [_P, _P, _I, _I]),
    "orc_convert_to_bytes": (None, [_P, _P, _LL]),
    "orc_sampling_density": (None, [_P, _P, _I, _I]),
    "orc_validate": (_I, [_P]),
    "orc_run": (_I, [_P, _P, _LL, ctypes.POINTER(_I)]),
    "orc_make_scene": (_LL, [_I, _U, _I, _F, _F, _P, _LL]),
    "orc_eval_scene": (None, [_P, _P, _I, _LL, _LL, _P]),
    "orc_eval_scene_f64": (None, [_P, _P, _I, _LL, _P]),
    "orc_eval_scene_grouped": (None, [_P, _P, _I, _LL, _I, _P]),
    "orc_qmc_image": (None, [_P, _I, _I, _I, _I, _U, _I, _P]),
    "orc_halton": (_F, [_I, _I]),
    "orc_eval_range": (None, [_P, _P, _LL, _LL]),
    "orc_recon_keys": (_LL, [_P, _I, _I, _I, _I, _P, _LL]),
    "orc_num_samples": (_LL, [_P]),
    "orc_num_leaves": (_LL, [_P]),
    "orc_new_samples": (_LL, [_P]),
    "orc_new_nodes": (_LL, [_P]),
    "orc_max_error": (_F, [_P]),
    "orc_get_state": (None, [_P, _P, _P, _P, _P, _P, _P]),
    "orc_set_state": (_I, [_P, _LL, _LL, _P, _P, _P, _P, _P, _P, _F]),
    "orc_set_values": (None, [_P, _LL, _LL, _P]),
    "orc_tea": (_U, [_U, _U, _I, _P, ctypes.POINTER(_U)]),
    "orc_tea4": (_U, [_U, _U]),
    "orc_lcg": (_U, [ctypes.POINTER(_U)]),
    "orc_rnd": (_F, [ctypes.POINTER(_U)]),
    "orc_powf": (_F, [_F, _F]),
    "orc_morton_cell": (None, [_LL, _I, _I, _I, _P, _P, _P]),
    "orc_largest_axis": (_I, [_P, _I]),
    "orc_sort5": (None, [_P, _P]),
    "orc_leaf_error": (_F, [_P, _I, _F]),
    "orc_make_color": (None, [_P, _P]),
}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        for k, (r, a) in _SIGS.items():
            f = getattr(lib, k)
            f.restype = r
            f.argtypes = a
        _lib = lib
    return _lib


def vp(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleGrid:
    """The oracle's KDGrid (same constructor arguments as the reference)."""

    def __init__(self, dim, max_samples, candidates_num, bits_per_dim, extra_img_bits, alpha,
                 scale_x, scale_y):
        self.lib = load()
        self.h = self.lib.orc_create(dim, int(max_samples), candidates_num, bits_per_dim,
                                     extra_img_bits, float(alpha), float(scale_x), float(scale_y))
        if not self.h:
            raise ValueError("orc_create rejected the configuration")
        self.dim = dim
        self.scale_x, self.scale_y = float(scale_x), float(scale_y)

    def initial_sampling(self):
        rc = self.lib.orc_initial_sampling(self.h)
        if rc:
            raise RuntimeError(f"orc_initial_sampling -> {rc}")

    def compute_errors(self):
        self.lib.orc_compute_errors(self.h)

    def adaptive_sampling(self):
        return self.lib.orc_adaptive_sampling(self.h)

    def sampling_pass(self):
        self.compute_errors()
        return self.adaptive_sampling()

    @property
    def n_samples(self):
        return self.lib.orc_num_samples(self.h)

    @property
    def n_leaves(self):
        return self.lib.orc_num_leaves(self.h)

    @property
    def new_samples(self):
        return self.lib.orc_new_samples(self.h)

    @property
    def new_nodes(self):
        return self.lib.orc_new_nodes(self.h)

    @property
    def max_error(self):
        return self.lib.orc_max_error(self.h)

    def state(self):
        n, l, d = self.n_samples, self.n_leaves, self.dim
        s = dict(coords=np.empty((n, d), np.float32), values=np.empty((n, 3), np.float32),
                 idx=np.empty((l, 4), np.int32), box=np.empty((l, 2 * d), np.float32),
                 err=np.empty(l, np.float32), seed=np.empty(l, np.uint32))
        self.lib.orc_get_state(self.h, vp(s["coords"]), vp(s["values"]), vp(s["idx"]),
                               vp(s["box"]), vp(s["err"]), vp(s["seed"]))
        return s

    def set_state(self, st, max_error=0.0):
        arrs = [np.ascontiguousarray(st[k], dt) for k, dt in
                (("coords", np.float32), ("values", np.float32), ("idx", np.int32),
                 ("box", np.float32), ("err", np.float32), ("seed", np.uint32))]
        rc = self.lib.orc_set_state(self.h, len(arrs[0]), len(arrs[2]), *[vp(a) for a in arrs],
                                    float(max_error))
        if rc:
            raise RuntimeError("orc_set_state failed")

    def set_values(self, first, values):
        v = np.ascontiguousarray(values, np.float32)
        self.lib.orc_set_values(self.h, int(first), len(v), vp(v))

    def eval_scene(self, params, first, count):
        """Evaluate the analytic scene for samples [first, first+count) in place."""
        self.lib.orc_eval_range(self.h, params, int(first), int(count))

    def run(self, params, target):
        p = ctypes.c_int(0)
        rc = self.lib.orc_run(self.h, params, int(target), ctypes.byref(p))
        if rc:
            raise RuntimeError(f"orc_run -> {rc}")
        return p.value

    def integrate(self, w, h):
        px = np.zeros((h, w, 4), np.float32)
        self.lib.orc_integrate(self.h, vp(px), w, h)
        return px

    def sampling_density(self, w, h):
        px = np.zeros((h, w, 4), np.float32)
        self.lib.orc_sampling_density(self.h, vp(px), w, h)
        return px

    def validate(self):
        return bool(self.lib.orc_validate(self.h))

    def recon_keys(self, w, h, tw, th):
        n = self.lib.orc_recon_keys(self.h, w, h, tw, th, None, 0)
        out = np.empty(n, np.uint64)
        self.lib.orc_recon_keys(self.h, w, h, tw, th, vp(out), n)
        return out

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_scene(kind, seed, dim, scale_x, scale_y):
    """Scene parameter block (bytes) of the seeded analytic scene."""
    n = load().orc_make_scene(kind, seed, dim, scale_x, scale_y, None, 0)
    buf = ctypes.create_string_buffer(n)
    load().orc_make_scene(kind, seed, dim, scale_x, scale_y, buf, n)
    return buf.raw


def convert_to_bytes(pixels):
    px = np.ascontiguousarray(pixels, np.float32).reshape(-1, 4)
    out = np.empty((len(px), 4), np.uint8)
    load().orc_convert_to_bytes(vp(px), vp(out), len(px))
    return out


def eval_scene(params, coords, dim):
    c = np.ascontiguousarray(coords, np.float32).reshape(-1, dim)
    out = np.empty((len(c), 3), np.float32)
    load().orc_eval_scene(params, vp(c), dim, 0, len(c), vp(out))
    return out


def eval_scene_f64(params, coords, dim):
    c = np.ascontiguousarray(coords, np.float64).reshape(-1, dim)
    out = np.empty((len(c), 3), np.float64)
    load().orc_eval_scene_f64(params, vp(c), dim, len(c), vp(out))
    return out


def eval_scene_grouped(params, coords, dim, fp64=False):
    """The integrand through the product's sphere groups (CPU build), to be
    held against eval_scene / eval_scene_f64."""
    t = np.float64 if fp64 else np.float32
    c = np.ascontiguousarray(coords, t).reshape(-1, dim)
    out = np.empty((len(c), 3), t)
    load().orc_eval_scene_grouped(params, vp(c), dim, len(c), 1 if fp64 else 0, vp(out))
    return out


def qmc_image(params, w, h, spp, first_sample=0, frame=0, fp64=False):
    """The plain QMC sampler's (H, W, 4) image (float32, or float64 if fp64)."""
    out = np.empty((h, w, 4), np.float64 if fp64 else np.float32)
    load().orc_qmc_image(params, w, h, spp, first_sample, frame, 1 if fp64 else 0, vp(out))
    return out


def halton(dim, i):
    return load().orc_halton(dim, i)


def powf(x, a):
    return load().orc_powf(float(x), float(a))


def tea(v0, v1, rounds, key):
    k = (ctypes.c_uint32 * 4)(*key)
    v1o = ctypes.c_uint32()
    v0o = load().orc_tea(v0, v1, rounds, k, ctypes.byref(v1o))
    return v0o, v1o.value
