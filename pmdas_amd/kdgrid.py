"""Python mirror of the reference's mdas::KDGrid<Point> (SDK/lib/mdas/kdgrid.h:10-107)
over the C ABI of libmdas.so.

Method names, argument meaning and call order follow the reference
(InitialSampling, ComputeErrors, AdaptiveSampling, SamplingPass, Integrate,
ConvertToBytes, SamplingDensity, Validate and the getters); errors raise
MdasError instead of being printed.  Extra helpers move state between host and
device for tests and tools.
"""
import ctypes

import numpy as np

from . import _lib
from . import hip
from ._lib import check


class Scene:
    """Seeded analytic integrand (replaces the OptiX raygen program)."""

    def __init__(self, kind, seed, dim, scale_x, scale_y):
        lib = _lib.load()
        h = ctypes.c_void_p()
        check(lib.mdas_scene_create(int(kind), ctypes.c_uint32(seed), int(dim), float(scale_x),
                                    float(scale_y), ctypes.byref(h)))
        self._h = h.value
        self.kind, self.seed, self.dim = kind, seed, dim

    @property
    def handle(self):
        return self._h

    def render_qmc(self, width, height, spp, first_sample=0, frame=0, fp64=False):
        """The plain QMC sampler's image (H, W, 4); FP64 evaluation if fp64."""
        dt = np.float64 if fp64 else np.float32
        buf = hip.DeviceBuffer(width * height * 4 * np.dtype(dt).itemsize)
        try:
            check(_lib.load().mdas_render_qmc(self._h, width, height, spp, first_sample, frame,
                                              1 if fp64 else 0, buf.ptr, None))
            hip.synchronize()
            return buf.download(dt, width * height * 4).reshape(height, width, 4)
        finally:
            buf.free()

    def params_bytes(self):
        lib = _lib.load()
        n = lib.mdas_scene_params(self._h, None, 0)
        buf = ctypes.create_string_buffer(n)
        lib.mdas_scene_params(self._h, buf, n)
        return buf.raw

    def close(self):
        if self._h:
            _lib.load().mdas_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many grids advanced together (one launch per kernel and pass for all
    images of a dimension; optionally replayed from a captured hipGraph)."""

    def __init__(self, grids, scenes):
        lib = _lib.load()
        n = len(grids)
        gh = (ctypes.c_void_p * n)(*[g.handle for g in grids])
        sh = (ctypes.c_void_p * n)(*[s.handle for s in scenes])
        h = ctypes.c_void_p()
        check(lib.mdas_batch_create(gh, sh, n, ctypes.byref(h)))
        self._h = h.value
        self.grids, self.scenes = list(grids), list(scenes)

    def run(self, targets, use_graph=True, stream=None):
        t = (ctypes.c_longlong * len(targets))(*[int(x) for x in targets])
        check(_lib.load().mdas_batch_run(self._h, t, stream.handle if stream else None,
                                         1 if use_graph else 0))

    def integrate(self, pixels, widths, heights, use_graph=True, stream=None):
        """Reconstruct image i into device float4 buffer pixels[i] (pointers or
        DeviceBuffers) of widths[i] x heights[i]; stream-ordered."""
        n = len(pixels)
        ptrs = (ctypes.c_void_p * n)(*[p.ptr if hasattr(p, "ptr") else p for p in pixels])
        w = (ctypes.c_int * n)(*[int(x) for x in widths])
        h = (ctypes.c_int * n)(*[int(x) for x in heights])
        check(_lib.load().mdas_batch_integrate(self._h, ptrs, w, h, stream.handle if stream else None,
                                               1 if use_graph else 0))

    def autotune(self, targets, stream=None):
        """Launch shape (workgroups per launch) of the fastest of the
        candidate runs; the batch keeps it."""
        t = (ctypes.c_longlong * len(targets))(*[int(x) for x in targets])
        c = ctypes.c_int(0)
        check(_lib.load().mdas_batch_autotune(self._h, t, stream.handle if stream else None,
                                              ctypes.byref(c)))
        return c.value

    def info(self):
        ms, blocks, groups = ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
        check(_lib.load().mdas_batch_info(self._h, ctypes.byref(ms), ctypes.byref(blocks),
                                          ctypes.byref(groups)))
        return dict(wall_ms=ms.value, pass_blocks=blocks.value, groups=groups.value)

    def close(self):
        if self._h:
            _lib.load().mdas_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class KDGrid:
    """mdas::KDGrid<PointN<dim>> (kdgrid.h:42-51 constructor arguments)."""

    def __init__(self, dim, max_samples, candidates_num, bits_per_dim, extra_img_bits, alpha,
                 scale_x, scale_y, log=None):
        lib = _lib.load()
        self._log = log.encode() if isinstance(log, str) else log
        cfg = _lib.MdasConfig(dim, int(max_samples), candidates_num, bits_per_dim, extra_img_bits,
                              float(alpha), float(scale_x), float(scale_y), self._log)
        h = ctypes.c_void_p()
        check(lib.mdas_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h.value
        self.dim = dim
        self.stream = None  # None = null stream, as the reference uses stream 0

    @property
    def handle(self):
        return self._h

    def _s(self):
        return self.stream.handle if self.stream is not None else None

    # -- the reference's entry points ------------------------------------
    def InitialSampling(self):
        check(_lib.load().mdas_initial_sampling(self._h, self._s()))

    def ComputeErrors(self):
        check(_lib.load().mdas_compute_errors(self._h, self._s()))

    def AdaptiveSampling(self):
        check(_lib.load().mdas_adaptive_sampling(self._h, self._s()))

    def SamplingPass(self):
        check(_lib.load().mdas_sampling_pass(self._h, self._s()))

    def Integrate(self, pixels, width, height):
        """pixels: device pointer to float4[width*height]."""
        check(_lib.load().mdas_integrate(self._h, pixels, width, height, self._s()))

    def IntegrateScatter(self, pixels, width, height):
        check(_lib.load().mdas_integrate_scatter(self._h, pixels, width, height, self._s()))

    def ConvertToBytes(self, pixels, pixels_bytes, width, height):
        check(_lib.load().mdas_convert_to_bytes(self._h, pixels, pixels_bytes, width, height,
                                                self._s()))

    def SamplingDensity(self, host_pixels, width, height):
        """host_pixels: numpy float32 array of width*height*4 (written)."""
        assert host_pixels.dtype == np.float32 and host_pixels.size == width * height * 4
        check(_lib.load().mdas_sampling_density(self._h, host_pixels.ctypes.data_as(ctypes.c_void_p),
                                                width, height))

    def Validate(self):
        v = ctypes.c_int(0)
        check(_lib.load().mdas_validate(self._h, ctypes.byref(v)))
        return bool(v.value)

    def GetSampleCoordinates(self):
        return _lib.load().mdas_get_sample_coordinates(self._h)

    def GetSampleValues(self):
        return _lib.load().mdas_get_sample_values(self._h)

    def GetCandidatesNum(self):
        return _lib.load().mdas_get_candidates_num(self._h)

    def GetMaxNodeSize(self):
        return _lib.load().mdas_get_max_node_size(self._h)

    def GetNewSamples(self):
        return _lib.load().mdas_get_new_samples(self._h)

    def GetNewNodes(self):
        return _lib.load().mdas_get_new_nodes(self._h)

    def GetNumberOfSamples(self):
        return _lib.load().mdas_get_number_of_samples(self._h)

    def GetNumberOfNodes(self):
        return _lib.load().mdas_get_number_of_nodes(self._h)

    def GetMaxSamples(self):
        return _lib.load().mdas_get_max_samples(self._h)

    def GetScaleX(self):
        return _lib.load().mdas_get_scale_x(self._h)

    def GetScaleY(self):
        return _lib.load().mdas_get_scale_y(self._h)

    def GetMaxError(self):
        return _lib.load().mdas_get_max_error(self._h)

    # -- additions ---------------------------------------------------------
    def EvalIntegrand(self, scene, offset, count):
        check(_lib.load().mdas_eval_integrand(self._h, scene.handle, int(offset), int(count),
                                              self._s()))

    def EvalNewSamples(self, scene):
        check(_lib.load().mdas_eval_new_samples(self._h, scene.handle, self._s()))

    def Run(self, scene, target_samples):
        check(_lib.load().mdas_run(self._h, scene.handle, int(target_samples), self._s()))
        return _lib.load().mdas_get_passes(self._h)

    def set_timing(self, on=True):
        check(_lib.load().mdas_set_timing(self._h, 1 if on else 0))

    def phase_times(self):
        t = _lib.MdasPhaseTimes()
        check(_lib.load().mdas_get_phase_times(self._h, ctypes.byref(t)))
        return {f: getattr(t, f) for f, _ in t._fields_}

    def kernel_stats(self):
        """{kernel: (launches, total_ms, algorithmic_bytes)} of the last run + integrate."""
        lib = _lib.load()
        n = lib.mdas_kernel_stats(self._h, -1, None, 0, None, None, None)
        out = {}
        for i in range(n):
            name = ctypes.create_string_buffer(64)
            la = ctypes.c_int()
            ms = ctypes.c_double()
            by = ctypes.c_double()
            check(lib.mdas_kernel_stats(self._h, i, name, 64, ctypes.byref(la), ctypes.byref(ms),
                                        ctypes.byref(by)))
            out[name.value.decode()] = (la.value, ms.value, by.value)
        return out

    def pass_stats(self, max_rows=4096):
        """Rows of (L, N, selected, split, selected_samples, stale, stale_values, 0)."""
        out = np.zeros((max_rows, 8), dtype=np.int32)
        n = check(_lib.load().mdas_get_pass_stats(self._h, out.ctypes.data_as(ctypes.c_void_p),
                                                  max_rows))
        return out[:n]

    def set_launch_shape(self, chunk_blocks, err_blocks):
        check(_lib.load().mdas_set_launch_shape(self._h, chunk_blocks, err_blocks))

    def set_recon_tile(self, tw, th):
        check(_lib.load().mdas_set_recon_tile(self._h, tw, th))

    def autotune_recon(self, pixels, width, height):
        """Pick the fastest reconstruction tile for the current state; returns (tw, th)."""
        tw, th = ctypes.c_int(), ctypes.c_int()
        check(_lib.load().mdas_autotune_recon(self._h, pixels, width, height, self._s(),
                                              ctypes.byref(tw), ctypes.byref(th)))
        return tw.value, th.value

    def download_state(self):
        """Host copies: coords (N,D), values (N,3), idx (L,4), box (L,2D), err (L), seed (L)."""
        lib = _lib.load()
        lib.mdas_device_synchronize()
        n, l, d = self.GetNumberOfSamples(), self.GetNumberOfNodes(), self.dim
        coords = np.empty((n, d), np.float32)
        values = np.empty((n, 3), np.float32)
        idx = np.empty((l, 4), np.int32)
        box = np.empty((l, 2 * d), np.float32)
        err = np.empty(l, np.float32)
        seed = np.empty(l, np.uint32)
        vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        check(lib.mdas_download_state(self._h, vp(coords), vp(values), vp(idx), vp(box), vp(err),
                                      vp(seed)))
        # counters may have moved if the state was changed on the device since
        n2, l2 = self.GetNumberOfSamples(), self.GetNumberOfNodes()
        assert (n, l) == (n2, l2), "state changed during download"
        return dict(coords=coords, values=values, idx=idx, box=box, err=err, seed=seed)

    def upload_state(self, coords, values, idx, box, err, seed, max_error=0.0):
        lib = _lib.load()
        n, l = len(coords), len(idx)
        arrs = [np.ascontiguousarray(a, dt) for a, dt in
                ((coords, np.float32), (values, np.float32), (idx, np.int32), (box, np.float32),
                 (err, np.float32), (seed, np.uint32))]
        vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        check(lib.mdas_upload_state(self._h, n, l, *[vp(a) for a in arrs], float(max_error)))

    def set_values(self, offset, values):
        v = np.ascontiguousarray(values, np.float32)
        check(_lib.load().mdas_set_values(self._h, int(offset), len(v),
                                          v.ctypes.data_as(ctypes.c_void_p)))

    def recon_keys(self):
        lib = _lib.load()
        n = ctypes.c_longlong(0)
        check(lib.mdas_get_recon_keys(self._h, None, 0, ctypes.byref(n)))
        out = np.empty(n.value, np.uint64)
        check(lib.mdas_get_recon_keys(self._h, out.ctypes.data_as(ctypes.c_void_p), n.value,
                                      ctypes.byref(n)))
        return out

    def integrate_image(self, width, height, scatter=False):
        """Integrate into a fresh device buffer and return (H, W, 4) float32."""
        buf = hip.DeviceBuffer(width * height * 16)
        try:
            if scatter:
                self.IntegrateScatter(buf.ptr, width, height)
            else:
                self.Integrate(buf.ptr, width, height)
            hip.synchronize()
            return buf.download(np.float32, width * height * 4).reshape(height, width, 4)
        finally:
            buf.free()

    # ------------------------------------------------ learned proposal --
    def set_learned(self, mode, mlp=None, rows=None, labels=None, cap_rows=0):
        """mode: 0 plain, LEARN_COLLECT | LEARN_SCORE bits; rows / labels are
        DeviceBuffers (or device pointers) for collection."""
        def ptr(b):
            return None if b is None else (b.ptr if hasattr(b, "ptr") else b)
        check(_lib.load().mdas_set_learned(self._h, mode, None if mlp is None else mlp.handle,
                                           ptr(rows), ptr(labels), cap_rows))

    def set_learned_exponent(self, kappa):
        check(_lib.load().mdas_set_learned_exponent(self._h, kappa))

    def set_learned_exponents(self, kappa_e, kappa_y):
        check(_lib.load().mdas_set_learned_exponents(self._h, kappa_e, kappa_y))

    def learned_rows(self):
        n = ctypes.c_longlong(0)
        check(_lib.load().mdas_get_learned_rows(self._h, ctypes.byref(n)))
        return n.value

    def leaf_features(self):
        """(features (L, 16), plain errors (L,)) after a learned-mode run."""
        L = self.GetNumberOfNodes()
        f = np.empty((L, _lib.FEATURES), np.float32)
        e = np.empty(L, np.float32)
        check(_lib.load().mdas_get_leaf_features(self._h, f.ctypes.data_as(ctypes.c_void_p),
                                                 e.ctypes.data_as(ctypes.c_void_p), L))
        return f, e

    def close(self):
        if self._h:
            _lib.load().mdas_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mlp:
    """The learned proposal's network (16 -> 64 -> 64 -> 1, GELU, BF16 on
    MFMA, FP32 accumulation) with its training on the GPU."""

    def __init__(self, seed=0):
        self._h = ctypes.c_void_p()
        check(_lib.load().mdas_mlp_create(seed, ctypes.byref(self._h)))

    @property
    def handle(self):
        return self._h

    def params(self):
        out = np.empty(_lib.MLP_FLOATS, np.float32)
        check(_lib.load().mdas_mlp_get_params(self._h, out.ctypes.data_as(ctypes.c_void_p),
                                              _lib.MLP_FLOATS))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        check(_lib.load().mdas_mlp_set_params(self._h, p.ctypes.data_as(ctypes.c_void_p), p.size))

    def fit_norm(self, rows_dev, n):
        check(_lib.load().mdas_mlp_fit_norm(self._h, rows_dev.ptr, n, None))

    def forward(self, X):
        """Host rows (n, 16) -> host outputs (n,), through device buffers."""
        X = np.ascontiguousarray(X, np.float32).reshape(-1, _lib.FEATURES)
        n = len(X)
        bx, by = hip.DeviceBuffer(X.nbytes), hip.DeviceBuffer(4 * n)
        try:
            bx.upload(X)
            check(_lib.load().mdas_mlp_forward(self._h, bx.ptr, n, by.ptr, None))
            hip.synchronize()
            return by.download(np.float32, n)
        finally:
            bx.free()
            by.free()

    def grad(self, X, T):
        """(gradient (MLP_PARAMS,), loss) of mean((y - T)^2) over host rows."""
        X = np.ascontiguousarray(X, np.float32).reshape(-1, _lib.FEATURES)
        T = np.ascontiguousarray(T, np.float32)
        n = len(X)
        bx, bt = hip.DeviceBuffer(X.nbytes), hip.DeviceBuffer(T.nbytes)
        g = np.empty(_lib.MLP_PARAMS, np.float32)
        loss = ctypes.c_float(0.0)
        try:
            bx.upload(X)
            bt.upload(T)
            check(_lib.load().mdas_mlp_grad(self._h, bx.ptr, bt.ptr, n,
                                            g.ctypes.data_as(ctypes.c_void_p), ctypes.byref(loss),
                                            None))
            return g, loss.value
        finally:
            bx.free()
            bt.free()

    def train(self, rows_dev, labels_dev, n, epochs, batch=65536, lr=1e-3, seed=0):
        """Adam on device rows; returns the mean loss of each epoch."""
        losses = np.zeros(epochs, np.float32)
        check(_lib.load().mdas_mlp_train(self._h, rows_dev.ptr, labels_dev.ptr, n, epochs, batch,
                                         lr, seed, losses.ctypes.data_as(ctypes.c_void_p), None))
        return losses

    def close(self):
        if self._h:
            _lib.load().mdas_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
