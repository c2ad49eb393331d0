"""ctypes binding of libmdas.so (the C ABI declared in include/mdas.h).

The library is the product: HIP kernels for gfx950 behind a C ABI.  There is
no CPU fallback -- if the shared library has not been built, importing the
binding raises ImportError with the build command.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmdas.so")

MDAS_OK = 0
MDAS_ERR_INVALID = -1
MDAS_ERR_CAPACITY = -2
MDAS_ERR_HIP = -3
MDAS_ERR_STATE = -4

SCENE_CONSTANT = 0
SCENE_DISKS_MB = 1
SCENE_SPHERES_DOF = 2
SCENE_SPHERES_MBDOF = 3
SCENE_AO = 4
SCENE_RAMP = 5

MLP_FLOATS = 5345
MLP_PARAMS = 5313
FEATURES = 16
LEARN_COLLECT = 1
LEARN_SCORE = 2
LEARN_IMAGE_LABEL = 4
LEARN_KEEP_FEATURES = 8


class MdasError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mdas error {code}: {msg}")
        self.code = code


class MdasConfig(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("max_samples", ctypes.c_longlong),
        ("candidates_num", ctypes.c_int),
        ("bits_per_dim", ctypes.c_int),
        ("extra_img_bits", ctypes.c_int),
        ("alpha", ctypes.c_float),
        ("scale_x", ctypes.c_float),
        ("scale_y", ctypes.c_float),
        ("log_path", ctypes.c_char_p),
    ]


class MdasPhaseTimes(ctypes.Structure):
    _fields_ = [
        ("initial_ms", ctypes.c_double),
        ("errors_ms", ctypes.c_double),
        ("adaptive_ms", ctypes.c_double),
        ("integrand_ms", ctypes.c_double),
        ("recon_ms", ctypes.c_double),
        ("sort_ms", ctypes.c_double),
        ("passes", ctypes.c_int),
        ("final_samples", ctypes.c_longlong),
        ("final_leaves", ctypes.c_longlong),
        ("bytes_initial", ctypes.c_double),
        ("bytes_errors", ctypes.c_double),
        ("bytes_adaptive", ctypes.c_double),
        ("bytes_recon", ctypes.c_double),
    ]


class MdasEnv(ctypes.Structure):
    _fields_ = [
        ("mdas", ctypes.c_int),
        ("samples", ctypes.c_float),
        ("scale_factor", ctypes.c_float),
        ("alpha", ctypes.c_float),
        ("bits_per_dim", ctypes.c_int),
        ("extra_img_bits", ctypes.c_int),
        ("candidates_num", ctypes.c_int),
        ("width", ctypes.c_int),
        ("height", ctypes.c_int),
        ("filename", ctypes.c_char * 512),
        ("scene_kind", ctypes.c_int),
        ("scene_seed", ctypes.c_uint32),
        ("dim", ctypes.c_int),
    ]


_H = ctypes.c_void_p
_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_LL = ctypes.c_longlong

# name -> (restype, argtypes)
SIGNATURES = {
    "mdas_create": (_I, [ctypes.POINTER(MdasConfig), ctypes.POINTER(_H)]),
    "mdas_destroy": (_I, [_H]),
    "mdas_initial_sampling": (_I, [_H, _P]),
    "mdas_compute_errors": (_I, [_H, _P]),
    "mdas_adaptive_sampling": (_I, [_H, _P]),
    "mdas_sampling_pass": (_I, [_H, _P]),
    "mdas_integrate": (_I, [_H, _P, _I, _I, _P]),
    "mdas_integrate_scatter": (_I, [_H, _P, _I, _I, _P]),
    "mdas_convert_to_bytes": (_I, [_H, _P, _P, _I, _I, _P]),
    "mdas_sampling_density": (_I, [_H, _P, _I, _I]),
    "mdas_validate": (_I, [_H, ctypes.POINTER(_I)]),
    "mdas_get_sample_coordinates": (_P, [_H]),
    "mdas_get_sample_values": (_P, [_H]),
    "mdas_get_dim": (_I, [_H]),
    "mdas_get_candidates_num": (_I, [_H]),
    "mdas_get_max_node_size": (_I, [_H]),
    "mdas_get_new_samples": (_I, [_H]),
    "mdas_get_new_nodes": (_I, [_H]),
    "mdas_get_number_of_samples": (_I, [_H]),
    "mdas_get_number_of_nodes": (_I, [_H]),
    "mdas_get_max_samples": (_I, [_H]),
    "mdas_get_scale_x": (_F, [_H]),
    "mdas_get_scale_y": (_F, [_H]),
    "mdas_get_max_error": (_F, [_H]),
    "mdas_scene_create": (_I, [_I, ctypes.c_uint32, _I, _F, _F, ctypes.POINTER(_H)]),
    "mdas_scene_destroy": (_I, [_H]),
    "mdas_scene_params": (_LL, [_H, _P, _LL]),
    "mdas_eval_integrand": (_I, [_H, _H, _I, _I, _P]),
    "mdas_eval_new_samples": (_I, [_H, _H, _P]),
    "mdas_render_qmc": (_I, [_H, _I, _I, _I, _I, ctypes.c_uint32, _I, _P, _P]),
    "mdas_run": (_I, [_H, _H, _LL, _P]),
    "mdas_get_passes": (_I, [_H]),
    "mdas_set_timing": (_I, [_H, _I]),
    "mdas_set_log_sink": (_I, [_H, _P, _P]),
    "mdas_integrate_noalloc": (_I, [_H, _P, _I, _I, _P]),
    "mdas_get_phase_times": (_I, [_H, ctypes.POINTER(MdasPhaseTimes)]),
    "mdas_get_pass_stats": (_I, [_H, _P, _I]),
    "mdas_kernel_stats": (_I, [_H, _I, ctypes.c_char_p, _I, ctypes.POINTER(_I),
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "mdas_set_recon_tile": (_I, [_H, _I, _I]),
    "mdas_set_launch_shape": (_I, [_H, _I, _I]),
    "mdas_autotune_recon": (_I, [_H, _P, _I, _I, _P, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "mdas_download_state": (_I, [_H, _P, _P, _P, _P, _P, _P]),
    "mdas_upload_state": (_I, [_H, _I, _I, _P, _P, _P, _P, _P, _P, _F]),
    "mdas_set_values": (_I, [_H, _I, _I, _P]),
    "mdas_get_recon_keys": (_I, [_H, _P, _LL, ctypes.POINTER(_LL)]),
    "mdas_batch_create": (_I, [_P, _P, _I, ctypes.POINTER(_H)]),
    "mdas_batch_destroy": (_I, [_H]),
    "mdas_batch_run": (_I, [_H, _P, _P, _I]),
    "mdas_batch_integrate": (_I, [_H, _P, _P, _P, _P, _I]),
    "mdas_batch_info": (_I, [_H, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_I),
                             ctypes.POINTER(_I)]),
    "mdas_batch_autotune": (_I, [_H, _P, _P, ctypes.POINTER(_I)]),
    "mdas_env_read": (_I, [ctypes.c_char_p, ctypes.POINTER(MdasEnv)]),
    "mdas_env_last_error": (ctypes.c_char_p, []),
    "mdas_env_config": (_I, [ctypes.POINTER(MdasEnv), _LL, ctypes.POINTER(MdasConfig)]),
    "mdas_mlp_create": (_I, [ctypes.c_uint32, ctypes.POINTER(_H)]),
    "mdas_mlp_destroy": (_I, [_H]),
    "mdas_mlp_set_params": (_I, [_H, _P, _I]),
    "mdas_mlp_get_params": (_I, [_H, _P, _I]),
    "mdas_mlp_fit_norm": (_I, [_H, _P, _LL, _P]),
    "mdas_mlp_forward": (_I, [_H, _P, _LL, _P, _P]),
    "mdas_mlp_grad": (_I, [_H, _P, _P, _LL, _P, ctypes.POINTER(_F), _P]),
    "mdas_mlp_train": (_I, [_H, _P, _P, _LL, _I, _I, _F, ctypes.c_uint32, _P, _P]),
    "mdas_set_learned": (_I, [_H, _I, _H, _P, _P, _LL]),
    "mdas_set_learned_exponent": (_I, [_H, _F]),
    "mdas_set_learned_exponents": (_I, [_H, _F, _F]),
    "mdas_get_learned_rows": (_I, [_H, ctypes.POINTER(_LL)]),
    "mdas_get_leaf_features": (_I, [_H, _P, _P, _LL]),
    "mdas_last_error": (ctypes.c_char_p, []),
    "mdas_device_synchronize": (_I, []),
}

_lib = None


def load():
    """Load libmdas.so once; raise ImportError if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C pmdas_amd` "
            "or __graft_entry__.build() (there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    """Raise MdasError for a negative status code."""
    if rc < 0:
        raise MdasError(rc, load().mdas_last_error().decode(errors="replace"))
    return rc
