"""ctypes binding of libsprgpu.so (the C ABI declared in include/sprgpu.h).

The shared library is built in-tree by core_amd/csrc/Makefile.  Loading it
fails loudly: there is no CPU fallback for any compute entry point.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsprgpu.so")

SPRG_OK = 0
SPRG_EINVAL = 1
SPRG_ENOMEM = 2
SPRG_EHIP = 3
SPRG_EPATCH = 4
SPRG_EUNSUPPORTED = 5
SPRG_EINTERNAL = 6

SPRG_FP64 = 0
SPRG_FP32R = 1
SPRG_FP64F = 2

_c_int = ctypes.c_int
_c_i64 = ctypes.c_int64
_c_dbl = ctypes.c_double
_vp = ctypes.c_void_p
_pp = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes)
SIGNATURES = {
    "sprg_last_error": (ctypes.c_char_p, []),
    "sprg_version": (ctypes.c_char_p, []),
    "sprg_ctx_create": (_c_int, [_c_int, _pp]),
    "sprg_ctx_destroy": (_c_int, [_vp]),
    "sprg_ctx_sync": (_c_int, [_vp]),
    "sprg_ctx_stream": (_vp, [_vp]),
    "sprg_malloc": (_c_int, [_vp, ctypes.c_size_t, _pp]),
    "sprg_free": (_c_int, [_vp, _vp]),
    "sprg_upload": (_c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "sprg_download": (_c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "sprg_mesh_create": (_c_int, [_vp, _c_i64, _c_i64, _c_i64, _vp, _vp, _vp, _c_int, _pp]),
    "sprg_mesh_destroy": (_c_int, [_vp]),
    "sprg_mesh_counts": (_c_int, [_vp, _vp]),
    "sprg_mesh_adjacency": (_c_int, [_vp, _c_int, _pp, _pp]),
    "sprg_mesh_phase_times": (_c_int, [_vp, _vp, _c_int]),
    "sprg_mesh_last_stats": (_c_int, [_vp, _vp, _c_int]),
    "sprg_mesh_autotune": (_c_int, [_vp, _c_int, _c_int, _vp, _c_int, _vp, _c_int]),
    "sprg_mesh_set_shape": (_c_int, [_vp, _vp, _c_int]),
    "sprg_grad_ip": (_c_int, [_vp, _c_int, _c_int, _vp, _c_int, _vp]),
    "sprg_patch_sets": (_c_int, [_vp, _c_int, _vp, _vp, _c_i64, _vp]),
    "sprg_recover": (_c_int, [_vp, _c_int, _c_int, _vp, _vp, _c_int]),
    "sprg_recover_noalloc": (_c_int, [_vp, _c_int, _c_int, _vp, _vp, _c_int]),
    "sprg_spr_size": (_c_int, [_vp, _c_int, _c_int, _vp, _c_dbl, _vp, _c_int, _vp]),
    "sprg_target_spr_size": (_c_int, [_vp, _c_int, _c_int, _vp, ctypes.c_uint64, _c_dbl, _c_dbl,
                                      _vp, _c_int, _vp]),
    "sprg_average_to_vertex": (_c_int, [_vp, _vp, _vp]),
    "sprg_batch_create": (_c_int, [_vp, _c_int, _vp, _pp]),
    "sprg_batch_destroy": (_c_int, [_vp]),
    "sprg_batch_size": (_c_int, [_vp, _c_int, _c_int, _vp, _vp, _c_dbl, _c_dbl, _c_dbl, _vp, _c_int,
                                 _vp]),
    "sprg_smb_counts": (_c_int, [ctypes.c_char_p, _vp, _vp]),
    "sprg_smb_read": (_c_int, [ctypes.c_char_p, _vp, _vp, _vp, _vp]),
    "sprg_mesh_load_smb": (_c_int, [_vp, ctypes.c_char_p, _c_int, _pp]),
    "sprg_box_counts": (_c_int, [_c_int, _vp, _vp, _vp]),
    "sprg_box_build": (_c_int, [_c_int, _c_int, _c_dbl, ctypes.c_uint64, _c_int, _vp, _vp, _vp,
                                _vp]),
}

_LIB = None


class SPRError(RuntimeError):
    """A nonzero status of the C ABI (the reference would abort instead)."""

    def __init__(self, code, msg):
        super().__init__("sprgpu error %d: %s" % (code, msg))
        self.code = code


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libsprgpu.so is not built (run make -C core_amd/csrc or "
                              "__graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc):
    if rc != 0:
        raise SPRError(rc, lib().sprg_last_error().decode(errors="replace"))
    return rc
