"""ctypes binding of libpisco.so (the extern "C" boundary of include/pisco.h).

The library is loaded from the package directory.  There is no fallback: if
libpisco.so is missing or cannot be loaded, importing this module raises, so
that nothing silently runs another implementation.

torch (when installed) is imported first so that the library binds to the HIP
runtime torch already loaded (both use the soname libamdhip64.so.7); device
pointers of torch tensors are then valid for the library.
"""
import ctypes
import os

try:  # share torch's HIP runtime if torch is present
    import torch  # noqa: F401
except Exception:  # pragma: no cover - torch is optional for host-only use
    torch = None

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpisco.so")

PISCO_OK = 0
PISCO_ERR_INVALID = 1
PISCO_ERR_STATE = 2
PISCO_ERR_MEMORY = 3
PISCO_ERR_DEVICE = 4

c_int, c_long, c_double, c_float, c_char, c_void_p, c_size_t = (
    ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_float, ctypes.c_char,
    ctypes.c_void_p, ctypes.c_size_t)
P = c_void_p  # all array arguments are passed as raw addresses


class RunConfig(ctypes.Structure):
    """CUDACONV::RunConfig (src/GpuConvolver/cuda/cudaConv2.h:13-26)."""
    _fields_ = [("nStreams", c_int), ("maxMemUsage", c_size_t), ("deviceId", c_int),
                ("gridSizeX", c_int), ("gridSizeY", c_int), ("blockSizeX", c_int),
                ("blockSizeY", c_int), ("ptgPerConv", c_int), ("pixelsPerDisc", c_int)]


class BatchPair(ctypes.Structure):
    """pisco_batch_pair (include/pisco.h)."""
    _fields_ = [("nside_beam", c_int), ("npix_beam", c_long), ("beam8", c_void_p),
                ("pa_offset", ctypes.c_float)]


# name -> (restype, argtypes); mirrors include/pisco.h one to one
SIGNATURES = {
    "pisco_last_error": (ctypes.c_char_p, []),
    "pisco_version": (c_int, [ctypes.c_char_p, c_int]),
    "pisco_create": (c_int, [ctypes.POINTER(RunConfig), ctypes.POINTER(c_void_p)]),
    "pisco_destroy": (None, [c_void_p]),
    "pisco_freeze_buffers": (c_int, [c_void_p, c_int]),
    "pisco_update_sky": (c_int, [c_void_p, c_int, P, P, P, P]),
    "pisco_update_beam": (c_int, [c_void_p, c_int, c_long, P, P]),
    "pisco_update_scan": (c_int, [c_void_p, c_long, P, P, P]),
    "pisco_exec_convolution": (c_int, [c_void_p, c_char, P, P]),
    "pisco_exec_adjoint_map": (c_int, [c_void_p, P, P, P, P, P]),
    "pisco_exec_adjoint_beam": (c_int, [c_void_p, P, P, P, P]),
    "pisco_set_tuning": (c_int, [c_void_p, ctypes.c_char_p, c_long]),
    "pisco_get_tuning": (c_int, [c_void_p, ctypes.c_char_p, ctypes.POINTER(c_long)]),
    "pisco_dev_pack_beam": (c_int, [c_void_p, c_void_p, c_long, P, P]),
    "pisco_dev_forward": (c_int, [c_void_p, c_void_p, c_int, P, P, P, c_int, c_long, P,
                                  c_long, P, P, P, P, P]),
    "pisco_dev_forward_noalloc": (c_int, [c_void_p, c_int, P, P, P, c_int, c_long, P, c_long, P,
                                          P, P, P, P]),
    "pisco_dev_forward_v": (c_int, [c_void_p, c_void_p, c_int, P, P, P, P, c_int, c_long, P,
                                    c_long, P, P, P, P, P]),
    "pisco_dev_adjoint_map": (c_int, [c_void_p, c_void_p, c_int, c_int, c_long, P, c_long,
                                      P, P, P, P, P, P, P, P]),
    "pisco_dev_adjoint_beam": (c_int, [c_void_p, c_void_p, c_int, P, P, P, c_int, c_long,
                                       c_long, P, P, P, P, P, P]),
    "pisco_batch_create": (c_int, [c_void_p, c_int, ctypes.POINTER(BatchPair),
                                   ctypes.POINTER(c_void_p)]),
    "pisco_batch_destroy": (None, [c_void_p]),
    "pisco_dev_batched_forward": (c_int, [c_void_p, c_void_p, c_void_p, c_int, P, P, P, c_long,
                                          P, P, P, P]),
    "pisco_dev_batched_adjoint_map": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_long,
                                              P, P, P, P, P, P, P]),
    "pisco_dev_batched_forward_pp": (c_int, [c_void_p, c_void_p, c_void_p, c_int, P, P, P,
                                             c_long, P, P, P, P]),
    "pisco_dev_batched_adjoint_map_pp": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_long,
                                                 P, P, P, P, P, P, P]),
    "pisco_skymap_create": (c_int, [c_int, c_int, ctypes.POINTER(c_void_p)]),
    "pisco_skymap_destroy": (None, [c_void_p]),
    "pisco_skymap_reset": (c_int, [c_void_p]),
    "pisco_skymap_set_chunk": (c_int, [c_void_p, c_long]),
    "pisco_skymap_accumulate": (c_int, [c_void_p, c_long, P, P, P, c_int, P, P]),
    "pisco_dev_skymap_accumulate": (c_int, [c_void_p, c_void_p, c_long, P, P, P, c_int, c_int,
                                            P, P]),
    "pisco_skymap_skipped": (c_int, [c_void_p, ctypes.POINTER(c_long)]),
    "pisco_skymap_solve": (c_int, [c_void_p, P, P, P, P, ctypes.POINTER(c_long)]),
    "pisco_dev_skymap_solve": (c_int, [c_void_p, c_void_p, P, P, P, P, ctypes.POINTER(c_long)]),
    "pisco_skymap_get_normal": (c_int, [c_void_p, P, P]),
    "pisco_dev_focal_plane_to_sky": (c_int, [c_void_p, c_void_p, c_int, P, P, P, c_long, P, P, P,
                                             c_int, P, P, P, P, P, P]),
    "pisco_focal_plane_to_sky": (c_int, [c_void_p, c_int, P, P, P, c_long, P, P, P, c_int, P,
                                         P, P]),
    "pisco_read_txt_columns": (c_int, [ctypes.c_char_p, c_int, c_long, P,
                                       ctypes.POINTER(c_long)]),
    "pisco_write_txt_columns": (c_int, [ctypes.c_char_p, c_int, c_long, P]),
    "pisco_dev_pix2ang": (c_int, [c_void_p, c_void_p, c_int, c_long, P, P, P]),
    "pisco_dev_ang2pix": (c_int, [c_void_p, c_void_p, c_int, c_long, P, P, P]),
    "pisco_dev_get_interpol": (c_int, [c_void_p, c_void_p, c_int, c_long, P, P, P, P]),
    "pisco_dev_query_disc": (c_int, [c_void_p, c_void_p, c_int, c_long, P, P, c_double, P,
                                     P, P]),
    "pisco_dev_match": (c_int, [c_void_p, c_void_p, c_int, c_int, c_long, P, P, P, P, P,
                                P, P]),
    "pisco_dev_match_mode": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_long, P, P,
                                     P, P, P, P, P]),
    "pisco_dev_libm": (c_int, [c_void_p, c_void_p, c_int, c_long, P, P, P]),
    "pisco_beam_rho_max": (c_double, [c_int, c_long]),
    "pisco_ang2pix": (c_long, [c_int, c_double, c_double]),
    "pisco_pix2ang": (c_int, [c_int, c_long, P, P]),
    "pisco_band_annulus": (c_long, [c_int, c_int, c_long, c_int, c_int, c_float, c_float, c_int,
                                    c_long, P]),
    "pisco_beam_gaussian_fields": (c_int, [c_int, c_long, c_double, c_double, c_double, P,
                                           P, P, P]),
    "pisco_beam_from_fields": (c_int, [c_long, P, P, P, P, P, P, P, P]),
    "pisco_polbeam_build": (c_int, [c_long, c_int, c_int, c_double, c_char, P, P]),
}


class PiscoError(RuntimeError):
    """An error status returned by libpisco; .code is the PISCO_ERR_* value."""

    def __init__(self, code, msg):
        super().__init__("pisco error %d: %s" % (code, msg))
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libpisco.so is not built (%s); run `python -m pisco4cmb_amd.build` "
            "first" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(rc):
    """Raise PiscoError for a non-zero status."""
    if rc != PISCO_OK:
        msg = lib.pisco_last_error()
        raise PiscoError(rc, msg.decode() if msg else "")
    return rc


def ptr_array(ptrs):
    """A C array of pointers (None -> NULL) from integer addresses."""
    arr = (c_void_p * len(ptrs))()
    for i, p in enumerate(ptrs):
        arr[i] = p if p else None
    return arr
