"""ctypes binding of the engine's C ABI (include/mhre.h).

Library resolution (first hit wins), mirroring the reference's one-build-per-
project workflow where ``make`` in a project directory produces the module
with that project's model.cuh compiled in:

1. ``$MHRE_LIB`` -- explicit path;
2. ``./libmhre.so`` in the current directory (a project built with
   ``python -m cuda_mcmc_amd.build --model-dir .``);
3. the package default ``cuda_mcmc_amd/lib/libmhre.so`` (ring+Gaussian
   project, native fused kernels for imodel 1 and 2).

There is no fallback: a missing or unloadable library raises.
"""
import ctypes as C
import os
import threading

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
DEFAULT_LIB = os.path.join(_PKG, "lib", "libmhre.so")

MHRE_MODEL_AUTO = 0
MHRE_MODEL_SGRA9 = 1
MHRE_MODEL_SGRA13 = 2
MHRE_MODEL_USER = 10
MHRE_MODEL_QUADSYS = 20

fp = C.POINTER(C.c_float)
ip = C.POINTER(C.c_int)
up = C.POINTER(C.c_ubyte)


class MhreFitSpec(C.Structure):
    _fields_ = [
        ("model", C.c_int), ("imodel", C.c_int),
        ("coor", fp), ("dat", fp), ("std2r", fp),
        ("icoor", ip), ("idat", ip),
        ("datm", fp), ("chi2m", fp),
        ("pdescr", ip), ("ivar", ip), ("invar", ip),
        ("ptotal", fp), ("pmint", fp), ("pmaxt", fp),
        ("pout", fp), ("tout", ip), ("chi2", fp),
        ("chi2c", fp), ("tcur", ip), ("flag", ip),
        ("n_acpt", ip), ("n_exch", ip), ("n_cnt", ip), ("n_hist", ip),
        ("beta", fp), ("pstp", fp), ("ptent", fp), ("ptentn", ip),
        ("pcur", fp), ("rndst", up),
        ("seed", C.c_ulonglong),
        ("ncoor", C.c_int), ("ndat", C.c_int), ("nicoor", C.c_int), ("nidat", C.c_int), ("ndatm", C.c_int),
        ("nptot", C.c_int), ("nprm", C.c_int), ("nadj", C.c_int), ("npass", C.c_int),
        ("nbeta", C.c_int), ("nseq", C.c_int), ("nburn", C.c_int), ("niter", C.c_int),
        ("adapt_target", C.c_float), ("init_chains", C.c_int), ("it0", C.c_int),
    ]


class MhreOptions(C.Structure):
    _fields_ = [
        ("waves_per_block", C.c_int), ("persistent", C.c_int), ("graph_iters", C.c_int),
        ("trace_iters", C.c_int), ("trace_dec", ip), ("trace_chi2", fp), ("trace_ex", ip),
        ("force_generic", C.c_int), ("waves_per_simd", C.c_int), ("streams", C.c_int),
        ("exchange_launch", C.c_int),
    ]


class MhreStats(C.Structure):
    _fields_ = [
        ("run_ms", C.c_double), ("chi2_evals", C.c_ulonglong), ("proposals", C.c_ulonglong),
        ("vis_evals", C.c_double), ("cph_evals", C.c_double), ("iters_done", C.c_int), ("nfits", C.c_int),
    ]


_libs = {}
_lock = threading.Lock()


def lib_path(path=None):
    if path:
        return os.path.abspath(path)
    env = os.environ.get("MHRE_LIB")
    if env:
        return os.path.abspath(env)
    local = os.path.abspath("libmhre.so")
    if os.path.isfile(local):
        return local
    return DEFAULT_LIB


def _bring_up_torch_runtime():
    """Load torch's HIP runtime before the engine library, when torch exists.

    torch ships its own libamdhip64 with the same soname as the system one
    the engine links.  Whichever is loaded first serves the whole process,
    and torch's GPU support does not start on a runtime other than its own,
    so the engine binds to torch's (the gradient interface hands torch
    device buffers to the engine).  Without torch this does nothing.
    """
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load(path=None):
    """Load (once) and return the engine library."""
    p = lib_path(path)
    with _lock:
        if p in _libs:
            return _libs[p]
        if not os.path.isfile(p):
            raise RuntimeError("MH-RE engine library not found at %s; build it with "
                               "`python -m cuda_mcmc_amd.build`" % p)
        _bring_up_torch_runtime()
        L = C.CDLL(p)
        _declare(L)
        _libs[p] = L
        return L


def _declare(L):
    L.mcmc_cuda.restype = C.c_int
    L.mcmc_cuda.argtypes = ([fp, fp, fp, ip, ip, fp, fp, ip, ip, ip, fp, fp, fp, fp, ip, fp, fp, ip, ip,
                             ip, ip, ip, ip, fp, fp, fp, ip, fp, up, C.c_ulonglong] + [C.c_int] * 14)
    L.calc_modchi2.restype = C.c_int
    L.calc_modchi2.argtypes = [fp, fp, fp, ip, ip, fp, fp, fp, fp, ip] + [C.c_int] * 9
    L.calc_model.restype = C.c_int
    L.calc_model.argtypes = [fp, ip, fp, fp] + [C.c_int] * 8
    L.reset_gpu.restype = C.c_int
    L.reset_gpu.argtypes = []
    L.list_gpus.restype = C.c_int
    L.list_gpus.argtypes = [C.c_char_p, C.c_int]
    L.mhre_last_error.restype = C.c_char_p
    L.mhre_version.restype = C.c_int
    L.mhre_user_model.restype = C.c_char_p
    L.mhre_native_eht.restype = C.c_int
    vp = C.c_void_p
    L.mhre_session_create.restype = C.c_int
    L.mhre_session_create.argtypes = [C.POINTER(MhreFitSpec), C.c_int, C.POINTER(MhreOptions), C.POINTER(vp)]
    L.mhre_session_run.restype = C.c_int
    L.mhre_session_run.argtypes = [vp, C.c_int]
    L.mhre_session_time_kernels.restype = C.c_int
    L.mhre_session_time_kernels.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.mhre_session_autotune.restype = C.c_int
    L.mhre_session_autotune.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.mhre_session_sync.restype = C.c_int
    L.mhre_session_sync.argtypes = [vp]
    L.mhre_session_download.restype = C.c_int
    L.mhre_session_download.argtypes = [vp, C.c_int]
    L.mhre_session_stats.restype = C.c_int
    L.mhre_session_stats.argtypes = [vp, C.POINTER(MhreStats)]
    L.mhre_session_fit_info.restype = C.c_int
    L.mhre_session_fit_info.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.c_int]
    L.mhre_session_destroy.restype = C.c_int
    L.mhre_session_destroy.argtypes = [vp]
    L.mhre_run_batch.restype = C.c_int
    L.mhre_run_batch.argtypes = [C.POINTER(MhreFitSpec), C.c_int, C.POINTER(MhreOptions), C.POINTER(MhreStats)]
    L.mhre_eht_data_create.restype = C.c_int
    L.mhre_eht_data_create.argtypes = [fp, fp, fp, ip, ip] + [C.c_int] * 6 + [C.POINTER(vp)]
    L.mhre_eht_chi2_grad.restype = C.c_int
    L.mhre_eht_chi2_grad.argtypes = [vp, vp, C.c_int, C.c_int, vp, vp, vp]
    L.mhre_eht_gn_system.restype = C.c_int
    L.mhre_eht_gn_system.argtypes = [vp, vp, C.c_int, C.c_int, vp, vp, vp, vp]
    L.mhre_eht_data_destroy.restype = C.c_int
    L.mhre_eht_data_destroy.argtypes = [vp]
    L.mhre_eht_data_info.restype = C.c_int
    L.mhre_eht_data_info.argtypes = [vp, C.POINTER(C.c_int), C.c_int]
    L.mhre_quad_jacobian.restype = C.c_int
    L.mhre_quad_jacobian.argtypes = [fp, fp, fp, ip, C.c_int, C.c_int, C.c_int, fp, C.c_int, fp, fp]
    L.mhre_quad_jacobian_noalloc.restype = C.c_int
    L.mhre_quad_jacobian_noalloc.argtypes = [fp, fp, fp, ip, C.c_int, C.c_int, C.c_int, fp, C.c_int, fp, fp]


class EngineError(RuntimeError):
    pass


def check(L, rc, what):
    if rc < 0:
        raise EngineError("%s failed (%d): %s" % (what, rc, L.mhre_last_error().decode(errors="replace")))
    return rc


def ptr(a, ctype):
    """Pointer to a C-contiguous numpy array (None -> NULL)."""
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    return a.ctypes.data_as(C.POINTER(ctype))


def need(a, dtype, ndim, name):
    """The reference's Cython signature checks (src/mcmc_interf.pyx:232-263) plus contiguity."""
    if not isinstance(a, np.ndarray):
        raise TypeError("%s must be a numpy array" % name)
    if a.dtype != np.dtype(dtype):
        raise TypeError("%s must have dtype %s, got %s" % (name, np.dtype(dtype), a.dtype))
    if ndim is not None and a.ndim != ndim:
        raise ValueError("%s must be %d-dimensional, got %d" % (name, ndim, a.ndim))
    if not a.flags.c_contiguous:
        raise ValueError("%s must be C-contiguous" % name)
    if not a.flags.writeable:
        raise ValueError("%s must be writeable" % name)
    return a
