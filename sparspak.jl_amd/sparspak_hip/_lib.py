"""ctypes binding of libsparspak_hip.so (declarations in include/sparspak_hip.h).

The library is built in-tree by ``sparspak.jl_amd/build.py`` (hipcc, gfx950).
Loading fails loudly when it is missing: there is no CPU fallback for the
numeric phase.
"""
import ctypes
import importlib.util
import os
import sys
import warnings

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsparspak_hip.so")

i64 = ctypes.c_int64
i64p = ctypes.POINTER(ctypes.c_int64)
dp = ctypes.POINTER(ctypes.c_double)
fp = ctypes.POINTER(ctypes.c_float)
vp = ctypes.c_void_p


class Stats(ctypes.Structure):
    _fields_ = [
        ("n", i64), ("nsuper", i64), ("nfront", i64), ("nlevel", i64),
        ("nnz_lnz", i64), ("nnz_unz", i64),
        ("first_fail_supernode", i64), ("refine_iters", i64),
        ("factor_flops", ctypes.c_double), ("solve_flops", ctypes.c_double),
        ("front_bytes", ctypes.c_double),
        ("t_upload_ms", ctypes.c_double), ("t_inmatrix_ms", ctypes.c_double),
        ("t_factor_ms", ctypes.c_double), ("t_solve_ms", ctypes.c_double),
        ("t_refine_ms", ctypes.c_double),
        ("backward_error", ctypes.c_double),
    ]


class KStat(ctypes.Structure):
    _fields_ = [("launches", i64), ("ms", ctypes.c_double), ("flops", ctypes.c_double),
                ("bytes", ctypes.c_double)]


class Tuning(ctypes.Structure):
    _fields_ = [("small_lmax", i64), ("block_cols", i64), ("upd_group", i64),
                ("small_wave_min", i64), ("factor_ms", ctypes.c_double)]


TUNING_KNOBS = ("small_lmax", "block_cols", "upd_group", "small_wave_min")

KCLASSES = ["assemble", "small_front", "large_init", "large_diag", "large_update",
            "large_writeback", "fwd_solve", "bwd_solve", "residual", "large_offdiag",
            "fwd_gemv", "bwd_gemv", "large_diag_border"]

# name -> (restype, argtypes)
_SIGS = {
    # host symbolic phase
    "spk_isstructuresymmetric": (i64, [i64, i64p, i64p]),
    "spk_symmetrized_nedges": (i64, [i64, i64p, i64p]),
    "spk_makestructuresymmetric": (i64, [i64, i64p, i64p, i64p, i64p]),
    "spk_mmd": (i64, [i64, i64p, i64p, i64p, i64p]),
    "spk_etree": (None, [i64, i64p, i64p, i64p, i64p, i64p]),
    "spk_postorder": (None, [i64, i64p, i64p, i64p, i64p]),
    "spk_colcounts": (i64, [i64, i64p, i64p, i64p, i64p, i64p, i64p]),
    "spk_supernodes": (i64, [i64, i64p, i64p, i64, i64p, i64p, i64p]),
    "spk_nonzero_indexs": (None, [i64, i64p, i64, i64p, i64p, i64p]),
    "spk_nonzero_indexs_spd": (None, [i64, i64p, i64, i64p, i64p]),
    "spk_symbfact": (i64, [i64, i64p, i64p, i64p, i64p, i64p, i64, i64p,
                           i64p, i64, i64p, i64p]),
    "spk_supernode_tree": (None, [i64, i64p, i64p, i64p, i64p, i64p]),
    "spk_front_maps": (i64, [i64, i64, i64p, i64p, i64p, i64p, i64p, i64p, i64p, i64p,
                             i64p]),
    "spk_amap": (i64, [i64, i64, i64, i64p, i64p, i64p, i64p, i64p, i64p, i64p, i64p,
                       i64p, i64, i64p]),
    # device numeric phase
    "spkhip_create": (i64, [ctypes.POINTER(vp), i64]),
    "spkhip_destroy": (i64, [vp]),
    "spkhip_symbolic_upload": (i64, [vp, i64, i64, i64, i64p, i64p, i64p,
                                     i64p, i64p, i64p, i64p, i64p, i64p,
                                     i64p]),
    "spkhip_inmatrix": (i64, [vp, dp]),
    "spkhip_inmatrix_device": (i64, [vp, vp]),
    "spkhip_factor": (i64, [vp, i64]),
    "spkhip_solve": (i64, [vp, i64, dp, i64]),
    "spkhip_solve_device": (i64, [vp, i64, vp, i64]),
    "spkhip_solve_device_noalloc": (i64, [vp, i64, vp, i64]),
    "spkhip_solve_refine": (i64, [vp, i64, dp, i64, ctypes.c_double, i64]),
    "spkhip_solve_refine_device": (i64, [vp, i64, vp, i64, ctypes.c_double,
                                         i64]),
    "spkhip_set_segments": (i64, [vp, i64, i64p]),
    "spkhip_set_value_sets": (i64, [vp, i64]),
    "spkhip_condest": (i64, [vp, dp, dp]),
    "spkhip_set_caller_stream": (i64, [vp, vp]),
    "spkhip_download": (i64, [vp, dp, dp, i64p]),
    "spkhip_download_f32": (i64, [vp, fp, fp, i64p]),
    "spkhip_stats": (i64, [vp, ctypes.POINTER(Stats)]),
    "spkhip_lufactor_f64": (i64, [i64, i64, i64p, i64p, i64p, i64p, i64p,
                                  dp, i64p, dp, i64p]),
    "spkhip_lufactor_f32": (i64, [i64, i64, i64p, i64p, i64p, i64p, i64p,
                                  fp, i64p, fp, i64p]),
    "spkhip_lulsolve_f64": (i64, [i64, i64p, i64p, i64p, i64p, dp, i64p,
                                  dp]),
    "spkhip_lulsolve_f32": (i64, [i64, i64p, i64p, i64p, i64p, fp, i64p,
                                  fp]),
    "spkhip_luusolve_f64": (i64, [i64, i64, i64p, i64p, i64p, i64p, dp,
                                  i64p, dp, dp]),
    "spkhip_luusolve_f32": (i64, [i64, i64, i64p, i64p, i64p, i64p, fp,
                                  i64p, fp, fp]),
    "spkhip_ldltfactor_f64": (i64, [i64, i64, i64p, i64p, i64p, i64p, i64p,
                                    dp]),
    "spkhip_ldltfactor_f32": (i64, [i64, i64, i64p, i64p, i64p, i64p, i64p,
                                    fp]),
    "spkhip_ldltsolve_f64": (i64, [i64, i64p, i64p, i64p, i64p, dp, dp]),
    "spkhip_ldltsolve_f32": (i64, [i64, i64p, i64p, i64p, i64p, fp, fp]),
    "spkhip_profile_enable": (i64, [vp, i64]),
    "spkhip_profile_get": (i64, [vp, i64, ctypes.POINTER(KStat)]),
    "spkhip_get_tuning": (i64, [vp, ctypes.POINTER(Tuning)]),
    "spkhip_set_tuning": (i64, [vp, ctypes.POINTER(Tuning)]),
    "spkhip_autotune": (i64, [vp, i64, i64, ctypes.POINTER(Tuning)]),
    "spkhip_device_count": (i64, []),
    "spkhip_version": (ctypes.c_char_p, []),
}

_lib = None


def _share_torch_hip_runtime():
    """One HIP runtime per process. PyTorch for ROCm bundles its own
    libamdhip64 (and ROCr), which its libraries reach under the unversioned
    name libamdhip64.so; this library asks for libamdhip64.so.7. If this
    library is loaded first, the dynamic linker maps /opt/rocm's runtime, a
    later `import torch` maps torch's copy beside it, and the second runtime
    to initialise finds no GPU. When torch is installed, its runtime is
    therefore mapped first (by path, without importing torch): its SONAME is
    libamdhip64.so.7, so this library binds to it, and torch later finds the
    same file already mapped."""
    if "torch" in sys.modules:
        return
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return
    if spec is None or not spec.submodule_search_locations:
        return
    for d in spec.submodule_search_locations:
        p = os.path.join(d, "lib", "libamdhip64.so")
        if os.path.exists(p):
            ctypes.CDLL(p, mode=ctypes.RTLD_GLOBAL)
            return


def lib():
    """Load libsparspak_hip.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with "
                "`python sparspak.jl_amd/build.py` (hipcc, gfx950)")
        _share_torch_hip_runtime()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def as_i64(a):
    """Contiguous int64 numpy array (no copy when already one)."""
    return np.ascontiguousarray(a, dtype=np.int64)


def p64(a):
    return a.ctypes.data_as(i64p)


def pd(a):
    return a.ctypes.data_as(dp)


def pf(a):
    return a.ctypes.data_as(fp)


class HipError(RuntimeError):
    pass


class RefinementWarning(RuntimeWarning):
    """Iterative refinement stopped above the requested backward error."""


def check(rc, what):
    """Translate a library return code into an exception (0 passes)."""
    rc = int(rc)
    if rc == 0:
        return rc
    if rc <= -1000:
        raise HipError(f"{what}: HIP error {-1000 - rc}")
    raise RuntimeError(f"{what}: error code {rc}")


def check_refine(rc, what, backward_error):
    """Return code of a refined solve: -5 (tolerance not reached within
    max_iter corrections; the solution is still returned) warns with the
    backward error reached, -6 (NaN or Inf in the solution) raises
    FloatingPointError, anything else as check()."""
    rc = int(rc)
    if rc == -5:
        warnings.warn(f"{what}: iterative refinement stopped at backward error "
                      f"{backward_error:.3e}, above the requested tolerance",
                      RefinementWarning, stacklevel=3)
        return rc
    if rc == -6:
        raise FloatingPointError(f"{what}: the solution is not finite (NaN or Inf)")
    return check(rc, what)
