"""ctypes binding of libmlcp_hip.so (the C-ABI declared in include/mlcp_hip.h).

There is no fallback: if the shared object is missing or cannot be loaded,
``lib()`` raises.  Build it with ``python -m mlcp_amd.build``.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmlcp_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "mlcp_hip.h")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i, _d, _v = C.c_int, C.c_double, C.c_void_p
_pi = C.POINTER(C.c_int)
_pd = C.POINTER(C.c_double)

_lib = None


def declared_symbols():
    """Function names declared in include/mlcp_hip.h."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", txt)) -
                  {"if", "defined"})


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built; run `python -m mlcp_amd.build`")
    L = C.CDLL(LIB_PATH)
    L.mlcp_hip_last_error.restype = C.c_char_p
    L.mlcp_hip_last_error.argtypes = []
    L.mlcp_hip_device_bytes.restype = C.c_long
    L.mlcp_hip_device_bytes.argtypes = []
    L.mlcp_hip_set_error_mode.restype = None
    L.mlcp_hip_set_error_mode.argtypes = [_i]
    # this host checks mlcp_hip_last_error() after every drop-in call, so a
    # failed intertwinecpp_/alscpp_ returns instead of ending the process
    L.mlcp_hip_set_error_mode(0)
    L.intertwinecpp_.restype = None
    L.intertwinecpp_.argtypes = [_dp, _dp, _ip, _pi, _pi, _pi, _ip, _ip, _dp,
                                 _dp, _pi, _pi, _pi, _dp, _pi, _pi, _pi, _pi,
                                 _pi, _pd]
    L.alscpp_.restype = None
    L.alscpp_.argtypes = [_dp, _dp, _ip, _pi, _pi, _pi, _pi]
    L.intwdevicemem_.restype = None
    L.intwdevicemem_.argtypes = [_pi] * 8 + [_ip]
    L.mlcp_hip_intertwine.restype = _i
    L.mlcp_hip_intertwine.argtypes = [_dp, _dp, _ip, _i, _i, _ip, _ip, _ip,
                                      _dp, _dp, _i, _i, _i, _dp, _i, _i, _i,
                                      _i, _d]
    L.mlcp_hip_session_create.restype = _i
    L.mlcp_hip_session_create.argtypes = [C.POINTER(_v), _dp, _dp, _ip, _i, _i,
                                          _ip, _ip, _ip, _dp, _dp, _i, _i, _i]
    L.mlcp_hip_session_cycle.restype = _i
    L.mlcp_hip_session_cycle.argtypes = [_v, _dp, _i, _i, _i, _i, _d]
    L.mlcp_hip_session_solve.restype = _i
    L.mlcp_hip_session_solve.argtypes = [_v, _dp, _dp, _i, _i, _i, _dp]
    L.mlcp_hip_session_download.restype = _i
    L.mlcp_hip_session_download.argtypes = [_v, _dp, _dp]
    L.mlcp_hip_session_upload.restype = _i
    L.mlcp_hip_session_upload.argtypes = [_v, _dp, _dp]
    L.mlcp_hip_session_phase_ms.restype = _i
    L.mlcp_hip_session_phase_ms.argtypes = [_v, _dp]
    L.mlcp_hip_session_destroy.restype = _i
    L.mlcp_hip_session_destroy.argtypes = [_v]
    L.mlcp_hip_session_save.restype = _i
    L.mlcp_hip_session_save.argtypes = [_v]
    L.mlcp_hip_session_restore.restype = _i
    L.mlcp_hip_session_restore.argtypes = [_v]
    L.mlcp_hip_session_probe.restype = _i
    L.mlcp_hip_session_probe.argtypes = [_v, _i, _i]
    L.mlcp_hip_session_probe_read.restype = _i
    L.mlcp_hip_session_probe_read.argtypes = [_v, _pd, C.POINTER(C.c_long)]
    L.mlcp_hip_session_power.restype = _i
    L.mlcp_hip_session_power.argtypes = [_v, _i, _i, _d]
    L.mlcp_hip_session_ortho.restype = _i
    L.mlcp_hip_session_ortho.argtypes = [_v, _i]
    L.mlcp_hip_session_prodhv.restype = _i
    L.mlcp_hip_session_prodhv.argtypes = [_v, _i, _dp, _dp]
    L.mlcp_hip_session_gep.restype = _i
    L.mlcp_hip_session_gep.argtypes = [_v, _i, _dp, _dp, _dp, _dp]
    L.mlcp_hip_session_update.restype = _i
    L.mlcp_hip_session_update.argtypes = [_v, _dp, _i]
    L.mlcp_hip_session_rayleigh.restype = _i
    L.mlcp_hip_session_rayleigh.argtypes = [_v, _dp]
    L.mlcp_hip_gep.restype = _i
    L.mlcp_hip_gep.argtypes = [_i, _dp, _dp, _dp, _dp]
    L.mlcp_hip_gep_noalloc.restype = _i
    L.mlcp_hip_gep_noalloc.argtypes = [_i, _v, _v, _v, _v, _v, _v]
    L.mlcp_hip_update_opers.restype = _i
    L.mlcp_hip_update_opers.argtypes = [_dp, _dp, _ip, _i, _i, _ip, _i, _ip,
                                        _ip, _dp, _i, _dp]
    L.mlcp_hip_session_set_form.restype = _i
    L.mlcp_hip_session_set_form.argtypes = [_v, _i, _d, _d]
    L.mlcp_hip_session_sweeps.restype = _i
    L.mlcp_hip_session_sweeps.argtypes = [_v, _ip]
    _fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    L.mlcp_mlp_create.restype = _i
    L.mlcp_mlp_create.argtypes = [C.POINTER(_v), _i, _i, _i, _i,
                                  C.c_ulonglong]
    L.mlcp_mlp_nparams.restype = C.c_long
    L.mlcp_mlp_nparams.argtypes = [_v]
    L.mlcp_mlp_params.restype = _i
    L.mlcp_mlp_params.argtypes = [_v, _fp, _i]
    L.mlcp_mlp_destroy.restype = _i
    L.mlcp_mlp_destroy.argtypes = [_v]
    L.mlcp_mlp_forward.restype = _i
    L.mlcp_mlp_forward.argtypes = [_v, _fp, _ip, _i, _fp]
    L.mlcp_mlp_loss_grad.restype = _i
    L.mlcp_mlp_loss_grad.argtypes = [_v, _fp, _ip, _fp, _i, _pd, _fp]
    L.mlcp_mlp_train.restype = _i
    L.mlcp_mlp_train.argtypes = [_v, _fp, _ip, _fp, _i, _i, _i, C.c_float,
                                 C.c_ulonglong, _dp]
    L.mlcp_hip_session_set_warmstart.restype = _i
    L.mlcp_hip_session_set_warmstart.argtypes = [_v, _v]
    L.mlcp_hip_session_tune_set.restype = _i
    L.mlcp_hip_session_tune_set.argtypes = [_v, _i, _i, _i]
    L.mlcp_hip_session_tune_get.restype = _i
    L.mlcp_hip_session_tune_get.argtypes = [_v, _ip]
    L.mlcp_hip_session_tune_key.restype = _i
    L.mlcp_hip_session_tune_key.argtypes = [_v, C.c_char_p, _i]
    _lib = L
    return L


def last_error() -> str:
    return lib().mlcp_hip_last_error().decode()


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed: {last_error()}")
