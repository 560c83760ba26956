"""ctypes access to the block driver (libmr_engine.so) that runs MuhRec's
ReconEngine block loop over the plugin boundary of libMuhRecHIPModules.so
(see include/muhrec_modules.h)."""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ENGINE_PATH = os.path.join(_HERE, "lib", "libmr_engine.so")
MODULES_PATH = os.path.join(_HERE, "lib", "libMuhRecHIPModules.so")

_lib = None


class EngineError(RuntimeError):
    pass


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(ENGINE_PATH):
            raise RuntimeError(f"{ENGINE_PATH} is missing: run imagingsuite_amd.build.build()")
        L = ctypes.CDLL(ENGINE_PATH)
        vp, sz, fp = ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_float)
        L.mre_engine_create.argtypes = [ctypes.POINTER(vp)] + [ctypes.c_char_p] * 6
        L.mre_engine_create.restype = ctypes.c_int
        L.mre_engine_run.argtypes = [vp, fp, sz, sz, sz, fp, fp, fp]
        L.mre_engine_run.restype = ctypes.c_int
        L.mre_engine_destroy.argtypes = [vp]
        L.mre_last_error.restype = ctypes.c_char_p
        L.mre_module_list.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, sz]
        L.mre_module_list.restype = ctypes.c_int
        L.mre_engine_phase_times.argtypes = [vp, ctypes.POINTER(ctypes.c_double), sz]
        L.mre_engine_phase_times.restype = ctypes.c_int
        L.mre_engine_set_references.argtypes = [vp, fp, fp, sz, sz, ctypes.c_float, ctypes.c_float]
        L.mre_engine_set_references.restype = ctypes.c_int
        L.mre_engine_set_doses.argtypes = [vp, fp, sz]
        L.mre_engine_set_doses.restype = ctypes.c_int
        _lib = L
    return _lib


def _kv(d: Optional[Dict[str, object]]) -> bytes:
    if not d:
        return b""
    return ";".join(f"{k}={v}" for k, v in d.items()).encode()


def module_list(application: str, module_so: str = MODULES_PATH) -> Dict[str, Dict[str, str]]:
    """GetModuleList of a module library as {name: {key: default}}."""
    buf = ctypes.create_string_buffer(1 << 16)
    rc = load().mre_module_list(module_so.encode(), application.encode(), buf, len(buf))
    if rc not in (0, -1):
        raise EngineError(load().mre_last_error().decode())
    out = {}
    for item in buf.value.decode().split(";"):
        if not item:
            continue
        name, _, rest = item.partition(":")
        out[name] = dict(kv.split("=", 1) for kv in rest.split(",") if kv)
    out["__rc__"] = rc
    return out


def device_handoffs(module_so: str = MODULES_PATH) -> int:
    """Stacks the module library's parallel-beam back-projectors took from
    ProjectionFilterSingle's device copy (keepondevice) since it was loaded."""
    L = ctypes.CDLL(module_so)   # the same loaded instance the engine opened
    L.mr_modules_device_handoffs.restype = ctypes.c_size_t
    return int(L.mr_modules_device_handoffs())


class Engine:
    """One pre-processing module + one back-projector, loaded by name."""

    def __init__(self, bp_module: str, bp_params: Dict[str, object], config: Dict[str, object],
                 preproc_module="", preproc_params=None, module_so: str = MODULES_PATH):
        """preproc_module: one module name or a list (the pre-processing
        chain, in order); preproc_params: its parameter dict, or a list of
        dicts, one per module."""
        L = load()
        self._h = ctypes.c_void_p()
        if isinstance(preproc_module, (list, tuple)):
            preproc_module = "|".join(preproc_module)
        if isinstance(preproc_params, (list, tuple)):
            pp = b"|".join(_kv(d) for d in preproc_params)
        else:
            pp = _kv(preproc_params)
        rc = L.mre_engine_create(ctypes.byref(self._h), module_so.encode(), preproc_module.encode(),
                                 pp, bp_module.encode(), _kv(bp_params), _kv(config))
        if rc != 0:
            raise EngineError(L.mre_last_error().decode())
        self.config = config

    def run(self, projections: np.ndarray, angles, weights, slices: int, out: Optional[np.ndarray] = None):
        """projections [n_proj, rows, width] -> volume [slices, width, width] (x fastest),
        written into ``out`` when given (C-contiguous float32 of that shape)."""
        p = np.ascontiguousarray(projections, np.float32)
        n, rows, width = p.shape
        a = np.ascontiguousarray(angles, np.float32)
        w = np.ascontiguousarray(weights, np.float32)
        if out is None:
            vol = np.zeros((slices, width, width), np.float32)
        else:
            if out.shape != (slices, width, width) or out.dtype != np.float32 or not out.flags.c_contiguous:
                raise ValueError("out must be a C-contiguous float32 array of shape (slices, width, width)")
            vol = out
        fp = ctypes.POINTER(ctypes.c_float)
        rc = load().mre_engine_run(self._h, p.ctypes.data_as(fp), n, rows, width, a.ctypes.data_as(fp),
                                   w.ctypes.data_as(fp), vol.ctypes.data_as(fp))
        if rc != 0:
            raise EngineError(load().mre_last_error().decode())
        return vol

    def set_references(self, dark: Optional[np.ndarray], flat: Optional[np.ndarray], dark_dose: float,
                       flat_dose: float):
        """FullLogNorm's averaged references ([rows, width] float32 or None)
        and doses, for every FullLogNorm module of the chain."""
        ref = dark if dark is not None else flat
        rows, width = ref.shape
        fp = ctypes.POINTER(ctypes.c_float)
        d = None if dark is None else np.ascontiguousarray(dark, np.float32)
        f = None if flat is None else np.ascontiguousarray(flat, np.float32)
        rc = load().mre_engine_set_references(self._h, None if d is None else d.ctypes.data_as(fp),
                                              None if f is None else f.ctypes.data_as(fp), rows, width,
                                              float(dark_dose), float(flat_dose))
        if rc != 0:
            raise EngineError(load().mre_last_error().decode())

    def set_doses(self, dose):
        """Per-projection doses (sent to the modules as 6-digit text)."""
        d = np.ascontiguousarray(dose, np.float32)
        rc = load().mre_engine_set_doses(self._h, d.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), d.size)
        if rc != 0:
            raise EngineError(load().mre_last_error().decode())

    PHASES = ("copy_in", "preprocess", "set_roi", "backproject", "get_slice")

    def phase_times(self) -> Dict[str, float]:
        """Wall time (ms) of the last run's phases, summed over slabs."""
        a = (ctypes.c_double * 5)()
        load().mre_engine_phase_times(self._h, a, 5)
        return dict(zip(self.PHASES, (round(v, 1) for v in a)))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            load().mre_engine_destroy(self._h)
            self._h = ctypes.c_void_p()

    __del__ = close
