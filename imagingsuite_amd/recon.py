"""Host-side mirror of MuhRec's hot-path modules over the HIP C ABI.

The classes keep the reference modules' names, parameter keys and argument
meanings:

* ``ProjectionFilterSingle`` -- frameworks/tomography/Preprocessing/
  StdPreprocModules/src/ProjectionFilter.cpp:78-118 (Configure, ProcessCore).
* ``BackProjector`` for ``MultiProjBPparallel`` / ``MultiProjBP`` /
  ``NearestNeighborBP`` -- frameworks/tomography/Backprojectors/
  StdBackProjectors/src/StdBackProjectorBase.cpp:120-201 (SetROI, Process(3D)).
* ``FDKbp_single`` -- frameworks/tomography/Backprojectors/FDKBackProjectors/
  src/fdkreconbase.cpp:91-277.

Data stays on the GPU: inputs and outputs are torch CUDA tensors (torch is
only the device-memory/stream plumbing).  Errors from the library raise
``MrError`` (the reference throws ReconException).
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import BpParams, FdkParams, MrError, check

BP_MODULES = {
    "MultiProjBPparallel": dict(interp=0, multiprojbp=0),
    "MultiProjBP": dict(interp=0, multiprojbp=1),
    "NearestNeighborBP": dict(interp=1, multiprojbp=0),
}


def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream) -> ctypes.c_void_p:
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def _order_after(stream, *tensors):
    """The back-projection calls return while their kernels still run on
    ``stream``.  When that is not the current stream, the tensors they read
    and write are marked as used on it (torch's caching allocator then keeps
    their memory until that work is done) and the current stream waits for
    it, so what the caller queues next (``.cpu()``, further kernels) sees the
    finished result."""
    if stream is None:
        return
    import torch
    cur = torch.cuda.current_stream()
    if stream == cur:
        return
    for t in tensors:
        t.record_stream(stream)
    cur.wait_stream(stream)


def _f32arr(x, n: int) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    if a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def _require_cuda_f32(t, ndim: int, name: str):
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a CUDA tensor")
    if t.dtype != torch.float32 or not t.is_contiguous() or t.dim() != ndim:
        raise TypeError(f"{name} must be a contiguous float32 tensor with {ndim} dims")


class ProjectionFilterSingle:
    """Ramp filter along detector rows (the reference's default pre-BP module)."""

    def __init__(self, width: int, filtertype: str = "Hamming", cutoff: float = 0.5,
                 order: float = 1.0, usebias: bool = True, biasweight: float = 0.1,
                 paddingdoubler: int = 2):
        L = _lib.load()
        ft = _lib.FILTER_TYPES.get(filtertype.lower())
        if ft is None:
            raise MrError(_lib.MR_ERR_INVALID, "Could not convert filter string to enum.")
        self.width = int(width)
        self._plan = ctypes.c_void_p()
        check(L.mr_filter_plan_create(ctypes.byref(self._plan), self.width, ft, cutoff, order,
                                      int(bool(usebias)), biasweight, paddingdoubler))
        fs, ni = ctypes.c_size_t(), ctypes.c_size_t()
        check(L.mr_filter_plan_info(self._plan, ctypes.byref(fs), ctypes.byref(ni), None))
        self.fft_size, self.n_insert = fs.value, ni.value
        H = np.zeros(self.fft_size // 2, np.float32)
        check(L.mr_filter_plan_info(self._plan, None, None,
                                    H.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        self.filter = H

    @property
    def kernel(self) -> str:
        """The kernel form this plan launches (mr_filter_plan_kernel)."""
        buf = ctypes.create_string_buffer(128)
        check(_lib.load().mr_filter_plan_kernel(self._plan, buf, 128))
        return buf.value.decode()

    def process(self, proj, stream=None):
        """Filters ``proj`` ([..., width] CUDA float32) in place and returns it."""
        import torch
        if not isinstance(proj, torch.Tensor) or not proj.is_cuda or proj.dtype != torch.float32 \
                or not proj.is_contiguous():
            raise TypeError("proj must be a contiguous float32 CUDA tensor")
        if proj.numel() == 0:
            raise MrError(_lib.MR_ERR_INVALID, "Empty projection image")
        if proj.shape[-1] != self.width:
            raise MrError(_lib.MR_ERR_INVALID, "row width does not match the configured ROI")
        check(_lib.load().mr_filter_rows(self._plan, _ptr(proj), proj.numel() // self.width,
                                         _stream(stream)))
        return proj

    def adjoint(self, grad, stream=None):
        """Applies the adjoint of ``process`` to ``grad`` ([..., width] CUDA
        float32) in place (gradients through the filter)."""
        _require_cuda_f32(grad, grad.dim(), "grad")
        if grad.shape[-1] != self.width:
            raise MrError(_lib.MR_ERR_INVALID, "row width does not match the configured ROI")
        check(_lib.load().mr_filter_rows_adjoint(self._plan, _ptr(grad), grad.numel() // self.width,
                                                 _stream(stream)))
        return grad

    def close(self):
        if getattr(self, "_plan", None) and self._plan.value:
            try:
                _lib.load().mr_filter_plan_destroy(self._plan)
            except (TypeError, AttributeError):   # interpreter shutdown
                pass
            self._plan = ctypes.c_void_p()

    __del__ = close


def _set_tilt(p, correct_tilt, tilt_angle, tilt_pivot, row0=None, tilt_rows=0):
    p.correct_tilt = int(bool(correct_tilt))
    p.tilt_angle, p.tilt_pivot = float(tilt_angle), float(tilt_pivot)
    if row0 is not None:
        p.tilt_row0 = int(row0)
    p.tilt_rows = int(tilt_rows)


class BackProjector:
    """Parallel-beam back-projector (MultiProjBPparallel, MultiProjBP, NearestNeighborBP).

    ``width``/``slices`` are the block's SizeU/SizeV, ``center`` is fCenter
    (ROI-relative), ``direction`` 0 = CW, 1 = CCW, ``buffer_size`` is
    ProjectionBufferSize (16 reproduces the reference exactly, including its
    dropped projections; 0 back-projects every projection).  ``fused`` folds
    the interpolation into fused multiply-adds (float rounding differs from
    the reference; otherwise results are bit-identical).

    ``correct_tilt`` runs the module's bCorrectTilt branch with
    ``tilt_angle`` (degrees) and ``tilt_pivot`` (detector row); ``row0`` is
    the block's first detector row (ProjectionInfo.roi[1]) and ``tilt_rows``
    the projection height that shrinks the circle mask (ProjectionInfo.nDims[1],
    default ``slices``)."""

    def __init__(self, name: str, width: int, slices: int, center: float, direction: int = 0,
                 buffer_size: int = 16, roi: Optional[Sequence[int]] = None, fused: bool = False,
                 correct_tilt: bool = False, tilt_angle: float = 0.0, tilt_pivot: float = 0.0,
                 row0: int = 0, tilt_rows: Optional[int] = None):
        if name not in BP_MODULES:
            raise MrError(_lib.MR_ERR_INVALID, f"unknown back-projector {name}")
        L = _lib.load()
        p = BpParams()
        p.width, p.slices, p.center, p.direction = int(width), int(slices), float(center), int(direction)
        p.interp = BP_MODULES[name]["interp"]
        p.multiprojbp = BP_MODULES[name]["multiprojbp"]
        p.buffer_size = int(buffer_size)
        p.fused = int(bool(fused))
        _set_tilt(p, correct_tilt, tilt_angle, tilt_pivot, row0, slices if tilt_rows is None else tilt_rows)
        p.use_roi = 0 if roi is None else 1
        if roi is not None:
            for i in range(4):
                p.roi[i] = int(roi[i])
        self.name, self.width, self.slices = name, int(width), int(slices)
        self._plan = ctypes.c_void_p()
        check(L.mr_bp_plan_create(ctypes.byref(self._plan), ctypes.byref(p)))

    @property
    def mask(self):
        f = np.zeros(self.width, np.uintp)
        s = np.zeros(self.width, np.uintp)
        szp = ctypes.POINTER(ctypes.c_size_t)
        check(_lib.load().mr_bp_plan_mask(self._plan, f.ctypes.data_as(szp), s.ctypes.data_as(szp)))
        return f.astype(np.int64), s.astype(np.int64)

    @property
    def kernel(self) -> str:
        """The kernel variant this plan launches (mr_bp_plan_kernel)."""
        buf = ctypes.create_string_buffer(128)
        check(_lib.load().mr_bp_plan_kernel(self._plan, buf, 128))
        return buf.value.decode()

    def process(self, proj, angles, weights, volume=None, accumulate: bool = False, stream=None):
        """Back-projects ``proj`` [n_proj, slices, width] (CUDA) into a
        [width, width, slices] ZXY volume (allocated when not given).  The
        kernels are enqueued on ``stream`` (default: the current stream) and
        the call returns without waiting; with another stream than the
        current one, the current stream is made to wait for them and the
        tensors are recorded on ``stream``."""
        import torch
        _require_cuda_f32(proj, 3, "proj")
        n_proj = proj.shape[0]
        if proj.shape[1] != self.slices or proj.shape[2] != self.width:
            raise MrError(_lib.MR_ERR_INVALID, "projection size does not match the ROI")
        a = _f32arr(angles, n_proj)
        w = _f32arr(weights, n_proj)
        if volume is None:
            volume = torch.empty((self.width, self.width, self.slices), dtype=torch.float32,
                                 device=proj.device)
            accumulate = False
        _require_cuda_f32(volume, 3, "volume")
        fp = ctypes.POINTER(ctypes.c_float)
        check(_lib.load().mr_bp_parallel(self._plan, _ptr(proj), n_proj, a.ctypes.data_as(fp),
                                         w.ctypes.data_as(fp), _ptr(volume), int(bool(accumulate)),
                                         _stream(stream)))
        _order_after(stream, proj, volume)
        return volume

    def adjoint(self, grad_volume, angles, weights, out=None, stream=None):
        """Adjoint of ``process``: a [width, width, slices] volume gradient ->
        [n_proj, slices, width] projection gradient (mr_bp_parallel_adjoint)."""
        import torch
        _require_cuda_f32(grad_volume, 3, "grad_volume")
        if tuple(grad_volume.shape) != (self.width, self.width, self.slices):
            raise MrError(_lib.MR_ERR_INVALID, "volume size does not match the ROI")
        a = np.ascontiguousarray(np.asarray(angles, dtype=np.float32))
        n_proj = a.size
        w = _f32arr(weights, n_proj)
        if out is None:
            out = torch.empty((n_proj, self.slices, self.width), dtype=torch.float32, device=grad_volume.device)
        _require_cuda_f32(out, 3, "out")
        fp = ctypes.POINTER(ctypes.c_float)
        check(_lib.load().mr_bp_parallel_adjoint(self._plan, _ptr(grad_volume), n_proj, a.ctypes.data_as(fp),
                                                 w.ctypes.data_as(fp), _ptr(out), _stream(stream)))
        return out

    def close(self):
        if getattr(self, "_plan", None) and self._plan.value:
            try:
                _lib.load().mr_bp_plan_destroy(self._plan)
            except (TypeError, AttributeError):   # interpreter shutdown
                pass
            self._plan = ctypes.c_void_p()

    __del__ = close


def fbp_parallel(filt: ProjectionFilterSingle, bp: BackProjector, proj, angles, weights,
                 volume, stream=None):
    """Filter ``proj`` in place and back-project it (one engine block).
    Enqueued on ``stream`` without a host wait, ordered before later work on
    the current stream (see BackProjector.process)."""
    _require_cuda_f32(proj, 3, "proj")
    _require_cuda_f32(volume, 3, "volume")
    n_proj = proj.shape[0]
    a = _f32arr(angles, n_proj)
    w = _f32arr(weights, n_proj)
    fp = ctypes.POINTER(ctypes.c_float)
    check(_lib.load().mr_fbp_parallel(filt._plan, bp._plan, _ptr(proj), n_proj,
                                      a.ctypes.data_as(fp), w.ctypes.data_as(fp), _ptr(volume),
                                      _stream(stream)))
    _order_after(stream, proj, volume)
    return volume


class FullLogNorm:
    """FullLogNorm (StdPreprocModules/src/NormPlugins.cpp:300-341, 411-558)
    in place on device stacks.  ``flat`` / ``dark`` are the averaged
    reference images (CUDA [rows, width] float32, or None when no open-beam /
    dark-current images are configured); ``flat_dose`` and ``dark_dose`` are
    the doses NormBase::ReferenceLoader computes.  ``uselut`` selects the
    look-up-table path (log of the count rounded to an integer in [0, 65535],
    mr_lognorm_lut)."""

    def __init__(self, flat=None, dark=None, flat_dose: float = 1.0, dark_dose: float = 0.0,
                 usenormregion: bool = True, uselut: bool = False, stream=None):
        _lib.load()
        self.dark = dark
        self.dark_dose = float(dark_dose)
        self.usenormregion = bool(usenormregion)
        self.uselut = bool(uselut)
        self.flat_log = None
        if flat is not None:
            _require_cuda_f32(flat, 2, "flat")
            if dark is not None:
                _require_cuda_f32(dark, 2, "dark")
            self.flat_log = flat.clone()
            check(_lib.load().mr_lognorm_reference(_ptr(self.flat_log), _ptr(dark) if dark is not None else None,
                                                   self.flat_log.numel(), float(flat_dose), _stream(stream)))
        elif dark is not None:
            _require_cuda_f32(dark, 2, "dark")

    def process(self, proj, dose=None, stream=None):
        """``proj`` CUDA [n_proj, rows, width], normalised in place; ``dose``
        the per-projection doses (the "dose" parameter list)."""
        _require_cuda_f32(proj, 3, "proj")
        n_proj = proj.shape[0]
        n_pix = proj.shape[1] * proj.shape[2]
        for ref in (self.flat_log, self.dark):
            if ref is not None and ref.numel() != n_pix:
                raise MrError(_lib.MR_ERR_INVALID, "reference image does not match the projection size")
        d = None
        if self.usenormregion:
            if dose is None:
                raise MrError(_lib.MR_ERR_INVALID, "usenormregion needs the per-projection dose list")
            d = _f32arr(dose, n_proj)
        fp = ctypes.POINTER(ctypes.c_float)
        fn = _lib.load().mr_lognorm_lut if self.uselut else _lib.load().mr_lognorm
        check(fn(_ptr(proj), n_proj, n_pix, _ptr(self.flat_log) if self.flat_log is not None else None,
                 _ptr(self.dark) if self.dark is not None else None,
                 d.ctypes.data_as(fp) if d is not None else None, self.dark_dose, _stream(stream)))
        return proj

    def process_and_filter(self, proj, filt: "ProjectionFilterSingle", dose=None, stream=None):
        """Normalisation and ``filt`` in one pass over ``proj`` (raw counts
        in, filtered -log projections out, in place): the same values as
        ``process`` followed by ``filt.process`` (mr_filter_rows_lognorm).
        The LUT path runs as the two steps."""
        _require_cuda_f32(proj, 3, "proj")
        n_proj, rows, width = proj.shape
        if width != filt.width:
            raise MrError(_lib.MR_ERR_INVALID, "row width does not match the configured ROI")
        if self.uselut:
            self.process(proj, dose, stream)
            return filt.process(proj, stream)
        for ref in (self.flat_log, self.dark):
            if ref is not None and ref.numel() != rows * width:
                raise MrError(_lib.MR_ERR_INVALID, "reference image does not match the projection size")
        d = None
        if self.usenormregion:
            if dose is None:
                raise MrError(_lib.MR_ERR_INVALID, "usenormregion needs the per-projection dose list")
            d = _f32arr(dose, n_proj)
        fp = ctypes.POINTER(ctypes.c_float)
        check(_lib.load().mr_filter_rows_lognorm(filt._plan, _ptr(proj), n_proj, rows,
                                                 _ptr(self.flat_log) if self.flat_log is not None else None,
                                                 _ptr(self.dark) if self.dark is not None else None,
                                                 d.ctypes.data_as(fp) if d is not None else None,
                                                 self.dark_dose, _stream(stream)))
        return proj


class FDKbp_single:
    """FDK cone-beam back-projector for one slab (SetROI(roi) + Process(3D)),
    float geometry (fdkbp_single.cpp).

    ``filter="shepp-logan"`` and ``cosine_weight=True`` select the filter
    extension, which the reference does not have: the Shepp-Logan ramp
    |k|/(P/2) sinc(k/P) instead of ramp_filter_tuned's roll-off, and the
    Feldkamp pre-weight SDD/sqrt(SDD^2+u^2+v^2) per detector pixel.

    ``correct_tilt`` gives slice k the centre of rotation
    tan(-tilt_angle) (window row 3 - k - tilt_pivot) + center and shrinks the
    mask by floor(tan|tilt_angle| tilt_rows), tilt_rows being
    ProjectionInfo.nDims[1]."""

    module = "FDKbp_single"

    def __init__(self, sod: float, sdd: float, pixel, voxel, ppoint, center: float,
                 roi: Sequence[int], projection_roi3: int, direction: int = 0,
                 matrix_roi: Optional[Sequence[int]] = None, fused: bool = False,
                 filter: str = "muhrec", cosine_weight: bool = False, correct_tilt: bool = False,
                 tilt_angle: float = 0.0, tilt_pivot: float = 0.0, tilt_rows: int = 0):
        L = _lib.load()
        p = FdkParams()
        _set_tilt(p, correct_tilt, tilt_angle, tilt_pivot, None, tilt_rows)
        p.module = _lib.FDK_MODULES[self.module]
        p.fused = int(bool(fused))
        if filter not in ("muhrec", "shepp-logan"):
            raise MrError(_lib.MR_ERR_INVALID, "FDK filter must be 'muhrec' or 'shepp-logan'")
        p.filter = 0 if filter == "muhrec" else 1
        p.cosine_weight = int(bool(cosine_weight))
        p.sod, p.sdd = sod, sdd
        p.pixel[0], p.pixel[1] = pixel
        p.voxel[0], p.voxel[1], p.voxel[2] = voxel
        p.ppoint[0], p.ppoint[1] = ppoint
        p.center, p.direction = center, direction
        for i in range(4):
            p.roi[i] = int(roi[i])
        p.projection_roi3 = int(projection_roi3)
        p.use_roi = 0 if matrix_roi is None else 1
        if matrix_roi is not None:
            for i in range(4):
                p.matrix_roi[i] = int(matrix_roi[i])
        self.nx = int(roi[2] - roi[0])
        self.nz = int(roi[3] - roi[1])
        self._plan = ctypes.c_void_p()
        check(L.mr_fdk_plan_create(ctypes.byref(self._plan), ctypes.byref(p)))

    @property
    def window(self):
        w = (ctypes.c_size_t * 4)()
        check(_lib.load().mr_fdk_plan_window(self._plan, w))
        return [int(v) for v in w]

    def process(self, proj, angles, weights, volume=None, stream=None):
        """``proj`` [n_proj, win_rows, width] CUDA (consumed as scratch) ->
        [nz, nx, nx] XYZ volume."""
        import torch
        _require_cuda_f32(proj, 3, "proj")
        n_proj, rows, width = proj.shape
        if width != self.nx:
            raise MrError(_lib.MR_ERR_INVALID, "projection width does not match the ROI")
        a = _f32arr(angles, n_proj)
        w = _f32arr(weights, n_proj)
        if volume is None:
            volume = torch.empty((self.nz, self.nx, self.nx), dtype=torch.float32, device=proj.device)
        _require_cuda_f32(volume, 3, "volume")
        fp = ctypes.POINTER(ctypes.c_float)
        check(_lib.load().mr_fdk_bp(self._plan, _ptr(proj), n_proj, rows, a.ctypes.data_as(fp),
                                    w.ctypes.data_as(fp), _ptr(volume), _stream(stream)))
        return volume

    def close(self):
        if getattr(self, "_plan", None) and self._plan.value:
            try:
                _lib.load().mr_fdk_plan_destroy(self._plan)
            except (TypeError, AttributeError):   # interpreter shutdown
                pass
            self._plan = ctypes.c_void_p()

    __del__ = close


class FDKbp(FDKbp_single):
    """The double-geometry FDK module (fdkbp.cpp): projection matrix, sampling
    positions and accumulation in double, normalised by SOD/SDD/nProj."""

    module = "FDKbp"


class BatchReconstructor:
    """Many parallel-beam stacks of uneven geometry in one persistent launch
    (mr_batch_*).  ``stacks`` is a list of dicts with keys proj (CUDA
    [n_proj, slices, width]), angles, weights, center and optionally
    filtertype, cutoff, direction, buffer_size; volumes are allocated here.
    ``fused`` folds the interpolation into FMAs (float rounding only);
    otherwise every stack is bit-identical to MultiProjBPparallel."""

    def __init__(self, stacks, tune: bool = True, fused: bool = False):
        import torch
        L = _lib.load()
        self._keep = []
        arr = (_lib.BatchStack * len(stacks))()
        self.volumes = []
        self.updates = 0
        for i, s in enumerate(stacks):
            proj = s["proj"]
            _require_cuda_f32(proj, 3, "proj")
            n, nv, nu = proj.shape
            a = _f32arr(s["angles"], n)
            w = _f32arr(s["weights"], n)
            vol = torch.empty((nu, nu, nv), dtype=torch.float32, device=proj.device)
            self._keep += [a, w, proj]
            self.volumes.append(vol)
            d = arr[i]
            d.width, d.slices, d.n_proj = nu, nv, n
            d.center, d.direction = float(s["center"]), int(s.get("direction", 0))
            d.filter_type = _lib.FILTER_TYPES[s.get("filtertype", "Hamming").lower()]
            d.cutoff, d.order = float(s.get("cutoff", 0.5)), float(s.get("order", 1.0))
            d.use_bias, d.bias_weight = int(s.get("usebias", True)), float(s.get("biasweight", 0.1))
            d.padding_doubler = int(s.get("paddingdoubler", 2))
            d.buffer_size = int(s.get("buffer_size", 16))
            d.d_proj = proj.data_ptr()
            d.angles = a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
            d.weights = w.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
            d.d_volume = vol.data_ptr()
        self._plan = ctypes.c_void_p()
        flags = (_lib.BATCH_TUNE if tune else 0) | (_lib.BATCH_FUSED if fused else 0)
        check(L.mr_batch_plan_create(ctypes.byref(self._plan), arr, len(stacks), flags))
        zt, kc, ni, upd = ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t(), ctypes.c_double()
        check(L.mr_batch_plan_info(self._plan, ctypes.byref(zt), ctypes.byref(kc), ctypes.byref(ni),
                                   ctypes.byref(upd)))
        self.slice_tile, self.proj_chunk, self.n_items, self.updates = zt.value, kc.value, ni.value, upd.value

    def run(self, stream=None):
        check(_lib.load().mr_batch_run(self._plan, _stream(stream)))

    def run_graph(self, stream=None):
        check(_lib.load().mr_batch_run_graph(self._plan, _stream(stream)))

    def run_sequential(self, stream=None):
        check(_lib.load().mr_batch_run_sequential(self._plan, _stream(stream)))

    def close(self):
        if getattr(self, "_plan", None) and self._plan.value:
            try:
                _lib.load().mr_batch_plan_destroy(self._plan)
            except (TypeError, AttributeError):   # interpreter shutdown
                pass
            self._plan = ctypes.c_void_p()

    __del__ = close


class LinearForwardProjector:
    """LinearForwardProjector::project (linearforwardprojector.cpp:15-143) on
    the GPU: ``project(slice, angles)`` takes a [M, M] slice (slice[y, x], the
    reference's TImage<float,2>) and returns proj [n_ang, M] (proj[t, r]); a
    [M, M, nz] MatrixZXY stack returns [n_ang, M, nz].  Bit-identical to the
    reference per slice."""

    def project(self, vol, angles, stream=None):
        import torch
        if not isinstance(vol, torch.Tensor) or not vol.is_cuda or vol.dtype != torch.float32:
            raise TypeError("vol must be a float32 CUDA tensor")
        if vol.dim() not in (2, 3) or vol.shape[0] != vol.shape[1]:
            raise MrError(_lib.MR_ERR_INVALID, "slice must be [M, M] or [M, M, nz]")
        v = vol.contiguous()
        M = int(v.shape[0])
        nz = 1 if v.dim() == 2 else int(v.shape[2])
        a = np.ascontiguousarray(angles, np.float32).ravel()
        out = torch.empty((a.size, M, nz), dtype=torch.float32, device=v.device)
        check(_lib.load().mr_linear_forward(_ptr(v), M, nz, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                            a.size, _ptr(out), _stream(stream)))
        return out[..., 0] if v.dim() == 2 else out


class IterativeReconstructor:
    """SIRT / Landweber / CGLS on the matched Joseph pair (mr_iter_*), the
    methods of IterativeMeth.txt (frameworks/tomography/Backprojectors/
    IterativeBackProj/src/IterativeMeth.txt:1-113).  Volume [nx, nx, nz]
    (MatrixZXY), sinogram [n_ang, n_det, nz]; nz % 4 == 0."""

    def __init__(self, nx: int, nz: int, n_det: int, angles, center: float, method: str = "sirt",
                 direction: int = 0):
        if method not in _lib.ITER_METHODS:
            raise MrError(_lib.MR_ERR_INVALID, "method must be sirt, landweber or cgls")
        L = _lib.load()
        self.nx, self.nz, self.n_det, self.method = nx, nz, n_det, method
        self.angles = np.ascontiguousarray(angles, np.float32)
        self._g = _lib.JosephGeom(nx, nz, n_det, center, direction)
        self._plan = ctypes.c_void_p()
        check(L.mr_iter_plan_create(ctypes.byref(self._plan), ctypes.byref(self._g),
                                    self.angles.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                    self.angles.size, _lib.ITER_METHODS[method]))

    def run(self, sino, x=None, iterations: int = 10, tau: float = 0.0, nonneg: Optional[bool] = None,
            residuals: bool = False, stream=None):
        """Iterates on ``x`` (zeros if None) in place; returns x, or (x,
        history of |b - A x|^2 before each update) with ``residuals``."""
        import torch
        _require_cuda_f32(sino, 3, "sino")
        if tuple(sino.shape) != (self.angles.size, self.n_det, self.nz):
            raise MrError(_lib.MR_ERR_INVALID, "sinogram shape does not match the plan")
        if x is None:
            x = torch.zeros((self.nx, self.nx, self.nz), dtype=torch.float32, device=sino.device)
        _require_cuda_f32(x, 3, "x")
        if tuple(x.shape) != (self.nx, self.nx, self.nz):
            raise MrError(_lib.MR_ERR_INVALID, "volume shape does not match the plan")
        hist = np.zeros(max(iterations, 1), np.float64) if residuals else None
        nn = -1 if nonneg is None else int(bool(nonneg))
        check(_lib.load().mr_iter_run(self._plan, _ptr(sino), _ptr(x), int(iterations), float(tau), nn,
                                      hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if residuals else None,
                                      _stream(stream)))
        return (x, hist[:iterations]) if residuals else x

    def close(self):
        if getattr(self, "_plan", None) and self._plan.value:
            try:
                _lib.load().mr_iter_plan_destroy(self._plan)
            except (TypeError, AttributeError):
                pass
            self._plan = ctypes.c_void_p()

    def __del__(self):
        self.close()


class _PinnedBuffer:
    def __init__(self, nbytes: int):
        self.ptr = ctypes.c_void_p()
        check(_lib.load().mr_host_alloc(ctypes.byref(self.ptr), max(int(nbytes), 1)))

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            try:
                _lib.load().mr_host_free(self.ptr)
            except (TypeError, AttributeError):
                pass
            self.ptr = ctypes.c_void_p()


def host_empty(shape, dtype=np.float32) -> np.ndarray:
    """A numpy array in pinned host memory (mr_host_alloc); FBPStream copies
    such arrays by DMA without staging.  The memory is freed with the array."""
    dt = np.dtype(dtype)
    n = int(np.prod(shape)) * dt.itemsize
    buf = _PinnedBuffer(n)
    raw = (ctypes.c_char * max(n, 1)).from_address(buf.ptr.value)
    a = np.frombuffer(raw, dtype=dt, count=int(np.prod(shape))).reshape(shape)
    # keep the allocation alive as long as any view of the array
    raw._mr_owner = buf
    return a


class FBPStream:
    """Streaming FBP of a host-resident stack (mr_fbp_stream_*): the engine's
    SliceBlock loop (ReconEngine::Run3DFull) with filter and back-projection on
    the GPU and the host copies of neighbouring blocks overlapped with the
    kernels.  ``run`` takes projections [n_proj, rows, width] (numpy float32)
    and returns slices [rows, width, width]; equal bit for bit to filtering and
    back-projecting each block on the device."""

    def __init__(self, width: int, rows: int, n_proj: int, center: float, slice_block: int = 32,
                 backprojector: str = "MultiProjBPparallel", direction: int = 0, buffer_size: int = 16,
                 fused: bool = False, filtertype: str = "Hamming", cutoff: float = 0.5, order: float = 1.0,
                 usebias: bool = True, biasweight: float = 0.1, paddingdoubler: int = 2,
                 correct_tilt: bool = False, tilt_angle: float = 0.0, tilt_pivot: float = 0.0,
                 row0: int = 0, tilt_rows: Optional[int] = None):
        if backprojector not in BP_MODULES:
            raise MrError(_lib.MR_ERR_INVALID, f"unknown back-projector {backprojector}")
        ft = _lib.FILTER_TYPES.get(filtertype.lower())
        if ft is None:
            raise MrError(_lib.MR_ERR_INVALID, "Could not convert filter string to enum.")
        p = BpParams()
        p.width, p.slices, p.center, p.direction = int(width), int(slice_block), float(center), int(direction)
        p.interp = BP_MODULES[backprojector]["interp"]
        p.multiprojbp = BP_MODULES[backprojector]["multiprojbp"]
        p.buffer_size = int(buffer_size)
        p.fused = int(bool(fused))
        # row0: the first slab's detector row; every slab of slice_block rows
        # is a block of its own, as the engine's SliceBlock loop makes it
        _set_tilt(p, correct_tilt, tilt_angle, tilt_pivot, row0, rows if tilt_rows is None else tilt_rows)
        self.width, self.rows, self.n_proj = int(width), int(rows), int(n_proj)
        self._s = ctypes.c_void_p()
        check(_lib.load().mr_fbp_stream_create(ctypes.byref(self._s), ctypes.byref(p), self.rows, self.n_proj,
                                               int(slice_block), ft, cutoff, order, int(bool(usebias)),
                                               biasweight, paddingdoubler))

    def run(self, proj: np.ndarray, angles, weights, out: Optional[np.ndarray] = None,
            serial: bool = False) -> np.ndarray:
        shape = (self.n_proj, self.rows, self.width)
        if not isinstance(proj, np.ndarray) or proj.dtype != np.float32 or proj.shape != shape \
                or not proj.flags.c_contiguous:
            raise MrError(_lib.MR_ERR_INVALID, f"proj must be a contiguous float32 array of shape {shape}")
        oshape = (self.rows, self.width, self.width)
        if out is None:
            out = np.empty(oshape, np.float32)
        if out.dtype != np.float32 or out.shape != oshape or not out.flags.c_contiguous:
            raise MrError(_lib.MR_ERR_INVALID, f"out must be a contiguous float32 array of shape {oshape}")
        a = _f32arr(angles, self.n_proj)
        w = _f32arr(weights, self.n_proj)
        fp = ctypes.POINTER(ctypes.c_float)
        check(_lib.load().mr_fbp_stream_run(self._s, proj.ctypes.data_as(ctypes.c_void_p), a.ctypes.data_as(fp),
                                            w.ctypes.data_as(fp), out.ctypes.data_as(ctypes.c_void_p),
                                            int(bool(serial))))
        return out

    def run_slabs(self, proj: np.ndarray, angles, weights, on_slab, out: Optional[np.ndarray] = None):
        """As ``run``, handing each slab to ``on_slab(first_slice, slices)``
        (``slices`` a [n, width, width] view of ``out``) as soon as it is
        complete, while the GPU works on the later slabs
        (mr_fbp_stream_run_slabs).  ``on_slab`` returning True stops the run
        (MrError with MR_ERR_ABORTED)."""
        shape = (self.n_proj, self.rows, self.width)
        if not isinstance(proj, np.ndarray) or proj.dtype != np.float32 or proj.shape != shape \
                or not proj.flags.c_contiguous:
            raise MrError(_lib.MR_ERR_INVALID, f"proj must be a contiguous float32 array of shape {shape}")
        oshape = (self.rows, self.width, self.width)
        if out is None:
            out = np.empty(oshape, np.float32)
        if out.dtype != np.float32 or out.shape != oshape or not out.flags.c_contiguous:
            raise MrError(_lib.MR_ERR_INVALID, f"out must be a contiguous float32 array of shape {oshape}")
        a = _f32arr(angles, self.n_proj)
        w = _f32arr(weights, self.n_proj)
        fp = ctypes.POINTER(ctypes.c_float)
        errors = []

        def cb(_user, first, n, _ptr):
            try:
                return 1 if on_slab(int(first), out[first:first + n]) else 0
            except Exception as ex:   # an exception must not cross the C ABI
                errors.append(ex)
                return 1

        fn = _lib.SLAB_FN(cb)
        rc = _lib.load().mr_fbp_stream_run_slabs(self._s, proj.ctypes.data_as(ctypes.c_void_p), a.ctypes.data_as(fp),
                                                 w.ctypes.data_as(fp), out.ctypes.data_as(ctypes.c_void_p), fn, None)
        if errors:
            raise errors[0]
        check(rc)
        return out

    def close(self):
        if getattr(self, "_s", None) and self._s.value:
            try:
                _lib.load().mr_fbp_stream_destroy(self._s)
            except (TypeError, AttributeError):
                pass
            self._s = ctypes.c_void_p()

    def __del__(self):
        self.close()

