"""ctypes binding of libtopefft.so (the C-ABI declared in include/topefft.h).

The structs below mirror the header field for field; `check_layout()`
compares their sizes with what the library reports.  The binding raises if
the compiled library is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "libtopefft.so")

C2C, R2C, FORWARD, INVERSE = 1, 0, 1, 0
TOPE_SUCCESS = 0
TOPE_ERROR_INVALID_SIZE = -1001
TOPE_ERROR_INVALID_ARG = -1002
TOPE_ERROR_NO_DEVICE = -1003
TOPE_ERROR_NOT_INITIALIZED = -1004
TOPE_ERROR_NO_KERNEL = -1005
TOPE_ERROR_UNSUPPORTED = -1006
TOPE_ERROR_SCRATCH_LIMIT = -1007
NOALLOC_MAX_ELEMENTS = 1 << 21


class TopeFFT(C.Structure):
    _fields_ = [("error", C.c_int), ("event", C.c_void_p), ("device", C.c_int),
                ("stream", C.c_void_p), ("impl", C.c_void_p),
                ("ret_num_platforms", C.c_uint), ("ret_num_devices", C.c_uint)]


_COMMON_TAIL = [("dataSize", C.c_size_t), ("totalMemory", C.c_uint64),
                ("totalPreKernel", C.c_uint64), ("totalKernel", C.c_uint64),
                ("dim", C.c_uint), ("impl", C.c_void_p)]


class TopePlan1D(C.Structure):
    _fields_ = [("x", C.c_int), ("log", C.c_int), ("type", C.c_int), ("radix", C.c_int),
                ("data", C.c_void_p), ("scratch", C.c_void_p), ("twiddle", C.c_void_p)] + _COMMON_TAIL


class TopePlan2D(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("logX", C.c_int), ("logY", C.c_int),
                ("type", C.c_int), ("radX", C.c_int), ("radY", C.c_int),
                ("data", C.c_void_p), ("scratch", C.c_void_p),
                ("twdX", C.c_void_p), ("twdY", C.c_void_p)] + _COMMON_TAIL


class TopePlan3D(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("z", C.c_int),
                ("logX", C.c_int), ("logY", C.c_int), ("logZ", C.c_int),
                ("type", C.c_int), ("radX", C.c_int), ("radY", C.c_int), ("radZ", C.c_int),
                ("data", C.c_void_p), ("scratch", C.c_void_p),
                ("twdX", C.c_void_p), ("twdY", C.c_void_p), ("twdZ", C.c_void_p)] + _COMMON_TAIL


class TopePlanBatch(C.Structure):
    _fields_ = [("n", C.c_int), ("count", C.c_int), ("precision", C.c_int), ("mode", C.c_int),
                ("data", C.c_void_p), ("dataSize", C.c_size_t), ("totalMemory", C.c_uint64),
                ("totalPreKernel", C.c_uint64), ("totalKernel", C.c_uint64), ("impl", C.c_void_p)]


class TopePlanMany(C.Structure):
    _fields_ = [("count", C.c_int), ("precision", C.c_int), ("method", C.c_int), ("reserved", C.c_int),
                ("data", C.c_void_p), ("dataSize", C.c_size_t), ("elements", C.c_longlong),
                ("totalMemory", C.c_uint64), ("totalPreKernel", C.c_uint64), ("totalKernel", C.c_uint64),
                ("impl", C.c_void_p)]


class TopePlanReal(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("z", C.c_int), ("dim", C.c_int),
                ("pitch", C.c_int), ("precision", C.c_int), ("data", C.c_void_p), ("real", C.c_void_p), ("dataSize", C.c_size_t), ("realSize", C.c_size_t),
                ("totalMemory", C.c_uint64), ("totalPreKernel", C.c_uint64), ("totalKernel", C.c_uint64),
                ("impl", C.c_void_p)]


MANY_PERSISTENT, MANY_PERSISTENT_GRAPH, MANY_PER_TRANSFORM, MANY_PER_TRANSFORM_GRAPH = 0, 1, 2, 3
MANY_METHOD_NAMES = {0: "persistent", 1: "persistent_graph", 2: "per_transform", 3: "per_transform_graph"}

PREC_F64, PREC_F32 = 0, 1
MODE_VALU, MODE_MFMA_BF16X3, MODE_MFMA_BF16X6, MODE_MFMA_FP32, MODE_MFMA_FP16X3 = 0, 1, 2, 3, 4
MODE_NAMES = {MODE_VALU: "valu", MODE_MFMA_BF16X3: "mfma_bf16x3", MODE_MFMA_BF16X6: "mfma_bf16x6",
              MODE_MFMA_FP32: "mfma_fp32", MODE_MFMA_FP16X3: "mfma_fp16x3"}

# name -> (restype, argtypes)
_F, _P1, _P2, _P3 = C.POINTER(TopeFFT), C.POINTER(TopePlan1D), C.POINTER(TopePlan2D), C.POINTER(TopePlan3D)
_DP = C.POINTER(C.c_double)
_PB = C.POINTER(TopePlanBatch)
_PM = C.POINTER(TopePlanMany)
_PR = C.POINTER(TopePlanReal)
SIGNATURES = {
    "topeFFTSetStream": (None, [_F, C.c_void_p]),
    "topePlanExecPtr": (None, [_F, C.c_void_p, C.c_void_p, C.c_int]),
    "topeRealPlanInit": (None, [_F, _PR, C.c_int, C.c_int, C.c_int, _DP]),
    "topeRealExec": (None, [_F, _PR, _DP, C.c_int]),
    "topeRealExecDevice": (None, [_F, _PR, C.c_int]),
    "topeRealExecPtr": (None, [_F, _PR, C.c_void_p, C.c_void_p, C.c_int]),
    "topeRealPlanInitEx": (None, [_F, _PR, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]),
    "topeRealExecEx": (None, [_F, _PR, C.c_void_p, C.c_int]),
    "topeRealDestroy": (None, [_F, _PR]),
    "topeFFTInit": (None, [_F]),
    "topeFFTDestroy": (None, [_F]),
    "tope1DPlanInit": (None, [_F, _P1, C.c_int, C.c_int, _DP]),
    "tope1DExec": (None, [_F, _P1, _DP, C.c_int]),
    "tope1DDestroy": (None, [_F, _P1]),
    "tope2DPlanInit": (None, [_F, _P2, C.c_int, C.c_int, C.c_int, _DP]),
    "tope2DExec": (None, [_F, _P2, _DP, C.c_int]),
    "tope2DDestroy": (None, [_F, _P2]),
    "tope3DPlanInit": (None, [_F, _P3, C.c_int, C.c_int, C.c_int, C.c_int, _DP]),
    "tope3DExec": (None, [_F, _P3, _DP, C.c_int]),
    "tope3DDestroy": (None, [_F, _P3]),
    "profileThis": (C.c_uint64, [C.c_void_p]),
    "tope1DExecDevice": (None, [_F, _P1, C.c_int]),
    "tope2DExecDevice": (None, [_F, _P2, C.c_int]),
    "tope3DExecDevice": (None, [_F, _P3, C.c_int]),
    "topePlanUpload": (None, [_F, C.c_void_p, C.c_void_p, C.c_size_t]),
    "topePlanDownload": (None, [_F, C.c_void_p, C.c_void_p, C.c_size_t]),
    "topeFFTSynchronize": (None, [_F]),
    "topePlanPassCount": (C.c_int, [C.c_void_p]),
    "topePlanPassTiming": (None, [C.c_void_p, C.c_int]),
    "topePlanPassTimes": (C.c_int, [_F, C.c_void_p, _DP, C.c_int]),
    "topePlanPassName": (C.c_char_p, [C.c_void_p, C.c_int]),
    "topePlanTimePass": (C.c_double, [_F, C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "topeStructSize": (C.c_size_t, [C.c_int]),
    "topeBatchPlanInit": (None, [_F, _PB, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]),
    "topeBatchExec": (None, [_F, _PB, C.c_void_p, C.c_int]),
    "topeBatchExecDevice": (None, [_F, _PB, C.c_int]),
    "topeBatchDestroy": (None, [_F, _PB]),
    "tope2DPlanInitEx": (None, [_F, _P2, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]),
    "tope3DPlanInitEx": (None, [_F, _P3, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]),
    "tope2DPlanInitMode": (None, [_F, _P2, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]),
    "tope3DPlanInitMode": (None, [_F, _P3, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]),
    "tope2DExecEx": (None, [_F, _P2, C.c_void_p, C.c_int]),
    "tope3DExecEx": (None, [_F, _P3, C.c_void_p, C.c_int]),
    "topeManyPlanInit": (None, [_F, _PM, C.c_int, C.POINTER(C.c_int), C.c_int, C.c_void_p]),
    "topeManySetMethod": (None, [_F, _PM, C.c_int]),
    "topeManyExec": (None, [_F, _PM, C.c_void_p, C.c_int]),
    "topeManyExecDevice": (None, [_F, _PM, C.c_int]),
    "topeManyExecPtr": (None, [_F, _PM, C.c_void_p, C.c_int]),
    "topeManyOffsets": (C.c_longlong, [_PM, C.POINTER(C.c_longlong)]),
    "topeManyTimeExec": (C.c_double, [_F, _PM, C.c_int]),
    "topeManyGetSchedule": (None, [_PM, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "topeManyTune": (C.c_int, [_F, _PM, C.c_char_p, C.c_int]),
    "topeManyDestroy": (None, [_F, _PM]),
    "topeErrorString": (C.c_char_p, [C.c_int]),
    "topePlanTwiddles": (C.c_longlong, [_F, C.c_void_p, C.c_int, _DP, C.c_longlong]),
    "topeAxisVariantName": (C.c_char_p, [C.c_int]),
    "topeExec1DNoAlloc": (None, [_F, C.c_void_p, C.c_int, C.c_int, C.c_int]),
}

_lib = None


def load(path=LIB_PATH):
    """Load the compiled library (raises OSError if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise OSError(f"libtopefft.so not built at {path}; run __graft_entry__.build()")
        lib = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check_layout():
    lib = load()
    mine = [C.sizeof(TopeFFT), C.sizeof(TopePlan1D), C.sizeof(TopePlan2D), C.sizeof(TopePlan3D),
            C.sizeof(TopePlanBatch), C.sizeof(TopePlanMany), C.sizeof(TopePlanReal)]
    theirs = [lib.topeStructSize(i) for i in range(7)]
    return mine, theirs


class TopeError(RuntimeError):
    def __init__(self, code, what):
        self.code = code
        super().__init__(f"{what}: {load().topeErrorString(code).decode()} ({code})")


def _dptr(a):
    return a.ctypes.data_as(_DP)


class Framework:
    """topeFFTInit / topeFFTDestroy."""

    def __init__(self):
        self.lib = load()
        self.f = TopeFFT()
        self.lib.topeFFTInit(C.byref(self.f))
        if self.f.error:
            raise TopeError(self.f.error, "topeFFTInit")

    def check(self, what):
        if self.f.error:
            raise TopeError(self.f.error, what)

    def synchronize(self):
        self.lib.topeFFTSynchronize(C.byref(self.f))
        self.check("topeFFTSynchronize")

    def set_stream(self, stream):
        """Enqueue later plan operations on a caller's hipStream_t (an int
        handle, e.g. torch.cuda.current_stream().cuda_stream); None restores
        the framework's own stream."""
        self.lib.topeFFTSetStream(C.byref(self.f), C.c_void_p(stream) if stream else None)
        self.check("topeFFTSetStream")

    def destroy(self):
        if self.f.impl:
            self.lib.topeFFTDestroy(C.byref(self.f))


class Plan:
    """A 1D/2D/3D C2C plan built through tope{1,2,3}DPlanInit.

    `x` is a complex128 array (or interleaved float64) of shape (nz, ny, nx)
    or flat; PlanInit uploads it, exec() transforms and downloads like the
    reference, exec_device() transforms on the device only."""

    def __init__(self, fw, dims, x, type_=C2C, precision=None, mode=None):
        self.fw, self.lib = fw, fw.lib
        self.dims = tuple(int(d) for d in dims)
        self.ndim = len(self.dims)
        n = int(np.prod(self.dims))
        self.precision = precision
        self.dtype = np.complex64 if precision == PREC_F32 else np.complex128
        host = np.ascontiguousarray(x, dtype=self.dtype).reshape(-1)
        if host.size != n:
            raise ValueError("input size does not match dims")
        self.n = n
        self.host = host.view(np.float32 if precision == PREC_F32 else np.float64)
        if precision is not None:
            hp = host.ctypes.data_as(C.c_void_p)
            if self.ndim == 2:
                self.t = TopePlan2D()
                if mode is None:
                    self.lib.tope2DPlanInitEx(C.byref(fw.f), C.byref(self.t), self.dims[0], self.dims[1], type_,
                                              precision, hp)
                else:
                    self.lib.tope2DPlanInitMode(C.byref(fw.f), C.byref(self.t), self.dims[0], self.dims[1], type_,
                                                precision, mode, hp)
            elif self.ndim == 3:
                self.t = TopePlan3D()
                if mode is None:
                    self.lib.tope3DPlanInitEx(C.byref(fw.f), C.byref(self.t), self.dims[0], self.dims[1],
                                              self.dims[2], type_, precision, hp)
                else:
                    self.lib.tope3DPlanInitMode(C.byref(fw.f), C.byref(self.t), self.dims[0], self.dims[1],
                                                self.dims[2], type_, precision, mode, hp)
            else:
                raise ValueError("precision plans are 2D or 3D (1D: BatchPlan)")
            fw.check(f"tope{self.ndim}DPlanInit{'Ex' if mode is None else 'Mode'}{self.dims}")
            return
        if self.ndim == 1:
            self.t = TopePlan1D()
            self.lib.tope1DPlanInit(C.byref(fw.f), C.byref(self.t), self.dims[0], type_, _dptr(self.host))
        elif self.ndim == 2:
            self.t = TopePlan2D()
            self.lib.tope2DPlanInit(C.byref(fw.f), C.byref(self.t), self.dims[0], self.dims[1], type_,
                                    _dptr(self.host))
        else:
            self.t = TopePlan3D()
            self.lib.tope3DPlanInit(C.byref(fw.f), C.byref(self.t), self.dims[0], self.dims[1],
                                    self.dims[2], type_, _dptr(self.host))
        fw.check(f"tope{self.ndim}DPlanInit{self.dims}")

    @property
    def impl(self):
        return self.t.impl

    def exec(self, direction=FORWARD, out=None):
        """Reference Exec: transform on the device and download into `out`
        (a fresh complex128 array if None)."""
        if out is None:
            out = np.empty(self.n, dtype=self.dtype)
        if self.precision is not None:
            fn = getattr(self.lib, f"tope{self.ndim}DExecEx")
            fn(C.byref(self.fw.f), C.byref(self.t), out.ctypes.data_as(C.c_void_p), direction)
            self.fw.check(f"tope{self.ndim}DExecEx")
            return out
        buf = out.view(np.float64)
        fn = getattr(self.lib, f"tope{self.ndim}DExec")
        fn(C.byref(self.fw.f), C.byref(self.t), _dptr(buf), direction)
        self.fw.check(f"tope{self.ndim}DExec")
        return out

    def exec_device(self, direction=FORWARD):
        fn = getattr(self.lib, f"tope{self.ndim}DExecDevice")
        fn(C.byref(self.fw.f), C.byref(self.t), direction)
        self.fw.check(f"tope{self.ndim}DExecDevice")

    def exec_ptr(self, device_ptr, direction=FORWARD):
        """Transform a caller-owned device array (the plan's layout) in place,
        asynchronously on the framework's stream."""
        self.lib.topePlanExecPtr(C.byref(self.fw.f), self.impl, C.c_void_p(device_ptr), direction)
        self.fw.check("topePlanExecPtr")

    def upload(self, x):
        h = np.ascontiguousarray(x, dtype=self.dtype).reshape(-1)
        self.lib.topePlanUpload(C.byref(self.fw.f), self.impl, h.ctypes.data_as(C.c_void_p), h.nbytes)
        self.fw.check("topePlanUpload")

    def download(self):
        out = np.empty(self.n, dtype=self.dtype)
        self.lib.topePlanDownload(C.byref(self.fw.f), self.impl, out.ctypes.data_as(C.c_void_p), out.nbytes)
        self.fw.check("topePlanDownload")
        return out

    def pass_count(self):
        return self.lib.topePlanPassCount(self.impl)

    def pass_names(self):
        return [self.lib.topePlanPassName(self.impl, i).decode() for i in range(self.pass_count())]

    def pass_timing(self, enable=True):
        self.lib.topePlanPassTiming(self.impl, 1 if enable else 0)

    def pass_times(self):
        n = self.pass_count()
        ms = np.zeros(max(n, 1))
        self.lib.topePlanPassTimes(C.byref(self.fw.f), self.impl, _dptr(ms), n)
        self.fw.check("topePlanPassTimes")
        return ms[:n]

    def time_pass(self, i, direction=FORWARD, reps=20):
        ms = self.lib.topePlanTimePass(C.byref(self.fw.f), self.impl, i, direction, reps)
        self.fw.check("topePlanTimePass")
        return ms

    def twiddles(self, axis):
        cnt = self.lib.topePlanTwiddles(C.byref(self.fw.f), self.impl, axis, None, 0)
        if cnt < 0:
            self.fw.check("topePlanTwiddles")
        buf = np.empty(2 * max(cnt, 1))
        got = self.lib.topePlanTwiddles(C.byref(self.fw.f), self.impl, axis, _dptr(buf), cnt)
        self.fw.check("topePlanTwiddles")
        return buf[: 2 * got].view(np.complex128)

    def destroy(self):
        if self.t.impl:
            fn = getattr(self.lib, f"tope{self.ndim}DDestroy")
            fn(C.byref(self.fw.f), C.byref(self.t))


class BatchPlan:
    """`count` length-`n` transforms (topeBatchPlanInit).  FP64 plans take
    complex128 input, FP32 plans complex64."""

    def __init__(self, fw, n, count, x, precision=PREC_F64, mode=MODE_VALU):
        self.fw, self.lib = fw, fw.lib
        self.n, self.count, self.precision, self.mode = n, count, precision, mode
        self.dtype = np.complex128 if precision == PREC_F64 else np.complex64
        host = np.ascontiguousarray(x, dtype=self.dtype).reshape(-1)
        if host.size != n * count:
            raise ValueError("input size does not match n * count")
        self.host = host
        self.t = TopePlanBatch()
        self.lib.topeBatchPlanInit(C.byref(fw.f), C.byref(self.t), n, count, precision, mode,
                                   host.ctypes.data_as(C.c_void_p))
        fw.check(f"topeBatchPlanInit(n={n}, count={count}, prec={precision}, mode={mode})")

    @property
    def impl(self):
        return self.t.impl

    def exec(self, direction=FORWARD):
        out = np.empty(self.n * self.count, dtype=self.dtype)
        self.lib.topeBatchExec(C.byref(self.fw.f), C.byref(self.t), out.ctypes.data_as(C.c_void_p), direction)
        self.fw.check("topeBatchExec")
        return out

    def exec_device(self, direction=FORWARD):
        self.lib.topeBatchExecDevice(C.byref(self.fw.f), C.byref(self.t), direction)
        self.fw.check("topeBatchExecDevice")

    def pass_count(self):
        return self.lib.topePlanPassCount(self.impl)

    def pass_names(self):
        return [self.lib.topePlanPassName(self.impl, i).decode() for i in range(self.pass_count())]

    def time_pass(self, i, direction=FORWARD, reps=20):
        ms = self.lib.topePlanTimePass(C.byref(self.fw.f), self.impl, i, direction, reps)
        self.fw.check("topePlanTimePass")
        return ms

    def destroy(self):
        if self.t.impl:
            self.lib.topeBatchDestroy(C.byref(self.fw.f), C.byref(self.t))


class RealPlan:
    """topeRealPlanInit / Exec / ExecDevice / Destroy: real array of dims
    (nx[, ny[, nz]]), x fastest; forward() gives the half spectrum of
    numpy's rfftn layout (nz, ny, nx//2 + 1), inverse() the real array."""

    def __init__(self, fw, dims, x, precision=None):
        self.fw, self.lib = fw, fw.lib
        d = tuple(int(v) for v in dims) + (1,) * (3 - len(dims))
        self.dims = d
        self.n = d[0] * d[1] * d[2]
        self.half = (d[0] // 2 + 1) * d[1] * d[2]
        self.precision = precision
        self.rdtype = np.float32 if precision == PREC_F32 else np.float64
        self.cdtype = np.complex64 if precision == PREC_F32 else np.complex128
        host = np.ascontiguousarray(x, dtype=self.rdtype).reshape(-1)
        if host.size != self.n:
            raise ValueError("input size does not match dims")
        self.t = TopePlanReal()
        if precision is None:
            self.lib.topeRealPlanInit(C.byref(fw.f), C.byref(self.t), d[0], d[1], d[2], _dptr(host))
        else:
            self.lib.topeRealPlanInitEx(C.byref(fw.f), C.byref(self.t), d[0], d[1], d[2], precision,
                                        host.ctypes.data_as(C.c_void_p))
        fw.check(f"topeRealPlanInit{d}")

    def _exec(self, out, direction):
        if self.precision is None:
            self.lib.topeRealExec(C.byref(self.fw.f), C.byref(self.t), _dptr(out.view(np.float64)), direction)
        else:
            self.lib.topeRealExecEx(C.byref(self.fw.f), C.byref(self.t), out.ctypes.data_as(C.c_void_p),
                                    direction)
        self.fw.check("topeRealExec(%s)" % ("FORWARD" if direction == FORWARD else "INVERSE"))
        return out

    def forward(self):
        return self._exec(np.empty(self.half, self.cdtype), FORWARD)

    def inverse(self):
        return self._exec(np.empty(self.n, self.rdtype), INVERSE)

    def exec_device(self, direction=FORWARD):
        self.lib.topeRealExecDevice(C.byref(self.fw.f), C.byref(self.t), direction)
        self.fw.check("topeRealExecDevice")

    def exec_ptr(self, real_ptr, spectrum_ptr, direction=FORWARD):
        """Transform caller-owned device arrays in the plan's device layout
        (real: realSize bytes; spectrum: dataSize bytes, rows at `pitch`),
        asynchronously on the framework's stream."""
        self.lib.topeRealExecPtr(C.byref(self.fw.f), C.byref(self.t), C.c_void_p(real_ptr),
                                 C.c_void_p(spectrum_ptr), direction)
        self.fw.check("topeRealExecPtr")

    def pass_count(self):
        return self.lib.topePlanPassCount(self.t.impl)

    def pass_names(self):
        return [self.lib.topePlanPassName(self.t.impl, i).decode() for i in range(self.pass_count())]

    def time_pass(self, i, direction=FORWARD, reps=10):
        ms = self.lib.topePlanTimePass(C.byref(self.fw.f), self.t.impl, i, direction, reps)
        self.fw.check("topePlanTimePass")
        return ms

    def destroy(self):
        if self.t.impl:
            self.lib.topeRealDestroy(C.byref(self.fw.f), C.byref(self.t))


class ManyPlan:
    """Packed batch of uneven power-of-two lengths (topeManyPlanInit)."""

    def __init__(self, fw, sizes, x, precision=PREC_F64):
        self.fw, self.lib = fw, fw.lib
        self.sizes = np.ascontiguousarray(sizes, dtype=np.int32)
        self.dtype = np.complex128 if precision == PREC_F64 else np.complex64
        host = np.ascontiguousarray(x, dtype=self.dtype).reshape(-1)
        if host.size != int(self.sizes.sum()):
            raise ValueError("packed input length does not match sum(sizes)")
        self.host = host
        self.t = TopePlanMany()
        self.lib.topeManyPlanInit(C.byref(fw.f), C.byref(self.t), len(self.sizes),
                                  self.sizes.ctypes.data_as(C.POINTER(C.c_int)), precision,
                                  host.ctypes.data_as(C.c_void_p))
        fw.check(f"topeManyPlanInit(count={len(self.sizes)})")

    def offsets(self):
        out = np.empty(len(self.sizes), dtype=np.int64)
        self.lib.topeManyOffsets(C.byref(self.t), out.ctypes.data_as(C.POINTER(C.c_longlong)))
        return out

    def set_method(self, m):
        self.lib.topeManySetMethod(C.byref(self.fw.f), C.byref(self.t), m)
        self.fw.check("topeManySetMethod")

    def exec(self, direction=FORWARD):
        out = np.empty(self.host.size, dtype=self.dtype)
        self.lib.topeManyExec(C.byref(self.fw.f), C.byref(self.t), out.ctypes.data_as(C.c_void_p), direction)
        self.fw.check("topeManyExec")
        return out

    def exec_device(self, direction=FORWARD):
        self.lib.topeManyExecDevice(C.byref(self.fw.f), C.byref(self.t), direction)
        self.fw.check("topeManyExecDevice")

    def exec_ptr(self, device_ptr, direction=FORWARD):
        """Transform a caller-owned device array of the packed layout in
        place, asynchronously on the plan's stream."""
        self.lib.topeManyExecPtr(C.byref(self.fw.f), C.byref(self.t), C.c_void_p(device_ptr), direction)
        self.fw.check("topeManyExecPtr")

    def time_exec(self, reps=10):
        """Mean device ms of one exec of the current method (HIP events)."""
        ms = self.lib.topeManyTimeExec(C.byref(self.fw.f), C.byref(self.t), reps)
        self.fw.check("topeManyTimeExec")
        return ms

    def schedule(self):
        r = (C.c_int * 16)()
        g = C.c_int()
        self.lib.topeManyGetSchedule(C.byref(self.t), r, C.byref(g))
        return list(r), g.value

    def tune(self, wisdom_path=None, reps=5):
        rc = self.lib.topeManyTune(C.byref(self.fw.f), C.byref(self.t),
                                   wisdom_path.encode() if wisdom_path else None, reps)
        self.fw.check("topeManyTune")
        return rc

    def destroy(self):
        if self.t.impl:
            self.lib.topeManyDestroy(C.byref(self.fw.f), C.byref(self.t))


def transform(x, dims, direction=FORWARD):
    """One-shot transform with the reference call sequence
    (Init, PlanInit, Exec, Destroy) -- test/2d.c:82-114."""
    fw = Framework()
    p = Plan(fw, dims, x)
    try:
        return p.exec(direction)
    finally:
        p.destroy()
        fw.destroy()


def axis_variants():
    """Names of every compiled axis-pass kernel variant, parsed into
    (name, precision, n, e, tb, col, split, dif2)."""
    lib = load()
    out = []
    i = 0
    while True:
        nm = lib.topeAxisVariantName(i)
        if nm is None:
            break
        nm = nm.decode()
        parts = nm.split("_")
        prec = PREC_F64 if parts[1] == "double" else PREC_F32
        out.append((nm, prec, int(parts[2][1:]), int(parts[3][1:]), int(parts[4][2:]), parts[5] == "col",
                    "split" in parts[6:], "dif2" in parts[6:]))
        i += 1
    return out
