"""caps_amd: MI355X (gfx950) Strassen-Winograd multiply with CAPS's interface.

Python mirror of the reference's entry points over the C ABI in
include/caps_amd.h (libcaps_amd.so, loaded with ctypes):

    MatDescriptor           struct MatDescriptor (matrix.h:8-16)
    multiply(A, B, C, desc, pattern=None)
                            void multiply(...) (multiply.h:13)
    verify_descriptor(desc) verifyDescriptor (matrix.cpp:3-7), raises
    num_entries_per_proc    numEntriesPerProc (matrix.cpp:9-12)
    distribute_from_1proc / collect_to_1proc
                            distributeFrom1Proc / collectTo1Proc with one
                            process (communication.cpp:657-659, 706-708)
    print_counters          printCounters (counters.cpp:61-74)

Host arrays are numpy float64; device arrays are torch tensors on a HIP
device (torch is only used for device memory and streams).  The library has
no CPU fallback: if libcaps_amd.so is missing or cannot be loaded, every call
raises CapsError.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcaps_amd.so")

SEVEN = 7  # library.h:6

MODE_FP64 = 0
MODE_OZAKI_INT8 = 1
MODE_CRT_INT8 = 2
MODE_CRT_BF16 = 3
MODE_AUTO = 99
FLAG_NO_GRAPH = 1
FLAG_REFERENCE_AUTO = 2
FLAG_TIMERS = 4
FLAG_NO_FUSION = 8
FLAG_SMALL_TILE = 16
DEFAULT_MODULI = 14
# phase timers (counters.h:13-18 order, then the split mode's phases)
TIMER_NAMES = ("comm", "add", "mul", "rearrange", "comm_summa", "split", "recon", "unused")
NUM_TIMERS = 8

_STATUS = {0: "OK", 1: "EINVAL", 2: "EPATTERN", 3: "ENOMEM", 4: "EHIP", 5: "ENODEV",
           6: "EUNSUPPORTED"}


class CapsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"caps error {_STATUS.get(code, code)}: {msg}")
        self.code = code


class MatDescriptor(ctypes.Structure):
    """struct MatDescriptor (matrix.h:8-16); DESC_SIZE = 7 ints."""
    _fields_ = [("lda", ctypes.c_int), ("nrec", ctypes.c_int), ("nproc", ctypes.c_int),
                ("nprocr", ctypes.c_int), ("nprocc", ctypes.c_int),
                ("nproc_summa", ctypes.c_int), ("bs", ctypes.c_int)]

    @classmethod
    def one_device(cls, lda: int, nrec: int, bs: int = 1) -> "MatDescriptor":
        return cls(lda, nrec, 1, 1, 1, 1, bs)


DESC_SIZE = 7
assert ctypes.sizeof(MatDescriptor) == DESC_SIZE * ctypes.sizeof(ctypes.c_int)


class Options(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("moduli", ctypes.c_int), ("flags", ctypes.c_int),
                ("cutoff", ctypes.c_int)]


class Counters(ctypes.Structure):
    _fields_ = [("additions", ctypes.c_longlong), ("multiplications", ctypes.c_longlong),
                ("leaf_products", ctypes.c_longlong), ("peak_workspace_mb", ctypes.c_double),
                ("pattern", ctypes.c_char * 256),
                ("timers_ms", ctypes.c_double * 8), ("timed_launches", ctypes.c_longlong * 8)]


class Candidate(ctypes.Structure):
    """caps_candidate: one configuration of the autotuner / planner."""
    _fields_ = [("mode", ctypes.c_int), ("moduli", ctypes.c_int), ("cutoff", ctypes.c_int),
                ("tile", ctypes.c_int), ("pattern", ctypes.c_char * 32),
                ("ms", ctypes.c_double), ("err", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {"mode": self.mode, "moduli": self.moduli, "cutoff": self.cutoff,
                "tile": self.tile, "pattern": self.pattern.decode(), "ms": self.ms,
                "err": self.err}


MASK_FP64 = 1 << MODE_FP64
MASK_OZAKI_INT8 = 1 << MODE_OZAKI_INT8
MASK_CRT_INT8 = 1 << MODE_CRT_INT8
MASK_CRT_BF16 = 1 << MODE_CRT_BF16


# every symbol include/caps_amd.h declares, with (restype, argtypes)
_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_D = ctypes.POINTER(MatDescriptor)
_O = ctypes.POINTER(Options)
SIGNATURES = {
    "caps_init": (ctypes.c_int, [ctypes.c_int]),
    "caps_last_error": (ctypes.c_char_p, []),
    "caps_verify_desc": (ctypes.c_int, [_D]),
    "caps_num_entries": (ctypes.c_longlong, [_D]),
    "caps_multiply": (ctypes.c_int, [_vp, _vp, _vp, _D, ctypes.c_char_p, _O]),
    "caps_multiply_device": (ctypes.c_int, [_vp, _vp, _vp, _D, ctypes.c_char_p, _O, _vp]),
    "caps_multiply_device_noalloc": (ctypes.c_int, [_vp, _vp, _vp, _D, ctypes.c_char_p, _O,
                                                    _vp]),
    "caps_multiply_batch_device": (ctypes.c_int, [ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp,
                                                  _vp]),
    "caps_to_layout": (ctypes.c_int, [_vp, _vp, _D]),
    "caps_from_layout": (ctypes.c_int, [_vp, _vp, _D]),
    "caps_to_layout_device": (ctypes.c_int, [_vp, _vp, _D, _vp]),
    "caps_from_layout_device": (ctypes.c_int, [_vp, _vp, _D, _vp]),
    "caps_get_counters": (ctypes.c_int, [ctypes.POINTER(Counters)]),
    "caps_reset_counters": (ctypes.c_int, []),
    "caps_set_available_memory": (ctypes.c_int, [ctypes.c_longlong]),
    "caps_release": (ctypes.c_int, []),
    "caps_reference_product": (ctypes.c_int, [_vp, _vp, _vp, _D, _vp]),
    "caps_normwise_error": (ctypes.c_int, [_vp, _vp, _D, ctypes.POINTER(ctypes.c_double), _vp]),
    "caps_measure": (ctypes.c_int, [_vp, _vp, _vp, _D, _vp, ctypes.c_int,
                                    ctypes.POINTER(Candidate), _vp]),
    "caps_autotune": (ctypes.c_int, [_vp, _vp, _vp, _D, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_int, ctypes.POINTER(Candidate), ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_int), ctypes.POINTER(Candidate), _vp]),
    "caps_plan_choice": (ctypes.c_int, [_D, _O, ctypes.POINTER(Candidate)]),
    "caps_planner_config": (ctypes.c_int, [ctypes.c_double, ctypes.c_int, ctypes.c_double]),
    "caps_tune_save": (ctypes.c_int, [ctypes.c_char_p]),
    "caps_tune_load": (ctypes.c_int, [ctypes.c_char_p]),
    "caps_enumerate_candidates": (ctypes.c_int, [_D, ctypes.c_int, ctypes.POINTER(Candidate),
                                                 ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "caps_candidate_features": (ctypes.c_int, [_D, ctypes.POINTER(Candidate), ctypes.c_double,
                                               _vp]),
    "caps_planner_use_model": (ctypes.c_int, [_vp, _vp, ctypes.c_int]),
    "caps_mlp_nparams": (ctypes.c_int, []),
    "caps_mlp_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_ulonglong]),
    "caps_mlp_destroy": (ctypes.c_int, [_vp]),
    "caps_mlp_get_params": (ctypes.c_int, [_vp, _vp, ctypes.c_int]),
    "caps_mlp_set_params": (ctypes.c_int, [_vp, _vp, ctypes.c_int]),
    "caps_mlp_forward": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp]),
    "caps_mlp_gradient": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, _vp, _vp]),
    "caps_mlp_train": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_float, ctypes.c_ulonglong, _vp]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Load libcaps_amd.so (raises CapsError if it is absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise CapsError(6, f"{path} not built; run `python -m caps_amd.build` "
                           "(there is no CPU fallback)")
    try:
        lib = ctypes.CDLL(path)
    except OSError as e:
        raise CapsError(6, f"cannot load {path}: {e}") from e
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc: int):
    if rc != 0:
        raise CapsError(rc, load().caps_last_error().decode(errors="replace"))


def _opts(mode: int, moduli: int, flags: int, cutoff: int = 0) -> Options:
    return Options(mode, moduli, flags, cutoff)


def _pattern(p):
    if p is None:
        return None
    return p.encode() if isinstance(p, str) else bytes(p)


# ------------------------------------------------------------------ descriptor
def verify_descriptor(desc: MatDescriptor) -> None:
    _check(load().caps_verify_desc(ctypes.byref(desc)))


def num_entries_per_proc(desc: MatDescriptor) -> int:
    return int(load().caps_num_entries(ctypes.byref(desc)))


# ------------------------------------------------------------------ multiply
def _host_ptr(a: np.ndarray, n2: int, writable: bool = False):
    if not isinstance(a, np.ndarray) or a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise CapsError(1, "host arrays must be C-contiguous float64 numpy arrays")
    if a.size != n2:
        raise CapsError(1, f"array has {a.size} entries, descriptor needs {n2}")
    if writable and not a.flags["WRITEABLE"]:
        raise CapsError(1, "output array is read-only")
    return a.ctypes.data


def multiply(A: np.ndarray, B: np.ndarray, C: np.ndarray, desc: MatDescriptor,
             pattern: str | None = None, mode: int = MODE_FP64, moduli: int = 0,
             flags: int = 0, cutoff: int = 0) -> None:
    """C = A * B on the device, all three in the CAPS layout (host memory).
    multiply.h:13: A and B are read, C is overwritten."""
    n2 = desc.lda * desc.lda
    o = _opts(mode, moduli, flags, cutoff)
    _check(load().caps_multiply(_host_ptr(A, n2), _host_ptr(B, n2), _host_ptr(C, n2, True),
                                ctypes.byref(desc), _pattern(pattern), ctypes.byref(o)))


def _dev_ptr(t, n2: int) -> int:
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float64:
        raise CapsError(1, "device arrays must be float64 torch tensors on the GPU")
    if not t.is_contiguous() or t.numel() != n2:
        raise CapsError(1, f"device array must be contiguous with {n2} entries")
    return t.data_ptr()


def _stream_ptr(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def multiply_device(A, B, C, desc: MatDescriptor, pattern: str | None = None,
                    mode: int = MODE_FP64, moduli: int = 0, flags: int = 0,
                    stream=None, cutoff: int = 0) -> None:
    """Enqueue C = A * B on `stream` (default: torch's current stream) for
    device-resident CAPS-layout tensors.  Does not synchronise."""
    n2 = desc.lda * desc.lda
    o = _opts(mode, moduli, flags, cutoff)
    _check(load().caps_multiply_device(_dev_ptr(A, n2), _dev_ptr(B, n2), _dev_ptr(C, n2),
                                       ctypes.byref(desc), _pattern(pattern), ctypes.byref(o),
                                       _stream_ptr(stream)))


NOALLOC_SCRATCH_MIB = 256  # CAPS_NOALLOC_SCRATCH_MIB


def multiply_device_noalloc(A, B, C, desc: MatDescriptor, pattern: str | None = None,
                            mode: int = MODE_FP64, moduli: int = 0, flags: int = 0,
                            stream=None, cutoff: int = 0) -> None:
    """multiply_device without device allocation: tables and workspace in the
    library's fixed per-device scratch (NOALLOC_SCRATCH_MIB MiB); calls that
    overlap on one device race."""
    n2 = desc.lda * desc.lda
    o = _opts(mode, moduli, flags, cutoff)
    _check(load().caps_multiply_device_noalloc(_dev_ptr(A, n2), _dev_ptr(B, n2), _dev_ptr(C, n2),
                                               ctypes.byref(desc), _pattern(pattern),
                                               ctypes.byref(o), _stream_ptr(stream)))


def multiply_batch_device(As, Bs, Cs, descs, patterns=None, mode: int = MODE_FP64,
                          moduli: int = 0, flags: int = 0, stream=None, cutoffs=None) -> None:
    """Enqueue independent multiplies C_i = A_i * B_i (device tensors, CAPS
    layout, per-item descriptors, patterns and cutoffs) as one batched plan
    replayed as one graph."""
    k = len(As)
    assert len(Bs) == k and len(Cs) == k and len(descs) == k
    P = ctypes.c_void_p * k
    pa = P(*[_dev_ptr(a, d.lda * d.lda) for a, d in zip(As, descs)])
    pb = P(*[_dev_ptr(b, d.lda * d.lda) for b, d in zip(Bs, descs)])
    pc = P(*[_dev_ptr(c, d.lda * d.lda) for c, d in zip(Cs, descs)])
    D = (MatDescriptor * k)(*descs)
    pats = None
    if patterns is not None:
        enc = [_pattern(p) for p in patterns]
        pats = (ctypes.c_char_p * k)(*enc)
    cut = list(cutoffs) if cutoffs is not None else [0] * k
    O = (Options * k)(*[_opts(mode, moduli, flags, c) for c in cut])
    _check(load().caps_multiply_batch_device(k, pa, pb, pc, D, pats, O, _stream_ptr(stream)))


# ------------------------------------------------------------------ layout
def distribute_from_1proc(desc: MatDescriptor, colmajor: np.ndarray) -> np.ndarray:
    """Column-major n x n (numpy, indexed [i, j]) -> CAPS-layout vector."""
    n = desc.lda
    src = np.ascontiguousarray(np.asfortranarray(colmajor, dtype=np.float64).ravel(order="F"))
    out = np.empty(n * n, dtype=np.float64)
    _check(load().caps_to_layout(_host_ptr(src, n * n), _host_ptr(out, n * n, True),
                                 ctypes.byref(desc)))
    return out


def collect_to_1proc(desc: MatDescriptor, caps: np.ndarray) -> np.ndarray:
    """CAPS-layout vector -> n x n numpy matrix."""
    n = desc.lda
    src = np.ascontiguousarray(caps, dtype=np.float64)
    out = np.empty(n * n, dtype=np.float64)
    _check(load().caps_from_layout(_host_ptr(src, n * n), _host_ptr(out, n * n, True),
                                   ctypes.byref(desc)))
    return out.reshape((n, n), order="F")


def to_layout_device(colmajor, caps, desc: MatDescriptor, stream=None) -> None:
    n2 = desc.lda * desc.lda
    _check(load().caps_to_layout_device(_dev_ptr(colmajor, n2), _dev_ptr(caps, n2),
                                        ctypes.byref(desc), _stream_ptr(stream)))


def from_layout_device(caps, colmajor, desc: MatDescriptor, stream=None) -> None:
    n2 = desc.lda * desc.lda
    _check(load().caps_from_layout_device(_dev_ptr(caps, n2), _dev_ptr(colmajor, n2),
                                          ctypes.byref(desc), _stream_ptr(stream)))


# ------------------------------------------------------------------ counters / memory
def counters() -> dict:
    c = Counters()
    _check(load().caps_get_counters(ctypes.byref(c)))
    return {"additions": c.additions, "multiplications": c.multiplications,
            "leaf_products": c.leaf_products, "peak_workspace_mb": c.peak_workspace_mb,
            "pattern": c.pattern.decode(),
            "timers_ms": {TIMER_NAMES[k]: c.timers_ms[k] for k in range(NUM_TIMERS)},
            "timed_launches": {TIMER_NAMES[k]: c.timed_launches[k] for k in range(NUM_TIMERS)}}


def reset_counters() -> None:
    _check(load().caps_reset_counters())


def set_available_memory(doubles: int) -> None:
    """setAvailableMemory (memory.cpp:31-33), in doubles; 0 restores the
    default budget (half of the free device memory)."""
    _check(load().caps_set_available_memory(int(doubles)))


def plan_choice(desc: MatDescriptor, mode: int = MODE_FP64, moduli: int = 0, flags: int = 0,
                cutoff: int = 0) -> dict:
    """What a call with pattern None and these options would run: leaf mode,
    moduli, cutoff and pattern (the automatic choice of multiplyInternal,
    multiply.cpp:43-66, under the memory budget)."""
    c = Candidate()
    o = _opts(mode, moduli, flags, cutoff)
    _check(load().caps_plan_choice(ctypes.byref(desc), ctypes.byref(o), ctypes.byref(c)))
    return {"mode": c.mode, "moduli": c.moduli, "cutoff": c.cutoff,
            "pattern": c.pattern.decode(), "predicted_ms": c.ms, "predicted_err": c.err}


def release() -> None:
    _check(load().caps_release())


def print_counters(t: float, desc: MatDescriptor) -> str:
    """The report of printCounters (counters.cpp:61-74) for one device."""
    c = counters()
    flops = c["additions"] + c["multiplications"]
    lda = desc.lda
    lines = [
        f"peak memory usage {c['peak_workspace_mb']:.2f} MB",
        f"Execution pattern {c['pattern']}",
        f"flops={flops} ({c['additions'] - c['multiplications']}+{2 * c['multiplications']})",
        f"time {t:f} actual Gflop/s {flops / 1e9 / t:f} "
        f"effective Gflop/s {2.0 * lda ** 3 / 1e9 / t:f}",
    ]
    text = "\n".join(lines)
    print(text)
    return text


# ------------------------------------------------------------------ autotuner
def candidates(desc: MatDescriptor, mode_mask: int) -> list:
    n = ctypes.c_int()
    _check(load().caps_enumerate_candidates(ctypes.byref(desc), mode_mask, None, 0, ctypes.byref(n)))
    arr = (Candidate * max(1, n.value))()
    _check(load().caps_enumerate_candidates(ctypes.byref(desc), mode_mask, arr, n.value,
                                            ctypes.byref(n)))
    return [arr[i] for i in range(n.value)]


def candidate_features(desc: MatDescriptor, cand: Candidate, spread: float) -> np.ndarray:
    f = np.empty(8, dtype=np.float32)
    _check(load().caps_candidate_features(ctypes.byref(desc), ctypes.byref(cand), spread,
                                          f.ctypes.data))
    return f


def reference_product(A, B, Ref, desc: MatDescriptor, stream=None) -> None:
    n2 = desc.lda * desc.lda
    _check(load().caps_reference_product(_dev_ptr(A, n2), _dev_ptr(B, n2), _dev_ptr(Ref, n2),
                                         ctypes.byref(desc), _stream_ptr(stream)))


def normwise_error(C, Ref, desc: MatDescriptor, stream=None) -> float:
    n2 = desc.lda * desc.lda
    e = ctypes.c_double()
    _check(load().caps_normwise_error(_dev_ptr(C, n2), _dev_ptr(Ref, n2), ctypes.byref(desc),
                                      ctypes.byref(e), _stream_ptr(stream)))
    return e.value


def measure(A, B, C, desc: MatDescriptor, cand: Candidate, Ref=None, reps: int = 3,
            stream=None) -> Candidate:
    n2 = desc.lda * desc.lda
    _check(load().caps_measure(_dev_ptr(A, n2), _dev_ptr(B, n2), _dev_ptr(C, n2),
                               ctypes.byref(desc), _dev_ptr(Ref, n2) if Ref is not None else None,
                               reps, ctypes.byref(cand), _stream_ptr(stream)))
    return cand


def autotune(A, B, C, desc: MatDescriptor, mode_mask: int, max_err: float, reps: int = 3,
             stream=None):
    """Exhaustive search; returns (best, all measured candidates)."""
    n2 = desc.lda * desc.lda
    cap = 256
    out = (Candidate * cap)()
    n = ctypes.c_int()
    best = Candidate()
    _check(load().caps_autotune(_dev_ptr(A, n2), _dev_ptr(B, n2), _dev_ptr(C, n2),
                                ctypes.byref(desc), mode_mask, max_err, reps, out, cap,
                                ctypes.byref(n), ctypes.byref(best), _stream_ptr(stream)))
    return best, [out[i] for i in range(n.value)]


def planner_config(max_err: float, mode_mask: int, exponent_spread: float = 10.0) -> None:
    _check(load().caps_planner_config(max_err, mode_mask, exponent_spread))


def tune_save(path: str) -> None:
    _check(load().caps_tune_save(path.encode()))


def tune_load(path: str) -> None:
    _check(load().caps_tune_load(path.encode()))


# ------------------------------------------------------------------ cost / error model
class MLP:
    """The planner's 3-layer perceptron (8 -> 64 -> 64 -> 2, ReLU, FP32) whose
    forward, backward and Adam kernels run on the GPU (caps_mlp_* in the C ABI)."""

    N_IN, N_OUT = 8, 2
    _installed = None

    def __init__(self, seed: int = 0):
        self._h = ctypes.c_void_p()
        _check(load().caps_mlp_create(ctypes.byref(self._h), seed))
        self.nparams = load().caps_mlp_nparams()

    def close(self):
        if self._h:
            load().caps_mlp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _f32(a, cols):
        a = np.ascontiguousarray(a, dtype=np.float32)
        assert a.ndim == 2 and a.shape[1] == cols
        return a

    def params(self) -> np.ndarray:
        out = np.empty(self.nparams, dtype=np.float32)
        _check(load().caps_mlp_get_params(self._h, out.ctypes.data, self.nparams))
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(load().caps_mlp_set_params(self._h, p.ctypes.data, self.nparams))

    def forward(self, X) -> np.ndarray:
        X = self._f32(X, self.N_IN)
        Y = np.empty((X.shape[0], self.N_OUT), dtype=np.float32)
        _check(load().caps_mlp_forward(self._h, X.ctypes.data, X.shape[0], Y.ctypes.data))
        return Y

    def gradient(self, X, T, M):
        X, T, M = self._f32(X, self.N_IN), self._f32(T, self.N_OUT), self._f32(M, self.N_OUT)
        g = np.empty(self.nparams, dtype=np.float32)
        loss = ctypes.c_float()
        _check(load().caps_mlp_gradient(self._h, X.ctypes.data, T.ctypes.data, M.ctypes.data,
                                        X.shape[0], g.ctypes.data, ctypes.byref(loss)))
        return g, float(loss.value)

    def use_as_planner(self, norm) -> None:
        """Install as the planner of MODE_AUTO calls (norm: 20 floats, see
        caps_planner_use_model)."""
        norm = np.ascontiguousarray(norm, dtype=np.float32)
        self._norm = norm
        _check(load().caps_planner_use_model(self._h, norm.ctypes.data, norm.size))
        MLP._installed = self  # installed models stay alive until cleared

    @staticmethod
    def clear_planner() -> None:
        _check(load().caps_planner_use_model(None, None, 0))
        MLP._installed = None

    def train(self, X, T, M, epochs=200, batch=256, lr=1e-3, seed=0) -> float:
        X, T, M = self._f32(X, self.N_IN), self._f32(T, self.N_OUT), self._f32(M, self.N_OUT)
        loss = ctypes.c_float()
        _check(load().caps_mlp_train(self._h, X.ctypes.data, T.ctypes.data, M.ctypes.data,
                                     X.shape[0], epochs, batch, lr, seed, ctypes.byref(loss)))
        return float(loss.value)
