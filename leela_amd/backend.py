"""ctypes binding of the C ABI in include/leela_amd.h.

This is the Python mirror of the reference's OpenCL_Network seam
(OpenCL.h:47-111): push_convolve / push_innerproduct / forward, batched.
The HIP library must be built (python -m leela_amd.build); there is no
fallback path — a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libleela_amd.so"
HEADER = PKG.parent / "include" / "leela_amd.h"

LZ_OK = 0
LZ_ERR_ARG = -1
LZ_ERR_HIP = -2
LZ_ERR_STATE = -3
LZ_ERR_NOMEM = -4
LZ_ERR_UNSUPPORTED = -5
NOALLOC_MAX_BOARDS = 2048

NET_POLICY, NET_VALUE = 0, 1
PREC_FP32, PREC_BF16, PREC_FP8, PREC_FP16 = 0, 1, 2, 3
PRECISIONS = {"fp32": PREC_FP32, "bf16": PREC_BF16, "fp8": PREC_FP8,
              "fp16": PREC_FP16}


class LeelaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None

# void (*lz_done_fn)(void *user, int status, const float *raw, const float *out)
DONE_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.POINTER(C.c_float),
                      C.POINTER(C.c_float))


def load_library():
    """Load libleela_amd.so. Raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: the HIP backend is not built "
            "(run `python -m leela_amd.build`)")
    # One HIP runtime per process: torch ships its own libamdhip64.so.7. Once
    # torch is loaded, the library's libamdhip64.so.7 dependency resolves to
    # that copy; loaded the other way round, torch would bring up a second
    # runtime and find no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(str(LIB_PATH))
    vp = C.c_void_p
    f32p = C.POINTER(C.c_float)
    L.lz_init.argtypes = [C.POINTER(vp), C.c_int]
    L.lz_push_convolve.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                   f32p, f32p]
    L.lz_push_innerproduct.argtypes = [vp, C.c_int, C.c_int, C.c_int, f32p,
                                       f32p]
    L.lz_push_activation.argtypes = [vp, C.c_int, C.c_int]
    L.lz_push_batchnorm.argtypes = [vp, C.c_int, C.c_int, f32p, f32p,
                                    C.c_float]
    L.lz_finalize.argtypes = [vp]
    L.lz_network_initialize.argtypes = [vp]
    L.lz_network_get.argtypes = []
    L.lz_network_get.restype = vp
    L.lz_workspace_create.argtypes = [vp, C.c_int, C.POINTER(vp)]
    L.lz_workspace_destroy.argtypes = [vp]
    L.lz_workspace_destroy.restype = None
    L.lz_forward_device_ws.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int, vp,
                                       vp, vp, vp]
    L.lz_set_softmax_temp.argtypes = [vp, C.c_float]
    L.lz_set_layered_max.argtypes = [vp, C.c_int]
    L.lz_forward.argtypes = [vp, C.c_int, C.c_int, C.c_int,
                             C.POINTER(C.c_uint8), f32p, f32p]
    L.lz_forward_device.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp,
                                    vp]
    L.lz_forward_device_noalloc.argtypes = [vp, C.c_int, C.c_int, C.c_int,
                                            vp, vp, vp, vp]
    L.lz_forward_timed.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp,
                                   vp, f32p, C.c_int, C.POINTER(C.c_int)]
    L.lz_forward_ragged.argtypes = [vp, C.c_int, C.c_int, C.c_int,
                                    C.POINTER(C.c_int32),
                                    C.POINTER(C.c_uint8), f32p, f32p]
    L.lz_forward_ragged_device.argtypes = [vp, C.c_int, C.c_int, C.c_int,
                                           C.POINTER(C.c_int32), vp, vp, vp,
                                           vp]
    L.lz_autotune.argtypes = [vp, C.c_int, C.c_int, C.c_char_p, C.c_size_t]
    L.lz_autotune_ragged.argtypes = [vp, C.c_int, C.c_int,
                                     C.POINTER(C.c_int32), C.c_char_p,
                                     C.c_size_t]
    L.lz_queue_benchmark.argtypes = [vp, C.c_int, C.POINTER(C.c_int32),
                                     C.POINTER(C.c_uint8),
                                     C.POINTER(C.c_uint8), C.c_int, C.c_int,
                                     C.POINTER(C.c_double)]
    L.lz_autotune_save.argtypes = [vp, C.c_char_p]
    L.lz_autotune_load.argtypes = [vp, C.c_char_p]
    L.lz_queue_start.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.lz_submit.argtypes = [vp, C.c_int, C.POINTER(C.c_uint8), DONE_FN, vp]
    L.lz_submit_side.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_uint8),
                                 DONE_FN, vp]
    L.lz_flush.argtypes = [vp]
    L.lz_join.argtypes = [vp]
    L.lz_queue_depth.argtypes = [vp]
    L.lz_queue_stats.argtypes = [vp, C.POINTER(C.c_int64)]
    L.lz_queue_stop.argtypes = [vp]
    L.lz_normal_fill.argtypes = [C.c_uint32, C.c_size_t, C.c_float, f32p]
    L.lz_device_name.argtypes = [vp]
    L.lz_device_name.restype = C.c_char_p
    L.lz_last_error.argtypes = [vp]
    L.lz_last_error.restype = C.c_char_p
    L.lz_destroy.argtypes = [vp]
    L.lz_destroy.restype = None
    _lib = L
    return L


def header_functions() -> list[str]:
    """Names of the functions include/leela_amd.h declares."""
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lz_\w+)\s*\(", text)))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a, ct=C.c_float):
    return a.ctypes.data_as(C.POINTER(ct))


def normal_fill(seed: int, n: int, stddev: float) -> np.ndarray:
    out = np.empty(n, np.float32)
    rc = load_library().lz_normal_fill(seed, n, stddev, _ptr(out))
    if rc != LZ_OK:
        raise LeelaError(rc, "lz_normal_fill")
    return out


class Backend:
    """One lz_ctx: device weights for the policy and value nets."""

    def __init__(self, device: int = 0):
        self._L = load_library()
        self._ctx = C.c_void_p()
        self._shared = False
        rc = self._L.lz_init(C.byref(self._ctx), device)
        if rc != LZ_OK:
            raise LeelaError(rc, self._L.lz_last_error(None).decode())

    @classmethod
    def default(cls) -> "Backend":
        """Network::get_Network: the process-wide context with the
        reference's layer tables and the synthetic weight set
        (lz_network_get); close() leaves it alive."""
        b = cls.__new__(cls)
        b._L = load_library()
        ptr = b._L.lz_network_get()
        if not ptr:
            raise LeelaError(LZ_ERR_STATE, "lz_network_get failed")
        b._ctx = C.c_void_p(ptr)
        b._shared = True
        return b

    def _check(self, rc: int, what: str):
        if rc != LZ_OK:
            raise LeelaError(rc, f"{what}: "
                             f"{self._L.lz_last_error(self._ctx).decode()}")

    @property
    def device_name(self) -> str:
        return self._L.lz_device_name(self._ctx).decode()

    def push_convolve(self, net: int, filt: int, cin: int, cout: int, w, b):
        w, b = _f32(w), _f32(b)
        self._check(self._L.lz_push_convolve(self._ctx, net, filt, cin, cout,
                                             _ptr(w), _ptr(b)),
                    "lz_push_convolve")

    def push_innerproduct(self, net: int, inputs: int, outputs: int, w, b):
        w, b = _f32(w), _f32(b)
        self._check(self._L.lz_push_innerproduct(self._ctx, net, inputs,
                                                 outputs, _ptr(w), _ptr(b)),
                    "lz_push_innerproduct")

    def push_activation(self, net: int, kind: int):
        """lz_push_activation: LZ_ACT_ELU 0, LZ_ACT_RELU 1, LZ_ACT_NONE 2
        for the last pushed conv."""
        self._check(self._L.lz_push_activation(self._ctx, net, kind),
                    "lz_push_activation")

    def push_batchnorm(self, net: int, means, variances, scale: float):
        """lz_push_batchnorm on the last pushed conv (Caffe BatchNorm
        blobs: means, variances, stored scale factor)."""
        m, v = _f32(means), _f32(variances)
        if m.size != v.size:
            raise ValueError("means and variances differ in length")
        self._check(self._L.lz_push_batchnorm(self._ctx, net, m.size, _ptr(m),
                                              _ptr(v), float(scale)),
                    "lz_push_batchnorm")

    def load_net(self, net: int, layers):
        """layers: (kind, filter, cin, cout, w, b) with kind 0=conv 1=FC."""
        for kind, filt, cin, cout, w, b in layers:
            if kind == 0:
                self.push_convolve(net, filt, cin, cout, w, b)
            else:
                self.push_innerproduct(net, cin, cout, w, b)

    def finalize(self):
        self._check(self._L.lz_finalize(self._ctx), "lz_finalize")

    def set_layered_max(self, max_boards: int):
        """lz_set_layered_max: batches of at most max_boards 19x19
        positions run the layered 16-bit form (0: never)."""
        self._check(self._L.lz_set_layered_max(self._ctx, max_boards),
                    "lz_set_layered_max")

    def set_temperature(self, t: float):
        self._check(self._L.lz_set_softmax_temp(self._ctx, t),
                    "lz_set_softmax_temp")

    def forward(self, net: int, precision: int, planes: np.ndarray):
        """planes uint8 [n][C][361] (C: the first conv's inputs, 32 for the
        reference's nets) -> (raw, out) float32."""
        planes = np.ascontiguousarray(planes, dtype=np.uint8)
        n = planes.shape[0]
        per = 361 if net == NET_POLICY else 1
        raw = np.zeros((n, per), np.float32)
        out = np.zeros((n, per), np.float32)
        self._check(self._L.lz_forward(self._ctx, net, precision, n,
                                       _ptr(planes, C.c_uint8), _ptr(raw),
                                       _ptr(out)), "lz_forward")
        return raw, out

    def forward_device(self, net: int, precision: int, n: int, d_planes: int,
                       d_raw: int, d_out: int, stream: int | None = None):
        """Device-pointer forward (ints are raw device addresses)."""
        self._check(self._L.lz_forward_device(self._ctx, net, precision, n,
                                              d_planes, d_raw, d_out, stream),
                    "lz_forward_device")

    def workspace(self, max_boards: int):
        """lz_workspace_create: a workspace for forward_device_ws (free it
        with free_workspace)."""
        w = C.c_void_p()
        self._check(self._L.lz_workspace_create(self._ctx, max_boards,
                                                C.byref(w)),
                    "lz_workspace_create")
        return w

    def free_workspace(self, w):
        self._L.lz_workspace_destroy(w)

    def forward_device_ws(self, ws, net: int, precision: int, n: int,
                          d_planes: int, d_raw: int, d_out: int,
                          stream: int | None = None):
        """lz_forward_device with the caller's workspace."""
        self._check(self._L.lz_forward_device_ws(self._ctx, ws, net,
                                                 precision, n, d_planes,
                                                 d_raw, d_out, stream),
                    "lz_forward_device_ws")

    def forward_device_noalloc(self, net: int, precision: int, n: int,
                               d_planes: int, d_raw: int, d_out: int,
                               stream: int | None = None):
        """lz_forward_device_noalloc: 16-bit / FP8 forms, n <= 2048, no
        allocation (static scratch array)."""
        self._check(self._L.lz_forward_device_noalloc(
            self._ctx, net, precision, n, d_planes, d_raw, d_out, stream),
            "lz_forward_device_noalloc")

    def forward_timed(self, net: int, precision: int, n: int, d_planes: int,
                      d_raw: int, d_out: int, stream: int | None = None):
        """Device-pointer forward; returns per-launch durations in ms."""
        ms = np.zeros(64, np.float32)
        cnt = C.c_int(0)
        self._check(self._L.lz_forward_timed(self._ctx, net, precision, n,
                                             d_planes, d_raw, d_out, stream,
                                             _ptr(ms), 64, C.byref(cnt)),
                    "lz_forward_timed")
        return ms[:cnt.value].copy()

    def forward_ragged(self, net: int, precision: int, boards):
        """boards: list of uint8 [32][S*S] plane stacks of any sides 5..19.
        Returns per-board (raw, out) arrays (policy S*S each, value 1)."""
        sides = np.array([int(round(np.sqrt(b.shape[-1]))) for b in boards],
                         np.int32)
        flat = np.concatenate([np.ascontiguousarray(b, np.uint8).ravel()
                               for b in boards])
        per = [s * s if net == NET_POLICY else 1 for s in sides]
        tot = int(sum(per))
        raw = np.zeros(tot, np.float32)
        out = np.zeros(tot, np.float32)
        self._check(self._L.lz_forward_ragged(
            self._ctx, net, precision, len(boards), _ptr(sides, C.c_int32),
            _ptr(flat, C.c_uint8), _ptr(raw), _ptr(out)), "lz_forward_ragged")
        offs = np.concatenate([[0], np.cumsum(per)])
        return ([raw[offs[i]:offs[i + 1]] for i in range(len(boards))],
                [out[offs[i]:offs[i + 1]] for i in range(len(boards))])

    def forward_ragged_device(self, net: int, precision: int, sides,
                              d_planes: int, d_raw: int, d_out: int,
                              stream: int | None = None):
        sides = np.ascontiguousarray(sides, np.int32)
        self._check(self._L.lz_forward_ragged_device(
            self._ctx, net, precision, len(sides), _ptr(sides, C.c_int32),
            d_planes, d_raw, d_out, stream), "lz_forward_ragged_device")

    def autotune(self, precision: int, n: int) -> str:
        """Times the fused kernels' launch shapes for batches of n and keeps
        the fastest; returns the timing report."""
        buf = C.create_string_buffer(8192)
        self._check(self._L.lz_autotune(self._ctx, precision, n, buf,
                                        len(buf)), "lz_autotune")
        return buf.value.decode()

    def autotune_ragged(self, precision: int, sides) -> str:
        """autotune for a ragged batch of boards of these sides."""
        sides = np.ascontiguousarray(sides, np.int32)
        buf = C.create_string_buffer(8192)
        self._check(self._L.lz_autotune_ragged(
            self._ctx, precision, len(sides), _ptr(sides, C.c_int32), buf,
            len(buf)), "lz_autotune_ragged")
        return buf.value.decode()

    def queue_benchmark(self, sides, policy_boards, value_boards,
                        threads: int = 16, rounds: int = 1) -> float:
        """lz_queue_benchmark over boards (lists of [32][S*S] arrays);
        returns positions per second (both nets)."""
        sides = np.ascontiguousarray(sides, np.int32)
        pp = np.concatenate([np.ascontiguousarray(b, np.uint8).ravel()
                             for b in policy_boards])
        vp_ = np.concatenate([np.ascontiguousarray(b, np.uint8).ravel()
                              for b in value_boards])
        out = C.c_double(0.0)
        rc = self._L.lz_queue_benchmark(
            self._ctx, len(sides), _ptr(sides, C.c_int32),
            _ptr(pp, C.c_uint8), _ptr(vp_, C.c_uint8), threads, rounds,
            C.byref(out))
        if rc != LZ_OK:
            raise LeelaError(rc, "lz_queue_benchmark")
        return out.value

    def autotune_save(self, path: str):
        self._check(self._L.lz_autotune_save(self._ctx, path.encode()),
                    "lz_autotune_save")

    def autotune_load(self, path: str):
        self._check(self._L.lz_autotune_load(self._ctx, path.encode()),
                    "lz_autotune_load")

    # ---- batched leaf queue (lz_queue_start / lz_submit / lz_join) ----

    def queue_start(self, precision: int, max_batch: int = 1024,
                    max_wait_us: int = 200):
        """Starts the dispatcher. Results are delivered to the callbacks
        given to submit(); keep the callback objects alive until join()."""
        self._check(self._L.lz_queue_start(self._ctx, precision, max_batch,
                                           max_wait_us), "lz_queue_start")

    def submit(self, net: int, planes: np.ndarray, done, user: int = 0):
        """planes uint8 [32][361]; done is a DONE_FN-wrapped callable."""
        planes = np.ascontiguousarray(planes, dtype=np.uint8)
        rc = self._L.lz_submit(self._ctx, net, _ptr(planes, C.c_uint8), done,
                               C.c_void_p(user))
        if rc != LZ_OK:
            raise LeelaError(rc, "lz_submit")

    def submit_side(self, net: int, side: int, planes: np.ndarray, done,
                    user: int = 0):
        """A board of any side 5..19: planes uint8 [32][side*side]."""
        planes = np.ascontiguousarray(planes, dtype=np.uint8)
        rc = self._L.lz_submit_side(self._ctx, net, side,
                                    _ptr(planes, C.c_uint8), done,
                                    C.c_void_p(user))
        if rc != LZ_OK:
            raise LeelaError(rc, "lz_submit_side")

    def flush(self):
        self._check(self._L.lz_flush(self._ctx), "lz_flush")

    def join(self):
        self._check(self._L.lz_join(self._ctx), "lz_join")

    def queue_depth(self) -> int:
        return int(self._L.lz_queue_depth(self._ctx))

    def queue_stats(self) -> dict:
        s = np.zeros(6, np.int64)
        self._check(self._L.lz_queue_stats(self._ctx, _ptr(s, C.c_int64)),
                    "lz_queue_stats")
        return {"batches": int(s[0]), "positions": int(s[1]),
                "per_bucket": {b: int(c) for b, c in
                               zip((64, 256, 1024, 2048), s[2:])}}

    def queue_stop(self):
        self._check(self._L.lz_queue_stop(self._ctx), "lz_queue_stop")

    def close(self):
        if self._ctx and not self._shared:
            self._L.lz_destroy(self._ctx)
        self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
