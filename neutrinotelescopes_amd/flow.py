"""Flow surrogate on the GPU: host mirror of the reference's RQNormFlow
(src/surrogate_models/extended_cascade_model.jl:53-139, 237-295) over the
libntflow kernels (include/ntflow.h, compiled into libntprop.so).

`FlowModel` holds the weights on the device; `FlowRows` holds a batch of
training or evaluation rows (features, time, count, flags) in HBM.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _abi

N_IN = _abi.NTF_N_IN
HIDDEN = _abi.NTF_HIDDEN
P_PAD = _abi.NTF_P_PAD
N_PARAMS = _abi.NTF_N_PARAMS
N_BINS = _abi.NTF_BINS
N_LAYERS = _abi.NTF_LAYERS
N_WEIGHTS = _abi.NTF_N_WEIGHTS
ROW_TIME, ROW_COUNT = _abi.NTF_ROW_TIME, _abi.NTF_ROW_COUNT

_SHAPES = [("W1", (N_IN, HIDDEN)), ("b1", (HIDDEN,)), ("W2", (HIDDEN, HIDDEN)), ("b2", (HIDDEN,)),
           ("W3", (HIDDEN, P_PAD)), ("b3", (P_PAD,))]


def unflatten(flat: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in _SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(flat[o:o + n]).reshape(shape)
        o += n
    assert o == N_WEIGHTS
    return out


def flatten(w: dict) -> np.ndarray:
    return np.concatenate([np.asarray(w[name], dtype=np.float32).reshape(-1) for name, _ in _SHAPES])


def init_weights(seed: int = 0, n_in: int = 5) -> np.ndarray:
    """Flux Dense defaults (Glorot-uniform weights, zero biases) for the
    n_in → 128 → 128 → 78 network; padding rows/columns are zero."""
    rng = np.random.default_rng(seed)

    def glorot(fan_in, fan_out):
        lim = np.sqrt(6.0 / (fan_in + fan_out))
        return rng.uniform(-lim, lim, (fan_in, fan_out))

    w = {"W1": np.zeros((N_IN, HIDDEN)), "b1": np.zeros(HIDDEN), "W2": glorot(HIDDEN, HIDDEN),
         "b2": np.zeros(HIDDEN), "W3": np.zeros((HIDDEN, P_PAD)), "b3": np.zeros(P_PAD)}
    w["W1"][:n_in] = glorot(n_in, HIDDEN)
    w["W3"][:, :N_PARAMS] = glorot(HIDDEN, N_PARAMS)
    return flatten(w)


@dataclass
class FlowRows:
    """A batch of rows in device memory (torch tensors on the GPU)."""
    x: "object"        # [n, 16] float32
    t: "object"        # [n] float32
    count: "object"    # [n] float32
    flags: "object"    # [n] uint8
    aux: "object" = None   # [n, 2] float32 (log λ offset, row weight) or None

    @property
    def n(self) -> int:
        return int(self.t.shape[0])

    @classmethod
    def from_numpy(cls, x, t=None, count=None, flags=None, aux=None, device="cuda:0"):
        import torch
        x = np.asarray(x, dtype=np.float32)
        n = x.shape[0]
        if x.shape[1] > N_IN:
            raise ValueError(f"at most {N_IN} features")
        xp = np.zeros((n, N_IN), np.float32)
        xp[:, :x.shape[1]] = x
        t = np.zeros(n, np.float32) if t is None else np.asarray(t, np.float32)
        count = np.zeros(n, np.float32) if count is None else np.asarray(count, np.float32)
        flags = np.full(n, ROW_TIME | ROW_COUNT, np.uint8) if flags is None else np.asarray(flags, np.uint8)
        dev = torch.device(device)
        a = None if aux is None else torch.from_numpy(np.ascontiguousarray(aux, np.float32).reshape(n, 2)).to(dev)
        return cls(torch.from_numpy(xp).to(dev), torch.from_numpy(t).to(dev), torch.from_numpy(count).to(dev),
                   torch.from_numpy(flags).to(dev), a)

    def slice(self, idx):
        return FlowRows(self.x[idx], self.t[idx], self.count[idx], self.flags[idx],
                        None if self.aux is None else self.aux[idx])

    def args(self):
        return (C.c_void_p(self.x.data_ptr()), C.c_void_p(self.t.data_ptr()), C.c_void_p(self.count.data_ptr()),
                C.c_void_p(self.flags.data_ptr()), None if self.aux is None else C.c_void_p(self.aux.data_ptr()),
                C.c_int64(self.n))


@dataclass
class Segments:
    """Contiguous segments of a row array split into chunks of bounded length
    (device int64 tensors): chunk c = [chunk_offsets[c], chunk_offsets[c+1]),
    segment s = chunks [segment_chunks[s], segment_chunks[s+1])."""
    chunk_offsets: "object"
    segment_chunks: "object"

    @staticmethod
    def tables(offsets, chunk: int = 8192):
        """(chunk_offsets, segment_chunks) as numpy for segment offsets."""
        offsets = np.asarray(offsets, np.int64)
        n_chunks = (np.diff(offsets) + chunk - 1) // chunk
        seg_chunks = np.concatenate([[0], np.cumsum(n_chunks)]).astype(np.int64)
        within = np.arange(seg_chunks[-1]) - np.repeat(seg_chunks[:-1], n_chunks)
        starts = np.repeat(offsets[:-1], n_chunks) + chunk * within
        return np.concatenate([starts, offsets[-1:]]).astype(np.int64), seg_chunks

    @classmethod
    def from_offsets(cls, offsets, chunk: int = 8192, device="cuda:0") -> "Segments":
        import torch
        co, sc = cls.tables(offsets, chunk)
        dev = torch.device(device)
        return cls(torch.from_numpy(co).to(dev), torch.from_numpy(sc).to(dev))


class FlowModel:
    """Conditional RQ-spline flow (arrival-time density + log expected count)
    with its weights in HBM."""

    def __init__(self, device: int = 0, weights: np.ndarray | None = None, seed: int = 0, n_in: int = 5):
        import torch
        self.torch = torch
        self.lib = _abi.load_library()
        self.device = device
        self.dev = torch.device(f"cuda:{device}")
        h = C.c_void_p()
        _abi.check(self.lib.ntf_create(device, C.byref(h)), "ntf_create")
        self.handle = h
        self.n_in = n_in
        self.set_weights(init_weights(seed, n_in) if weights is None else weights)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.ntf_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass

    def _stream(self):
        return C.c_void_p(self.torch.cuda.current_stream(self.dev).cuda_stream)

    def set_weights(self, flat: np.ndarray):
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        if flat.size != N_WEIGHTS:
            raise ValueError("weight vector has the wrong size")
        _abi.check(self.lib.ntf_set_weights(self.handle, flat.ctypes.data_as(C.c_void_p)), "ntf_set_weights")

    def get_weights(self) -> np.ndarray:
        self.torch.cuda.synchronize(self.dev)
        out = np.empty(N_WEIGHTS, np.float32)
        _abi.check(self.lib.ntf_get_weights(self.handle, out.ctypes.data_as(C.c_void_p)), "ntf_get_weights")
        return out

    def forward(self, rows: FlowRows, want_params: bool = True):
        """Network outputs [n, 80] (or None) and per-row losses [n] (device)."""
        torch = self.torch
        p = torch.empty((rows.n, P_PAD), dtype=torch.float32, device=self.dev) if want_params else None
        loss = torch.empty(rows.n, dtype=torch.float32, device=self.dev)
        _abi.check(self.lib.ntf_forward(self.handle, *rows.args(),
                                        C.c_void_p(p.data_ptr()) if p is not None else None,
                                        C.c_void_p(loss.data_ptr()), self._stream()), "ntf_forward")
        return p, loss

    def loss_grad(self, rows: FlowRows, scale: float = 1.0):
        """(Σ row loss, scale·Σ ∂loss/∂weights) as host numpy."""
        torch = self.torch
        g = torch.empty(N_WEIGHTS, dtype=torch.float32, device=self.dev)
        loss = torch.empty(1, dtype=torch.float64, device=self.dev)
        _abi.check(self.lib.ntf_loss_grad(self.handle, *rows.args(), C.c_float(scale), C.c_void_p(g.data_ptr()),
                                          C.c_void_p(loss.data_ptr()), self._stream()), "ntf_loss_grad")
        return float(loss.cpu()[0]), g.cpu().numpy()

    def pair_rows(self, sources, modules, axes, norm_mean, norm_std, scale_knots, log_yield_1gev: float):
        """Rows (x [n, 16], aux [n, 2]) of every (source, module, PMT) pair on
        the device; sources [n_src, 8], modules [n_mod, 3], axes [n_pmt, 3]
        are device tensors."""
        torch = self.torch
        n = sources.shape[0] * modules.shape[0] * axes.shape[0]
        x = torch.empty((n, N_IN), dtype=torch.float32, device=self.dev)
        aux = torch.empty((n, 2), dtype=torch.float32, device=self.dev)
        norm = np.ascontiguousarray(np.concatenate([norm_mean, norm_std]), np.float32)
        knots = np.ascontiguousarray(scale_knots, np.float32)
        _abi.check(self.lib.ntf_pair_rows(self.handle, C.c_void_p(sources.data_ptr()), sources.shape[0],
                                          C.c_void_p(modules.data_ptr()), modules.shape[0],
                                          C.c_void_p(axes.data_ptr()), axes.shape[0], norm.ctypes.data_as(C.c_void_p),
                                          knots.ctypes.data_as(C.c_void_p), knots.shape[0], C.c_float(log_yield_1gev),
                                          C.c_void_p(x.data_ptr()), C.c_void_p(aux.data_ptr()), self._stream()),
                   "ntf_pair_rows")
        return x, aux

    def time_logpdf(self, params, pair, t):
        """logpdf of times t under the outputs `params` [n_pairs, 80] of pair[i]."""
        torch = self.torch
        out = torch.empty(t.shape[0], dtype=torch.float32, device=self.dev)
        _abi.check(self.lib.ntf_time_logpdf(self.handle, C.c_void_p(params.data_ptr()), C.c_int64(params.shape[0]),
                                            C.c_void_p(pair.data_ptr()), C.c_void_p(t.data_ptr()),
                                            C.c_int64(t.shape[0]), C.c_void_p(out.data_ptr()), self._stream()),
                   "ntf_time_logpdf")
        return out

    def segment_sum(self, values, segments: "Segments", out, sign: float = 1.0, skip=None,
                    accumulate: bool = False):
        """out[s] (=/+=) sign·Σ of segment s of `values` (skip != 0 left out)."""
        _abi.check(self.lib.ntf_segment_sum(self.handle, C.c_void_p(values.data_ptr()),
                                            None if skip is None else C.c_void_p(skip.data_ptr()),
                                            C.c_void_p(segments.chunk_offsets.data_ptr()),
                                            C.c_int64(segments.chunk_offsets.shape[0] - 1),
                                            C.c_void_p(segments.segment_chunks.data_ptr()), C.c_int64(out.shape[0]),
                                            C.c_float(sign), int(accumulate), C.c_void_p(out.data_ptr()),
                                            self._stream()), "ntf_segment_sum")
        return out

    def train_step(self, rows: FlowRows, lr: float, loss_out=None):
        """One Adam step on the mean loss of `rows`; `loss_out` (device float64
        tensor, optional) receives the loss before the update."""
        _abi.check(self.lib.ntf_train_step(self.handle, *rows.args(), C.c_float(lr),
                                           C.c_void_p(loss_out.data_ptr()) if loss_out is not None else None,
                                           self._stream()), "ntf_train_step")

    def train_step_gather(self, rows: FlowRows, idx, lr: float, loss_out=None):
        """train_step on the batch rows.slice(idx) (idx: device int64 tensor of
        row indices) without building it: the kernel reads the rows in place."""
        import torch
        if idx.dtype != torch.int64 or not idx.is_contiguous() or idx.device != self.dev:
            raise ValueError("idx: a contiguous int64 tensor on the model's device")
        args = rows.args()
        _abi.check(self.lib.ntf_train_step_gather(self.handle, *args[:5], args[5], C.c_void_p(idx.data_ptr()),
                                                  C.c_int64(idx.numel()), C.c_float(lr),
                                                  C.c_void_p(loss_out.data_ptr()) if loss_out is not None else None,
                                                  self._stream()), "ntf_train_step_gather")
