"""Device-resident electron kinetic engine driving libbz.so.

PyTorch supplies device memory and the stream; every stage runs in the
hand-written HIP kernels of libbz.so.  The particle state lives in two SoA
buffers (ping-pong); the live count, step and time live in the library's
device clock, so a step is enqueued without any host synchronisation and can
be captured in a hipGraph.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib
from . import physics as ph

try:
    import torch
except ImportError as exc:  # pragma: no cover
    raise ImportError("boltzmann_amd needs PyTorch (ROCm) for device memory") from exc


def _tptr(t: Optional["torch.Tensor"]):
    return None if t is None else C.c_void_p(t.data_ptr())


class DeviceState:
    """SoA particle arrays (x, vx, vy, vz, w) of one precision on the device."""

    def __init__(self, capacity: int, dtype, device):
        self.arrays = [torch.zeros(capacity, dtype=dtype, device=device) for _ in range(5)]
        self.struct = _lib.State(*[a.data_ptr() for a in self.arrays])

    def __getitem__(self, k):
        return self.arrays[k]


class KineticEngine:
    """One run of the electron kinetic step on one GPU."""

    def __init__(self, cfg: ph.RunConfig, device: int = 0, capacity: Optional[int] = None,
                 rng_mode: int = _lib.BZ_RNG_PHILOX, max_keys: Optional[int] = None):
        if not torch.cuda.is_available():
            raise _lib.BzError("no GPU visible: the electron step runs only in the HIP kernels")
        self.lib = _lib.load()
        self.cfg = cfg
        self.dev = torch.device("cuda", device)
        torch.cuda.set_device(self.dev)
        self.tdtype = torch.float64 if cfg.precision == _lib.BZ_F64 else torch.float32
        self.capacity = int(capacity or max(1024, int(cfg.n_particles * cfg.capacity_factor)))
        nn = cfg.n_nodes
        self.n_keys_max = int(max_keys or nn * max(cfg.nb, 1))
        self.ctx = C.c_void_p()
        _lib.check(self.lib.bz_ctx_create(C.byref(self.ctx), device, cfg.precision,
                                          self.capacity, max(nn, 2), max(self.n_keys_max, 1)),
                   "bz_ctx_create")
        self.stream = torch.cuda.current_stream(self.dev)
        _lib.check(self.lib.bz_set_stream(self.ctx, C.c_void_p(self.stream.cuda_stream)))
        self.bufs = [DeviceState(self.capacity, self.tdtype, self.dev) for _ in range(2)]
        self.cur = 0
        # collision model
        self.table = None
        self.xs = None
        if cfg.xs_mode == _lib.BZ_XS_TABLE:
            self.table = ph.build_table(cfg.xsec, cfg.gas_density_m3, cfg.np_dtype)
            t = self.table
            self._tab_dev = [torch.from_numpy(a).to(self.dev) for a in
                             (t.energy, t.sig.ravel(), t.slope.ravel())]
            self._hint_dev = torch.from_numpy(t.hint).to(self.dev)
            self.xs = _lib.XSec(self._tab_dev[0].data_ptr(), self._tab_dev[1].data_ptr(),
                                self._tab_dev[2].data_ptr(), self._hint_dev.data_ptr(),
                                t.energy.size, t.hint.size, t.hint_base, t.hint_bits)
        self.params = ph.push_params(cfg, rng_mode, self.table)
        # grid quantities
        N = cfg.n_cells if cfg.dim == 1 else 1
        f64 = dict(dtype=torch.float64, device=self.dev)
        self.moments = torch.zeros(3 * nn, **f64)
        self.E_t = torch.zeros(max(N, 1), dtype=self.tdtype, device=self.dev)
        if cfg.dim == 0:
            self.E_t.fill_(cfg.e_field0_vcm())
        if cfg.dim == 1:
            self.fp = ph.field_params(cfg)
            self.ni = torch.full((N + 1,), cfg.density_cm3, **f64)
            self.ns = torch.zeros(N + 1, **f64)
            self.Te = torch.zeros(N + 1, **f64)
            self.g2 = torch.zeros(N + 1, **f64)
            self.V = torch.zeros(max(N - 1, 1), **f64)
            self.E = torch.zeros(N, **f64)
            self.chol = torch.from_numpy(ph.poisson_cholesky(N, cfg.dx).ravel()).to(self.dev)

    # ------------------------------------------------------------------ state
    @property
    def state(self) -> DeviceState:
        return self.bufs[self.cur]

    def load_state(self, arrays: Sequence[np.ndarray], step: int = 0, t: float = 0.0) -> None:
        n = int(arrays[0].size)
        if n > self.capacity:
            raise _lib.BzError(f"{n} particles exceed capacity {self.capacity}")
        for dst, src in zip(self.state.arrays, arrays):
            dst[:n].copy_(torch.from_numpy(np.ascontiguousarray(src)).to(self.tdtype))
        torch.cuda.synchronize(self.dev)
        _lib.check(self.lib.bz_clock_set(self.ctx, n, step, t), "bz_clock_set")

    def clock(self):
        n = C.c_int64()
        s = C.c_uint64()
        t = C.c_double()
        cnt = (C.c_uint64 * _lib.BZ_NCOUNTERS)()
        _lib.check(self.lib.bz_clock_get(self.ctx, C.byref(n), C.byref(s), C.byref(t), cnt))
        return n.value, s.value, t.value, list(cnt)

    def count(self) -> int:
        return self.clock()[0]

    def counters(self) -> dict:
        cnt = self.clock()[3]
        return {k: int(cnt[i]) for i, k in enumerate(_lib.COUNTER_NAMES)}

    def fetch_state(self):
        n = self.count()
        return tuple(a[:n].cpu().numpy() for a in self.state.arrays)

    # ------------------------------------------------------------------ stages
    def push(self, R: Optional["torch.Tensor"] = None, G: Optional["torch.Tensor"] = None):
        """Fused collision + push + deposit + compaction into the other buffer."""
        p = self.params
        if R is not None:
            p = _lib.Params.from_buffer_copy(self.params)
            p.rng_mode = _lib.BZ_RNG_STREAM
            p.r_stride = R.shape[1]
        src, dst = self.bufs[self.cur], self.bufs[1 - self.cur]
        E = self.E_t if self.cfg.dim == 1 else None
        _lib.check(self.lib.bz_push_mcc(
            self.ctx, C.byref(src.struct), C.byref(dst.struct), _tptr(E), C.byref(p),
            C.byref(self.xs) if self.xs is not None else None, _tptr(R), _tptr(G),
            _tptr(self.moments)), "bz_push_mcc")
        self.cur = 1 - self.cur

    def deposit(self):
        _lib.check(self.lib.bz_deposit(self.ctx, C.byref(self.state.struct), C.byref(self.params),
                                       _tptr(self.moments)), "bz_deposit")

    def field_step(self, advance: bool = True):
        if self.cfg.dim != 1:
            if advance:
                _lib.check(self.lib.bz_clock_advance(self.ctx, self.cfg.dt))
            return
        _lib.check(self.lib.bz_field_step(
            self.ctx, C.byref(self.fp), _tptr(self.moments), _tptr(self.ni), _tptr(self.ns),
            _tptr(self.Te), _tptr(self.g2), _tptr(self.V), _tptr(self.E), _tptr(self.E_t),
            _tptr(self.chol), 1 if advance else 0), "bz_field_step")

    def initialize_fields(self):
        """Initial deposit and Poisson solve at t (boltzmann.py:363-399)."""
        self.deposit()
        if self.cfg.dim == 1:
            self.field_step(advance=False)

    def step(self):
        self.push()
        self.field_step(advance=True)

    def synchronize(self):
        _lib.check(self.lib.bz_synchronize(self.ctx))

    def close(self):
        if self.ctx:
            self.lib.bz_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass
