"""Device-resident electron kinetic engine driving libbz.so.

PyTorch supplies device memory and the stream; every stage runs in the
hand-written HIP kernels of libbz.so.  The particle state lives in two SoA
buffers; the slot count, step and time live in the library's device clock,
so a step is enqueued without any host synchronisation and can be captured
in a hipGraph.

1-D runs push in place: an absorbed particle leaves a hole (w = 0) and the
secondaries are appended after the slots in use.  Resampling rewrites the
state without holes; otherwise the state is compacted every
`compact_every` steps (holes come only from absorption at the walls, ~0.02 %
of the electrons per step in the discharges, so 64 steps leave ~1 %).  0-D
runs compact in every push.
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
                 rng_mode: int = _lib.BZ_RNG_PHILOX, max_keys: Optional[int] = None,
                 compact_every: int = 64, push_keys: Optional[bool] = None,
                 inplace_resample: Optional[bool] = None,
                 compact_resampled_every: Optional[int] = None,
                 fill_holes: Optional[bool] = None, push_keys16: Optional[bool] = None):
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
        self.inplace = cfg.dim == 1
        self.compact_every = int(compact_every)
        self._since_compact = 0
        # 1-D runs resample in place; the holes a pass leaves are compacted
        # away before the push after every `compact_resampled_every` passes
        self.inplace_resample = self.inplace if inplace_resample is None else \
            bool(inplace_resample) and self.inplace
        if compact_resampled_every is None:
            compact_resampled_every = 4 if cfg.resample_every == 1 else 1
        self.compact_resampled_every = int(compact_resampled_every)
        self._resampled_since_compact = 0
        # compaction of the in-place runs: fill the holes below the live
        # count from the live slots above it (bz_compact_inplace), or copy a
        # stable filter to the other buffer (bz_compact)
        self.fill_holes = self.inplace_resample if fill_holes is None else \
            bool(fill_holes) and self.inplace
        # captured blocks of steps: (state buffer, steps) -> graph id
        self._graphs = {}
        self._cap_stream = None
        # directly launched 1-D steps (a resampling check every step): the
        # field step of step i runs on a side stream beside the resampling
        # of step i + 1, which does not read it (run() only; see _run_overlapped)
        self.overlap_field = cfg.dim == 1
        self._field_stream = None
        # collision model
        self.table = None
        self.xs = None
        if cfg.xs_mode != _lib.BZ_XS_FIT:
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
        # binning / resampling
        self.nb = max(cfg.nb, 1)
        self.n_keys = nn * self.nb
        i32 = dict(dtype=torch.int32, device=self.dev)
        self.counts = torch.zeros(self.n_keys + 1, **i32)
        self.offsets = torch.zeros(self.n_keys + 2, **i32)
        self.perm = torch.zeros(self.capacity, **i32)
        fr = cfg.bin_fractions if len(cfg.bin_fractions) == self.nb else ph.bin_fractions(self.nb)
        tg = np.array([int(f * cfg.target_ppc) for f in fr] * nn, dtype=np.int32)
        self.targets = torch.from_numpy(tg).to(self.dev)
        self.targets_host = tg
        thr = cfg.bin_thresholds if len(cfg.bin_thresholds) == self.nb - 1 else \
            ph.bin_thresholds(self.nb)
        self.thr = (C.c_double * 8)(*thr)
        self.policy = _lib.ResamplePolicy(cfg.split_below, cfg.merge_at, cfg.merge_keep,
                                          cfg.trigger_at, cfg.min_wt, cfg.split_to, 0, 0, 0, 0)
        # the 1-D push writes the binning keys of the next sort (4 bytes per
        # slot) when a sort follows every push: resampling every step with no
        # trigger to decline it
        if push_keys is None:
            always = not self.policy.ref_trigger and self.policy.trigger_at <= 0
            push_keys = (cfg.dim == 1 and cfg.target_ppc > 0 and cfg.resample_every == 1
                         and always)
        self.push_keys = bool(push_keys) and cfg.dim == 1
        _lib.check(self.lib.bz_set_push_keys(self.ctx, self.nb if self.push_keys else 0,
                                             self.thr), "bz_set_push_keys")
        # the in-place resampling reads the keys three times: 16-bit words
        # where they fit (bz_bin_sort, the key-order path, takes 32-bit ones)
        if push_keys16 is None:
            push_keys16 = self.inplace_resample
        self.push_keys16 = bool(push_keys16) and self.push_keys and nn * max(cfg.nb, 1) < 65535
        _lib.check(self.lib.bz_set_push_keys16(self.ctx, 1 if self.push_keys16 else 0),
                   "bz_set_push_keys16")

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
        """Slots in use (1-D: live particles and holes)."""
        return self.clock()[0]

    def live(self) -> int:
        """Live particles (w > 0); synchronises."""
        n = self.count()
        return int((self.state.arrays[4][:n] > 0).sum().item())

    def counters(self) -> dict:
        cnt = self.clock()[3]
        return {k: int(cnt[i]) for i, k in enumerate(_lib.COUNTER_NAMES)}

    def fetch_slots(self):
        """All slots in use, holes (w = 0) included: the arrays the next push
        reads (its random stream is keyed by slot index)."""
        n = self.count()
        return tuple(a[:n].cpu().numpy() for a in self.state.arrays)

    def fetch_state(self):
        """Live particles in slot order (holes dropped): survivors, then the
        secondaries of the last pushes in parent order."""
        n = self.count()
        arrs = [a[:n].cpu().numpy() for a in self.state.arrays]
        live = arrs[4] > 0
        if live.all():
            return tuple(arrs)
        return tuple(np.ascontiguousarray(a[live]) for a in arrs)

    # ------------------------------------------------------------------ stages
    def push(self, R: Optional["torch.Tensor"] = None, G: Optional["torch.Tensor"] = None):
        """Fused collision + push + deposit: in place with births appended
        (1-D, the other buffer stages the births), or compacting into the
        other buffer (0-D)."""
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
        if self.inplace:
            self._since_compact += 1
        else:
            self.cur = 1 - self.cur

    def compact(self):
        """Remove the holes (w = 0): in place by hole filling, or as a
        stable filter into the other buffer."""
        if self.fill_holes:
            _lib.check(self.lib.bz_compact_inplace(self.ctx, C.byref(self.state.struct)),
                       "bz_compact_inplace")
        else:
            src, dst = self.bufs[self.cur], self.bufs[1 - self.cur]
            _lib.check(self.lib.bz_compact(self.ctx, C.byref(src.struct), C.byref(dst.struct)),
                       "bz_compact")
            self.cur = 1 - self.cur
        self._since_compact = 0
        self._resampled_since_compact = 0

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

    def bin_sort(self):
        """Keys (cell, energy bin), per-key counts/offsets and the stable permutation."""
        _lib.check(self.lib.bz_bin_sort(self.ctx, C.byref(self.state.struct),
                                        C.byref(self.params), self.thr, self.nb,
                                        _tptr(self.counts), _tptr(self.offsets),
                                        _tptr(self.perm)), "bz_bin_sort")

    def trigger(self, counts: np.ndarray) -> bool:
        """Resample trigger on host counts (reference _check_for_resampling)."""
        c = counts[:self.n_keys].astype(np.int64)
        t = self.targets_host.astype(np.int64)
        ok = (t > 0) & (c > 0)
        if not self.policy.ref_trigger and self.policy.trigger_at <= 0:
            return True
        if self.policy.ref_trigger:
            last = (np.arange(self.n_keys) % self.nb) == self.nb - 1
            return bool(np.any(ok & last & (c >= 3.0 * t)))
        return bool(np.any(ok & (c >= self.policy.trigger_at * t)))

    def bin_count(self):
        """Per-key counts alone (the trigger's input), no sort."""
        _lib.check(self.lib.bz_bin_count(self.ctx, C.byref(self.state.struct),
                                         C.byref(self.params), self.thr, self.nb,
                                         _tptr(self.counts)), "bz_bin_count")

    def resample(self, force: bool = False) -> bool:
        """One resampling check.  1-D runs resample in place (the reference's
        layout: merged members become holes, clones and merged particles are
        appended; the trigger is evaluated on the device, so nothing is read
        back).  0-D runs count, and when triggered sort and split / merge into
        the other buffer in key order.  Returns whether the pass was enqueued
        (in place: always; the device decides whether it fires)."""
        if self.inplace_resample:
            pol = self.policy
            if force:
                pol = _lib.ResamplePolicy.from_buffer_copy(self.policy)
                pol.ref_trigger = 0
                pol.trigger_at = 0.0
            _lib.check(self.lib.bz_resample_inplace(
                self.ctx, C.byref(self.state.struct), C.byref(self.params), self.thr, self.nb,
                _tptr(self.targets), self.n_keys, C.byref(pol)), "bz_resample_inplace")
            self._resampled_since_compact += 1
            return True
        always = not self.policy.ref_trigger and self.policy.trigger_at <= 0
        # the counts are read back (a host synchronisation) only when the
        # trigger can decline; the sort runs only for a pass that happens
        if not force and not always:
            self.bin_count()
            if not self.trigger(self.counts.cpu().numpy()):
                return False
        self.bin_sort()
        src, dst = self.bufs[self.cur], self.bufs[1 - self.cur]
        _lib.check(self.lib.bz_resample(self.ctx, C.byref(src.struct), C.byref(dst.struct),
                                        C.byref(self.params), _tptr(self.counts),
                                        _tptr(self.offsets), _tptr(self.perm),
                                        _tptr(self.targets), self.n_keys,
                                        C.byref(self.policy)), "bz_resample")
        self.cur = 1 - self.cur
        self._since_compact = 0  # the resampled state holds no holes
        return True

    def _compact_due(self) -> bool:
        if not self.inplace:
            return False
        if self.compact_every and self._since_compact >= self.compact_every:
            return True
        return bool(self.compact_resampled_every and
                    self._resampled_since_compact >= self.compact_resampled_every)

    def initialize_fields(self):
        """Initial deposit, moment update and Poisson solve at t
        (boltzmann.py:363-399): ni gains the (zero) ionization source and Te
        and E follow from the deposit; the ions are not advanced."""
        self.deposit()
        if self.cfg.dim == 1:
            fluid = self.fp.fluid_on
            self.fp.fluid_on = 0
            try:
                self.field_step(advance=False)
            finally:
                self.fp.fluid_on = fluid

    def step(self, index: Optional[int] = None):
        """One kinetic step in the reference's order (boltzmann.py:570-678):
        resampling check every `resample_every` steps (then the compaction
        when due, so the push writes the keys of the state it leaves), push +
        collisions + deposit, then the grid step."""
        k = self.cfg.resample_every
        if k and self.cfg.target_ppc > 0 and index is not None and index % k == 0:
            self.resample()
        if self._compact_due():
            self.compact()
        self.push()
        self.field_step(advance=True)

    def run(self, index: int, n_steps: int, graphs: bool = True) -> int:
        """n_steps steps from step `index`: the sequence step(index),
        step(index + 1), ... makes, with the pushes and field steps between
        two resampling checks or compactions replayed from a captured
        hipGraph (one per block length and state buffer), so the host
        enqueues one launch per block instead of two per step.  Returns the
        next step index."""
        k = self.cfg.resample_every if self.cfg.target_ppc > 0 else 0
        # a check every step leaves one push per block: launched directly,
        # which also keeps the keys that push writes for the next sort
        if not graphs or not self.inplace or k == 1:
            if self.overlap_field and self.inplace and self.inplace_resample and k == 1:
                return self._run_overlapped(index, n_steps)
            for _ in range(n_steps):
                self.step(index)
                index += 1
            return index
        end = index + n_steps
        while index < end:
            if k and index % k == 0:
                self.resample()
            if self._compact_due():
                self.compact()
            m = end - index
            if k:
                m = min(m, k - index % k)
            if self.compact_every:
                m = max(1, min(m, self.compact_every - self._since_compact))
            self._replay(m)
            index += m
        return index

    def _run_overlapped(self, index: int, n_steps: int) -> int:
        """step(index), step(index + 1), ... with the grid step of each on a
        side stream: after the push the clock advances on the engine's
        stream (the time and step the field step would advance), the field
        step (fluid ions, Poisson: no time advance of its own) follows on the
        side stream and the next step's resampling check and compaction run
        beside it (they read the keys and the state, not the fields, and
        key their draws on the advanced step); the next push waits for the
        field.  Every kernel sees the inputs of step(); the engine's stream
        waits for the last field step before this returns."""
        if self._field_stream is None:
            self._field_stream = torch.cuda.Stream(self.dev)
        fs = self._field_stream
        a_ptr = C.c_void_p(self.stream.cuda_stream)
        f_ptr = C.c_void_p(fs.cuda_stream)
        pushed = torch.cuda.Event()
        fielded = None
        k = self.cfg.resample_every
        try:
            for _ in range(n_steps):
                if k and self.cfg.target_ppc > 0 and index % k == 0:
                    self.resample()
                if self._compact_due():
                    self.compact()
                if fielded is not None:
                    self.stream.wait_event(fielded)
                self.push()
                _lib.check(self.lib.bz_clock_advance(self.ctx, self.cfg.dt), "bz_clock_advance")
                pushed.record(self.stream)
                fs.wait_event(pushed)
                _lib.check(self.lib.bz_set_stream(self.ctx, f_ptr))
                try:
                    self.field_step(advance=False)
                finally:
                    _lib.check(self.lib.bz_set_stream(self.ctx, a_ptr))
                fielded = torch.cuda.Event()
                fielded.record(fs)
                index += 1
        finally:
            if fielded is not None:
                self.stream.wait_event(fielded)
        return index

    def _replay(self, m: int):
        key = (self.cur, m)
        gid = self._graphs.get(key)
        if gid is None:
            gid = self._capture(m)
            self._graphs[key] = gid
        _lib.check(self.lib.bz_graph_launch(self.ctx, gid), "bz_graph_launch")
        self._since_compact += m

    def _capture(self, m: int) -> int:
        """Capture m pushes + field steps on a side stream (capture needs a
        created stream); the graph is replayed on the engine's stream."""
        if self._cap_stream is None:
            self._cap_stream = torch.cuda.Stream(self.dev)
        since = self._since_compact
        _lib.check(self.lib.bz_set_stream(self.ctx, C.c_void_p(self._cap_stream.cuda_stream)))
        gid = C.c_int32(-1)
        try:
            _lib.check(self.lib.bz_graph_begin(self.ctx), "bz_graph_begin")
            try:
                for _ in range(m):
                    self.push()
                    self.field_step(advance=True)
            finally:
                _lib.check(self.lib.bz_graph_end(self.ctx, C.byref(gid)), "bz_graph_end")
        finally:
            self._since_compact = since
            _lib.check(self.lib.bz_set_stream(self.ctx, C.c_void_p(self.stream.cuda_stream)))
        return gid.value

    def precapture(self) -> int:
        """Capture every block `run` can replay (each block length up to the
        resampling period, or the compaction period without resampling, on
        either state buffer) before a timed run, so that the run itself only
        replays.  Returns the number of graphs held."""
        k = self.cfg.resample_every if self.cfg.target_ppc > 0 else 0
        if not self.inplace or k == 1:
            return len(self._graphs)  # every step is launched directly
        limit = k if k else self.compact_every
        if self.compact_every:
            limit = min(limit, self.compact_every)
        cur = self.cur
        try:
            for c in (0, 1):
                self.cur = c
                for m in range(1, max(limit, 1) + 1):
                    if (c, m) not in self._graphs:
                        self._graphs[(c, m)] = self._capture(m)
        finally:
            self.cur = cur
        return len(self._graphs)

    def set_collision_counts(self, on: bool) -> None:
        """Turn the 1-D push's collision counters on or off (captured step
        blocks are dropped: their pushes carry the setting)."""
        _lib.check(self.lib.bz_set_collision_counts(self.ctx, 1 if on else 0),
                   "bz_set_collision_counts")
        self.drop_graphs()

    def collision_counts(self, reset: bool = True) -> "torch.Tensor":
        """Per node since the last reset: sum w of the elastic collisions, sum
        w of the excitations, number of ionizations ([3, n_nodes], FP64 on
        the device)."""
        nn = self.cfg.n_nodes
        out = torch.empty(3 * nn, dtype=torch.float64, device=self.dev)
        _lib.check(self.lib.bz_collision_counts(self.ctx, _tptr(out), nn, 1 if reset else 0),
                   "bz_collision_counts")
        return out.view(3, nn)

    def drop_graphs(self):
        """Forget the captured blocks (after a change of the push knobs)."""
        for gid in self._graphs.values():
            _lib.check(self.lib.bz_graph_destroy(self.ctx, gid), "bz_graph_destroy")
        self._graphs = {}

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
