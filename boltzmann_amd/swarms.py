"""Batched 0-D electron swarms on one GPU.

Hundreds of independent swarms, each at its own reduced field E/N and with
its own particle count, advance together: one fused launch steps every
instance in place (bz_swarm_step; a 0-D particle has no position, so x holds
its instance number and secondaries appended at the end inherit it), and a
captured hipGraph replays many steps with no host involvement.  Population
control (bz_swarm_control) sorts the slots by instance and keeps each swarm
near its initial size with the ES-WPNC merge / split, which conserves
weight, momentum and energy; afterwards instance k owns the slot range
[off[k], off[k+1]).  Per instance the run yields the mean energy,
the mobility, the rate coefficients of every collision channel and the
EEDF, the inputs of the learned closure.

The reference has no batched mode: its 0-D test problems are single runs
made by editing the driver (README.md:7-8).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import torch

from . import _lib
from . import physics as ph


@dataclass
class SwarmConfig:
    name: str = "c2_batched_256_swarms"
    n_inst: int = 256
    en_td: Optional[tuple] = None        # explicit E/N per instance (Td)
    en_td_lo: float = 1.0
    en_td_hi: float = 1000.0
    counts: Optional[tuple] = None       # explicit particle counts
    count_lo: int = 10_000
    count_hi: int = 500_000
    gas_density_m3: float = 3.2e22
    te_ev: float = 1.0
    dt: float = 1e-12
    precision: int = _lib.BZ_F32
    control_every: int = 50
    seed: int = 12345
    capacity_factor: float = 2.5
    xsec: ph.XSecModel = field(default_factory=ph.XSecModel)

    def fields_td(self) -> np.ndarray:
        if self.en_td is not None:
            return np.asarray(self.en_td, dtype=np.float64)
        return np.logspace(math.log10(self.en_td_lo), math.log10(self.en_td_hi), self.n_inst)

    def particle_counts(self) -> np.ndarray:
        if self.counts is not None:
            return np.asarray(self.counts, dtype=np.int64)
        rng = np.random.default_rng(self.seed)
        u = rng.uniform(math.log(self.count_lo), math.log(self.count_hi), self.n_inst)
        return np.exp(u).astype(np.int64)

    def run_config(self) -> ph.RunConfig:
        return ph.RunConfig(name=self.name, dim=0, n_particles=int(self.particle_counts().sum()),
                            precision=self.precision, xs_mode=_lib.BZ_XS_TABLE, dt=self.dt,
                            gas_density_m3=self.gas_density_m3, te_ev=self.te_ev,
                            seed=self.seed, xsec=self.xsec)


def c2() -> SwarmConfig:
    """256 swarms, E/N log-spaced over 1..1000 Td, counts log-uniform in
    [1e4, 5e5] from the fixed seed (about 32M electrons in total)."""
    return SwarmConfig()


def _tptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


class SwarmBatch:
    def __init__(self, cfg: SwarmConfig, device: int = 0, capacity: Optional[int] = None):
        if not torch.cuda.is_available():
            raise _lib.BzError("no GPU visible: the swarms run only in the HIP kernels")
        self.lib = _lib.load()
        self.cfg = cfg
        self.dev = torch.device("cuda", device)
        torch.cuda.set_device(self.dev)
        self.n_inst = len(cfg.fields_td())
        self.counts0 = cfg.particle_counts()
        if self.counts0.size != self.n_inst:
            raise ValueError("one particle count per instance is needed")
        self.total0 = int(self.counts0.sum())
        self.capacity = int(capacity or max(1024, self.total0 * cfg.capacity_factor))
        rc = cfg.run_config()
        self.tdtype = torch.float64 if cfg.precision == _lib.BZ_F64 else torch.float32
        self.ctx = C.c_void_p()
        _lib.check(self.lib.bz_ctx_create(C.byref(self.ctx), device, cfg.precision, self.capacity,
                                          2, self.n_inst), "bz_ctx_create")
        # a created stream (graph capture cannot use the null stream)
        self.stream = torch.cuda.Stream(device=self.dev)
        _lib.check(self.lib.bz_set_stream(self.ctx, C.c_void_p(self.stream.cuda_stream)))
        z = dict(dtype=self.tdtype, device=self.dev)
        self.bufs = [[torch.zeros(self.capacity, **z) for _ in range(5)] for _ in range(2)]
        self.structs = [_lib.State(*[a.data_ptr() for a in b]) for b in self.bufs]
        self.cur = 0
        self.table = ph.build_table(cfg.xsec, cfg.gas_density_m3, rc.np_dtype)
        t = self.table
        self._tab = [torch.from_numpy(a.ravel()).to(self.dev) for a in (t.energy, t.sig, t.slope)]
        self._hint = torch.from_numpy(t.hint).to(self.dev)
        self.xs = _lib.XSec(self._tab[0].data_ptr(), self._tab[1].data_ptr(),
                            self._tab[2].data_ptr(), self._hint.data_ptr(), t.energy.size,
                            t.hint.size, t.hint_base, t.hint_bits)
        self.params = ph.push_params(rc, _lib.BZ_RNG_PHILOX, t)
        self.en_td = cfg.fields_td()
        self.e_vcm = self.en_td * ph.TD * cfg.gas_density_m3 / 100.0
        f64 = dict(dtype=torch.float64, device=self.dev)
        off = np.concatenate([[0], np.cumsum(self.counts0)]).astype(np.int64)
        self.inst_off = torch.from_numpy(off).to(self.dev)
        self.inst_E = torch.from_numpy(self.e_vcm).to(self.dev).to(self.tdtype)
        self.moments = torch.zeros(self.n_inst * 8, **f64)
        self.acc = torch.zeros(self.n_inst * 8, **f64)
        self.targets = torch.from_numpy(self.counts0.astype(np.int32)).to(self.dev)
        self.counts_ws = torch.zeros(self.n_inst + 1, dtype=torch.int32, device=self.dev)
        self.offsets_ws = torch.zeros(self.n_inst + 2, dtype=torch.int32, device=self.dev)
        self.policy = _lib.ResamplePolicy(0.5, 1.5, 1.0, 1.5, 0.0, 1.0, 0, 0, 0, 0)
        # regrouping only: no instance is split or merged
        self.group_policy = _lib.ResamplePolicy(0.0, 1e300, 1.0, 1e300, 0.0, 1.0, 0, 0, 0, 0)
        self.acc_steps = 0
        self.graphs = {}
        self.steps_done = 0
        torch.cuda.synchronize(self.dev)  # set-up copies above ran on the current stream

    # ------------------------------------------------------------------
    def load_initial(self):
        """Isotropic Maxwellian at te_ev, weight 1, in every instance."""
        rng = np.random.default_rng(self.cfg.seed + 1)
        n = self.total0
        sv = math.sqrt(self.cfg.te_ev * ph.Q_E / ph.M_E)
        with torch.cuda.stream(self.stream):
            for k in (1, 2, 3):
                v = rng.normal(0.0, sv, n)
                self.bufs[self.cur][k][:n].copy_(torch.from_numpy(v).to(self.tdtype))
            self.bufs[self.cur][4][:n].fill_(1.0)
            inst = np.repeat(np.arange(self.n_inst), self.counts0)
            self.bufs[self.cur][0][:n].copy_(torch.from_numpy(inst).to(self.tdtype))
        self.stream.synchronize()
        _lib.check(self.lib.bz_clock_set(self.ctx, n, 0, 0.0))

    def _step_call(self, accumulate: bool):
        # in place (the other buffer stages the secondaries); the step also
        # advances the device clock by dt
        src, dst = self.structs[self.cur], self.structs[1 - self.cur]
        _lib.check(self.lib.bz_swarm_step(self.ctx, C.byref(src), C.byref(dst),
                                          C.byref(self.params), C.byref(self.xs), self.n_inst,
                                          None, _tptr(self.inst_E),
                                          _tptr(self.moments), _tptr(self.acc),
                                          1 if accumulate else 0), "bz_swarm_step")

    def _control_call(self, policy=None):
        src, dst = self.structs[self.cur], self.structs[1 - self.cur]
        pol = self.policy if policy is None else policy
        _lib.check(self.lib.bz_swarm_control(self.ctx, C.byref(src), C.byref(dst),
                                             C.byref(self.params), self.n_inst,
                                             _tptr(self.inst_off), _tptr(self.targets),
                                             _tptr(self.counts_ws), _tptr(self.offsets_ws),
                                             C.byref(pol)), "bz_swarm_control")
        self.cur = 1 - self.cur

    def step(self, accumulate: bool = False):
        self._step_call(accumulate)
        if accumulate:
            self.acc_steps += 1
        self.steps_done += 1

    def control(self):
        self._control_call()

    def group(self):
        """Sort the slots by instance without splitting or merging anything,
        so that offsets() describes the state (what eedf() reads)."""
        self._control_call(self.group_policy)

    def capture(self, n_steps: int, accumulate: bool, with_control: bool = True,
                control_every: int = 0) -> int:
        """Capture n_steps steps into a hipGraph.

        A step works in place; every control swaps the ping-pong buffers, so
        the captured sequence must hold an even number of controls to leave
        the state where the next replay expects it.  control_every > 0 puts
        a population control after every control_every steps; otherwise
        with_control appends two controls (the second one finds every
        instance inside its band and only copies)."""
        if control_every:
            if n_steps % control_every:
                raise ValueError("n_steps must be a multiple of control_every")
            plan = ([0] * control_every + [1]) * (n_steps // control_every)
        else:
            plan = [0] * n_steps + ([1, 1] if with_control else [])
        if sum(plan) % 2:
            raise ValueError("the captured sequence must hold an even number of controls")
        cur0 = self.cur
        _lib.check(self.lib.bz_graph_begin(self.ctx))
        try:
            for op in plan:
                if op:
                    self._control_call()
                else:
                    self._step_call(accumulate)
        finally:
            gid = C.c_int32(-1)
            _lib.check(self.lib.bz_graph_end(self.ctx, C.byref(gid)))
        self.cur = cur0
        self.graphs[gid.value] = (n_steps, accumulate, cur0)
        return gid.value

    def replay(self, gid: int, times: int = 1):
        n_steps, accumulate, cur0 = self.graphs[gid]
        if self.cur != cur0:
            # the state sits in the other buffer (an odd number of controls
            # since the capture): one regroup moves it back
            self.group()
        for _ in range(times):
            _lib.check(self.lib.bz_graph_launch(self.ctx, gid))
            self.steps_done += n_steps
            if accumulate:
                self.acc_steps += n_steps

    def run(self, steps: int, accumulate: bool = False):
        """Eager stepping with population control every control_every steps."""
        for s in range(steps):
            self.step(accumulate)
            if self.cfg.control_every and (self.steps_done % self.cfg.control_every == 0):
                self.control()

    def reset_accumulators(self):
        with torch.cuda.stream(self.stream):
            self.acc.zero_()
        self.acc_steps = 0

    def synchronize(self):
        _lib.check(self.lib.bz_synchronize(self.ctx))

    # ------------------------------------------------------------------
    def count(self) -> int:
        n = C.c_int64()
        _lib.check(self.lib.bz_clock_get(self.ctx, C.byref(n), None, None, None))
        return n.value

    def counters(self) -> dict:
        cnt = (C.c_uint64 * _lib.BZ_NCOUNTERS)()
        _lib.check(self.lib.bz_clock_get(self.ctx, None, None, None, cnt))
        return {k: int(cnt[i]) for i, k in enumerate(_lib.COUNTER_NAMES)}

    def offsets(self) -> np.ndarray:
        """Slot range of every instance as the last control() / group() left
        it (steps since then append secondaries after the last range)."""
        self.synchronize()
        return self.inst_off.cpu().numpy()

    def instance_counts(self) -> np.ndarray:
        """Live particles per instance, from the instance numbers in x."""
        x, _, _, _, w = self.state()
        return np.bincount(x[w > 0].astype(np.int64), minlength=self.n_inst)

    def state(self):
        self.synchronize()
        n = self.count()
        return tuple(a[:n].cpu().numpy() for a in self.bufs[self.cur])

    def results(self) -> dict:
        """Per-instance transport data from the accumulated moments:
        mean energy (eV), mobility (m^2/(V s)), rate coefficients (m^3/s).

        A channel with frequency nu is selected in a step with probability
        P_null nu / nu_max = nu * pc, so the rate coefficient is
        (event weight) / (N_gas * pc * weight), which is sigma<v> exactly in
        expectation (pc -> dt as nu_max dt -> 0)."""
        self.synchronize()
        a = self.acc.cpu().numpy().reshape(self.n_inst, 8)
        W = np.maximum(a[:, 0], 1e-300)
        steps = max(self.acc_steps, 1)
        N = self.cfg.gas_density_m3
        dt = self.params.pc
        e_vm = self.e_vcm * 100.0
        with np.errstate(divide="ignore", invalid="ignore"):
            mob = np.where(e_vm > 0, np.abs(a[:, 2] / W) / e_vm, np.nan)
        return {
            "en_td": self.en_td,
            "mean_energy": a[:, 1] / W,
            "mobility": mob,
            "k_elastic": a[:, 3] / (N * dt * W),
            "k_excitation": a[:, 4] / (N * dt * W),
            "k_ionization": a[:, 5] / (N * dt * W),
            "births": a[:, 6],
            "steps": steps,
        }

    def eedf(self, n_bins: int = 200, e_lo: float = 0.01, e_hi: float = 1000.0):
        """EEDF per instance on log-spaced bins, normalised to unit integral (1/eV).
        The state is regrouped by instance first (group(): a stable sort, no
        particle changes)."""
        self.group()
        with torch.cuda.stream(self.stream):
            out = torch.zeros(self.n_inst * n_bins, dtype=torch.float64, device=self.dev)
        s = self.structs[self.cur]
        _lib.check(self.lib.bz_swarm_eedf(self.ctx, C.byref(s), self.n_inst, _tptr(self.inst_off),
                                          n_bins, e_lo, e_hi, self.params.kev, _tptr(out)),
                   "bz_swarm_eedf")
        self.synchronize()
        h = out.cpu().numpy().reshape(self.n_inst, n_bins)
        edges = np.logspace(math.log10(e_lo), math.log10(e_hi), n_bins + 1)
        width = np.diff(edges)
        tot = np.maximum(h.sum(axis=1, keepdims=True), 1e-300)
        return edges, h / tot / width

    def close(self):
        if self.ctx:
            self.lib.bz_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass
