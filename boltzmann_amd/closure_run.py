"""Training data for the learned closure, and the closure-accelerated 1-D run.

Training data.  Batched 0-D swarms over the C2 reduced-field range (256
instances, E/N log-spaced 1..1000 Td) start from several temperatures, so
the samples cover mean energies away from the equilibrium one at each E/N.
Every window of steps yields one sample per instance: the rate
coefficients k_j = <sigma_j v> integrated over the instance's EEDF (every
electron contributes, so the estimate is far less noisy than counting
events), the mean energy of the same EEDF, the window-mean reduced mobility
mu N = |<v_x>| / (E/N), and a log-uniform electron density (the swarms have
no electron-electron collisions: the rates do not depend on it, and the
network has to learn that).

Closure-accelerated run.  The kinetic update runs once per n_sub fluid /
Poisson sub-steps and advances the particles through the whole interval,
in n_kin pushes of h = n_sub * dt / n_kin with free-flight collisions
(exact at any push length) and a Poisson solve on the new electron density
between them; n_kin keeps omega_pe h below 2, the leapfrog's stability
bound (a single push of n_sub * dt at the discharge's density is not
stable).  Between kinetic updates the learned rate supplies the ionization source of
the sub-steps from the local E/N of each sub-step and the last kinetic
mean energy and density.  At each kinetic update the kinetic ionization
estimate for the interval replaces the closure's (the kinetic step stays
the reference), and the difference is the closure's rate error,
accumulated in FP64 on the device.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from typing import Optional

import numpy as np
import torch

from . import _lib
from . import physics as ph
from .closure import ClosureNet, features, targets
from .engine import KineticEngine


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def swarm_training_set(n_samples: int = 1_000_000, per_instance: int = 4000,
                       window: int = 20, temps=(0.3, 3.0, 15.0), n_bins: int = 1000,
                       control_every: int = 5, seed: int = 12345, log=None):
    """(x, t) training arrays from batched swarms; see the module docstring."""
    from .swarms import SwarmBatch, SwarmConfig
    xs = ph.XSecModel()
    edges = np.logspace(-2, 3, n_bins + 1)
    centers = np.sqrt(edges[:-1] * edges[1:])
    sig = xs.sigma(centers)                                  # [3, nb] m^2
    speed = np.sqrt(2.0 * centers * ph.Q_E / ph.M_E)          # m/s
    sv = (sig * speed).T                                      # [nb, 3]
    rng = np.random.default_rng(seed)
    per_temp = int(math.ceil(n_samples / (256 * len(temps))))
    X, T = [], []
    for ti, te in enumerate(temps):
        cfg = SwarmConfig(name=f"closure_train_{te}", counts=(per_instance,) * 256, te_ev=te,
                          seed=seed + 101 * ti, capacity_factor=4.0)
        b = SwarmBatch(cfg)
        b.load_initial()
        gid = b.capture(window, accumulate=True, with_control=False)
        N = cfg.gas_density_m3
        for w in range(per_temp):
            b.reset_accumulators()
            b.replay(gid)
            r = b.results()
            _, f = b.eedf(n_bins, 0.01, 1000.0)
            frac = f * np.diff(edges)                          # weight fraction per bin
            k = frac @ sv                                      # [256, 3] m^3/s
            me = frac @ centers
            ne = 10.0 ** rng.uniform(14.0, 19.0, b.n_inst)
            X.append(features(b.en_td, me, ne))
            T.append(targets(k[:, 2], k[:, 1], r["mobility"] * N))
            if control_every and (w + 1) % control_every == 0:
                b.control()
            if log and (w + 1) % 200 == 0:
                log(f"{te} eV: window {w + 1} of {per_temp}")
                b.control()
        if log:
            log(f"swarm samples at {te} eV: {per_temp * b.n_inst}")
        b.close()
    x = np.concatenate(X)[:n_samples].astype(np.float32)
    t = np.concatenate(T)[:n_samples].astype(np.float32)
    return x, t


def train_closure(x, t, epochs: int = 30, batch: int = 8192, lr: float = 3e-3, seed: int = 0,
                  holdout: float = 0.05, log=None):
    """Train a ClosureNet; returns (net, report) with held-out errors of the
    trained and of the untrained network (the ablation)."""
    n = x.shape[0]
    rng = np.random.default_rng(seed)
    idx = rng.permutation(n)
    nh = max(1, int(holdout * n))
    ho, tr = idx[:nh], idx[nh:]
    net = ClosureNet(seed=seed)
    net.fit_normalisation(x[tr], t[tr])
    before = _errors(net, x[ho], t[ho])
    t0 = time.perf_counter()
    hist = net.train(x[tr], t[tr], epochs=epochs, batch=batch, lr=lr, seed=seed,
                     log_every=(epochs // 5 or 1) if log else 0)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    after = _errors(net, x[ho], t[ho])
    report = {"samples": int(n), "train": int(tr.size), "holdout": int(nh), "epochs": epochs,
              "train_seconds": secs, "loss_first_epoch": hist[0], "loss_last_epoch": hist[-1],
              "holdout_untrained": before, "holdout_trained": after}
    return net, report


def _errors(net: ClosureNet, x, t):
    """Held-out errors: loss, and the median / 95th-percentile relative
    error of k_ion where k_ion > 1e-18 m^3/s (ionization that matters)."""
    y = net.predict(x)
    loss = net.loss(x, t)
    m = t[:, 0] > -18.0
    rel = np.abs(10.0 ** (y[m, 0].astype(np.float64) - t[m, 0]) - 1.0) if m.any() else np.zeros(1)
    rel_exc = np.abs(10.0 ** (y[m, 1].astype(np.float64) - t[m, 1]) - 1.0) if m.any() \
        else np.zeros(1)
    return {"loss": float(loss), "k_ion_rel_err_median": float(np.median(rel)),
            "k_ion_rel_err_p95": float(np.percentile(rel, 95)),
            "k_exc_rel_err_median": float(np.median(rel_exc))}


class ClosureDischarge:
    """A 1-D discharge advanced with the closure between kinetic updates."""

    def __init__(self, cfg: ph.RunConfig, net_params: dict, n_sub: int = 8,
                 resample_every: Optional[int] = None, n_kin: Optional[int] = None):
        if cfg.dim != 1 or cfg.xs_mode != _lib.BZ_XS_FLIGHT:
            raise ValueError("the closure run needs a 1-D discharge with free-flight collisions")
        self.cfg = cfg
        self.n_sub = n_sub
        if n_kin is None:
            # pushes per kinetic update: omega_pe h <= 1.8 at 1.5 times the
            # initial density (the bulk density of these discharges grows by
            # some 30 %)
            wpe = 56.4 * math.sqrt(1.5 * cfg.density_cm3 * 1e6)  # rad/s, n in m^-3
            n_kin = max(1, math.ceil(wpe * n_sub * cfg.dt / 1.8))
            while n_sub % n_kin:
                n_kin += 1
        self.n_kin = int(n_kin)
        self.eng = KineticEngine(cfg)
        # kinetic push length h (free flights: exact at any length)
        self.kparams = ph.push_params(cfg.with_(dt=n_sub * cfg.dt / self.n_kin),
                                      _lib.BZ_RNG_PHILOX, self.eng.table)
        # Poisson only between the pushes of one update (no time advance)
        self.fp_poisson = _lib.FieldParams.from_buffer_copy(self.eng.fp)
        self.fp_poisson.fluid_on = 0
        self.fp_poisson.skip_moments = 1
        self.net = ClosureNet(ctx=self.eng.ctx)
        self.net.norm = {k: np.asarray(v, np.float64) for k, v in net_params["norm"].items()}
        self.net.set_parameters(net_params["master"])
        nn = cfg.n_nodes
        f64 = dict(dtype=torch.float64, device=self.eng.dev)
        self.ksrc = torch.zeros(nn, **f64)
        self.src = torch.zeros(nn, **f64)
        self.acc = torch.zeros(nn, **f64)
        self.tally = torch.zeros(8, **f64)
        self.resample_every = cfg.resample_every if resample_every is None else resample_every
        self.macro = 0

    def load(self, state):
        self.eng.load_state(state)
        self.eng.initialize_fields()

    def kinetic_update(self):
        """The kinetic step of one macro step: the resampling check, then
        n_kin pushes over n_sub * dt (in place, births appended) with the
        field re-solved on the kinetic density between them; moments row 1
        holds the ionization of the whole interval."""
        e = self.eng
        k = self.resample_every
        if k and self.cfg.target_ppc > 0 and self.macro % k == 0:
            e.resample()
        nn = self.cfg.n_nodes
        for kk in range(self.n_kin):
            if kk:
                _lib.check(e.lib.bz_field_step(
                    e.ctx, C.byref(self.fp_poisson), _p(e.moments), _p(e.ni), _p(e.ns),
                    _p(e.Te), _p(e.g2), _p(e.V), _p(e.E), _p(e.E_t), _p(e.chol), 0),
                    "bz_field_step")
            src_, dst = e.bufs[e.cur], e.bufs[1 - e.cur]
            _lib.check(e.lib.bz_push_mcc(e.ctx, C.byref(src_.struct), C.byref(dst.struct),
                                         _p(e.E_t), C.byref(self.kparams), C.byref(e.xs), None,
                                         None, _p(e.moments)), "bz_push_mcc")
            if self.n_kin > 1:
                if kk == 0:
                    self.ksrc.copy_(e.moments[nn:2 * nn])
                else:
                    self.ksrc += e.moments[nn:2 * nn]
            e._since_compact += 1
            if e.compact_every and e._since_compact >= e.compact_every:
                e.compact()
        if self.n_kin > 1:
            # the update's ionization source: every push of it
            e.moments[nn:2 * nn].copy_(self.ksrc)

    def sub_step(self):
        """One fluid / Poisson sub-step of dt driven by the learned
        ionization source (self.src; self.acc accumulates it)."""
        e = self.eng
        net = self.net.struct()
        _lib.check(e.lib.bz_closure_source(e.ctx, C.byref(net), self.cfg.n_cells, _p(e.E),
                                           _p(e.moments), self.cfg.gas_density_m3, self.cfg.dt,
                                           _p(self.src), _p(self.acc)), "bz_closure_source")
        _lib.check(e.lib.bz_field_substep(e.ctx, C.byref(e.fp), _p(e.moments), _p(self.src),
                                          _p(e.ni), _p(e.ns), _p(e.Te), _p(e.g2), _p(e.V),
                                          _p(e.E), _p(e.E_t), _p(e.chol)), "bz_field_substep")

    def commit(self):
        """The kinetic ionization of the interval replaces the learned one;
        the difference is tallied (the closure's rate error)."""
        e = self.eng
        _lib.check(e.lib.bz_closure_commit(e.ctx, self.cfg.n_nodes, _p(e.moments), _p(self.acc),
                                           _p(e.ni), _p(e.g2), _p(e.Te), _p(self.tally)),
                   "bz_closure_commit")
        self.macro += 1

    def macro_step(self):
        self.kinetic_update()
        for _ in range(self.n_sub):
            self.sub_step()
        self.commit()

    def report(self) -> dict:
        j = self.tally.cpu().numpy()
        kin, pred = j[0], j[1]
        return {"kinetic_updates": int(j[3]),
                "ionization_kinetic": float(kin), "ionization_closure": float(pred),
                "ionization_rate_error": float(abs(pred - kin) / kin) if kin > 0 else None,
                "ionization_rate_error_per_update": float(j[2] / kin) if kin > 0 else None,
                "mean_ne_sum": float(j[4] / max(j[3], 1)),
                "mean_energy_ev": float(j[5] / j[4]) if j[4] > 0 else None}

    def close(self):
        self.eng.close()
