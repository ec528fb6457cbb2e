"""BoltzmannDSMC: the reference's driver surface on the MI355X engine.

Constructor and run() keep the reference's arguments (1bin/boltzmann.py:23,
:565) plus the binning strategy (1, 3 or 5 energy bins), which the reference
selects by directory.  Physics settings are the reference's (boltzmann.py:46-78):
0.1 Torr argon, L = 2.54 cm, 100 V at 13.56 MHz, argon fit cross sections,
300 K gas (with the reference's electron-mass variance), the Liu initial
condition (boltzmann.py:242-333) and the ES-WPNC policy of _do_resampling /
_check_for_resampling.  The step order is run()'s: resampling check every
resample_freq steps, push + collisions + deposit + compaction, host moment
update, heavies, Poisson, cycle averages and output every in_spc steps.
Grid quantities stay on the device; host copies are made at output.

restart_cyclenum > 0 starts from the dump files of that cycle
(1bin/boltzmann.py:157-239, unreachable there because the prescribed
initial condition is hard-wired, quirk Q15): the particles with w != 0 of
e_x0/e_v0/e_v1/e_v2/e_wt, the ion density of ni, ns = 0, then the initial
deposit and field solve as for the prescribed start.  The clock resumes at
step restart_cyclenum * in_spc (time = that many steps of dt; the
reference's reader keeps time 0).

fix_quirks=True changes three diagnostics of the reference (left as they
are by default): the collision counters are deposited every push, cycle
averaged and written (Q10: collct is never reset and never written,
1bin/boltzmann.py:310-312,618-630); the cycle averages include the output
step (Q18: spc - 1 steps summed, divided by spc, :681,694-698); and the
metastable average is written from ns (Q19: data_write is handed ni_CA for
ns_CA, :705).
"""
from __future__ import annotations

import math
import sys

import numpy as np
import torch

from . import _lib
from . import physics as ph
from .engine import KineticEngine

P_TORR = 0.1
L_CM = 2.54
V0 = 100.0
FREQ = 13.56e6
KTM = 1.5 * 1.17255 * 10 ** 11  # variance per component at 1 eV (boltzmann.py:263)


def reference_config(Nc: int, gridpts: int, bnf: int, in_dt: float, nb: int = 1,
                     seed: int = 12345) -> ph.RunConfig:
    N = gridpts - 1
    fr = ph.bin_fractions(nb)
    # per-bin targets int(frac * Nc / N) (1bin/boltzmann.py:147, 3bin :147-149, 5bin :147-151)
    tt = [int(f * Nc / N) for f in fr]
    return ph.RunConfig(
        name=f"reference_{nb}bin", dim=1, n_particles=sum(tt) * (N + 1),
        precision=_lib.BZ_F64, xs_mode=_lib.BZ_XS_FIT, nb=nb, n_cells=N, length_cm=L_CM,
        bnf=bnf, dt=(1.0 / FREQ) / in_dt, gas_density_m3=1e6 * 3.22e16 * P_TORR,
        rf_voltage=V0, rf_freq=FREQ, target_ppc=sum(tt), resample_every=1, trigger_at=3.0,
        merge_at=3.0, split_below=0.5, merge_keep=0.6, split_to=2.0, min_wt=1e-1,
        bin_fractions=fr, bin_thresholds=ph.bin_thresholds(nb), gas_temp_k=300.0, seed=seed,
        capacity_factor=4.0)


def read_restart(name: str, cycle: int, rank: int = 0):
    """The particles and ion density a run dumped at cycle `cycle` (the
    files of data_write, 1bin/kernels/data_write_1D.py:9-75, read as
    1bin/boltzmann.py:168-197 reads them): raw float64 e_x0, e_v0, e_v1,
    e_v2, e_wt of the rank, the rows with w != 0 kept; ni of the cycle.
    Returns ((x, vx, vy, vz, w), ni)."""
    cyc = "_cycle" + str(int(cycle)) + ".dat"
    cyc_rk = "_cycle" + str(int(cycle)) + "_mpi" + str(int(rank)) + ".dat"
    ni = np.fromfile(name + "ni" + cyc, dtype=np.float64)
    arrs = [np.fromfile(name + key + cyc_rk, dtype=np.float64)
            for key in ("e_x0", "e_v0", "e_v1", "e_v2", "e_wt")]
    if len({a.size for a in arrs}) != 1:
        raise ValueError(f"restart files of cycle {cycle} differ in length")
    keep = arrs[4] != 0
    return tuple(np.ascontiguousarray(a[keep]) for a in arrs), ni


def liu_initial_state(cfg: ph.RunConfig, seed: int, num_nodes: int = 1):
    """The reference's prescribed initial condition (IC_flag == 1,
    1bin/boltzmann.py:242-333): per node ii, ppc particles at the node
    position (+dx/4 at the first node, -dx/4 at the last), velocities
    N(0, 0.5 * kTm) per component (Te = 0.5 eV, kTm the 1 eV variance,
    :253,:267,:293-295), weight ne(x_ii) / ppc / num_nodes (:297) with
    ne(x) = P (1e7 + 1e9 (1-x/L)^2 (x/L)^2) (:246), ppc the sum of the
    per-bin targets (3bin :282).  Then the rows are shuffled by one
    permutation applied to every array (:315-318, 'shuffling for
    efficiency'); the reference's permutation is CuPy's seeded shuffle, this
    one numpy's (a uniformly random permutation either way).  Returns the
    state and ne at the nodes."""
    N = cfg.n_cells
    rng = np.random.default_rng(seed)
    xs = np.linspace(0, cfg.length_cm, N + 1)
    ne = P_TORR * (1e7 + 1e9 * (1 - xs / cfg.length_cm) ** 2 * (xs / cfg.length_cm) ** 2)
    dx = cfg.dx
    ppc = max(cfg.target_ppc, 1)
    x, v, w = [], [], []
    for ii in range(N + 1):
        pos = xs[ii] + (dx / 4 if ii == 0 else (-dx / 4 if ii == N else 0.0))
        x.append(np.full(ppc, pos))
        v.append(rng.normal(0.0, math.sqrt(0.5 * KTM), size=(3, ppc)))
        w.append(np.full(ppc, ne[ii] / ppc / num_nodes))
    x = np.concatenate(x)
    v = np.concatenate(v, axis=1)
    w = np.concatenate(w)
    perm = rng.permutation(x.size)
    return (x[perm], v[0][perm], v[1][perm], v[2][perm], w[perm]), ne


class BoltzmannDSMC:
    def __init__(self, num_steps, name, restart_cyclenum, Nc, gridpts, bnf, num_nodes, pyk,
                 resample_freq, in_dt, in_spc, nbins: int = 1, ref_rescale: bool = False,
                 seed: int = 12345, device: int = 0, fix_quirks: bool = False, rank: int = 0):
        self.pyk = bool(pyk)
        self.fix_quirks = bool(fix_quirks)
        self.restart_cyclenum = int(restart_cyclenum or 0)
        self.cfg = reference_config(int(Nc), int(gridpts), int(bnf), float(in_dt), int(nbins),
                                    seed)
        self.cfg = self.cfg.with_(resample_every=int(resample_freq))
        self.N = self.cfg.n_cells
        self.dx = self.cfg.dx
        self.L = self.cfg.length_cm
        self.dt_big = self.cfg.dt
        self.steps_per_cycle = int(in_spc)
        self.num_nodes = int(num_nodes)
        self.curr_t = 0.0
        step0 = 0
        if self.restart_cyclenum > 0:
            state, ni0 = read_restart(name, self.restart_cyclenum, rank)
            if ni0.size != self.N + 1:
                raise ValueError(f"ni of cycle {self.restart_cyclenum} has {ni0.size} nodes, "
                                 f"the grid {self.N + 1}")
            step0 = self.restart_cyclenum * self.steps_per_cycle
            self.curr_t = step0 * self.dt_big
        else:
            state, ni0 = liu_initial_state(self.cfg, seed, self.num_nodes)
        self.eng = KineticEngine(self.cfg, device=device)
        tt = np.array([int(f * int(Nc) / self.N) for f in self.cfg.bin_fractions],
                      dtype=np.int32)
        self.eng.targets_host = np.tile(tt, self.N + 1)
        self.eng.targets.copy_(torch.from_numpy(self.eng.targets_host))
        # the reference's own behaviours: the dead ionizer's flag at node 0
        # and the (e*w)/w binning energy
        self.eng.params.ref_flags = _lib.BZ_REF_DEAD_IONIZER | _lib.BZ_REF_ENERGY_ROUNDTRIP
        pol = self.eng.policy
        pol.ref_trigger = 1
        pol.max_ppc = int(10.0 * max(self.eng.targets_host.max(), 1))
        pol.ref_rescale = 1 if ref_rescale else 0
        self.eng.ni.copy_(torch.from_numpy(ni0))
        self.eng.load_state(state, step=step0, t=self.curr_t)
        self.eng.initialize_fields()
        if self.fix_quirks:
            self.eng.set_collision_counts(True)
        # the reference adds the (zero) initial ionization source and sets Te
        self.Nc_list = [self.eng.count()]
        nn = self.N + 1
        f64 = dict(dtype=torch.float64, device=self.eng.dev)
        self._ca = {k: torch.zeros(nn if k not in ("E", "V") else (self.N if k == "E" else
                                                                     self.N - 1), **f64)
                    for k in ("ne", "ni", "ns", "Te", "E", "V")}
        if self.fix_quirks:
            self._ca.update({k: torch.zeros(nn, **f64) for k in ("coll_el", "coll_exc",
                                                                 "coll_ion")})

    # reference-named views (host copies)
    @property
    def ne_ar(self):
        return self.eng.moments[:self.N + 1].cpu().numpy()

    @property
    def ni_ar(self):
        return self.eng.ni.cpu().numpy()

    @property
    def Te_ar(self):
        return self.eng.Te.cpu().numpy()

    @property
    def E_ar(self):
        return self.eng.E.cpu().numpy()

    def run(self, num_steps, name, restart_cyclenum=0, verbose: bool = True):
        e = self.eng
        nn = self.N + 1
        cyclestep = 0
        if verbose:
            print("INIT NC = ", self.Nc_list[0], flush=True)
        for bigstep in range(int(num_steps)):
            if verbose and bigstep % 100 == 0:
                print("Step ", bigstep, ". Nc = ", e.count(), flush=True)
            e.step(bigstep)
            self.curr_t += self.dt_big
            cyclestep += 1
            write = (bigstep + 1) % self.steps_per_cycle == 0
            # the reference sums only the steps before the output step (Q18)
            if not write or self.fix_quirks:
                self._ca["ne"] += e.moments[:nn]
                self._ca["ni"] += e.ni
                self._ca["ns"] += e.ns
                self._ca["Te"] += e.Te
                self._ca["E"] += e.E
                self._ca["V"] += e.V
                if self.fix_quirks:
                    cc = e.collision_counts(reset=True)
                    self._ca["coll_el"] += cc[0]
                    self._ca["coll_exc"] += cc[1]
                    self._ca["coll_ion"] += cc[2]
            if write:
                cyclestep = 0
                for v in self._ca.values():
                    v /= self.steps_per_cycle
                self.data_write(bigstep, name, restart_cyclenum)
                for v in self._ca.values():
                    v.zero_()
                e.g2.zero_()
        self.Nc_list = [e.count()]
        return self

    def data_write(self, bigstep, name, restart_cyclenum=0, rank: int = 0):
        """Raw float64 files in the reference's naming (kernels/data_write_1D.py:3-75)."""
        e = self.eng
        k = str(int((bigstep + 1) / self.steps_per_cycle) + restart_cyclenum)
        cyc = "_cycle" + k + ".dat"
        cyc_rk = "_cycle" + k + "_mpi" + str(rank) + ".dat"
        nn = self.N + 1
        grids = {"ne": e.moments[:nn], "ni": e.ni, "Te": e.Te, "E": e.E, "ionprod_total": e.g2}
        if rank == 0:
            for key, arr in grids.items():
                arr.cpu().numpy().astype(np.float64).tofile(name + key + cyc)
            for key in ("ne", "ni", "Te", "E"):
                self._ca[key].cpu().numpy().tofile(name + key + "_CA" + cyc)
            if self.fix_quirks:
                # the metastable average itself (Q19) and the collision counters (Q10)
                for key in ("ns", "coll_el", "coll_exc", "coll_ion"):
                    self._ca[key].cpu().numpy().tofile(name + key + "_CA" + cyc)
        st = e.fetch_state()
        for key, arr in (("e_x0", st[0]), ("e_v0", st[1]), ("e_v1", st[2]), ("e_v2", st[3]),
                         ("e_wt", st[4])):
            np.asarray(arr, dtype=np.float64).tofile(name + key + cyc_rk)
        print("Writing data for cycle ", (bigstep + 1) / self.steps_per_cycle, flush=True)
        sys.stdout.flush()


__all__ = ["BoltzmannDSMC", "reference_config", "liu_initial_state", "read_restart"]
