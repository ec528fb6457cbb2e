"""Physical constants, run configurations and the host-side set-up of the
electron kinetic step: derived push constants, the tabulated cross sections
with their lookup hints, the Cholesky factor of the Poisson operator and the
initial particle state.

Units follow the reference (1bin/boltzmann.py:46-78): positions in cm,
velocities in m/s, E in V/cm, energies in eV, densities (and particle
weights) in cm^-3, gas density for the collision frequency in m^-3.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace
from typing import Optional

import numpy as np

from . import _lib

# Constants written as the reference writes them (electron_kernel_numba.py:14-17,
# boltzmann.py:56-58), so the derived doubles carry identical bits.
Q_E = 1.602 * 10 ** (-19)
M_E = 9.10938 * 10 ** (-31)
M_N_AR = 6.6 * 10 ** (-26)
J_EV = 6.241509 * (10 ** 18)
EPS_E_VCM = 5.5263e5  # permittivity in e/(V cm)
TD = 1e-21  # 1 Townsend in V m^2

E_ION = 15.76
E_EXC = 11.55


@dataclass
class XSecModel:
    """Synthetic argon-like cross sections tabulated on log-spaced energies."""
    n: int = 1000
    e_min: float = 0.01
    e_max: float = 1000.0
    sigma_el: float = 1e-19
    mass_ratio: float = 1.36e-5
    e_exc: float = E_EXC
    sigma_exc_peak: float = 1e-20
    e_exc_peak: float = 30.0
    e_ion: float = E_ION
    sigma_ion0: float = 3e-20

    def energies(self) -> np.ndarray:
        return np.logspace(math.log10(self.e_min), math.log10(self.e_max), self.n)

    def sigma(self, e: np.ndarray) -> np.ndarray:
        """[3, len(e)] cross sections in m^2: elastic, excitation, ionization."""
        e = np.asarray(e, dtype=np.float64)
        s = np.zeros((3, e.size))
        s[0] = self.sigma_el
        rise = self.sigma_exc_peak * (e - self.e_exc) / (self.e_exc_peak - self.e_exc)
        fall = self.sigma_exc_peak * self.e_exc_peak / np.maximum(e, 1e-300)
        s[1] = np.where(e < self.e_exc, 0.0, np.where(e <= self.e_exc_peak, rise, fall))
        r = e / self.e_ion
        with np.errstate(divide="ignore", invalid="ignore"):
            ion = self.sigma_ion0 * (1.0 - 1.0 / r) * np.log(r) / r
        s[2] = np.where(e <= self.e_ion, 0.0, ion)
        return s


@dataclass
class RunConfig:
    """One workload of the electron kinetic step."""
    name: str
    dim: int                      # 0: 0-D swarm, 1: 1-D bounded discharge
    n_particles: int
    precision: int = _lib.BZ_F32
    xs_mode: int = _lib.BZ_XS_TABLE
    nb: int = 1                   # energy bins per cell: 1, 3 or 5
    n_cells: int = 1              # N (1-D)
    length_cm: float = 2.54
    bnf: int = 4
    dt: float = 1e-12
    steps: int = 1
    gas_density_m3: float = 3.22e21
    e_over_n_td: float = 100.0    # 0-D field
    te_ev: float = 1.0            # initial Maxwellian temperature
    density_cm3: float = 1e9      # 1-D initial uniform density
    rf_voltage: float = 100.0
    rf_freq: float = 13.56e6
    target_ppc: int = 0           # resampling target per cell (0: none)
    resample_every: int = 0       # 0: never
    trigger_at: float = 3.0      # <= 0: every check fires
    merge_at: float = 3.0
    split_below: float = 0.5
    merge_keep: float = 0.6
    split_to: float = 2.0
    min_wt: float = 0.1
    bin_fractions: tuple = (1.0,)
    bin_thresholds: tuple = ()
    gas_temp_k: float = 0.0       # 0: cold gas
    seed: int = 12345
    fluid_on: bool = True
    poisson_on: bool = True
    ion_scheme: int = 0           # 0 reference central fluxes, 1 Scharfetter-Gummel sub-cycled
    push_order: int = _lib.BZ_PUSH_REF  # BZ_PUSH_LEAPFROG: kick then drift
    max_substeps: int = 256
    capacity_factor: float = 2.0
    xsec: XSecModel = field(default_factory=XSecModel)

    def with_(self, **kw) -> "RunConfig":
        return replace(self, **kw)

    @property
    def n_nodes(self) -> int:
        return self.n_cells + 1 if self.dim == 1 else 1

    @property
    def dx(self) -> float:
        return self.length_cm / self.n_cells

    @property
    def np_dtype(self):
        return np.float64 if self.precision == _lib.BZ_F64 else np.float32

    def e_field0_vcm(self) -> float:
        # E = (E/N) * N_gas in V/m, then V/cm
        return self.e_over_n_td * TD * self.gas_density_m3 / 100.0


def bin_thresholds(nb: int) -> tuple:
    """Energy bin edges of the reference's strategies (bin_particles.py)."""
    return {1: (), 3: (10.0, 15.76), 5: (5.0, 10.0, 15.76, 20.0)}[nb]


def bin_fractions(nb: int) -> tuple:
    """Per-bin share of the cell target (3bin/boltzmann.py:147-149, 5bin :147-151)."""
    return {1: (1.0,), 3: (0.7, 0.2, 0.1), 5: (0.4, 0.3, 0.2, 0.07, 0.03)}[nb]


# ---------------------------------------------------------------------------
# Derived constants of the push
# ---------------------------------------------------------------------------
def push_params(cfg: RunConfig, rng_mode: int = _lib.BZ_RNG_PHILOX,
                table: Optional["XSecTable"] = None) -> _lib.Params:
    p = _lib.Params()
    dt = cfg.dt
    p.dt = dt
    p.xfac = 100 * dt
    p.afac = dt * (Q_E / M_E)
    p.kev = 0.5 * M_E * J_EV
    p.jm = J_EV * M_E
    if cfg.xs_mode == _lib.BZ_XS_FIT:
        p.fel = math.sqrt(1 - (2 * M_E / M_N_AR))
    else:
        p.fel = math.sqrt(1 - 2 * cfg.xsec.mass_ratio)
    p.L = cfg.length_cm
    n = cfg.n_cells if cfg.dim == 1 else 1
    p.dx = cfg.length_cm / n
    p.half_dx = 0.5 * p.dx
    p.L_dx = cfg.length_cm + p.dx
    p.dep_M = float(cfg.bnf * (n + 1))
    p.nn = cfg.gas_density_m3
    p.e_ion = cfg.xsec.e_ion if cfg.xs_mode != _lib.BZ_XS_FIT else E_ION
    p.e_exc = cfg.xsec.e_exc
    if table is not None:
        p.nu_max = table.nu_max
        p.p_null = -math.expm1(-table.nu_max * dt)
        p.pc = p.p_null / table.nu_max
    if cfg.gas_temp_k > 0:
        # reference: variance Tg[eV] * 1.5 * 1.17255e11 (boltzmann.py:556-561)
        p.gas_sigma = math.sqrt((cfg.gas_temp_k / 11606.0) * 1.5 * 1.17255e11)
    p.e_field0 = cfg.e_field0_vcm() if cfg.dim == 0 else 0.0
    p.seed = cfg.seed
    p.dim = cfg.dim
    p.xs_mode = cfg.xs_mode
    p.rng_mode = rng_mode
    p.n_cells = n
    p.n_nodes = n + 1 if cfg.dim == 1 else 1
    p.bnf = cfg.bnf
    p.r_stride = 0
    p.push_order = cfg.push_order
    return p


# ---------------------------------------------------------------------------
# Tabulated cross sections (sigma * N_gas) with bucket hints
# ---------------------------------------------------------------------------
@dataclass
class XSecTable:
    energy: np.ndarray   # [n] state precision
    sig: np.ndarray      # [3, n]
    slope: np.ndarray    # [3, n]
    hint: np.ndarray     # int32 [n_hint]
    hint_base: int
    hint_bits: int
    nu_max: float


def _bucket(values: np.ndarray, bits: int) -> np.ndarray:
    v = np.ascontiguousarray(values)
    if v.dtype == np.float64:
        return (v.view(np.uint64) >> np.uint64(52 - bits)).astype(np.int64)
    return (v.view(np.uint32) >> np.uint32(23 - bits)).astype(np.int64)


def build_table(model: XSecModel, gas_density_m3: float, dtype, hint_bits: int = 6) -> XSecTable:
    e64 = model.energies()
    s64 = model.sigma(e64) * gas_density_m3  # sigma*N, 1/m
    e = e64.astype(dtype)
    s = s64.astype(dtype)
    ed = e.astype(np.float64)
    sd = s.astype(np.float64)
    slope = np.zeros_like(sd)
    slope[:, :-1] = (sd[:, 1:] - sd[:, :-1]) / (ed[1:] - ed[:-1])
    slope = slope.astype(dtype)
    # bucket -> lowest candidate segment
    b_lo = int(_bucket(e[:1], hint_bits)[0])
    b_hi = int(_bucket(e[-1:], hint_bits)[0])
    nh = b_hi - b_lo + 1
    buckets = np.arange(b_lo, b_hi + 1, dtype=np.int64)
    if dtype == np.float64:
        lo_vals = (buckets.astype(np.uint64) << np.uint64(52 - hint_bits)).view(np.float64)
    else:
        lo_vals = (buckets.astype(np.uint32) << np.uint32(23 - hint_bits)).view(np.float32)
    lo_vals = np.maximum(lo_vals, e[0])
    hint = np.searchsorted(e, lo_vals, side="right") - 1
    hint = np.clip(hint, 0, model.n - 2).astype(np.int32)
    # each bucket must span at most 8 segments (the device scan is bounded)
    nxt = np.empty(nh, dtype=np.int64)
    nxt[:-1] = hint[1:]
    nxt[-1] = model.n - 2
    if np.max(nxt - hint) > 7:
        raise ValueError("hint buckets too coarse for the table spacing")
    # null-collision bound: on segment j, sum_c max(s_c[j], s_c[j+1]) * v(E[j+1])
    v = np.sqrt(2.0 * ed / (J_EV * M_E))
    smax = np.maximum(sd[:, :-1], sd[:, 1:]).sum(axis=0)
    nu_max = float(np.max(smax * v[1:])) * (1.0 + 1e-6)
    return XSecTable(energy=np.ascontiguousarray(e), sig=np.ascontiguousarray(s),
                     slope=np.ascontiguousarray(slope), hint=np.ascontiguousarray(hint),
                     hint_base=b_lo, hint_bits=hint_bits, nu_max=nu_max)


# ---------------------------------------------------------------------------
# Poisson operator: banded Cholesky of (1/dx^2) tridiag(-1, 2, -1), N-1 rows
# ---------------------------------------------------------------------------
def poisson_cholesky(n_cells: int, dx: float) -> np.ndarray:
    """[2, N-1] (diag, upper band) of U with A = U^T U, in the order of the
    unblocked LAPACK band factorisation: u_jj = sqrt(a_jj - u_{j-1,j}^2),
    u_{j,j+1} = a_{j,j+1} * (1/u_jj) (boltzmann.py:744-755 uses scipy's
    cholesky_banded of the same matrix)."""
    m = n_cells - 1
    a_d = 2.0 / (dx ** 2)
    a_o = -1.0 / (dx ** 2)
    diag = np.zeros(m)
    low = np.zeros(m)
    prev = 0.0
    for j in range(m):
        ajj = a_d - prev * prev
        if ajj <= 0.0:
            raise ValueError("Poisson operator not positive definite")
        ujj = math.sqrt(ajj)
        diag[j] = ujj
        if j + 1 < m:
            prev = a_o * (1.0 / ujj)
            low[j + 1] = prev
    return np.ascontiguousarray(np.stack([diag, low]))


def field_params(cfg: RunConfig) -> _lib.FieldParams:
    fp = _lib.FieldParams()
    fp.dx = cfg.dx
    fp.dt = cfg.dt
    fp.eps = EPS_E_VCM
    fp.v0 = cfg.rf_voltage
    fp.freq = cfg.rf_freq
    nn_cm3 = cfg.gas_density_m3 * 1e-6
    # boltzmann.py:49-55: coefficients scale as 1/N_gas
    fp.D_i = 2.07e18 / nn_cm3
    fp.D_s = 2.42e18 / nn_cm3
    fp.mu_i = 4.65e19 / nn_cm3
    fp.n = cfg.n_cells
    fp.fluid_on = 1 if cfg.fluid_on else 0
    fp.poisson_on = 1 if cfg.poisson_on else 0
    fp.ion_scheme = cfg.ion_scheme
    fp.max_substeps = cfg.max_substeps
    return fp


# ---------------------------------------------------------------------------
# Initial state (set-up only, not timed)
# ---------------------------------------------------------------------------
def initial_state(cfg: RunConfig, n: Optional[int] = None, seed: Optional[int] = None):
    """SoA arrays (x, vx, vy, vz, w) of an isotropic Maxwellian at te_ev.

    0-D: all particles at x = L/2 with weight 1.  1-D: uniform density
    cfg.density_cm3, particles placed uniformly in (0, L) and sorted by x,
    weight = density * L / (n * dx) * dx ... i.e. n_per_node = n/(N+1) so
    that the node deposit equals the density (reference convention
    w = n(x)/ppc, boltzmann.py:297)."""
    n = cfg.n_particles if n is None else n
    rng = np.random.default_rng(cfg.seed if seed is None else seed)
    sigma_v = math.sqrt(cfg.te_ev * Q_E / M_E)
    v = rng.normal(0.0, sigma_v, size=(3, n))
    if cfg.dim == 0:
        x = np.full(n, 0.5 * cfg.length_cm)
        w = np.ones(n)
    else:
        x = np.sort(rng.uniform(0.0, cfg.length_cm, size=n))
        # keep particles strictly inside the domain
        x = np.clip(x, 1e-6 * cfg.dx, cfg.length_cm * (1 - 1e-9))
        ppn = n / cfg.n_nodes
        w = np.full(n, cfg.density_cm3 / ppn)
    dt = cfg.np_dtype
    return tuple(np.ascontiguousarray(a, dtype=dt) for a in (x, v[0], v[1], v[2], w))
