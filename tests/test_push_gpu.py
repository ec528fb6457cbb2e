"""Parity of the fused HIP electron step with the CPU oracle, through the C ABI.

Decisions (channel, survival, birth, canonical order) must match exactly;
per-particle floating point within 1e-12 (FP64) / 1e-6 (FP32) relative to the
velocity scale; node moments within 1e-12 / 1e-6 relative.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

KEV = 0.5 * O.M_E * O.J_EV


def _cfg(ph, dim, prec, xs, n, n_cells=48, gas=3.2e22, dt=2e-11, **kw):
    return ph.RunConfig(name="parity", dim=dim, n_particles=n, precision=prec, xs_mode=xs,
                        n_cells=n_cells if dim == 1 else 1, length_cm=2.5, dt=dt,
                        gas_density_m3=gas, **kw)


def _state(n, dtype, dim, L=2.5, seed=0, dead_frac=0.02):
    rng = np.random.default_rng(seed)
    e = np.exp(rng.uniform(np.log(0.05), np.log(80.0), n))
    sp = np.sqrt(e / KEV)
    d = rng.normal(size=(3, n))
    d /= np.linalg.norm(d, axis=0)
    v = sp * d
    x = rng.uniform(1e-4 * L, L * (1 - 1e-4), n) if dim == 1 else np.full(n, 0.5 * L)
    w = rng.uniform(0.5, 2.0, n)
    w[rng.uniform(size=n) < dead_frac] = 0.0
    return tuple(np.ascontiguousarray(a, dtype=dtype) for a in (x, v[0], v[1], v[2], w))


def _setup(cfg, table):
    return O.Setup(dim=cfg.dim, dt=cfg.dt, xs_mode=cfg.xs_mode, n_cells=cfg.n_cells,
                   length_cm=cfg.length_cm, bnf=cfg.bnf, gas_density_m3=cfg.gas_density_m3,
                   nu_max=table.nu_max if table is not None else 0.0,
                   e_field0=cfg.e_field0_vcm() if cfg.dim == 0 else 0.0,
                   push_order=cfg.push_order)


def _tab(table):
    return None if table is None else (table.energy, table.sig, table.slope)


def _compare_state(got, want, dtype, dim=1, L=2.5):
    """x is compared in 1-D only: 0-D runs carry no position.  Velocities
    are compared relative to the speed; positions relative to |x| plus a
    thousandth of the gap, since x0 + dt v can cancel next to a wall and an
    ulp of the scattered v then is a large fraction of the tiny x."""
    assert got[0].size == want[0].size
    rt = 1e-12 if dtype == np.float64 else 1e-6
    vs = np.sqrt(sum(np.asarray(a, dtype=np.float64) ** 2 for a in want[1:4])) + 1.0
    for k in range(0 if dim == 1 else 1, 5):
        g = np.asarray(got[k], dtype=np.float64)
        w_ = np.asarray(want[k], dtype=np.float64)
        if k in (1, 2, 3):
            scale = vs
        elif k == 0:
            scale = np.abs(w_) + 1e-3 * L
        else:
            scale = np.abs(w_) + 1e-30
        bad = np.abs(g - w_) > rt * scale
        assert not bad.any(), (k, np.flatnonzero(bad)[:5], g[bad][:5], w_[bad][:5])


CASES = [
    ("table-f32-1d", 1, 4, 1),
    ("table-f64-1d", 1, 8, 1),
    ("table-f32-0d", 0, 4, 1),
    ("table-f64-0d", 0, 8, 1),
    ("fit-f64-1d", 1, 8, 0),
    ("fit-f32-1d", 1, 4, 0),
    ("fit-f64-0d", 0, 8, 0),
    ("direct-f32-1d", 1, 4, 3),
    ("direct-f64-1d", 1, 8, 3),
    ("direct-f64-0d", 0, 8, 3),
]


# the update order of the push only matters where positions move (1-D)
PUSH_CASES = [(*c, rng, 0) for c in CASES for rng in ("stream", "philox")] + \
    [(*c, "philox", 1) for c in CASES if c[1] == 1]


@pytest.mark.parametrize("name,dim,prec,xs,rng_mode,order", PUSH_CASES,
                         ids=[f"{c[0]}-{c[4]}-{'leapfrog' if c[5] else 'ref-order'}"
                              for c in PUSH_CASES])
def test_push_parity(gpu, name, dim, prec, xs, rng_mode, order):
    import torch
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine
    n = 30000
    gas = 3.2e24 if xs == 1 else 3.22e23
    cfg = _cfg(ph, dim, prec, xs, n, gas=gas, push_order=order)
    eng = KineticEngine(cfg, capacity=2 * n + 64)
    dt_np = cfg.np_dtype
    st = _state(n, dt_np, dim, seed=sum(map(ord, name)))
    eng.load_state(st)
    E = np.random.default_rng(1).uniform(-400, 400, cfg.n_cells).astype(dt_np)
    if dim == 1:
        eng.E_t.copy_(torch.from_numpy(E))
    seed = cfg.seed
    if rng_mode == "stream":
        R = np.random.default_rng(2).uniform(0, 1, size=(7, n)).astype(dt_np)
        eng.push(R=torch.from_numpy(R).to(eng.dev))
    else:
        R = O.collision_uniforms(seed, 0, n, dt_np)
        eng.push()
    got = eng.fetch_state()
    mom = eng.moments.cpu().numpy().reshape(3, -1)
    cnt = eng.counters()
    want, wmom, wcnt = O.push_step(_setup(cfg, eng.table), st, E if dim == 1 else None, R,
                                   table=_tab(eng.table), capacity=2 * n + 64)
    assert cnt["births"] == wcnt[0] and cnt["ionization"] == wcnt[4]
    assert cnt["absorbed"] == wcnt[1]
    assert cnt["elastic"] == wcnt[2] and cnt["excitation"] == wcnt[3] and cnt["null"] == wcnt[5]
    assert wcnt[4] > 0 and wcnt[2] > 0
    _compare_state(got, want, dt_np, dim)
    rt = 1e-12 if prec == 8 else 1e-6
    np.testing.assert_allclose(mom, wmom, rtol=rt, atol=rt * np.abs(wmom).max())
    eng.close()


QUEUE_CASES = [(xs, prec, order) for xs in (1, 3) for prec in (4, 8) for order in (0, 1)]


@pytest.mark.parametrize("xs,prec,order", QUEUE_CASES,
                         ids=[f"{'table' if c[0] == 1 else 'direct'}-f{8 * c[1]}-"
                              f"{'leapfrog' if c[2] else 'ref-order'}" for c in QUEUE_CASES])
def test_push_parity_queued(gpu, xs, prec, order):
    """The candidate queue of the 1-D push (forced on, one block per CU so
    that every wave walks several tiles and its batches cross tiles) gives
    the oracle's decisions, births in parent order and moments."""
    import torch
    from boltzmann_amd import _lib
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine
    n = 1_000_000
    gas = 3.2e24 if xs == 1 else 3.22e23
    cfg = _cfg(ph, 1, prec, xs, n, gas=gas, push_order=order)
    eng = KineticEngine(cfg, capacity=2 * n + 64)
    _lib.check(eng.lib.bz_set_push_queue(eng.ctx, 1))
    _lib.check(eng.lib.bz_set_push_grid(eng.ctx, 1))
    _lib.check(eng.lib.bz_set_push_shape(eng.ctx, 3))
    dt_np = cfg.np_dtype
    st = _state(n, dt_np, 1, seed=17 + xs + prec + order)
    eng.load_state(st)
    E = np.random.default_rng(1).uniform(-400, 400, cfg.n_cells).astype(dt_np)
    eng.E_t.copy_(torch.from_numpy(E))
    R = O.collision_uniforms(cfg.seed, 0, n, dt_np)
    eng.push()
    got = eng.fetch_state()
    mom = eng.moments.cpu().numpy().reshape(3, -1)
    cnt = eng.counters()
    s = _setup(cfg, eng.table)
    s.nthreads = 8
    want, wmom, wcnt = O.push_step(s, st, E, R, table=_tab(eng.table), capacity=2 * n + 64)
    assert cnt["births"] == wcnt[0] and cnt["ionization"] == wcnt[4]
    assert cnt["absorbed"] == wcnt[1]
    assert cnt["elastic"] == wcnt[2] and cnt["excitation"] == wcnt[3] and cnt["null"] == wcnt[5]
    assert wcnt[4] > 100 and wcnt[2] > 1000
    _compare_state(got, want, dt_np, 1)
    rt = 1e-12 if prec == 8 else 1e-6
    np.testing.assert_allclose(mom, wmom, rtol=rt, atol=rt * np.abs(wmom).max())
    eng.close()


FLIGHT_CASES = [("flight-f32-1d", 1, 4), ("flight-f64-1d", 1, 8), ("flight-f32-0d", 0, 4),
                ("flight-f64-0d", 0, 8)]


FLIGHT_ORDERS = [(*c, 0) for c in FLIGHT_CASES] + [(*c, 1) for c in FLIGHT_CASES if c[1] == 1]


@pytest.mark.parametrize("name,dim,prec,order", FLIGHT_ORDERS,
                         ids=[f"{c[0]}-{'leapfrog' if c[3] else 'ref-order'}"
                              for c in FLIGHT_ORDERS])
def test_flight_parity(gpu, name, dim, prec, order):
    """Free-flight collisions (about 12 flights per particle and step) over
    two steps: event counts, births, absorptions bit-exact; states and
    moments to rounding."""
    import torch
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine
    n = 20000
    cfg = _cfg(ph, dim, prec, 2, n, gas=3.2e23, push_order=order)
    eng = KineticEngine(cfg, capacity=3 * n + 64)
    dt_np = cfg.np_dtype
    st = _state(n, dt_np, dim, seed=sum(map(ord, name)))
    eng.load_state(st)
    E = np.random.default_rng(1).uniform(-400, 400, cfg.n_cells).astype(dt_np)
    if dim == 1:
        eng.E_t.copy_(torch.from_numpy(E))
    cur = st
    for step in range(2):
        eng.push()
        _lib_step = eng.lib.bz_clock_advance(eng.ctx, cfg.dt)
        assert _lib_step == 0
        got = eng.fetch_state()
        mom = eng.moments.cpu().numpy().reshape(3, -1)
        cnt = eng.counters()
        s = _setup(cfg, eng.table)
        s.seed, s.step = cfg.seed, step
        want, wmom, wcnt = O.push_step(s, cur, E if dim == 1 else None, None,
                                       table=_tab(eng.table), capacity=3 * n + 64)
        if step == 0:
            # ~12 flights per particle (nu_max dt), most ending in null events
            assert wcnt[5] > 5 * n and wcnt[2] > n // 2 and wcnt[4] > 0
            base = dict(cnt)
            for k in base:
                base[k] = 0
        assert cnt["births"] - base["births"] == wcnt[0]
        assert cnt["absorbed"] - base["absorbed"] == wcnt[1]
        assert cnt["elastic"] - base["elastic"] == wcnt[2]
        assert cnt["excitation"] - base["excitation"] == wcnt[3]
        assert cnt["null"] - base["null"] == wcnt[5]
        _compare_state(got, want, dt_np, dim)
        rt = 1e-12 if prec == 8 else 1e-6
        np.testing.assert_allclose(mom, wmom, rtol=rt, atol=rt * np.abs(wmom).max())
        base = dict(cnt)
        # the next step reads the slots as they lie (1-D: holes included;
        # the random stream is keyed by slot index)
        cur = eng.fetch_slots()
        np.testing.assert_array_equal(cur[4][cur[4] > 0], want[4])
    eng.close()


@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4])
def test_schedules_and_shapes_agree(gpu, shape):
    """Every tile shape and grid gives bit-identical slots and counters: the
    1-D push at any blocks-per-CU setting (against shape 0 at the full
    grid), the compacting 0-D push under the static (cooperative) and the
    ticket schedule.  The FP64 node moments agree to rounding (blocks sum
    different particles, in LDS atomic order)."""
    import torch
    from boltzmann_amd import _lib
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine
    n = 70_000

    def run(dim, shape_, sched, bpc):
        cfg = _cfg(ph, dim, 4, 1, n, gas=3.2e24)
        eng = KineticEngine(cfg, capacity=2 * n + 64)
        _lib.check(eng.lib.bz_set_push_shape(eng.ctx, shape_))
        _lib.check(eng.lib.bz_set_push_schedule(eng.ctx, sched))
        _lib.check(eng.lib.bz_set_push_grid(eng.ctx, bpc))
        eng.load_state(_state(n, np.float32, dim, seed=77))
        if dim == 1:
            eng.E_t.copy_(torch.linspace(-300, 300, cfg.n_cells))
        for _ in range(3):
            eng.push()
            if dim == 1:
                eng.field_step(advance=False)
            _lib.check(eng.lib.bz_clock_advance(eng.ctx, cfg.dt))
        out = (eng.fetch_slots(), eng.moments.cpu().numpy(), eng.counters())
        eng.close()
        return out

    for dim, runs in ((1, [(0, 0, 0), (shape, 0, 0), (shape, 0, 1), (shape, 0, 3)]),
                      (0, [(shape, 0, 0), (shape, 1, 0)])):
        (a, ma, ca), *rest = [run(dim, *r) for r in runs]
        for b, mb, cb in rest:
            for u, v in zip(a, b):
                np.testing.assert_array_equal(u, v)
            np.testing.assert_allclose(mb, ma, rtol=1e-14, atol=1e-14 * np.abs(ma).max())
            assert cb == ca


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("dim", [0, 1])
def test_deposit_parity(gpu, prec, dim):
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine
    n = 100_000
    cfg = _cfg(ph, dim, prec, 1, n, n_cells=300)
    eng = KineticEngine(cfg, capacity=n + 16)
    st = _state(n, cfg.np_dtype, dim, seed=3)
    eng.load_state(st)
    eng.deposit()
    mom = eng.moments.cpu().numpy().reshape(3, -1)
    want = O.deposit(_setup(cfg, eng.table), st)
    np.testing.assert_allclose(mom, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    eng.close()


def test_empty_and_capacity(gpu):
    import torch
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine
    cfg = _cfg(ph, 1, 4, 1, 16, gas=3.2e24)
    eng = KineticEngine(cfg, capacity=4096)
    eng.load_state(tuple(np.zeros(0, dtype=np.float32) for _ in range(5)))
    eng.push()
    assert eng.count() == 0
    # in place: the births are appended after the n slots; those beyond the
    # capacity are dropped (and counted), the rest keep parent order
    for n in (4096, 4092):
        st = _state(n, np.float32, 1, seed=9, dead_frac=0.0)
        eng.load_state(st)
        eng.lib.bz_counters_reset(eng.ctx)
        R = np.random.default_rng(3).uniform(0, 1, size=(7, n)).astype(np.float32)
        eng.push(R=torch.from_numpy(R).to(eng.dev))
        c = eng.counters()
        want, _, wcnt = O.push_step(_setup(cfg, eng.table), st, np.zeros(48, np.float32), R,
                                    table=_tab(eng.table), capacity=2 * n + 16)
        births = int(wcnt[0])
        kept = max(0, min(births, 4096 - n))
        assert births > kept  # the capacity is reached
        assert eng.count() == n + kept
        assert c["overflow"] == births - kept and c["births"] == births
        n_keep = want[0].size - births + kept
        _compare_state(eng.fetch_state(), tuple(a[:n_keep] for a in want), np.float32)
    eng.close()


@pytest.mark.parametrize("prec", [8, 4])
def test_trajectory_0d(gpu, prec):
    """Many steps of a 0-D swarm (Philox stream): counts stay identical to the
    oracle step by step; moments track within tolerance."""
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine
    n = 4000
    cfg = _cfg(ph, 0, prec, 1, n, gas=3.2e22, dt=1e-12, e_over_n_td=600.0, te_ev=5.0)
    eng = KineticEngine(cfg, capacity=4 * n)
    st = ph.initial_state(cfg)
    eng.load_state(st)
    setup = _setup(cfg, eng.table)
    cur = st
    steps = 40 if prec == 8 else 10
    for s in range(steps):
        eng.push()
        eng.field_step()
        R = O.collision_uniforms(cfg.seed, s, cur[0].size, cfg.np_dtype)
        cur, wmom, _ = O.push_step(setup, cur, None, R, table=_tab(eng.table),
                                   capacity=4 * n)
        assert eng.count() == cur[0].size, s
    got = eng.fetch_state()
    rt = 1e-9 if prec == 8 else 1e-4
    for k in range(1, 5):  # 0-D: no position
        np.testing.assert_allclose(got[k], cur[k], rtol=rt, atol=rt * np.abs(cur[k]).max())
    eng.close()


@pytest.mark.parametrize("N,scheme,emax", [(256, 0, 50.0), (256, 1, 800.0), (4096, 1, 2000.0),
                                            (4096, 0, 5.0)])
def test_field_parity(gpu, N, scheme, emax):
    """Moment update, heavies (reference central fluxes or the sub-cycled
    Scharfetter-Gummel fluxes) and the Poisson solve, up to 4096 cells."""
    import torch
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine
    cfg = _cfg(ph, 1, 8, 1, 100, n_cells=N, gas=3.22e21, dt=1.844e-10, rf_voltage=200.0,
               ion_scheme=scheme)
    eng = KineticEngine(cfg, capacity=1024)
    rng = np.random.default_rng(7)
    ne = rng.uniform(0.8e9, 1.2e9, N + 1)
    nsrc = rng.uniform(0, 1e4, N + 1)
    nrg = ne * rng.uniform(1, 5, N + 1)
    mom = np.concatenate([ne, nsrc, nrg])
    E0 = rng.uniform(-emax, emax, N)
    eng.moments.copy_(torch.from_numpy(mom))
    eng.E.copy_(torch.from_numpy(E0))
    ni0 = eng.ni.cpu().numpy().copy()
    eng.load_state(tuple(np.zeros(1) for _ in range(5)), t=3.3e-9)
    eng.field_step(advance=True)
    fp = eng.fp
    chol = ph.poisson_cholesky(N, cfg.dx)
    # the oracle takes the reference's scipy factor
    from scipy.linalg import cholesky_banded
    Ac = np.zeros((2, N - 1))
    Ac[0, 1:] = -1.0 / cfg.dx ** 2
    Ac[1, :] = 2.0 / cfg.dx ** 2
    c = cholesky_banded(Ac)
    chol_ref = np.stack([c[1], np.concatenate([[0.0], c[0, 1:]])])
    out = O.field_step(N, fp.dx, fp.dt, fp.eps, fp.v0, fp.freq, fp.mu_i, fp.D_i, fp.D_s, 1, 1,
                       3.3e-9 + cfg.dt, mom, ni0, np.zeros(N + 1), np.zeros(N + 1),
                       np.zeros(N + 1), np.zeros(N - 1), E0.copy(), chol_ref,
                       ion_scheme=scheme, max_substeps=fp.max_substeps)
    if scheme == 1:
        assert out[0].min() >= 0
    names = ("ni", "ns", "Te", "g2", "V", "E")
    for nm, want in zip(names, out):
        got = getattr(eng, nm).cpu().numpy()
        np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-10 * (np.abs(want).max() + 1e-300),
                                   err_msg=nm)
    assert chol is not None
    _, step, t, _ = eng.clock()
    assert step == 1 and t == pytest.approx(3.3e-9 + cfg.dt, rel=1e-15)
    eng.close()


@pytest.mark.parametrize("prec,xs", [(8, 0), (8, 3), (4, 3)])
def test_push_warm_gas_philox(gpu, prec, xs):
    """Warm gas (300 K, quirk Q11: the variance uses the electron mass) in
    Philox mode: the device's own gas-velocity draw (Box-Muller, purpose 1)
    equals the oracle's O.gas_velocities fed to the oracle push as the
    reference's gas_v array (1bin/boltzmann.py:556-561,589)."""
    import torch
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine
    n = 30000
    cfg = _cfg(ph, 1, prec, xs, n, gas=3.22e23, gas_temp_k=300.0, push_order=1)
    eng = KineticEngine(cfg, capacity=2 * n + 64)
    assert eng.params.gas_sigma > 0
    dt_np = cfg.np_dtype
    st = _state(n, dt_np, 1, seed=77)
    eng.load_state(st)
    E = np.random.default_rng(4).uniform(-300, 300, cfg.n_cells).astype(dt_np)
    eng.E_t.copy_(torch.from_numpy(E))
    eng.push()
    got = eng.fetch_state()
    mom = eng.moments.cpu().numpy().reshape(3, -1)
    cnt = eng.counters()
    R = O.collision_uniforms(cfg.seed, 0, n, dt_np)
    G = O.gas_velocities(cfg.seed, 0, n, dt_np, eng.params.gas_sigma)
    want, wmom, wcnt = O.push_step(_setup(cfg, eng.table), st, E, R, G=G,
                                   table=_tab(eng.table), capacity=2 * n + 64)
    assert cnt["births"] == wcnt[0] and cnt["ionization"] == wcnt[4]
    assert cnt["elastic"] == wcnt[2] and cnt["absorbed"] == wcnt[1]
    _compare_state(got, want, dt_np, 1)
    rt = 1e-12 if prec == 8 else 1e-6
    np.testing.assert_allclose(mom, wmom, rtol=rt, atol=rt * np.abs(wmom).max())
    eng.close()


def test_push_dead_ionizer_flag(gpu):
    """BZ_REF_DEAD_IONIZER: the device adds the reference's unscaled flag of
    each primary that ionized and left to ni_src at node 0, as the oracle
    does (post_processing_1D.py:48-55)."""
    from boltzmann_amd import _lib
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine
    n, L, N = 4000, 2.5, 25
    rng = np.random.default_rng(12)
    x = np.concatenate([rng.uniform(L - 1e-6, L - 1e-7, n // 2), rng.uniform(0.5, 2.0, n // 2)])
    sp = math.sqrt(120.0 / KEV)
    vx = np.concatenate([np.full(n // 2, sp), np.full(n // 2, -1e3)])
    st = (x, vx, np.zeros(n), np.zeros(n), np.full(n, 0.7))
    cfg = _cfg(ph, 1, 8, 0, n, n_cells=N, gas=3e23, dt=1e-10)
    for flags in (0, _lib.BZ_REF_DEAD_IONIZER):
        eng = KineticEngine(cfg, capacity=2 * n + 64)
        eng.params.ref_flags = flags
        eng.load_state(st)
        eng.push()
        mom = eng.moments.cpu().numpy().reshape(3, -1)
        R = O.collision_uniforms(cfg.seed, 0, n, np.float64)
        s = O.Setup(dim=1, dt=cfg.dt, xs_mode=0, n_cells=N, length_cm=L, gas_density_m3=3e23,
                    ref_flags=flags)
        _, wmom, wcnt = O.push_step(s, st, np.zeros(N), R)
        np.testing.assert_allclose(mom, wmom, rtol=1e-12, atol=1e-12 * np.abs(wmom).max())
        if flags:
            assert mom[1][0] >= 10  # dead ionizers happened and were counted
        eng.close()


COLL_CASES = [("table-f64", 8, 1, False), ("direct-f32", 4, 3, False),
              ("direct-f32-queued", 4, 3, True), ("flight-f64", 8, 2, False)]


@pytest.mark.parametrize("name,prec,xs,queued", COLL_CASES, ids=[c[0] for c in COLL_CASES])
def test_collision_counters_match_the_oracle(gpu, name, prec, xs, queued):
    """bz_collision_counts after two 1-D pushes equals the oracle's counters
    of the two steps (the reference's collct deposit: sum w of elastic and
    of excitation collisions, the ionization count, at the primary's node
    after the push, node 0 when absorbed): counts exact, weights within
    1e-12 / 1e-6; a read with reset leaves zeros; counters off add nothing."""
    import torch
    from boltzmann_amd import _lib
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine
    n = 200_000 if queued else 30_000
    gas = 3.2e24 if xs == 1 else 3.22e23
    cfg = _cfg(ph, 1, prec, xs, n, gas=gas, push_order=1)
    eng = KineticEngine(cfg, capacity=3 * n + 64)
    if queued:
        _lib.check(eng.lib.bz_set_push_queue(eng.ctx, 1))
        _lib.check(eng.lib.bz_set_push_grid(eng.ctx, 1))
        _lib.check(eng.lib.bz_set_push_shape(eng.ctx, 3))
    dt_np = cfg.np_dtype
    st = _state(n, dt_np, 1, seed=len(name))
    eng.load_state(st)
    E = np.random.default_rng(1).uniform(-400, 400, cfg.n_cells).astype(dt_np)
    eng.E_t.copy_(torch.from_numpy(E))
    eng.push()  # counters off: nothing is counted
    assert not eng.collision_counts(reset=True).cpu().numpy().any()
    eng.set_collision_counts(True)
    s = _setup(cfg, eng.table)
    s.seed = cfg.seed
    want = np.zeros((3, cfg.n_nodes))
    for _ in range(2):
        src = eng.fetch_slots()
        _, step, _, _ = eng.clock()
        R = None if xs == 2 else O.collision_uniforms(cfg.seed, step, src[0].size, dt_np)
        s.step = step
        _, _, _, coll = O.push_step(s, src, E, R, table=_tab(eng.table),
                                    capacity=3 * n + 64, collisions=True)
        want += coll
        eng.push()
    got = eng.collision_counts(reset=True).cpu().numpy()
    np.testing.assert_array_equal(got[2], want[2])
    rt = 1e-12 if prec == 8 else 1e-6
    np.testing.assert_allclose(got[:2], want[:2], rtol=rt, atol=rt * np.abs(want[:2]).max())
    assert want[0].sum() > 0 and want[2].sum() > 0 and (xs == 1 or want[1].sum() > 0)
    assert not eng.collision_counts(reset=False).cpu().numpy().any()
    eng.close()
