"""GPU binning / radix sort / ES-WPNC resampling against the oracle.

Keys, counts, offsets and the permutation must be identical; the resampled
state has the same count and order, selections identical (fixed-point CDF),
floating point within 1e-12 (FP64) or 2e-6 (FP32) of the velocity scale."""
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

KEV = 0.5 * O.M_E * O.J_EV


def _state(counts_per_cell, N, L, dtype, seed=0, drift=0.0, e_ev=8.0):
    rng = np.random.default_rng(seed)
    dx = L / N
    xs, vs, ws = [], [], []
    for c, m in enumerate(counts_per_cell):
        lo = max((c - 0.5) * dx, 1e-6)
        hi = min((c + 0.5) * dx, L - 1e-6)
        xs.append(rng.uniform(lo, hi, m))
        v = rng.normal(0, math.sqrt(e_ev / KEV / 3), size=(3, m))
        v[0] += drift
        vs.append(v)
        ws.append(rng.uniform(0.5, 2.0, m))
    x = np.concatenate(xs)
    v = np.concatenate(vs, axis=1)
    w = np.concatenate(ws)
    p = rng.permutation(x.size)
    return tuple(np.ascontiguousarray(a[p], dtype=dtype) for a in (x, v[0], v[1], v[2], w))


def _engine(N, prec, nb, target, n, **kw):
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine
    cfg = ph.RunConfig(name="rs", dim=1, n_particles=n, precision=prec, nb=nb, n_cells=N,
                       length_cm=2.5, dt=1e-11, target_ppc=target,
                       bin_fractions=ph.bin_fractions(nb), bin_thresholds=ph.bin_thresholds(nb),
                       **kw)
    return cfg, KineticEngine(cfg, capacity=4 * n + 4096)


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("nb", [1, 3, 5])
def test_bin_sort_parity(gpu, prec, nb):
    from boltzmann_amd import physics as ph
    N = 300 if nb < 5 else 60
    rng = np.random.default_rng(nb)
    cells = list(rng.integers(0, 400, N + 1))
    cfg, eng = _engine(N, prec, nb, 100, int(sum(cells)))
    st = _state(cells, N, 2.5, cfg.np_dtype, seed=nb, e_ev=12.0)
    w = st[4].copy()
    w[::23] = 0
    st = st[:4] + (w,)
    eng.load_state(st)
    eng.bin_sort()
    s = O.Setup(dim=1, dt=cfg.dt, n_cells=N, length_cm=2.5)
    c, o, p = O.bin_sort(s, nb, ph.bin_thresholds(nb), st)
    np.testing.assert_array_equal(eng.counts.cpu().numpy(), c)
    np.testing.assert_array_equal(eng.offsets.cpu().numpy(), o)
    np.testing.assert_array_equal(eng.perm[:st[0].size].cpu().numpy().view(np.uint32), p)
    eng.close()


@pytest.mark.parametrize("prec,nb,N", [(8, 3, 300), (4, 5, 4096)])
def test_bin_count_matches_sort(gpu, prec, nb, N):
    """The trigger's counts without the sort equal the oracle's per-key
    counts; 4096 cells x 5 bins exceed the per-block LDS histogram and take
    the global-atomic path."""
    from boltzmann_amd import physics as ph
    rng = np.random.default_rng(7 + nb)
    cells = list(rng.integers(0, 60 if N > 1000 else 400, N + 1))
    cfg, eng = _engine(N, prec, nb, 100, int(sum(cells)))
    st = _state(cells, N, 2.5, cfg.np_dtype, seed=nb, e_ev=12.0)
    w = st[4].copy()
    w[::19] = 0
    st = st[:4] + (w,)
    eng.load_state(st)
    eng.bin_count()
    s = O.Setup(dim=1, dt=cfg.dt, n_cells=N, length_cm=2.5)
    c, _, _ = O.bin_sort(s, nb, ph.bin_thresholds(nb), st)
    np.testing.assert_array_equal(eng.counts.cpu().numpy()[:c.size], c)
    eng.close()


RS_CASES = [
    ("f64-1bin-drift", 8, 1, 3e5),
    ("f64-3bin", 8, 3, 0.0),
    ("f32-3bin", 4, 3, 0.0),
    ("f32-5bin", 4, 5, 1e5),
]


@pytest.mark.parametrize("name,prec,nb,drift", RS_CASES, ids=[c[0] for c in RS_CASES])
def test_resample_parity(gpu, name, prec, nb, drift):
    from boltzmann_amd import physics as ph
    N, target = 40, 300
    rng = np.random.default_rng(len(name))
    cells = [int(v) for v in rng.choice([0, 20, 60, 150, 300, 700, 1500], N + 1)]
    cfg, eng = _engine(N, prec, nb, target, int(sum(cells)), trigger_at=0.0, merge_at=2.0)
    st = _state(cells, N, 2.5, cfg.np_dtype, seed=3, drift=drift, e_ev=9.0)
    eng.load_state(st, step=5)
    fired = eng.resample()
    assert fired
    got = eng.fetch_state()
    cnt = eng.counters()
    s = O.Setup(dim=1, dt=cfg.dt, n_cells=N, length_cm=2.5)
    pol = O.policy(split_below=cfg.split_below, merge_at=cfg.merge_at, merge_keep=cfg.merge_keep,
                   trigger_at=cfg.trigger_at, min_wt=cfg.min_wt, split_to=cfg.split_to)
    want, wc = O.resample(s, pol, nb, ph.bin_thresholds(nb), eng.targets_host, cfg.seed, 5, st,
                          capacity=eng.capacity)
    assert got[0].size == want[0].size
    assert cnt["split"] == wc[8] and cnt["merge"] == wc[9] and cnt["resample"] == 1
    assert wc[8] > 0 and wc[9] > 0
    np.testing.assert_array_equal(got[0], want[0])  # positions are copied: same selection
    rt = 1e-12 if prec == 8 else 2e-6
    vs = np.sqrt(sum(want[k].astype(np.float64) ** 2 for k in (1, 2, 3))) + 1
    for k in (1, 2, 3):
        assert np.all(np.abs(got[k].astype(np.float64) - want[k]) <= rt * vs), k
    np.testing.assert_allclose(got[4], want[4], rtol=rt)
    # conservation of the GPU result (total weight, momentum, energy)
    if prec == 8:
        for a, b in ((got, st),):
            Wa, Wb = a[4].sum(), b[4].sum()
            assert abs(Wa - Wb) <= 1e-12 * Wb
            Ea = np.sum(a[4] * (a[1] ** 2 + a[2] ** 2 + a[3] ** 2))
            Eb = np.sum(b[4] * (b[1] ** 2 + b[2] ** 2 + b[3] ** 2))
            assert abs(Ea - Eb) <= 1e-12 * Eb
            Pa, Pb = np.sum(a[4] * a[1]), np.sum(b[4] * b[1])
            assert abs(Pa - Pb) <= 1e-12 * math.sqrt(Wb * Eb)
    eng.close()


def test_step_with_resampling_runs(gpu):
    """A few full steps of a small c1-like discharge with resampling each step."""
    from boltzmann_amd import configs
    cfg = configs.c1().with_(n_particles=60_000, n_cells=32, target_ppc=1000, trigger_at=0.0,
                             resample_every=1)
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine
    eng = KineticEngine(cfg)
    st = ph.initial_state(cfg)
    eng.load_state(st)
    eng.initialize_fields()
    w0 = float(st[4].astype(np.float64).sum())
    assert eng.resample()
    w1 = float(eng.fetch_state()[4].astype(np.float64).sum())
    assert abs(w1 - w0) <= 1e-5 * w0  # FP32 weights: storage rounding only
    s = O.Setup(dim=1, dt=cfg.dt, n_cells=cfg.n_cells, length_cm=cfg.length_cm)
    pol = O.policy(split_below=cfg.split_below, merge_at=cfg.merge_at, merge_keep=cfg.merge_keep,
                   trigger_at=cfg.trigger_at, min_wt=cfg.min_wt, split_to=cfg.split_to)
    want, _ = O.resample(s, pol, cfg.nb, cfg.bin_thresholds, eng.targets_host, cfg.seed, 0, st,
                         capacity=eng.capacity)
    assert eng.count() == want[0].size
    for i in range(1, 5):
        eng.step(i)
    n = eng.count()
    assert 0 < n < 2 * cfg.n_particles
    c = eng.counters()
    assert c["resample"] == 5 and c["overflow"] == 0
    eng.close()


def test_run_with_graphs_matches_steps(gpu):
    """run() replays captured blocks of pushes + field steps between the
    resampling checks and compactions; the run equals step() called step by
    step, bit for bit (resampling fires and swaps the state buffers, and the
    compactions fall inside the check interval)."""
    import torch
    from boltzmann_amd import configs
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine
    cfg = configs.c1().with_(n_particles=200_000, n_cells=64, target_ppc=1000, resample_every=5)
    st = ph.initial_state(cfg)
    engs = []
    for _ in range(2):
        e = KineticEngine(cfg, compact_every=7)
        e.load_state(st)
        e.initialize_fields()
        engs.append(e)
    a, b = engs
    for i in range(23):
        a.step(i)
    assert b.run(0, 23) == 23
    assert len(b._graphs) >= 3
    torch.cuda.synchronize()
    ca, cb = a.clock(), b.clock()
    assert ca[:3] == cb[:3] and ca[3] == cb[3]
    assert a.counters()["resample"] > 0 and a.counters()["merge"] > 0
    for u, v in zip(a.fetch_slots(), b.fetch_slots()):
        np.testing.assert_array_equal(u, v)
    # the FP64 node sums are added in LDS atomics within a block, in the
    # order its waves arrive: equal to rounding, not bit for bit
    np.testing.assert_allclose(a.moments.cpu().numpy(), b.moments.cpu().numpy(), rtol=1e-13)
    np.testing.assert_allclose(a.E.cpu().numpy(), b.E.cpu().numpy(), rtol=1e-12, atol=1e-9)
    a.close()
    b.close()
