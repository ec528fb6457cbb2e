"""GPU binning / radix sort / key-order ES-WPNC resampling against the oracle.

Keys, counts, offsets and the permutation must be identical; the resampled
state has the same count and order, selections identical (fixed-point CDF),
floating point within 1e-12 (FP64) or 1e-6 (FP32) relative."""
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


def _engine(N, prec, nb, target, n, inplace=False, **kw):
    """An engine whose resample() is the key-order bz_resample (inplace=False)
    or the in-place bz_resample_inplace."""
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine
    cfg = ph.RunConfig(name="rs", dim=1, n_particles=n, precision=prec, nb=nb, n_cells=N,
                       length_cm=2.5, dt=1e-11, target_ppc=target,
                       bin_fractions=ph.bin_fractions(nb), bin_thresholds=ph.bin_thresholds(nb),
                       **kw)
    return cfg, KineticEngine(cfg, capacity=4 * n + 4096, inplace_resample=inplace)


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
    rt = 1e-12 if prec == 8 else 1e-6
    sp = np.sqrt(sum(want[k].astype(np.float64) ** 2 for k in (1, 2, 3)))
    for k in (1, 2, 3):
        g, wv = got[k].astype(np.float64), want[k].astype(np.float64)
        assert np.all(np.abs(g - wv) <= rt * sp), k
    g, wv = got[4].astype(np.float64), want[4].astype(np.float64)
    assert np.all(np.abs(g - wv) <= rt * np.abs(wv))
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
    want, _ = O.resample_inplace(s, pol, cfg.nb, cfg.bin_thresholds, eng.targets_host, cfg.seed,
                                 0, st, capacity=eng.capacity)
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
    step, bit for bit (resampling fires, and each pass is followed by a
    compaction in place)."""
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
    assert len(b._graphs) >= 2
    torch.cuda.synchronize()
    ca, cb = a.clock(), b.clock()
    assert ca[:3] == cb[:3] and ca[3] == cb[3]
    assert a.counters()["resample"] > 0 and a.counters()["merge"] > 0
    for u, v in zip(a.fetch_slots(), b.fetch_slots()):
        np.testing.assert_array_equal(u, v)
    # fixed-point node sums: bit for bit
    np.testing.assert_array_equal(a.moments.cpu().numpy(), b.moments.cpu().numpy())
    np.testing.assert_array_equal(a.E.cpu().numpy(), b.E.cpu().numpy())
    a.close()
    b.close()


@pytest.mark.parametrize("prec", [8, 4])
def test_bin_keys_energy_roundtrip_flag(gpu, prec):
    """BZ_REF_ENERGY_ROUNDTRIP: keys from (e*w)/w in the state precision, one
    ulp below the 10 eV threshold, equal the oracle's, on the sort path and
    on the keys the push writes (bz_set_push_keys)."""
    from boltzmann_amd import _lib
    from boltzmann_amd import physics as ph
    dt = np.float64 if prec == 8 else np.float32
    kev = dt(KEV)
    below = np.nextafter(dt(10.0), dt(0.0))
    # a velocity whose energy, as the kernels evaluate it, is one ulp below
    # 10 eV: vx near sqrt(10/kev), vy tuning the sum below an ulp of vx^2
    v0 = dt(math.sqrt(10.0 / KEV))
    vx_c = [v0]
    for _ in range(300):
        vx_c.append(np.nextafter(vx_c[-1], dt(1e9)))
        vx_c.insert(0, np.nextafter(vx_c[0], dt(0.0)))
    vx_c = np.array(vx_c, dtype=dt)
    found = None
    for vy in np.arange(0, 3000, 5, dtype=dt):
        e = kev * ((vx_c * vx_c + vy * vy) + dt(0) * dt(0))
        hit = np.nonzero(e == below)[0]
        if hit.size:
            found = (vx_c[hit[0]], vy)
            break
    assert found is not None
    vx, vy = found
    N, nb, m = 4, 3, 6000
    rng = np.random.default_rng(9)
    w = rng.uniform(0.1, 50.0, m).astype(dt)
    st = (np.full(m, 1.0, dt), np.full(m, vx, dt), np.full(m, vy, dt), np.zeros(m, dt), w)
    cfg, eng = _engine(N, prec, nb, 100, m)
    eng.params.ref_flags = _lib.BZ_REF_ENERGY_ROUNDTRIP
    eng.load_state(st)
    eng.bin_sort()
    s = O.Setup(dim=1, dt=cfg.dt, n_cells=N, length_cm=2.5, ref_flags=2)
    c, o, p = O.bin_sort(s, nb, ph.bin_thresholds(nb), st)
    np.testing.assert_array_equal(eng.counts.cpu().numpy(), c)
    np.testing.assert_array_equal(eng.perm[:m].cpu().numpy().view(np.uint32), p)
    node = int((1.0 + 0.5 * 2.5 / N) / (2.5 / N))
    assert 0 < c[node * nb + 1] < m  # the round trip moved some particles
    eng.close()


def test_run_overlapped_field_matches_steps(gpu):
    """A discharge that resamples every step runs its steps directly, with
    each step's field step on a side stream beside the next step's
    resampling check (KineticEngine._run_overlapped); the run equals step()
    called step by step, bit for bit: slots, node moments, field, clock
    (C4's layout scaled: 5 bins, keys written by the push, compaction by
    hole filling every 4th pass)."""
    import torch
    from boltzmann_amd import configs
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine
    cfg = configs.c4().with_(n_particles=300_000, n_cells=256, target_ppc=1200)
    st = ph.initial_state(cfg)
    engs = []
    for _ in range(2):
        e = KineticEngine(cfg)
        e.load_state(st)
        e.initialize_fields()
        engs.append(e)
    a, b = engs
    n = 13
    for i in range(n):
        a.step(i)
    assert b.run(0, n) == n
    assert b._field_stream is not None  # the overlapped path ran
    torch.cuda.synchronize()
    ca, cb = a.clock(), b.clock()
    assert ca[:3] == cb[:3] and ca[3] == cb[3]
    assert a.counters()["merge"] > 0 and a.counters()["split"] > 0
    for u, v in zip(a.fetch_slots(), b.fetch_slots()):
        np.testing.assert_array_equal(u, v)
    np.testing.assert_array_equal(a.moments.cpu().numpy(), b.moments.cpu().numpy())
    np.testing.assert_array_equal(a.E.cpu().numpy(), b.E.cpu().numpy())
    np.testing.assert_array_equal(a.ni.cpu().numpy(), b.ni.cpu().numpy())
    a.close()
    b.close()
