"""In-place ES-WPNC resampling (the reference's layout) against the oracle.

bz_resample_inplace leaves untouched segments where they are, divides a
split particle's weight in place and appends its clones, turns a merged
segment's members into holes and appends its K particles
(1bin/boltzmann.py:806-877).  Every slot, holes included, must equal the
oracle's: counts and selections exactly (positions are copied), floating
point within 1e-12 (FP64) or 1e-6 (FP32) relative.  The caller-supplied
uniform stream is checked against the reference's own formulas with numpy:
merge selections searchsorted(cumsum(w/W), (U+k)/K) (:845-848) and split
draws searchsorted(cumsum(p), u, 'right') with p = w/sum(w) over
w > min_wt (CuPy's choice with p, :812-814)."""
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

KEV = 0.5 * O.M_E * O.J_EV


def _state(counts_per_cell, N, L, dtype, seed=0, drift=0.0, e_ev=8.0):
    """Particles per cell, weights uniform in [0.5, 2), slots shuffled."""
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


def _engine(N, prec, nb, target, n, capacity=None, inplace_resample=True, **kw):
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine
    cfg = ph.RunConfig(name="rsip", dim=1, n_particles=n, precision=prec, nb=nb, n_cells=N,
                       length_cm=2.5, dt=1e-11, target_ppc=target,
                       bin_fractions=ph.bin_fractions(nb), bin_thresholds=ph.bin_thresholds(nb),
                       **kw)
    return cfg, KineticEngine(cfg, capacity=capacity or 4 * n + 4096,
                              inplace_resample=inplace_resample)


def _policy(cfg, max_ppc=0):
    return O.policy(split_below=cfg.split_below, merge_at=cfg.merge_at,
                    merge_keep=cfg.merge_keep, trigger_at=cfg.trigger_at, min_wt=cfg.min_wt,
                    split_to=cfg.split_to, max_ppc=max_ppc)


def _compare_slots(got, want, prec):
    assert got[0].size == want[0].size
    np.testing.assert_array_equal(got[0], want[0])  # positions are copied: same selections
    np.testing.assert_array_equal(got[4] > 0, want[4] > 0)  # the same holes
    # per particle: velocity components relative to the speed, weight to itself
    rt = 1e-12 if prec == 8 else 1e-6
    sp = np.sqrt(sum(want[k].astype(np.float64) ** 2 for k in (1, 2, 3)))
    for k in (1, 2, 3):
        g, wv = got[k].astype(np.float64), want[k].astype(np.float64)
        assert np.all(np.abs(g - wv) <= rt * sp), k
    g, wv = got[4].astype(np.float64), want[4].astype(np.float64)
    assert np.all(np.abs(g - wv) <= rt * np.abs(wv))


def _moments(st):
    x, vx, vy, vz, w = [a.astype(np.float64) for a in st]
    live = w > 0
    W = w[live].sum()
    P = np.array([np.sum(w[live] * v[live]) for v in (vx, vy, vz)])
    E = np.sum(w[live] * (vx[live] ** 2 + vy[live] ** 2 + vz[live] ** 2))
    return W, P, E


IP_CASES = [
    ("f64-1bin-drift", 8, 1, 3e5),
    ("f64-3bin", 8, 3, 0.0),
    ("f32-3bin", 4, 3, 0.0),
    ("f32-5bin", 4, 5, 1e5),
]


@pytest.mark.parametrize("name,prec,nb,drift", IP_CASES, ids=[c[0] for c in IP_CASES])
def test_resample_inplace_parity(gpu, name, prec, nb, drift):
    from boltzmann_amd import physics as ph
    N, target = 40, 300
    rng = np.random.default_rng(len(name))
    cells = [int(v) for v in rng.choice([0, 20, 60, 150, 300, 700, 1500], N + 1)]
    cfg, eng = _engine(N, prec, nb, target, int(sum(cells)), trigger_at=0.0, merge_at=2.0)
    st = _state(cells, N, 2.5, cfg.np_dtype, seed=3, drift=drift, e_ev=9.0)
    w = st[4].copy()
    w[::31] = 0  # holes in the input are skipped and kept
    st = st[:4] + (w,)
    eng.load_state(st, step=5)
    eng.resample()
    got = eng.fetch_slots()
    cnt = eng.counters()
    s = O.Setup(dim=1, dt=cfg.dt, n_cells=N, length_cm=2.5)
    want, wc = O.resample_inplace(s, _policy(cfg), nb, ph.bin_thresholds(nb), eng.targets_host,
                                  cfg.seed, 5, st, capacity=eng.capacity)
    assert cnt["split"] == wc[8] and cnt["merge"] == wc[9] and cnt["resample"] == 1
    assert wc[8] > 0 and wc[9] > 0
    _compare_slots(got, want, prec)
    # untouched slots keep their bits, the layout is the reference's
    n = st[0].size
    W0, P0, E0 = _moments(st)
    W1, P1, E1 = _moments(got)
    if prec == 8:
        assert abs(W1 - W0) <= 1e-12 * W0
        assert abs(E1 - E0) <= 1e-12 * E0
        assert np.all(np.abs(P1 - P0) <= 1e-12 * math.sqrt(W0 * E0))
    else:
        # FP32 states: each merged particle stores its FP64 weight and
        # velocity rounded to FP32, so a merged segment keeps W, P and E to
        # the FP32 rounding, 2^-24 relative at most (the FP64 sums of FP32
        # values cannot in general equal the FP64 totals: 1e-12 is out of
        # reach for FP32 storage); the whole state by the merged fraction
        eps = 2.0 ** -24
        assert abs(W1 - W0) <= eps * W0
        assert abs(E1 - E0) <= 2 * eps * E0
        assert np.all(np.abs(P1 - P0) <= 2 * eps * math.sqrt(W0 * E0))
    # the compaction after the pass: the oracle's hole filling
    eng.compact()
    live = eng.fetch_slots()
    for a, b in zip(live, O.fill_holes(got)):
        np.testing.assert_array_equal(a, b)
    assert got[0].size >= n
    eng.close()


@pytest.mark.parametrize("inplace", [True, False], ids=["inplace", "key_order"])
def test_resample_many_keys_and_large_segments(gpu, inplace):
    """More keys than the split / merge grids have blocks (300 cells x 5 bins
    = 1505 keys, each block walks several), a merge segment above the
    merge's LDS CDF (6144) clipped by max_ppc, and a split segment above the
    split's LDS arrays (3968): the workspace paths of both layouts (in place,
    and the key-order output of bz_resample), against the oracle, with FP64
    conservation to 1e-12."""
    from boltzmann_amd import physics as ph
    N, nb = 300, 5
    rng = np.random.default_rng(11)
    cells = [int(v) for v in rng.choice([0, 3, 40, 90, 400], N + 1)]
    cells[17] = 9000   # a large merge segment (all of cell 17 in one bin)
    cells[200] = 4400  # a large split segment
    n = int(sum(cells))
    cfg, eng = _engine(N, 8, nb, 100, n, trigger_at=0.0, merge_at=2.0,
                       inplace_resample=inplace)
    st = list(_state(cells, N, 2.5, np.float64, seed=12, e_ev=2.0))
    # pin cells 17 and 200 to one energy bin each: 2 eV, every bin-0 speed
    dx = 2.5 / N
    cell = ((st[0] + 0.5 * dx) / dx).astype(int)
    for c in (17, 200):
        m = cell == c
        sp = math.sqrt(2.0 / KEV / 3)
        st[1][m], st[2][m], st[3][m] = sp, -sp, sp * (1 + 0.1 * rng.random(m.sum()))
    st = tuple(st)
    tg = eng.targets_host.copy()
    tg[17 * nb + 0] = 3000    # 9000 >= 2 x 3000: merge, clipped to max_ppc
    tg[200 * nb + 0] = 10000  # 4400 < 0.5 x 10000: split
    eng.targets_host = tg
    import torch
    eng.targets.copy_(torch.from_numpy(tg))
    eng.policy.max_ppc = 7000
    eng.load_state(st, step=9)
    eng.resample()
    got = eng.fetch_slots()
    s = O.Setup(dim=1, dt=cfg.dt, n_cells=N, length_cm=2.5)
    oracle_resample = O.resample_inplace if inplace else O.resample
    want, wc = oracle_resample(s, _policy(cfg, max_ppc=7000), nb, ph.bin_thresholds(nb), tg,
                               cfg.seed, 9, st, capacity=eng.capacity)
    assert wc[8] > 0 and wc[9] > 0
    c = eng.counters()
    assert c["split"] == wc[8] and c["merge"] == wc[9]
    _compare_slots(got, want, 8)
    W0, P0, E0 = _moments(st)
    W1, P1, E1 = _moments(got)
    assert abs(W1 - W0) <= 1e-12 * W0
    assert abs(E1 - E0) <= 1e-12 * E0
    assert np.all(np.abs(P1 - P0) <= 1e-12 * math.sqrt(W0 * E0))
    if inplace:
        # the clipped merge leaves members 7000.. of cell 17's segment as they were
        assert np.sum((got[4][:n] == 0) & (st[4] > 0)) >= 7000
    eng.close()


@pytest.mark.parametrize("prec", [8, 4], ids=["f64", "f32"])
def test_resample_inplace_keys_beyond_the_lds_histogram(gpu, prec):
    """8200 cells x 5 bins = 41005 keys: more than the key histogram holds in
    LDS (40960 counters), so the slots are counted with global atomics; 41
    key chunks in the plan scans.  Slot count not a multiple of the 16-byte
    key vectors, holes in the input.  Counts, selections and every slot
    against the oracle."""
    from boltzmann_amd import physics as ph
    N, nb, target = 8200, 5, 40
    rng = np.random.default_rng(41)
    cells = [int(v) for v in rng.choice([0, 5, 15, 40, 120], N + 1)]
    n = int(sum(cells))
    if n % 8 == 0:
        cells[3] += 1
        n += 1
    cfg, eng = _engine(N, prec, nb, target, n, trigger_at=0.0, merge_at=2.0)
    assert (N + 1) * nb > 40960
    st = _state(cells, N, 2.5, cfg.np_dtype, seed=42, e_ev=9.0)
    w = st[4].copy()
    w[::29] = 0
    st = st[:4] + (w,)
    eng.load_state(st, step=7)
    eng.resample()
    got = eng.fetch_slots()
    cnt = eng.counters()
    s = O.Setup(dim=1, dt=cfg.dt, n_cells=N, length_cm=2.5)
    want, wc = O.resample_inplace(s, _policy(cfg), nb, ph.bin_thresholds(nb), eng.targets_host,
                                  cfg.seed, 7, st, capacity=eng.capacity)
    assert wc[8] > 0 and wc[9] > 0
    assert cnt["split"] == wc[8] and cnt["merge"] == wc[9] and cnt["resample"] == 1
    _compare_slots(got, want, prec)
    eng.close()


def _members(st, dx, nb, thr, key, n_nodes):
    x, vx, vy, vz, w = [a.astype(np.float64) for a in st]
    cell = np.clip(((x + 0.5 * dx) / dx).astype(int), 0, n_nodes - 1)
    e = KEV * (vx * vx + vy * vy + vz * vz)
    b = np.zeros(x.size, dtype=int)
    for t in thr:
        b += e >= t
    k = cell * nb + b
    return np.nonzero((k == key) & (w > 0))[0]


def test_resample_stream_matches_reference_formulas(gpu):
    """Caller-supplied uniforms: the GPU equals the oracle given the same
    stream, and the selections equal the reference's numpy formulas.  A
    mismatch is allowed only where a threshold lies within 4 ulp of a CDF
    value (the reference's FP cumsum against the exact fixed-point CDF);
    none occurs here."""
    import torch

    from boltzmann_amd import _lib
    from boltzmann_amd import physics as ph
    N, nb, target = 24, 3, 200
    rng = np.random.default_rng(5)
    cells = [int(v) for v in rng.choice([30, 60, 700, 1100], N + 1)]
    n = int(sum(cells))
    cfg, eng = _engine(N, 8, nb, target, n, trigger_at=0.0, merge_at=2.0)
    st = _state(cells, N, 2.5, np.float64, seed=21, e_ev=9.0)
    n_keys = (N + 1) * nb
    stride = 2 * target + 8
    U = rng.random((n_keys, stride))
    Ud = torch.from_numpy(U.ravel()).to(eng.dev)
    eng.policy.uniforms = Ud.data_ptr()
    eng.policy.u_stride = stride
    eng.load_state(st, step=2)
    eng.resample()
    got = eng.fetch_slots()
    eng.policy.uniforms = None
    s = O.Setup(dim=1, dt=cfg.dt, n_cells=N, length_cm=2.5)
    thr = ph.bin_thresholds(nb)
    want, wc = O.resample_inplace(s, _policy(cfg), nb, thr, eng.targets_host, cfg.seed, 2, st,
                                  capacity=eng.capacity, uniforms=U)
    _compare_slots(got, want, 8)
    assert wc[8] > 0 and wc[9] > 0
    # the reference's formulas, key by key, on the appended particles
    tg = eng.targets_host
    dx = 2.5 / N
    pos = {float(v): i for i, v in enumerate(st[0])}
    off = st[0].size
    mism = checked = 0
    for key in range(n_keys):
        ids = _members(st, dx, nb, thr, key, N + 1)
        c, t = ids.size, int(tg[key])
        if t <= 0:
            continue
        w = st[4][ids]
        if 0 < c < 0.5 * t:
            draws = int(2.0 * t) - c
            hw = w > cfg.min_wt
            if not hw.any() or draws <= 0:
                continue
            p = np.where(hw, w, 0.0)
            cdf = np.cumsum(p / p.sum())
            j = np.searchsorted(cdf, U[key, :draws], side="right")
            ref = np.bincount(np.minimum(j, c - 1), minlength=c)
            # in place: drawn particles keep their slot with w/(c+1)
            reps = np.rint(st[4][ids] / got[4][ids]).astype(int) - 1
            d = np.nonzero(ref != reps)[0]
            for q in d:
                near = np.min(np.abs(U[key, :draws][:, None] - cdf[[q]][None, :]))
                assert near < 4 * np.spacing(1.0), (key, q)
            mism += d.size
            checked += c
            off += draws
        elif c >= 2.0 * t:
            K = int(0.6 * t)
            W = w.sum()
            cdf = np.cumsum(w / W)
            R = (1.0 / K) * (U[key, :K] + np.arange(K))
            j = np.searchsorted(cdf, R)  # side 'left', as the reference calls it
            sel = np.array([pos[float(v)] for v in got[0][off:off + K]])
            d = np.nonzero(ids[np.minimum(j, c - 1)] != sel)[0]
            for q in d:
                assert np.min(np.abs(cdf - R[q])) < 4 * np.spacing(1.0), (key, q)
            mism += d.size
            checked += K
            off += K
    assert checked > 1000
    assert mism == 0
    eng.close()
    assert _lib.BZ_F64 == 8
