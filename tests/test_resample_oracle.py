"""Pins of the oracle's binning and ES-WPNC resampling: brute force on the
keys, conservation invariants, and the stratified-selection bounds."""
import math

import numpy as np
import pytest

from oracle import oracle as O

KEV = 0.5 * O.M_E * O.J_EV


def _cell_state(counts_per_cell, N, L, dtype, seed=0, drift=0.0, e_ev=6.0):
    """Particles placed inside node cells with per-cell counts; optional drift."""
    rng = np.random.default_rng(seed)
    dx = L / N
    xs, vs, ws = [], [], []
    for c, m in enumerate(counts_per_cell):
        lo = max((c - 0.5) * dx, 1e-6)
        hi = min((c + 0.5) * dx, L - 1e-6)
        xs.append(rng.uniform(lo, hi, m))
        sv = math.sqrt(e_ev / KEV / 3)
        v = rng.normal(0, sv, size=(3, m))
        v[0] += drift
        vs.append(v)
        ws.append(rng.uniform(0.5, 2.0, m))
    x = np.concatenate(xs)
    v = np.concatenate(vs, axis=1)
    w = np.concatenate(ws)
    p = rng.permutation(x.size)
    return tuple(np.ascontiguousarray(a[p], dtype=dtype) for a in (x, v[0], v[1], v[2], w))


def _keys_numpy(st, N, L, nb, thr):
    x, vx, vy, vz, w = [a.astype(np.float64) for a in st]
    dx = L / N
    cell = np.clip(((x + 0.5 * dx) / dx).astype(int), 0, N)
    e = KEV * (st[1].astype(np.float64) ** 2 + st[2].astype(np.float64) ** 2 +
               st[3].astype(np.float64) ** 2)
    b = np.zeros(x.size, dtype=int)
    for t in thr:
        b += e >= t
    key = cell * nb + b
    key[~(w > 0)] = (N + 1) * nb
    return key


@pytest.mark.parametrize("nb,thr", [(1, ()), (3, (10.0, 15.76)), (5, (5.0, 10.0, 15.76, 20.0))])
def test_bin_sort_brute_force(nb, thr):
    N, L = 20, 2.5
    st = _cell_state([300] * (N + 1), N, L, np.float64, seed=nb, e_ev=12.0)
    w = st[4].copy()
    w[::17] = 0.0
    st = st[:4] + (w,)
    s = O.Setup(dim=1, dt=1e-11, n_cells=N, length_cm=L)
    counts, offsets, perm = O.bin_sort(s, nb, thr, st)
    key = _keys_numpy(st, N, L, nb, thr)
    np.testing.assert_array_equal(counts, np.bincount(key, minlength=(N + 1) * nb + 1))
    np.testing.assert_array_equal(perm, np.argsort(key, kind="stable"))
    assert offsets[0] == 0 and offsets[-1] == key.size


POL = dict(split_below=0.5, merge_at=2.0, merge_keep=0.6, trigger_at=2.0, min_wt=0.1,
           split_to=2.0)


def _segment_sums(st, key, n_keys):
    x, vx, vy, vz, w = [a.astype(np.float64) for a in st]
    W = np.bincount(key, weights=w, minlength=n_keys + 1)
    P = [np.bincount(key, weights=w * v, minlength=n_keys + 1) for v in (vx, vy, vz)]
    E = np.bincount(key, weights=w * (vx * vx + vy * vy + vz * vz), minlength=n_keys + 1)
    return W, P, E


@pytest.mark.parametrize("drift", [0.0, 3e5])
def test_resample_conserves_per_segment(drift):
    """Merge and split conserve weight, momentum and energy of every
    (cell, bin) segment to 1e-12 (FP64); counts follow the policy."""
    N, L, target = 10, 2.5, 100
    cells = [30, 400, 100, 260, 10, 1, 0, 500, 90, 45, 220]
    st = _cell_state(cells, N, L, np.float64, seed=4, drift=drift)
    s = O.Setup(dim=1, dt=1e-11, n_cells=N, length_cm=L)
    nb = 1
    targets = np.full(N + 1, target, dtype=np.int32)
    pol = O.policy(**POL)
    out, cnt = O.resample(s, pol, nb, (), targets, 7, 3, st)
    assert cnt[10] == 1  # fired
    k0 = _keys_numpy(st, N, L, nb, ())
    k1 = _keys_numpy(out, N, L, nb, ())
    assert np.all(np.diff(k1) >= 0)  # output in key order
    W0, P0, E0 = _segment_sums(st, k0, N + 1)
    W1, P1, E1 = _segment_sums(out, k1, N + 1)
    np.testing.assert_allclose(W1, W0, rtol=1e-12)
    for a, b in zip(P1, P0):
        scale = np.sqrt(W0 * E0) + 1e-300
        assert np.all(np.abs(a - b) <= 1e-12 * scale)
    np.testing.assert_allclose(E1, E0, rtol=1e-12)
    n1 = np.bincount(k1, minlength=N + 2)[:N + 1]
    for c, m, got in zip(range(N + 1), cells, n1):
        if m >= 2 * target:
            assert got == int(0.6 * target), c
        elif 0 < m < 0.5 * target:
            assert got == int(2 * target), c
        else:
            assert got == m, c


def test_reference_rescale_loses_energy_with_drift():
    """The reference rescale sqrt(T/T_hat) from total energies
    (boltzmann.py:838,861-869) conserves energy only without drift; the
    central-moment rescale conserves it with drift."""
    N, L = 2, 2.5
    st = _cell_state([0, 600, 0], N, L, np.float64, seed=9, drift=8e5)
    s = O.Setup(dim=1, dt=1e-11, n_cells=N, length_cm=L)
    targets = np.full(N + 1, 100, dtype=np.int32)
    E0 = np.sum(st[4] * (st[1] ** 2 + st[2] ** 2 + st[3] ** 2))
    errs = {}
    for ref in (0, 1):
        out, _ = O.resample(s, O.policy(**POL, ref_rescale=ref), 1, (), targets, 1, 1, st)
        E1 = np.sum(out[4] * (out[1] ** 2 + out[2] ** 2 + out[3] ** 2))
        errs[ref] = abs(E1 - E0) / E0
    assert errs[0] < 1e-12
    assert errs[1] > 1e-6


def test_stratified_selection_bounds():
    """Stratified sampling picks particle j between floor(K w_j/W)-1 and
    ceil(K w_j/W)+1 times (each stratum holds one draw)."""
    N, L, target = 1, 2.5, 1000
    st = _cell_state([5000, 0], N, L, np.float64, seed=2)
    w = np.exp(np.random.default_rng(3).normal(0, 1.0, st[0].size))
    # distinct speeds identify the selected parents
    sp = np.arange(1, w.size + 1) * 1.0
    st = (st[0], sp, np.zeros_like(sp), np.zeros_like(sp), w)
    s = O.Setup(dim=1, dt=1e-11, n_cells=N, length_cm=L)
    out, _ = O.resample(s, O.policy(**{**POL, "ref_rescale": 0}), 1, (),
                        np.full(2, target, np.int32), 5, 0, st)
    K = int(0.6 * target)
    assert out[0].size == K
    # parents from x (positions are copied unchanged)
    parent = {float(xv): j for j, xv in enumerate(st[0])}
    sel = np.array([parent[float(xv)] for xv in out[0]])
    cnt = np.bincount(sel, minlength=w.size)
    expect = K * w / w.sum()
    assert np.all(cnt >= np.floor(expect) - 1) and np.all(cnt <= np.ceil(expect) + 1)
    assert np.all(np.diff(sel) >= 0)  # strata in order


def test_split_weights_and_copies():
    N, L, target = 1, 2.5, 200
    st = _cell_state([40, 0], N, L, np.float64, seed=6)
    w = st[4].copy()
    w[:5] = 0.05  # below min_wt: never split
    st = st[:4] + (w,)
    s = O.Setup(dim=1, dt=1e-11, n_cells=N, length_cm=L)
    out, cnt = O.resample(s, O.policy(**{**POL, "trigger_at": 0.0}), 1, (),
                          np.full(2, target, np.int32), 3, 0, st)
    assert out[0].size == 2 * target
    # each original is followed by its identical clones, weight divided by copies+1
    order = np.argsort(_keys_numpy(st, N, L, 1, ()), kind="stable")
    j = 0
    for i in order:
        reps = 1
        while j + reps < out[0].size and out[0][j + reps] == st[0][i] and \
                out[1][j + reps] == st[1][i]:
            reps += 1
        assert out[0][j] == st[0][i]
        np.testing.assert_allclose(out[4][j:j + reps], st[4][i] / reps, rtol=1e-15)
        if st[4][i] <= 0.1:
            assert reps == 1
        j += reps
    assert j == out[0].size


def test_reference_trigger_quirk():
    """ref_trigger: only the last bin's max >= 3*target fires (the final
    if/else of _check_for_resampling overrides the earlier tests)."""
    N, L = 4, 2.5
    st = _cell_state([10, 500, 10, 10, 10], N, L, np.float64, seed=1, e_ev=3.0)
    s = O.Setup(dim=1, dt=1e-11, n_cells=N, length_cm=L)
    nb, thr = 3, (10.0, 15.76)
    targets = np.tile(np.array([70, 20, 10], np.int32), N + 1)
    _, cnt_ref = O.resample(s, O.policy(**POL, ref_trigger=1), nb, thr, targets, 1, 0, st)
    _, cnt_new = O.resample(s, O.policy(**POL), nb, thr, targets, 1, 0, st)
    assert cnt_ref[10] == 0  # low-energy crowding alone does not fire the reference check
    assert cnt_new[10] == 1


def test_split_draws_follow_the_weights_chi2():
    """The split's categorical draws land on particle j with probability
    w_j / sum(w) over w > min_wt, and never on a light particle
    (cp.random.choice(HW_ids, ..., p=w/sum(w)), 1bin/boltzmann.py:809-814):
    copies per particle pooled over 40 passes (steps) against the expected
    counts, chi-square.  A uniform choice would fail it."""
    from scipy import stats
    N, L, target = 1, 2.5, 4000
    st = _cell_state([120, 0], N, L, np.float64, seed=31)
    w = np.random.default_rng(8).choice([0.05, 0.3, 1.0, 3.0, 9.0], st[4].size)
    st = st[:4] + (w,)
    s = O.Setup(dim=1, dt=1e-11, n_cells=N, length_cm=L)
    pol = O.policy(**{**POL, "trigger_at": 0.0})
    tg = np.full(2, target, np.int32)
    n = w.size
    copies = np.zeros(n)
    passes = 40
    for step in range(passes):
        out, cnt = O.resample_inplace(s, pol, 1, (), tg, 5, step, st)
        assert cnt[8] == 1
        # in place: slot j keeps w / (copies + 1)
        copies += np.rint(w / out[4][:n]) - 1
    draws = (2 * target - n) * passes
    heavy = w > 0.1
    assert np.all(copies[~heavy] == 0)
    p = w[heavy] / w[heavy].sum()
    chi2, pv = stats.chisquare(copies[heavy], draws * p)
    assert pv > 1e-3, (chi2, pv)
    # the same counts against a uniform choice are rejected
    assert stats.chisquare(copies[heavy], np.full(p.size, draws / p.size)).pvalue < 1e-6


def test_inplace_layout_is_the_reference_layout():
    """In-place resampling leaves untouched segments where they are, keeps a
    split particle in its slot with w/(c+1) followed (at the end) by its
    clones, and turns merged members into holes with the K merged particles
    appended; after the compaction the live particles are those of the
    key-order resampler, as a multiset, with equal totals."""
    N, L, nb, thr = 6, 2.5, 3, (10.0, 15.76)
    st = _cell_state([30, 500, 900, 60, 10, 700, 5], N, L, np.float64, seed=2, e_ev=8.0)
    s = O.Setup(dim=1, dt=1e-11, n_cells=N, length_cm=L)
    targets = np.tile(np.array([210, 60, 30], np.int32), N + 1)
    pol = O.policy(**{**POL, "trigger_at": 0.0})
    ip, ci = O.resample_inplace(s, pol, nb, thr, targets, 4, 1, st)
    ko, ck = O.resample(s, pol, nb, thr, targets, 4, 1, st)
    assert np.array_equal(ci, ck) and ci[8] > 0 and ci[9] > 0
    n = st[0].size
    key = _keys_numpy(st, N, L, nb, thr)
    counts = np.bincount(key, minlength=(N + 1) * nb + 1)
    for j in range(n):
        k = key[j]
        t = targets[k] if k < targets.size else 0
        c = counts[k]
        if t and c >= POL["merge_at"] * t:
            assert ip[4][j] == 0.0  # merged member: a hole
        elif t and 0 < c < 0.5 * t:
            assert ip[4][j] > 0 and ip[4][j] <= st[4][j]  # split in place
        else:
            for a, b in zip(ip, st):
                assert a[j] == b[j]  # untouched
    live = ip[4] > 0
    a = np.stack([v[live] for v in ip]).T
    b = np.stack(list(ko)).T
    assert a.shape == b.shape
    # the same rows: lexicographic sort of both
    np.testing.assert_array_equal(a[np.lexsort(a.T[::-1])], b[np.lexsort(b.T[::-1])])


def test_fill_holes_brute_force():
    """Hole filling against a pair-by-pair construction on small inputs: the
    live rows are kept as a multiset, the first L slots are all live, live
    rows below L never move, and the k-th hole takes the k-th live row above
    L; no holes and all holes included."""
    rng = np.random.default_rng(4)
    for n, p_hole in ((0, 0.3), (1, 1.0), (7, 0.0), (50, 0.3), (200, 0.9), (333, 0.05)):
        w = np.where(rng.uniform(size=n) < p_hole, 0.0, rng.uniform(0.5, 2.0, n))
        st = tuple(rng.normal(size=n) for _ in range(4)) + (w,)
        out = O.fill_holes(st)
        L = int((w > 0).sum())
        assert out[4].size == L and np.all(out[4] > 0)
        rows = list(zip(*st))
        want = list(rows[:L])
        tail = [r for r in rows[L:] if r[4] > 0]
        k = 0
        for j in range(L):
            if not (rows[j][4] > 0):
                want[j] = tail[k]
                k += 1
        assert k == len(tail)
        got = list(zip(*out))
        assert got == want
        assert sorted(r[4] for r in got) == sorted(r[4] for r in rows if r[4] > 0)
