"""Pins of the CPU oracle against closed forms, invariants, published
known-answer vectors and brute force (the reference ships no tests or
fixtures for this path)."""
import math

import numpy as np
import pytest

from oracle import oracle as O

# Random123 known-answer vectors for philox4x32_10 (kat_vectors file of the
# Random123 distribution, Salmon et al. SC'11).
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_kat(ctr, key, expect):
    assert O.philox(ctr, key) == expect


def _u53(a, b):
    return ((a >> 5) * 67108864 + (b >> 6)) / 2.0 ** 53


def test_collision_uniform_layout():
    """Column 0 (the decision uniform) from the purpose-5 draw shared by the
    particles 1024 slots apart (4 in FP32, 2 in FP64), columns 1,2,3,5 from
    the particle's own purpose-0 draws; particles chosen across group and
    word boundaries."""
    seed, step = 0x1234_5678_9ABC, 7
    idx = [0, 1, 1023, 1024, 2047, 2048, 3071, 3072, 4095, 4096, 5121, 9000]
    n = max(idx) + 1
    R64 = O.collision_uniforms(seed, step, n, np.float64)
    R32 = O.collision_uniforms(seed, step, n, np.float32)
    key = (seed & 0xFFFFFFFF, seed >> 32)
    for i in idx:
        b = [O.philox((i, blk, step, 0), key) for blk in range(2)]
        g2, h = ((i >> 11) << 10) | (i & 1023), (i >> 10) & 1
        d2 = O.philox((g2, 0, step, 5), key)
        want64 = [_u53(d2[2 * h], d2[2 * h + 1]), _u53(b[0][0], b[0][1]), _u53(b[0][2], b[0][3]),
                  _u53(b[1][0], b[1][1]), _u53(b[1][2], b[1][3])]
        assert list(R64[[0, 1, 2, 3, 5], i]) == want64
        g4, q = ((i >> 12) << 10) | (i & 1023), (i >> 10) & 3
        d4 = O.philox((g4, 0, step, 5), key)
        want32 = [(d4[q] >> 8) / 2.0 ** 24] + [(b[0][k] >> 8) / 2.0 ** 24 for k in range(4)]
        assert list(R32[[0, 1, 2, 3, 5], i].astype(np.float64)) == want32
    assert np.all(R64 >= 0) and np.all(R64 < 1)
    # the decision uniforms of a group's members are different words: no two
    # of 4096 consecutive FP32 particles repeat one
    assert np.unique(R32[0, :4096]).size > 4090


def test_uniform_moments():
    R = O.collision_uniforms(99, 3, 200_000, np.float64)[0]
    assert abs(R.mean() - 0.5) < 4 * math.sqrt(1 / 12 / R.size)
    assert abs(R.var() - 1 / 12) < 5e-3 / 12


# ---------------------------------------------------------------------------
def _state(n, dtype, dim, L=2.5, e_ev=5.0, seed=1):
    rng = np.random.default_rng(seed)
    sv = math.sqrt(e_ev * 1.602e-19 / 9.10938e-31 / 3)
    v = rng.normal(0, sv, size=(3, n))
    x = rng.uniform(0.01 * L, 0.99 * L, n) if dim == 1 else np.full(n, 0.5 * L)
    w = rng.uniform(0.5, 2.0, n)
    return tuple(np.ascontiguousarray(a, dtype=dtype) for a in (x, v[0], v[1], v[2], w))


def _table(dtype, gas=3.2e22):
    from boltzmann_amd import physics as ph
    t = ph.build_table(ph.XSecModel(), gas, dtype)
    return t, (t.energy, t.sig, t.slope)


KEV = 0.5 * O.M_E * O.J_EV


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_free_push_closed_form(dtype, order):
    """R0 = 1-eps: no collision.  Reference order (0): x' = x + 100 dt v,
    vx' = vx - dt q/m 100 E[int(x/dx)]; leapfrog (1): vx' first, then
    x' = x + 100 dt vx'."""
    n, N, L, dt = 2000, 64, 2.5, 2e-11
    t, tab = _table(dtype)
    s = O.Setup(dim=1, dt=dt, n_cells=N, length_cm=L, nu_max=t.nu_max, gas_density_m3=3.2e22,
                push_order=order)
    st = _state(n, dtype, 1, L)
    E = np.linspace(-50, 80, N).astype(dtype)
    R = np.full((7, n), 0.9999, dtype=dtype)
    out, mom, cnt = O.push_step(s, st, E, R, table=tab)
    x, vx, vy, vz, w = [a.astype(np.float64) for a in st]
    cell = np.clip((x / (L / N)).astype(int), 0, N - 1)
    vxn = vx - dt * (O.Q_E / O.M_E) * (100 * E.astype(np.float64)[cell])
    xn = x + 100 * dt * (vxn if order == 1 else vx)
    keep = (xn > 0) & (xn < L)
    assert out[0].size == keep.sum()
    assert cnt[1] == (~keep).sum() and cnt[5] == n and cnt[0] == 0
    rt = 1e-12 if dtype == np.float64 else 2e-6
    np.testing.assert_allclose(out[0], xn[keep], rtol=rt)
    kick = np.abs(dt * (O.Q_E / O.M_E) * (100 * E.astype(np.float64)[cell]))[keep]
    assert np.all(np.abs(out[1] - vxn[keep]) <= rt * (np.abs(vx[keep]) + kick) + 1e-9)
    np.testing.assert_array_equal(out[2], st[2][keep])
    np.testing.assert_array_equal(out[4], st[4][keep])


@pytest.mark.parametrize("order", [0, 1])
def test_uniform_field_trajectory(order):
    """k collisionless steps in a uniform field a: v_k = v0 - k a dt and
    x_k = x0 + 100 dt (k v0 - a dt S), S = k(k-1)/2 for the reference order
    (position first) and k(k+1)/2 for the leapfrog (velocity first)."""
    n, N, L, dt, k = 64, 16, 2.5, 1e-12, 25
    s = O.Setup(dim=1, dt=dt, n_cells=N, length_cm=L, gas_density_m3=1e-30,
                nu_max=1e-30, push_order=order)
    rng = np.random.default_rng(3)
    x0 = rng.uniform(1.0, 1.5, n)
    v0 = rng.normal(0, 1e5, n)
    st = (x0.copy(), v0.copy(), np.zeros(n), np.zeros(n), np.ones(n))
    E = np.full(N, 7.0)
    t = _table(np.float64, gas=1e-30)[1]
    R = np.full((7, n), 0.9999)
    for _ in range(k):
        st, _, _ = O.push_step(s, st, E, R, table=t)
    a = (O.Q_E / O.M_E) * 100 * 7.0
    S = k * (k - 1) / 2 if order == 0 else k * (k + 1) / 2
    np.testing.assert_allclose(st[1], v0 - k * a * dt, rtol=1e-12, atol=1e-6)
    np.testing.assert_allclose(st[0], x0 + 100 * dt * (k * v0 - a * dt * S), rtol=1e-12)


@pytest.mark.parametrize("xs_mode", [1, 2])
def test_births_last_layout(xs_mode):
    """The 1-D layout (survivors in slot order, then every secondary in
    parent order, as the reference's electron_kernel appends them at Nc and
    its compaction keeps them after the survivors) is a reordering of the
    parent-adjacent layout: survivors are the records not preceded by an
    ionizing parent's record."""
    n = 6000
    t, tab = _table(np.float64, gas=3.2e24)
    kw = dict(dim=1, dt=2e-12, xs_mode=xs_mode, n_cells=16, length_cm=2.5, nu_max=t.nu_max,
              gas_density_m3=3.2e24)
    rng = np.random.default_rng(4)
    sp = math.sqrt(40.0 / KEV)
    d = rng.normal(size=(3, n))
    d /= np.linalg.norm(d, axis=0)
    st = (rng.uniform(0.2, 2.3, n), *(sp * d), rng.uniform(0.5, 2, n))
    E = np.linspace(-30, 30, 16)
    R = rng.uniform(0, 1, size=(7, n))
    a, ma, ca = O.push_step(O.Setup(**kw, births_last=False), st, E, R, table=tab)
    b, mb, cb = O.push_step(O.Setup(**kw, births_last=True), st, E, R, table=tab)
    assert np.array_equal(ca, cb) and cb[0] > 50
    np.testing.assert_array_equal(ma, mb)
    births = int(cb[0])
    n_surv = a[0].size - births
    # secondaries start at the parent's pre-push (or collision) x with its
    # weight; in the adjacent layout they are the records whose weight and
    # position repeat the previous parent's, so split by the births_last
    # layout instead: its tail is exactly the births
    tail = np.stack([v[n_surv:] for v in b])
    head = np.stack([v[:n_surv] for v in b])
    A = np.stack(a)
    # every record of the births_last layout appears in the adjacent layout
    # in the same relative order within each group
    def subseq(sub, full):
        j = 0
        for col in range(full.shape[1]):
            if j < sub.shape[1] and np.array_equal(full[:, col], sub[:, j]):
                j += 1
        return j == sub.shape[1]
    assert subseq(head, A) and subseq(tail, A)
    assert head.shape[1] + tail.shape[1] == A.shape[1]


@pytest.mark.parametrize("gas", [3.2e21, 3.2e23])
def test_direct_collision_probability(gas):
    """Direct model (xs_mode 3): a mono-energetic beam at 20 eV collides
    with probability 1 - exp(-nu_t dt), nu_t = N v (s_el + s_exc + s_ion),
    split between channels in proportion to the frequencies, also where
    nu_t dt is of order one (there a single null-collision draw with the
    table's nu_max would undercount).  Uniforms on a regular grid make the
    counts exact up to one per channel."""
    from boltzmann_amd import physics as ph
    n, dt, e0 = 100_000, 2e-11, 20.0
    t, tab = _table(np.float64, gas=gas)
    s = O.Setup(dim=0, dt=dt, xs_mode=3, gas_density_m3=gas, nu_max=t.nu_max)
    sp = math.sqrt(e0 / KEV)
    d = np.random.default_rng(2).normal(size=(3, n))
    d /= np.linalg.norm(d, axis=0)
    st = (np.zeros(n), *(sp * d), np.ones(n))
    R = np.random.default_rng(3).uniform(0, 1, size=(7, n))
    R[0] = (np.arange(n) + 0.5) / n
    _, _, cnt = O.push_step(s, st, None, R, table=tab, capacity=2 * n)
    sig = ph.XSecModel().sigma(np.array([e0]))[:, 0] * gas
    nu = sig * sp
    p = -math.expm1(-nu.sum() * dt)
    want = n * p * nu / nu.sum()
    got = [int(cnt[2]), int(cnt[3]), int(cnt[4])]
    for g, w in zip(got, want):
        assert abs(g - w) <= 2 + 1e-4 * w, (got, want)
    # the null-collision draw with the table's nu_max would give n * nu dt-ish
    # counts capped by p_null: far off when nu_max dt is large
    assert p > 0.01


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_elastic_scatter_geometry(dtype):
    """R0 = 0 forces the first channel: speed scales by sqrt(1-2 m_e/M),
    cos(chi) = 2 R1 - 1 and phi = 2 pi R2 (electron_kernel_numba.py:197-209)."""
    n = 3000
    t, tab = _table(dtype)
    s = O.Setup(dim=0, dt=1e-12, nu_max=t.nu_max, gas_density_m3=3.2e22, e_field0=0.0)
    st = _state(n, dtype, 0)
    rng = np.random.default_rng(5)
    R = rng.uniform(0, 1, size=(7, n)).astype(dtype)
    R[0] = 0.0
    out, mom, cnt = O.push_step(s, st, None, R, table=tab)
    assert cnt[2] == n and out[0].size == n
    v0 = np.sqrt(sum(a.astype(np.float64) ** 2 for a in st[1:4]))
    v1 = np.sqrt(sum(a.astype(np.float64) ** 2 for a in out[1:4]))
    rt = 1e-13 if dtype == np.float64 else 1e-6
    np.testing.assert_allclose(v1, v0 * math.sqrt(1 - 2 * 1.36e-5), rtol=rt)
    np.testing.assert_allclose(out[3] / v1, 2 * R[1].astype(np.float64) - 1, atol=1e-6)
    phi = np.mod(np.arctan2(out[2].astype(np.float64), out[1].astype(np.float64)), 2 * np.pi)
    d = np.abs(phi - 2 * np.pi * R[2].astype(np.float64))
    d = np.minimum(d, 2 * np.pi - d)
    sin_chi = np.sqrt(np.maximum(0, 1 - (2 * R[1].astype(np.float64) - 1) ** 2))
    ok = sin_chi > 1e-3
    assert np.all(d[ok] < 1e-4)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_channel_energy_balance(dtype):
    """At 40 eV with random R: every output has the energy of exactly one
    channel; ionization pairs sit together, share (e - 15.76)/2 each, and
    the secondary keeps the parent's pre-push x and weight."""
    n = 20000
    t, tab = _table(dtype, gas=3.2e24)
    s = O.Setup(dim=1, dt=1e-12, n_cells=16, length_cm=2.5, nu_max=t.nu_max,
                gas_density_m3=3.2e24, births_last=False)
    rng = np.random.default_rng(11)
    sp = math.sqrt(40.0 / KEV)
    dirs = rng.normal(size=(3, n))
    dirs /= np.linalg.norm(dirs, axis=0)
    st = (rng.uniform(0.5, 2.0, n), *(sp * dirs), rng.uniform(0.5, 2, n))
    st = tuple(np.ascontiguousarray(a, dtype=dtype) for a in st)
    E = np.zeros(16, dtype=dtype)
    R = rng.uniform(0, 1, size=(7, n)).astype(dtype)
    out, mom, cnt = O.push_step(s, st, E, R, table=tab)
    e0 = KEV * np.sum(np.stack(st[1:4]).astype(np.float64) ** 2, axis=0)
    e1 = KEV * np.sum(np.stack(out[1:4]).astype(np.float64) ** 2, axis=0)
    rt = 1e-9 if dtype == np.float64 else 1e-5
    # walk outputs in canonical order
    j = 0
    counts = dict(el=0, ex=0, io=0, none=0)
    for i in range(n):
        e = e0[i]
        f = e1[j]
        if abs(f - e * (1 - 2 * 1.36e-5)) <= rt * e and abs(f - e) > rt * e:
            counts["el"] += 1
            j += 1
        elif abs(f - e) <= rt * e:
            counts["none"] += 1
            j += 1
        elif abs(f - (e - 11.55)) <= rt * e:
            counts["ex"] += 1
            j += 1
        elif abs(f - 0.5 * (e - 15.76)) <= rt * e:
            counts["io"] += 1
            assert abs(e1[j + 1] - 0.5 * (e - 15.76)) <= rt * e
            assert out[0][j + 1] == st[0][i] and out[4][j + 1] == st[4][i]
            j += 2
        else:
            raise AssertionError(f"particle {i}: energy {f} matches no channel (from {e})")
    assert j == out[0].size
    assert counts["el"] == cnt[2] and counts["ex"] == cnt[3] and counts["io"] == cnt[4]
    assert cnt[0] == cnt[4]
    # channel frequencies follow P_null * nu_c / nu_max (binomial bounds)
    sig = np.array([np.interp(40.0, t.energy.astype(np.float64), t.sig[c].astype(np.float64))
                    for c in range(3)])
    nu = sig * sp
    pn = 1 - math.exp(-t.nu_max * 1e-12)
    for c, key in enumerate(("el", "ex", "io")):
        p = pn * nu[c] / t.nu_max
        assert abs(counts[key] - n * p) < 5 * math.sqrt(n * p * (1 - p)) + 1


def test_fit_cross_sections_threshold():
    """Reference fit model: the elastic decision flips exactly at
    R0 = P sigma_el / sigma_t computed from electron_kernel_numba.py:115-195."""
    n = 400
    rng = np.random.default_rng(3)
    e = np.exp(rng.uniform(np.log(0.5), np.log(14.0), n))  # below ionization
    sp = np.sqrt(e / KEV)
    st = (np.full(n, 1.0), sp, np.zeros(n), np.zeros(n), np.ones(n))
    nn, dt = 3.22e21, 1e-11
    # elastic fit, as the reference writes it
    a = [-0.02704763, -0.08847237, -0.37608274, -1.4874467, -0.9870356, 14.28063581,
         -2.12069169, -0.3585636, 1.25262128, -2.28332905, -1.47508661, -0.11090525]
    b = [-0.23720051, -0.6084786, -1.8167778, -4.82420619, -4.2206026, 18.92275458,
         0.7555105, 0.79246666, -0.40002029, 4.89566076, 2.82263476, -0.77506081]
    c = [-19.67900951, -20.24111992, -21.51414308, -23.55745478, -23.44715988, -14.70556113,
         -19.71943671, -19.71260274, -19.48909003, -21.46789173, -20.1928456, -17.81752307]
    brk = [0.000780396, 0.006098667, 0.0476269, 0.1015574, 0.1943527, 0.2995781, 1.097156,
           2.339524, 4.988691, 10.63765, 22.68322]
    ee = KEV * sp * sp
    seg = np.searchsorted(brk, ee, side="right")
    L = np.log10(ee)
    lg = (np.take(a, seg) * L + np.take(b, seg)) * L + np.take(c, seg)
    sig = 10.0 ** lg * nn
    P = 1 - np.exp(-dt * np.sqrt(sp * sp) * sig)
    thr = sig * (P * (1.0 / sig))
    s = O.Setup(dim=0, dt=dt, xs_mode=0, gas_density_m3=nn)
    for factor, expect_el in ((1 - 1e-9, True), (1 + 1e-9, False)):
        R = np.zeros((7, n))
        R[0] = thr * factor
        R[1:] = 0.3
        out, mom, cnt = O.push_step(s, st, None, R)
        assert cnt[2] == (n if expect_el else 0)


def test_deposit_brute_force():
    """Node deposit = histogram over int(((x+dx/2)/(L+dx)) * bnf*(N+1)) // bnf."""
    n, N, L, bnf = 5000, 37, 2.54, 4
    st = _state(n, np.float64, 1, L, seed=8)
    s = O.Setup(dim=1, dt=1e-12, n_cells=N, length_cm=L, bnf=bnf)
    mom = O.deposit(s, st)
    x, vx, vy, vz, w = st
    dx = L / N
    fine = (((x + 0.5 * dx) / (L + dx)) * (bnf * (N + 1))).astype(int)
    node = fine // bnf
    e = KEV * (vx * vx + vy * vy + vz * vz)
    ne = np.bincount(node, weights=w, minlength=N + 1)
    nrg = np.bincount(node, weights=w * e, minlength=N + 1)
    np.testing.assert_allclose(mom[0], ne, rtol=1e-13)
    np.testing.assert_allclose(mom[2], nrg, rtol=1e-13)
    assert np.all(mom[1] == 0)
    # nearest-node rule agrees with int((x + dx/2)/dx) away from rounding ties
    near = ((x + 0.5 * dx) / dx).astype(int)
    assert np.mean(near == node) > 0.999


def test_threaded_push_matches_sequential():
    n = 50_000
    t, tab = _table(np.float64)
    st = _state(n, np.float64, 1, 2.5, seed=21)
    E = np.linspace(-100, 100, 32)
    R = O.collision_uniforms(5, 0, n, np.float64)
    s1 = O.Setup(dim=1, dt=5e-11, n_cells=32, length_cm=2.5, nu_max=t.nu_max,
                 gas_density_m3=3.2e22, nthreads=1)
    s4 = O.Setup(**{**s1.__dict__, "nthreads": 4})
    a, ma, ca = O.push_step(s1, st, E, R, table=tab)
    b, mb, cb = O.push_step(s4, st, E, R, table=tab)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)
    np.testing.assert_allclose(ma, mb, rtol=1e-13)
    np.testing.assert_array_equal(ca, cb)


# ---------------------------------------------------------------------------
def test_poisson_dense_and_manufactured():
    """Oracle Poisson vs a dense solve of (1/dx^2) tridiag(-1,2,-1) V = rhs,
    and the closed-form parabola of a uniform charge between grounded ends."""
    from scipy.linalg import cholesky_banded
    N, L = 50, 2.54
    dx = L / N
    Ac = np.zeros((2, N - 1))
    Ac[0, 1:] = -1.0 / dx ** 2
    Ac[1, :] = 2.0 / dx ** 2
    c = cholesky_banded(Ac)
    chol = np.stack([c[1], np.concatenate([[0.0], c[0, 1:]])])
    rng = np.random.default_rng(2)
    ne = rng.uniform(0.5e9, 1.5e9, N + 1)
    ni = rng.uniform(0.5e9, 1.5e9, N + 1)
    mom = np.stack([ne, np.zeros(N + 1), np.zeros(N + 1)]).ravel()
    eps, v0, f, t = 5.5263e5, 100.0, 13.56e6, 1.3e-8
    out = O.field_step(N, dx, 1e-11, eps, v0, f, 1, 1, 1, 0, 1, t, mom, ni.copy(),
                       np.zeros(N + 1), np.zeros(N + 1), np.zeros(N + 1), np.zeros(N - 1),
                       np.zeros(N), chol)
    V, E = out[4], out[5]
    A = (np.diag(2 * np.ones(N - 1)) - np.diag(np.ones(N - 2), 1) - np.diag(np.ones(N - 2), -1)) / dx ** 2
    rhs = (ni[1:N] - ne[1:N]) / eps
    vl = v0 * math.sin(2 * math.pi * f * t)
    rhs[0] += vl / dx ** 2
    np.testing.assert_allclose(V, np.linalg.solve(A, rhs), rtol=1e-9, atol=1e-9 * np.abs(V).max())
    Vf = np.concatenate([[vl], V, [0.0]])
    np.testing.assert_allclose(E, -(Vf[1:] - Vf[:-1]) / dx, rtol=1e-12, atol=1e-9)
    # uniform charge rho, grounded ends: V(x) = rho/(2 eps) x (L - x)
    rho = 1e8
    ni2 = np.full(N + 1, rho)
    mom2 = np.zeros(3 * (N + 1))
    out2 = O.field_step(N, dx, 1e-11, eps, 0.0, f, 1, 1, 1, 0, 1, 0.0, mom2, ni2, np.zeros(N + 1),
                        np.zeros(N + 1), np.zeros(N + 1), np.zeros(N - 1), np.zeros(N), chol)
    xs = dx * np.arange(1, N)
    np.testing.assert_allclose(out2[4], rho / (2 * eps) * xs * (L - xs), rtol=1e-9)


def test_fluid_mass_balance():
    """Explicit drift-diffusion: interior mass changes only by the edge fluxes."""
    from boltzmann_amd import physics as ph
    N, L = 40, 2.5
    dx = L / N
    rng = np.random.default_rng(4)
    ni = rng.uniform(0.9e9, 1.1e9, N + 1)
    E = rng.uniform(-30, 30, N)
    chol = ph.poisson_cholesky(N, dx)
    mu, Di, Ds, dt = 14.4, 0.64, 0.75, 1e-9
    ni_b = ni.copy()
    ni_b[0] = ni_b[1] * ((Di / dx - 0.5 * mu * max(E[0], 0)) / (Di / dx + 0.5 * mu * max(E[0], 0)))
    ni_b[N] = ni_b[N - 1] * ((Di / dx + 0.5 * mu * min(E[N - 1], 0)) / (Di / dx - 0.5 * mu * min(E[N - 1], 0)))
    J = mu * 0.5 * (ni_b[:-1] + ni_b[1:]) * E - Di * (ni_b[1:] - ni_b[:-1]) / dx
    out = O.field_step(N, dx, dt, 5.5263e5, 0.0, 1.0, mu, Di, Ds, 1, 0, 0.0, np.zeros(3 * (N + 1)),
                       ni.copy(), np.zeros(N + 1), np.zeros(N + 1), np.zeros(N + 1),
                       np.zeros(N - 1), E.copy(), chol)
    ni2 = out[0]
    d_mass = np.sum(ni2[1:N] - ni_b[1:N]) * dx
    mass = np.sum(np.abs(ni_b)) * dx
    assert abs(d_mass - (-dt * (J[N - 1] - J[0]))) <= 1e-12 * mass
    # and the interior update is the conservative difference of those fluxes
    np.testing.assert_allclose(ni2[1:N], ni_b[1:N] - dt * (J[1:] - J[:-1]) / dx, rtol=1e-12)


def test_moment_update():
    N = 8
    ne = np.arange(1, N + 2, dtype=float)
    nsrc = np.linspace(0, 1, N + 1)
    nrg = 3 * ne
    nrg[0] = 0
    ne[0] = 0
    mom = np.concatenate([ne, nsrc, nrg])
    ni = np.ones(N + 1)
    g2 = np.zeros(N + 1)
    out = O.field_step(N, 0.1, 1e-12, 1, 0, 1, 1, 1, 1, 0, 0, 0.0, mom, ni, np.zeros(N + 1),
                       np.zeros(N + 1), g2, np.zeros(N - 1), np.zeros(N), np.zeros(2 * (N - 1)))
    np.testing.assert_allclose(out[0], 1 + nsrc)
    assert out[2][0] == 0.0  # nan_to_num
    np.testing.assert_allclose(out[2][1:], 2.0)
    np.testing.assert_allclose(out[3], nsrc)


def _heavies(N, dx, dt, mu, Di, Ds, ni, E, scheme, ns=None, max_sub=0):
    out = O.field_step(N, dx, dt, 5.5263e5, 0.0, 1.0, mu, Di, Ds, 1, 0, 0.0,
                       np.zeros(3 * (N + 1)), ni.copy(),
                       np.zeros(N + 1) if ns is None else ns.copy(), np.zeros(N + 1),
                       np.zeros(N + 1), np.zeros(N - 1), E.copy(), np.zeros(2 * (N - 1)),
                       ion_scheme=scheme, max_substeps=max_sub)
    return out[0], out[1]


def test_sg_heavies_reduce_to_reference_without_field():
    """E = 0: B(0) = 1, the Scharfetter-Gummel flux is the central diffusion
    flux, and one sub-step suffices, so both schemes agree."""
    N, dx = 50, 0.05
    rng = np.random.default_rng(7)
    ni = rng.uniform(0.5, 1.5, N + 1) * 1e9
    ns = rng.uniform(0.5, 1.5, N + 1) * 1e8
    E = np.zeros(N)
    a, sa = _heavies(N, dx, 1e-9, 14.4, 0.64, 0.75, ni, E, 0, ns)
    b, sb = _heavies(N, dx, 1e-9, 14.4, 0.64, 0.75, ni, E, 1, ns)
    np.testing.assert_allclose(b, a, rtol=1e-13)
    np.testing.assert_allclose(sb, sa, rtol=1e-13)


def test_sg_heavies_positive_where_central_scheme_fails():
    """Cell Peclet number 50: the reference's central update produces
    negative densities; the sub-cycled Scharfetter-Gummel update stays
    non-negative and conserves the interior mass when both walls push ions
    inwards (zero wall flux)."""
    N, dx, mu, Di = 64, 0.01, 1.44e4, 643.0
    x = np.arange(N) / (N - 1)
    E = 500.0 * (1.0 - 2.0 * x)                   # E[0] > 0, E[N-1] < 0
    ni = np.full(N + 1, 1e9)
    ni[:N // 4] = 1e6                               # a steep front facing the drift
    dt = 2e-10
    c, _ = _heavies(N, dx, dt, mu, Di, 1.0, ni, E, 0)
    assert c[1:N].min() < 0
    s, _ = _heavies(N, dx, dt, mu, Di, 1.0, ni, E, 1)
    assert s.min() >= 0
    m0 = ni[1:N].sum()
    assert abs(s[1:N].sum() - m0) <= 1e-12 * m0


def test_sg_heavies_drift_speed():
    """Uniform field, Pe >> 1: a pulse drifts at mu E (upwind limit) and its
    mass is conserved while it is far from the walls."""
    N, dx, mu, Di, E0 = 400, 0.01, 1.44e4, 1.0, 50.0
    xc = (np.arange(N + 1)) * dx
    ni = np.exp(-((xc - 1.0) / 0.1) ** 2) * 1e9
    E = np.full(N, E0)
    dt = 0.2 * dx / (mu * E0)                       # 0.2 cells per step
    n = ni.copy()
    steps = 200
    for _ in range(steps):
        n, _ = _heavies(N, dx, dt, mu, Di, 1.0, n, E, 1)
    shift = (np.sum(xc * n) / np.sum(n)) - (np.sum(xc * ni) / np.sum(ni))
    assert abs(shift / (mu * E0 * dt * steps) - 1) < 1e-3
    assert abs(n[1:N].sum() / ni[1:N].sum() - 1) < 1e-12


def test_sg_heavies_substep_count():
    """The number of sub-steps follows the explicit bound
    h (a_i + b_{i-1}) / dx <= 1: a field step twice the bound equals two
    manual half steps."""
    N, dx, mu, Di = 32, 0.01, 1.44e4, 643.0
    rng = np.random.default_rng(9)
    E = rng.uniform(-300, 300, N)
    ni = rng.uniform(0.5, 1.5, N + 1) * 1e9
    Bf = lambda v: np.where(np.abs(v) < 1e-4, 1 - v / 2 + v * v / 12, v / np.expm1(v))
    pe = mu * E * dx / Di
    a = Di / dx * Bf(-pe)
    b = Di / dx * Bf(pe)
    rate = max(np.max((a[1:] + b[:-1]) / dx), 2 * 1.0 / dx ** 2)
    dt = 2.0 / rate * 0.999                         # exactly two sub-steps
    full, _ = _heavies(N, dx, dt, mu, Di, 1.0, ni, E, 1)
    half, _ = _heavies(N, dx, dt / 2, mu, Di, 1.0, ni, E, 1)
    half, _ = _heavies(N, dx, dt / 2, mu, Di, 1.0, half, E, 1)
    np.testing.assert_allclose(full, half, rtol=1e-13)


def _flight_setup(dim, dt, gas, n_cells=1, e_field0=0.0, dtype=np.float64):
    from boltzmann_amd import physics as ph
    tab = ph.build_table(ph.XSecModel(), gas, dtype)
    s = O.Setup(dim=dim, dt=dt, xs_mode=2, n_cells=n_cells, length_cm=2.5,
                gas_density_m3=gas, nu_max=tab.nu_max, e_field0=e_field0, seed=77, step=3)
    return s, (tab.energy, tab.sig, tab.slope), tab


def test_free_flight_elastic_frequency_is_exact_for_long_steps():
    """Mono-energetic 5 eV electrons (below every inelastic threshold), no
    field: the number of real (elastic) collisions per particle and step is
    Poisson(N sigma v dt) whatever nu_max dt is -- here nu_max dt ~ 9 and
    nu dt = 2, where one null-collision draw per step would cap it at 1."""
    n = 20000
    KEV = 0.5 * O.M_E * O.J_EV
    v = np.sqrt(5.0 / KEV)
    gas = 3.22e21
    nu = gas * 1e-19 * v
    dt = 2.0 / nu
    s, tab, t = _flight_setup(0, dt, gas)
    assert t.nu_max * dt > 5
    d = np.random.default_rng(3).normal(size=(3, n))
    d /= np.linalg.norm(d, axis=0)
    st = (np.zeros(n), *(v * d), np.ones(n))
    out, mom, cnt = O.push_step(s, st, None, None, table=tab, capacity=2 * n)
    mean = cnt[2] / n
    assert abs(mean / 2.0 - 1) < 0.015, mean          # 0.5% statistical error
    # each elastic event scales the speed by sqrt(1 - 2 m/M) exactly
    fel = math.sqrt(1 - 2 * 1.36e-5)
    sp = np.sqrt(out[1] ** 2 + out[2] ** 2 + out[3] ** 2)
    k = np.log(sp / v) / np.log(fel)
    assert np.all(np.abs(k - np.round(k)) < 1e-6)
    assert int(np.round(k).sum()) == cnt[2]
    assert cnt[4] == 0 and cnt[3] == 0


def test_free_flight_ballistic_without_gas():
    """Negligible gas: one flight spans the step, x' = x + 100 dt vx and
    vx' = vx - dt (q/m) 100 E, the single-collision step's push."""
    n = 1000
    s, tab, _ = _flight_setup(1, 1e-10, 1e3, n_cells=10)
    rng = np.random.default_rng(4)
    x = rng.uniform(1.0, 1.5, n)
    v = rng.normal(0, 1e5, (3, n))
    E = np.full(10, 50.0)
    out, _, cnt = O.push_step(s, (x, *v, np.ones(n)), E, None, table=tab, capacity=2 * n)
    assert cnt[2] + cnt[3] + cnt[4] == 0
    np.testing.assert_allclose(out[0], x + 100 * 1e-10 * v[0], rtol=1e-13)
    a = (1e-10 * (O.Q_E / O.M_E)) / 1e-10
    np.testing.assert_allclose(out[1], v[0] - (a * (100 * 50.0)) * 1e-10, rtol=1e-12)


def test_free_flight_reduces_to_null_collision_rates():
    """Hot electrons with ionization: event totals match the single-draw
    null-collision step when nu_max dt is small (both sample the same rates)."""
    n = 40000
    KEV = 0.5 * O.M_E * O.J_EV
    rng = np.random.default_rng(5)
    e = rng.uniform(20, 60, n)
    d = rng.normal(size=(3, n))
    d /= np.linalg.norm(d, axis=0)
    v = np.sqrt(e / KEV) * d
    gas = 3.22e21
    s, tab, t = _flight_setup(0, 1e-12, gas)
    st = (np.zeros(n), *v, np.ones(n))
    _, _, c_ff = O.push_step(s, st, None, None, table=tab, capacity=2 * n)
    s1 = O.Setup(dim=0, dt=1e-12, xs_mode=1, gas_density_m3=gas, nu_max=t.nu_max)
    _, _, c_nc = O.push_step(s1, st, None, O.collision_uniforms(5, 0, n, np.float64), table=tab,
                             capacity=2 * n)
    for k in (2, 3, 4):
        a, b = int(c_ff[k]), int(c_nc[k])
        assert abs(a - b) <= 5 * math.sqrt(a + b) + 5, (k, a, b)


# The reference's argon fits as electron_kernel_numba.py:67-171 writes them:
# log10 sigma = (a L + b) L + c, L = log10 e, one (a, b, c) per interval.
REF_EL = [(-0.02704763, -0.23720051, -19.67900951), (-0.08847237, -0.6084786, -20.24111992),
          (-0.37608274, -1.8167778, -21.51414308), (-1.4874467, -4.82420619, -23.55745478),
          (-0.9870356, -4.2206026, -23.44715988), (14.28063581, 18.92275458, -14.70556113),
          (-2.12069169, 0.7555105, -19.71943671), (-0.3585636, 0.79246666, -19.71260274),
          (1.25262128, -0.40002029, -19.48909003), (-2.28332905, 4.89566076, -21.46789173),
          (-1.47508661, 2.82263476, -20.1928456), (-0.11090525, -0.77506081, -17.81752307)]
REF_EL_BRK = [0.000780396, 0.006098667, 0.0476269, 0.1015574, 0.1943527, 0.2995781, 1.097156,
              2.339524, 4.988691, 10.63765, 22.68322]
REF_ION = [(-2149239.99337822, 5150451.29014456, -3085665.64874862),
           (-227229.83440255, 545171.14839945, -327016.69508374),
           (-20173.68563579, 48599.13254342, -29290.96100603),
           (-1838.90888364, 4491.58838596, -2763.82866252),
           (-186.98923927, 477.04784782, -324.75732578),
           (-24.48843901, 69.43965607, -69.145292411),
           (-4.33101714, 14.47124519, -31.66224458),
           (-0.68345667, 2.6053539, -21.99823511)]
REF_ION_BRK = [15.77742, 15.81964, 15.96196, 16.44162, 18.05837, 23.50771, 41.87498]
E_ION = 15.76


def _ref_fit(e):
    """numpy restatement of the interval sums (electron_kernel_numba.py:115-174):
    each interval's indicator times its polynomial, then 10**; ionization
    zero below 15.76 eV and 10**0 = 1 at exactly 15.76 eV, where no interval
    holds (the first starts strictly above it)."""
    e = np.asarray(e, dtype=np.float64)
    L = np.log10(e)
    lg0 = np.zeros_like(e)
    edges = [-np.inf] + REF_EL_BRK + [np.inf]
    for k, (a, b, c) in enumerate(REF_EL):
        ind = (edges[k] <= e) & (e < edges[k + 1]) if k else (e < edges[1])
        lg0 = lg0 + ind * ((a * L + b) * L + c)
    lg2 = np.zeros_like(e)
    iedges = [E_ION] + REF_ION_BRK + [np.inf]
    for k, (a, b, c) in enumerate(REF_ION):
        ind = (iedges[k] < e) & (e < iedges[k + 1]) if k == 0 else \
            (iedges[k] <= e) & (e < iedges[k + 1])
        lg2 = lg2 + ind * ((a * L + b) * L + c)
    s0 = 10.0 ** lg0
    s2 = np.where(e < E_ION, 0.0, 10.0 ** lg2)
    return s0, s2


def test_fit_cross_sections_all_intervals():
    """The oracle's fits against the restatement over 1e-4 .. 1000 eV: every
    elastic interval (12) and ionization interval (8), both sides of every
    breakpoint, and the reference's sigma_ion = 1 m^2 at e = 15.76 eV exactly
    (quirk Q13)."""
    nn = 3.22e21
    rng = np.random.default_rng(17)
    pts = [np.exp(rng.uniform(np.log(1e-4), np.log(1e3), 4000))]
    for b in REF_EL_BRK + REF_ION_BRK:
        pts.append(np.array([b * (1 - 1e-12), b, b * (1 + 1e-12), b * 0.99, b * 1.01]))
    pts.append(np.array([E_ION, np.nextafter(E_ION, 0), np.nextafter(E_ION, 100), 15.765]))
    e = np.concatenate(pts)
    s = O.Setup(dim=0, dt=1e-12, xs_mode=0, gas_density_m3=nn)
    s_el, s_ion = O.fit_sigma(s, e)
    r0, r2 = _ref_fit(e)
    np.testing.assert_allclose(s_el, r0 * nn, rtol=1e-12)
    np.testing.assert_allclose(s_ion, r2 * nn, rtol=1e-12, atol=0)
    # every interval was visited
    assert len(np.unique(np.searchsorted(REF_EL_BRK, e, side="right"))) == 12
    above = e[e > E_ION]
    assert len(np.unique(np.searchsorted(REF_ION_BRK, above, side="right"))) == 8
    # Q13 and the threshold
    q = O.fit_sigma(s, np.array([E_ION, np.nextafter(E_ION, 0)]))[1]
    assert q[0] == nn and q[1] == 0.0
    # a wrong coefficient anywhere would show: perturb one interval of the
    # restatement and the comparison fails there
    bad = list(REF_ION)
    bad[3] = (bad[3][0], bad[3][1] * (1 + 1e-7), bad[3][2])
    saved = REF_ION[:]
    REF_ION[:] = bad
    try:
        assert not np.allclose(_ref_fit(e)[1] * nn, s_ion, rtol=1e-12)
    finally:
        REF_ION[:] = saved


def test_fit_cross_sections_fp32_follow_fp64():
    """FP32 fits (the polynomial in FP32, 10** in FP32) stay within FP32
    rounding of the FP64 restatement, interval by interval."""
    nn = 3.22e21
    e = np.exp(np.random.default_rng(3).uniform(np.log(1e-3), np.log(900.0), 3000))
    s = O.Setup(dim=0, dt=1e-12, xs_mode=0, gas_density_m3=nn)
    e32 = e.astype(np.float32)
    a, b = O.fit_sigma(s, e32)
    ed = e32.astype(np.float64)
    r0, r2 = _ref_fit(ed)
    L = np.abs(np.log10(ed))

    def bound(coef, brk, lo):
        # FP32 rounding of the polynomial: a few ulps of its largest term,
        # times ln 10 in sigma (the ionization fits cancel terms of ~1e6)
        k = np.searchsorted(brk, ed, side="right")
        t = np.array([abs(coef[j][0]) * L[i] ** 2 + abs(coef[j][1]) * L[i] + abs(coef[j][2])
                      for i, j in enumerate(np.minimum(k, len(coef) - 1))])
        return 4 * math.log(10) * 2.0 ** -24 * t + 1e-6

    assert np.all(np.abs(a / nn / r0 - 1) <= bound(REF_EL, REF_EL_BRK, None))
    m = r2 > 0
    assert np.all(np.abs(b[m] / nn / r2[m] - 1) <= bound(REF_ION, REF_ION_BRK, None)[m])


def test_gas_velocity_draw_is_normal():
    """The gas velocities of the Philox stream (Box-Muller, purpose 1): each
    component N(0, sigma^2) (Kolmogorov-Smirnov), uncorrelated with the
    others, independent of the collision stream, and deterministic in
    (seed, step, index).  The reference draws multivariate_normal(0, Eg*I)
    (1bin/boltzmann.py:589)."""
    from scipy import stats
    n, sigma = 200_000, 1.5e4
    G = O.gas_velocities(12345, 7, n, np.float64, sigma)
    for c in range(3):
        assert abs(G[c].mean()) < 5 * sigma / math.sqrt(n)
        assert abs(G[c].var() / sigma ** 2 - 1) < 5 * math.sqrt(2 / n)
        assert stats.kstest(G[c] / sigma, "norm").pvalue > 1e-3
    cc = np.corrcoef(G)
    assert np.all(np.abs(cc - np.eye(3)) < 5 / math.sqrt(n))
    R = O.collision_uniforms(12345, 7, n, np.float64)
    assert abs(np.corrcoef(G[0], R[0])[0, 1]) < 5 / math.sqrt(n)
    np.testing.assert_array_equal(G, O.gas_velocities(12345, 7, n, np.float64, sigma))
    assert not np.array_equal(G, O.gas_velocities(12345, 8, n, np.float64, sigma))
    G32 = O.gas_velocities(12345, 7, n, np.float32, sigma)
    assert abs(G32[2].var() / sigma ** 2 - 1) < 5 * math.sqrt(2 / n)


def test_dead_ionizer_flag_at_node_0():
    """ref_flags 1: a primary that ionizes and leaves in the same step adds
    its unscaled flag 1 to ni_src at node 0 (the reference reduces every row,
    the dead one parked at x = 1e-8, post_processing_1D.py:48-55).  Particles
    packed against the right wall move out at 120 eV with no field: without
    a collision they leave.  The ones that ionized are the births born there;
    an ionized primary keeps e_ej = (120 - 15.76)/2 eV in a random direction,
    so the survivors at the wall with that energy are the ionizers that
    stayed, and the rest left: exactly that many units appear at node 0."""
    n, L, N = 4000, 2.5, 25
    rng = np.random.default_rng(12)
    x = np.concatenate([rng.uniform(L - 1e-6, L - 1e-7, n // 2), rng.uniform(0.5, 2.0, n // 2)])
    sp = math.sqrt(120.0 / KEV)
    vx = np.concatenate([np.full(n // 2, sp), np.full(n // 2, -1e3)])
    st = (x, vx, np.zeros(n), np.zeros(n), np.full(n, 0.7))
    R = O.collision_uniforms(3, 0, n, np.float64)
    s0 = O.Setup(dim=1, dt=1e-10, xs_mode=0, n_cells=N, length_cm=L, gas_density_m3=3e23)
    s1 = O.Setup(dim=1, dt=1e-10, xs_mode=0, n_cells=N, length_cm=L, gas_density_m3=3e23,
                 ref_flags=1)
    out0, m0, c0 = O.push_step(s0, st, np.zeros(N), R)
    out1, m1, c1 = O.push_step(s1, st, np.zeros(N), R)
    nb_ = int(c1[0])
    births_at_wall = int(np.sum(out1[0][out1[0].size - nb_:] > L - 1e-6))
    surv = tuple(a[:out1[0].size - nb_] for a in out1)
    e_surv = KEV * (surv[1] ** 2 + surv[2] ** 2 + surv[3] ** 2)
    e_ej = 0.5 * (KEV * sp * sp - 15.76)
    stayed = int(np.sum((surv[0] > L - 0.1) & (np.abs(e_surv / e_ej - 1) < 1e-9)))
    assert births_at_wall > 10 and stayed > 10
    assert m1[1][0] - m0[1][0] == births_at_wall - stayed
    m1[1][0] = m0[1][0]
    np.testing.assert_array_equal(m0, m1)
    for a, b in zip(out0, out1):
        np.testing.assert_array_equal(a, b)


def test_energy_roundtrip_binning_moves_threshold_energies():
    """ref_flags 2: the binning energy is (e*w)/w, the value the reference's
    tosum[:,3] holds when it bins (post_processing_1D.py:53,62).  An energy
    one ulp below the 10 eV threshold rounds up to 10 for some weights:
    those particles move to the next bin, and the keys follow the
    round-tripped energy exactly."""
    N, L, nb, thr = 4, 2.5, 3, (10.0, 15.76)
    below = np.nextafter(10.0, 0.0)
    v = math.sqrt(10.0 / KEV)
    for _ in range(200):  # the speed whose energy is one ulp below 10 eV
        if KEV * v * v == below:
            break
        v = np.nextafter(v, 0.0 if KEV * v * v > below else 1e9)
    assert KEV * v * v == below
    rng = np.random.default_rng(4)
    m = 4000
    w = rng.uniform(0.1, 50.0, m)
    st = (np.full(m, 1.0), np.full(m, v), np.zeros(m), np.zeros(m), w)
    s0 = O.Setup(dim=1, dt=1e-11, n_cells=N, length_cm=L)
    s2 = O.Setup(dim=1, dt=1e-11, n_cells=N, length_cm=L, ref_flags=2)
    c0, _, _ = O.bin_sort(s0, nb, thr, st)
    c2, _, _ = O.bin_sort(s2, nb, thr, st)
    node = int((1.0 + 0.5 * L / N) / (L / N))
    er = (below * w) / w
    flips = int(np.sum(er >= 10.0))
    assert flips > 10  # about 1 % of the weights
    assert c0[node * nb] == m and c0[node * nb + 1] == 0
    assert c2[node * nb] == m - flips and c2[node * nb + 1] == flips


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_exponential_within_two_ulp_of_libm(dtype):
    """The oracle's own e^v - 1 and e^v (the device evaluates the same
    operations, so the collision probabilities agree bit for bit) against
    numpy's, over the probabilities' range v = -nu dt in [-60, 0] and a few
    positive values: within 2 ulp, exact at 0."""
    rng = np.random.default_rng(5)
    v = np.concatenate([-np.logspace(-9, np.log10(60.0), 20000), rng.uniform(-1, 1, 5000),
                        np.linspace(0, 20, 500), [0.0, -np.log(2.0), np.log(2.0) / 2]])
    v = v.astype(dtype)
    em1, ex = O.expm1_exp(v)
    assert em1[v == 0].max() == 0.0 and ex[v == 0].min() == 1.0
    v64 = v.astype(np.float64)
    for got, want in ((em1, np.expm1(v64)), (ex, np.exp(v64))):
        ulp = np.spacing(np.abs(want).astype(dtype)).astype(np.float64)
        err = np.abs(got.astype(np.float64) - want) / ulp
        assert err.max() <= 2.0, (dtype, err.max(), v[np.argmax(err)])


@pytest.mark.parametrize("xs_mode", [1, 3, 2])
def test_collision_counters_sum_to_the_channel_counts(xs_mode):
    """The collision counters of the oracle's push (the reference's collct
    deposit): with unit weights the elastic and excitation rows sum to the
    step's elastic and excitation event counts and the ionization row to
    the ionizations; with every particle at one node the counters sit at
    that node, or at node 0 for the absorbed."""
    n, N, L = 40_000, 32, 2.5
    t, tab = _table(np.float64, gas=3.2e23)
    for spread in (True, False):
        rng = np.random.default_rng(7)
        x = rng.uniform(0.05 * L, 0.95 * L, n) if spread else np.full(n, 10.3 * L / N)
        sv = math.sqrt(20.0 / KEV / 3)
        st = (x, *rng.normal(0, sv, size=(3, n)), np.ones(n))
        s = O.Setup(dim=1, dt=2e-11, xs_mode=xs_mode, n_cells=N, length_cm=L, bnf=4,
                    gas_density_m3=3.2e23, nu_max=t.nu_max)
        R = None if xs_mode == 2 else O.collision_uniforms(3, 0, n, np.float64)
        E = np.zeros(N)
        _, _, cnt, coll = O.push_step(s, st, E, R, table=tab, collisions=True)
        assert coll[0].sum() == cnt[2] and coll[1].sum() == cnt[3] and coll[2].sum() == cnt[4]
        assert cnt[2] > 100 and cnt[4] > 10
        if not spread:
            # the electrons move less than a cell: node 10 (x = 10.3 dx)
            assert set(np.flatnonzero(coll.sum(0))) <= {0, 10}
