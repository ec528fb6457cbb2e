"""The reference-surface stage functions (boltzmann_amd.kernels) and the
BoltzmannDSMC driver, called the way the reference's run() calls them."""
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

KEV = 0.5 * O.M_E * O.J_EV


def test_reduction_brute_force(gpu):
    import torch
    from boltzmann_amd.kernels import reduction as red
    rng = np.random.default_rng(0)
    RM, bnf, r, ptcl = 301, 4, 4, 200_000
    M = bnf * RM
    dx = rng.uniform(0, 1, ptcl)
    dv = rng.normal(size=(ptcl, r))
    dv[rng.uniform(size=ptcl) < 0.1] = 0.0
    dev = torch.device("cuda", 0)
    dw = torch.full((M, r), 7.0, dtype=torch.float64, device=dev)
    red.reduction(r, 0, M, RM, dw, torch.from_numpy(dv).to(dev), torch.from_numpy(dx).to(dev),
                  ptcl)
    k = (dx * M).astype(int) // bnf
    want = np.stack([np.bincount(k, weights=dv[:, j], minlength=RM) for j in range(r)], axis=1)
    got = dw.cpu().numpy()
    np.testing.assert_allclose(got[:RM], want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    assert np.all(got[RM:] == 0.0)
    # signed fixed-point sums: the same bits on every call (the reference's
    # FP64 atomics are not reproducible)
    dw2 = torch.full((M, r), -3.0, dtype=torch.float64, device=dev)
    red.reduction(r, 0, M, RM, dw2, torch.from_numpy(dv).to(dev), torch.from_numpy(dx).to(dev),
                  ptcl)
    np.testing.assert_array_equal(dw2.cpu().numpy(), got)


class _CAI:
    """A minimal device array of another library: only
    __cuda_array_interface__ (version 3) over a torch tensor's memory."""

    def __init__(self, t):
        self._t = t  # keeps the memory alive
        typestr = {"torch.float64": "<f8", "torch.int64": "<i8", "torch.float32": "<f4"}
        self.__cuda_array_interface__ = {
            "shape": tuple(t.shape), "typestr": typestr[str(t.dtype)],
            "data": (t.data_ptr(), False), "version": 3, "strides": None}


@pytest.mark.parametrize("wrap", [False, True], ids=["torch", "cuda_array_interface"])
def test_electron_kernel_and_post_processing(gpu, wrap):
    """electron_kernel (reference fits, FP64) + post_processing against the
    oracle fed the same uniforms and gas velocities, on torch tensors and on
    foreign arrays that only expose __cuda_array_interface__ (used in place:
    the results appear in the memory they describe); the collision counters
    equal the oracle's."""
    import torch
    from boltzmann_amd import kernels as K
    rng = np.random.default_rng(1)
    N, n, cap = 199, 40_000, 120_000
    dev = torch.device("cuda", 0)
    e = np.exp(rng.uniform(np.log(0.1), np.log(60.0), n))
    d = rng.normal(size=(3, n))
    d /= np.linalg.norm(d, axis=0)
    v = np.sqrt(e / KEV) * d
    data = np.zeros((cap, 6))
    tosum = np.zeros((cap, 4))
    data[:n, 0] = rng.uniform(0.001, 2.539, n)
    data[:n, 3:6] = v.T
    tosum[:n, 0] = rng.uniform(1e3, 2e3, n)
    gas = rng.normal(0, 6.7e4, size=(cap, 3))
    E = rng.uniform(-60, 60, N)
    d_data = torch.from_numpy(data).to(dev)
    d_tosum = torch.from_numpy(tosum).to(dev)
    d_R = torch.zeros((cap, 7), dtype=torch.float64, device=dev)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    nn_gas = 1e6 * 3.22e16 * 0.1
    dt = (1 / 13.56e6) / 1e5
    K.set_deposit_bins(4)
    w_ = _CAI if wrap else (lambda t: t)
    K.electron_kernel(n, w_(d_data), w_(d_tosum), None, w_(torch.from_numpy(E).to(dev)),
                      w_(d_R), w_(d_R), None, None, nn_gas, 1, dt, w_(cnt), None, None,
                      w_(torch.from_numpy(gas).to(dev)), 1024, 64)
    nni = int(cnt[0])
    RN = N + 1
    ne, nrg = np.zeros(RN), np.zeros(RN)
    cg = [np.zeros(RN) for _ in range(4)]
    res = K.post_processing(None, None, None, None, None, None, w_(d_data), w_(d_tosum), None,
                            True, n, n, nni, ne, np.zeros(RN), np.zeros(RN), nrg, *cg, 4 * RN, RN,
                            2.54, 2.54 / N, None)
    nnew = res[-1]
    R = d_R[:n].cpu().numpy().T.copy()
    G = gas[:n].T.copy()
    s = O.Setup(dim=1, dt=dt, xs_mode=0, n_cells=N, length_cm=2.54, bnf=4, gas_density_m3=nn_gas)
    st = (data[:n, 0], data[:n, 3], data[:n, 4], data[:n, 5], tosum[:n, 0])
    want, wmom, wcnt, wcoll = O.push_step(s, st, E, R, G=G, capacity=cap, collisions=True)
    assert nni == int(wcnt[0]) and nnew == want[0].size
    # collision counters: elastic sum w, (no excitation in the fits), ionizations
    np.testing.assert_allclose(res[4], wcoll[0], rtol=1e-12, atol=1e-12 * wcoll[0].max())
    assert not res[5].any() and not wcoll[1].any()
    np.testing.assert_array_equal(res[6], wcoll[2])
    assert wcoll[0].sum() > 0  # (ionizations are rare at this step: 0 to a few)
    got = d_data[:nnew].cpu().numpy()
    gw = d_tosum[:nnew].cpu().numpy()
    vs = np.sqrt(want[1] ** 2 + want[2] ** 2 + want[3] ** 2) + 1
    np.testing.assert_allclose(got[:, 0], want[0], rtol=1e-12)
    for k, col in ((1, 3), (2, 4), (3, 5)):
        assert np.all(np.abs(got[:, col] - want[k]) <= 1e-12 * vs)
    np.testing.assert_array_equal(gw[:, 0], want[4])
    np.testing.assert_allclose(res[0], wmom[0], rtol=1e-12, atol=1e-12 * wmom[0].max())
    np.testing.assert_allclose(res[1], wmom[1], rtol=1e-12, atol=1e-12 * wmom[0].max())
    np.testing.assert_allclose(res[3], wmom[2], rtol=1e-12, atol=1e-12 * wmom[2].max())
    # uniforms handed back are the Philox stream of the stage's step (the
    # stage session counts its calls)
    cols = [0, 1, 2, 3, 5]
    assert any(np.array_equal(R[cols], O.collision_uniforms(12345, k, n, np.float64)[cols])
               for k in range(4))


def test_bin_particles_3bin(gpu):
    import torch
    from boltzmann_amd import kernels as K
    from boltzmann_amd import physics as ph
    rng = np.random.default_rng(2)
    N, n, cap = 50, 30_000, 40_000
    dev = torch.device("cuda", 0)
    tosum = np.zeros((cap, 4))
    tosum[:n, 0] = rng.uniform(0.5, 2.0, n)
    tosum[::31, 0] = 0.0
    tosum[:n, 3] = rng.exponential(8.0, n)
    cells = np.zeros(cap, dtype=np.int64)
    cells[:n] = rng.integers(0, N + 1, n)
    maxp = [400, 200, 300]  # HE, ME, LE
    lists = [torch.zeros((m, N + 1), dtype=torch.int64, device=dev) for m in maxp]
    inds = [torch.zeros(N + 1, dtype=torch.int64, device=dev) for _ in maxp]
    K.bin_particles(torch.from_numpy(tosum).to(dev), *lists, *inds,
                    torch.from_numpy(cells).to(dev), n, N, *maxp, 64)
    # brute force in the reference's bin order
    e = tosum[:n, 3].astype(np.float64)
    b = np.where(e < 10.0, 2, np.where(e < 15.76, 1, 0))  # index into (HE, ME, LE)
    alive = tosum[:n, 0] > 0
    for q in range(3):
        for c in range(N + 1):
            ids = np.flatnonzero(alive & (b == q) & (cells[:n] == c))
            assert int(inds[q][c]) == ids.size
            m = min(ids.size, maxp[q])
            np.testing.assert_array_equal(lists[q][:m, c].cpu().numpy(), ids[:m])
    _ = ph


def test_bin_particles_5bin_at_the_thresholds(gpu):
    """Energies exactly at each 5-bin threshold and one ulp on either side
    (the reference's `<` compares on tosum[:,3]: at a threshold the particle
    goes up a bin), a NaN energy (listed nowhere), cells -1 and N+1 (listed
    nowhere), dead rows, a list full before its cell is (write_inds keeps
    the full count); the arrays handed over through
    __cuda_array_interface__ only."""
    import torch
    from boltzmann_amd import kernels as K
    N = 6
    dev = torch.device("cuda", 0)
    thr = [5.0, 10.0, 15.76, 20.0]
    es = []
    for t in thr:
        es += [np.nextafter(t, 0.0), t, np.nextafter(t, 100.0)]
    es += [0.0, 1e-300, 3.0, 50.0, np.nan, 7.5, 7.5, 7.5]
    n = len(es)
    tosum = np.zeros((n + 8, 4))
    tosum[:n, 0] = 1.0
    tosum[:n, 3] = es
    tosum[3, 0] = 0.0  # dead
    cells = np.zeros(n + 8, dtype=np.int64)
    cells[:n] = np.arange(n) % (N + 1)
    cells[5] = -1
    cells[6] = N + 1
    cells[n - 3:n] = 2  # with row 2 (just above 5 eV): four ME1 rows in cell 2
    maxp = [3, 1, 3, 3, 3]  # HE, ME1 (full before its cell is), ME2, ME3, LE
    lists = [torch.full((m, N + 1), -7, dtype=torch.int64, device=dev) for m in maxp]
    inds = [torch.zeros(N + 1, dtype=torch.int64, device=dev) for _ in maxp]
    C = _CAI
    K.bin_particles(C(torch.from_numpy(tosum).to(dev)), *[C(a) for a in lists],
                    *[C(a) for a in inds], C(torch.from_numpy(cells).to(dev)), n, N, *maxp, 64)
    e = tosum[:n, 3]
    with np.errstate(invalid="ignore"):
        # index into (HE, ME1, ME2, ME3, LE): LE < 5 <= ME1 < 10 <= ME2 < 15.76 <= ME3 < 20 <= HE
        q = np.select([e < 5.0, e < 10.0, e < 15.76, e < 20.0, e >= 20.0], [4, 1, 2, 3, 0], -1)
    ok = (tosum[:n, 0] > 0) & (cells[:n] >= 0) & (cells[:n] <= N)
    for b in range(5):
        for c in range(N + 1):
            ids = np.flatnonzero(ok & (q == b) & (cells[:n] == c))
            assert int(inds[b][c]) == ids.size, (b, c)
            m = min(ids.size, maxp[b])
            got = lists[b][:, c].cpu().numpy()
            np.testing.assert_array_equal(got[:m], ids[:m])
            assert np.all(got[m:] == -7)  # untouched past the count
    assert int(inds[1][2]) == 4 and int(lists[1][0, 2]) == 2


def test_ef_and_heavies_kernels(gpu):
    from scipy.linalg import cholesky_banded
    from boltzmann_amd import kernels as K
    N, L = 120, 2.54
    dx = L / N
    Ac = np.zeros((2, N - 1))
    Ac[0, 1:] = -1.0 / dx ** 2
    Ac[1, :] = 2.0 / dx ** 2
    c = cholesky_banded(Ac)
    diag = c[1].copy()
    low = np.concatenate([[0.0], c[0, 1:]])
    rng = np.random.default_rng(3)
    ne = rng.uniform(0.5e8, 1e8, N + 1)
    ni = rng.uniform(0.5e8, 1e8, N + 1)
    E = np.zeros(N)
    V = np.zeros(N - 1)
    t = 2.2e-8
    Vo, Eo = K.EF_kernel(E, ne, ni, t, V, np.zeros(N - 1), diag, low, np.zeros(N + 1), N, dx,
                         5.5263e5, 100.0, 13.56e6)
    mom = np.concatenate([ne, np.zeros(N + 1), np.zeros(N + 1)])
    want = O.field_step(N, dx, 1e-11, 5.5263e5, 100.0, 13.56e6, 1, 1, 1, 0, 1, t, mom, ni.copy(),
                        np.zeros(N + 1), np.zeros(N + 1), np.zeros(N + 1), np.zeros(N - 1),
                        np.zeros(N), np.stack([diag, low]))
    np.testing.assert_allclose(Vo, want[4], rtol=1e-10, atol=1e-10 * np.abs(want[4]).max())
    np.testing.assert_allclose(Eo, want[5], rtol=1e-10, atol=1e-10 * np.abs(want[5]).max())
    E2 = rng.uniform(-20, 20, N)
    ni_h, ns_h, flag = K.heavies_kernel_fluid(0, E2, ni.copy(), np.zeros(N + 1), np.zeros(N),
                                              np.zeros(N + 1), np.zeros(N + 1), np.zeros(N), dx,
                                              1e-10, 144.0, 6.4, 7.5, N)
    want = O.field_step(N, dx, 1e-10, 5.5263e5, 0.0, 1.0, 144.0, 6.4, 7.5, 1, 0, 0.0,
                        np.zeros(3 * (N + 1)), ni.copy(), np.zeros(N + 1), np.zeros(N + 1),
                        np.zeros(N + 1), np.zeros(N - 1), E2.copy(), np.zeros(2 * (N - 1)))
    np.testing.assert_allclose(ni_h, want[0], rtol=1e-12)
    assert flag == 0


@pytest.mark.parametrize("nbins", [1, 3, 5])
def test_boltzmann_driver_runs(gpu, nbins, tmp_path):
    from boltzmann_amd.boltzmann import BoltzmannDSMC
    name = str(tmp_path / "run_")
    b = BoltzmannDSMC(30, name, 0, 60_000, 101, 4, 1, False, 10, 1e4, 15, nbins=nbins)
    n0 = b.Nc_list[0]
    b.run(30, name, 0, verbose=False)
    n1 = b.Nc_list[0]
    assert n0 > 0 and n1 > 0
    c = b.eng.counters()
    assert c["resample"] >= 0 and c["overflow"] == 0
    for key, size in (("ne", 101), ("ni", 101), ("Te", 101), ("E", 100), ("ionprod_total", 101)):
        arr = np.fromfile(name + key + "_cycle2.dat")
        assert arr.size == size
    x = np.fromfile(name + "e_x0_cycle2_mpi0.dat")
    assert x.size > 0 and np.all((x > 0) & (x < 2.54))
    assert math.isfinite(float(np.fromfile(name + "Te_CA_cycle1.dat").sum()))


def test_boltzmann_restart_and_quirk_fixes(gpu, tmp_path):
    """A run writes cycle 2; a run restarted from cycle 2 holds exactly the
    dumped particles (w != 0) and ion density, resumes the clock at step
    2 * spc and writes cycle 3.  With fix_quirks the ns average and the
    collision-counter averages are written, the counters are positive where
    electrons collide and the ionization average is a count per step."""
    import torch
    from boltzmann_amd.boltzmann import BoltzmannDSMC
    name = str(tmp_path / "run_")
    spc = 8
    a = BoltzmannDSMC(2 * spc, name, 0, 40_000, 65, 4, 1, False, 5, 1e4, spc,
                      fix_quirks=True)
    a.run(2 * spc, name, 0, verbose=False)
    for key in ("ns", "coll_el", "coll_exc", "coll_ion"):
        arr = np.fromfile(name + key + "_CA_cycle2.dat")
        assert arr.size == 65 and np.all(np.isfinite(arr))
    el = np.fromfile(name + "coll_el_CA_cycle2.dat")
    assert el.sum() > 0
    x = np.fromfile(name + "e_x0_cycle2_mpi0.dat")
    w = np.fromfile(name + "e_wt_cycle2_mpi0.dat")
    ni = np.fromfile(name + "ni_cycle2.dat")
    b = BoltzmannDSMC(spc, name, 2, 40_000, 65, 4, 1, False, 5, 1e4, spc)
    n, step, t, _ = b.eng.clock()
    assert step == 2 * spc and abs(t - 2 * spc * b.dt_big) <= 1e-15 * t
    st = b.eng.fetch_state()
    keep = w != 0
    assert n == int(keep.sum())
    np.testing.assert_array_equal(st[0], x[keep])
    np.testing.assert_array_equal(st[4], w[keep])
    np.testing.assert_array_equal(b.eng.ni.cpu().numpy(), ni)
    torch.cuda.synchronize()
    b.run(spc, name, 2, verbose=False)
    assert np.fromfile(name + "ne_cycle3.dat").size == 65
    assert b.Nc_list[0] > 0


def test_electron_kernel_sub_steps_and_setup_deposit(gpu):
    """post_processing of rows that were not pushed (the set-up deposit,
    boltzmann.py:363-371) against the oracle's compaction and deposit; then
    electron_kernel with dt_ratio = 2 (two pushes of dt/2, the rows compacted
    between them) against the oracle's push twice with its compaction in
    between: rows, births, the second push's moments and the collision
    counters summed over both pushes."""
    import ctypes as C
    import torch
    from boltzmann_amd import _lib
    from boltzmann_amd import kernels as K
    from boltzmann_amd.kernels import _session as S
    rng = np.random.default_rng(5)
    N, n, cap = 63, 20_000, 80_000
    RN = N + 1
    dev = torch.device("cuda", 0)
    e = np.exp(rng.uniform(np.log(0.5), np.log(60.0), n))
    d = rng.normal(size=(3, n))
    d /= np.linalg.norm(d, axis=0)
    v = np.sqrt(e / KEV) * d
    data = np.zeros((cap, 6))
    tosum = np.zeros((cap, 4))
    data[:n, 0] = rng.uniform(0.001, 2.539, n)
    data[:n, 3:6] = v.T
    tosum[:n, 0] = rng.uniform(1e3, 2e3, n)
    tosum[:n:17, 0] = 0.0  # dead rows: compacted away
    E = rng.uniform(-300, 300, N)
    nn_gas = 1e6 * 3.22e16 * 0.1 * 50  # a dense gas: collisions in both sub-steps
    dt = 2.0e-11
    K.set_deposit_bins(4)
    d_data = torch.from_numpy(data).to(dev)
    d_tosum = torch.from_numpy(tosum).to(dev)

    def post(Nc, nni):
        return K.post_processing(None, None, None, None, None, None, d_data, d_tosum, None,
                                 True, Nc, Nc, nni, np.zeros(RN), np.zeros(RN), np.zeros(RN),
                                 np.zeros(RN), *[np.zeros(RN) for _ in range(4)], 4 * RN, RN,
                                 2.54, 2.54 / N, None)

    res = post(n, 0)
    live = tosum[:n, 0] > 0
    s = O.Setup(dim=1, dt=dt / 2, xs_mode=0, n_cells=N, length_cm=2.54, bnf=4,
                gas_density_m3=nn_gas)
    st = tuple(np.ascontiguousarray(a[live]) for a in
               (data[:n, 0], data[:n, 3], data[:n, 4], data[:n, 5], tosum[:n, 0]))
    wdep = O.deposit(s, st)
    m = res[-1]
    assert m == int(live.sum())
    np.testing.assert_allclose(res[0], wdep[0], rtol=1e-12, atol=1e-12 * wdep[0].max())
    np.testing.assert_allclose(res[1], wdep[1], rtol=1e-12, atol=1e-12 * wdep[0].max())
    np.testing.assert_allclose(res[3], wdep[2], rtol=1e-12, atol=1e-12 * wdep[2].max())
    np.testing.assert_array_equal(d_tosum[:m, 0].cpu().numpy(), st[4])
    np.testing.assert_array_equal(d_data[:m, 0].cpu().numpy(), st[0])
    assert not d_tosum[m:n].any()

    # the stage session (shared by the stages of one device) names the
    # Philox step of each push; its collision counters start from zero
    sess = S.session(0, _lib.BZ_F64, cap, RN)
    drained = torch.empty(3 * RN, dtype=torch.float64, device=dev)
    _lib.check(sess.lib.bz_collision_counts(sess.ctx, C.c_void_p(drained.data_ptr()), RN, 1))
    step0 = sess.step
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    K.electron_kernel(m, d_data, d_tosum, None, torch.from_numpy(E).to(dev), None, None, None,
                      None, nn_gas, 2, dt, cnt, None, None, None)
    nni = int(cnt[0])
    cur = st
    births = 0
    coll = np.zeros((3, RN))
    for k in range(2):
        R = O.collision_uniforms(12345, step0 + k, cur[0].size, np.float64)
        cur, wmom, wcnt, wc = O.push_step(s, cur, E, R, capacity=4 * cap, collisions=True)
        births += int(wcnt[0])
        coll += wc
    assert nni == births > 0
    res = post(m, nni)
    want = cur
    assert res[-1] == want[0].size
    got = d_data[:res[-1]].cpu().numpy()
    np.testing.assert_array_equal(d_tosum[:res[-1], 0].cpu().numpy(), want[4])
    vs = np.sqrt(want[1] ** 2 + want[2] ** 2 + want[3] ** 2) + 1
    np.testing.assert_allclose(got[:, 0], want[0], rtol=1e-12, atol=1e-12)
    for kk, col in ((1, 3), (2, 4), (3, 5)):
        assert np.all(np.abs(got[:, col] - want[kk]) <= 1e-12 * vs)
    np.testing.assert_allclose(res[0], wmom[0], rtol=1e-12, atol=1e-12 * wmom[0].max())
    np.testing.assert_allclose(res[3], wmom[2], rtol=1e-12, atol=1e-12 * wmom[2].max())
    np.testing.assert_allclose(res[4], coll[0], rtol=1e-12, atol=1e-12 * coll[0].max())
    np.testing.assert_array_equal(res[6], coll[2])
    assert coll[0].sum() > 0
