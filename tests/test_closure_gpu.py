"""The closure on the grid: source kernel against the FP64 oracle network,
the closure-driven field sub-step and the kinetic commit against the field
oracle, and the trained-vs-untrained ablation on swarm data."""
import math

import numpy as np
import pytest

from oracle import mlp as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _grid(N=200, seed=0):
    rng = np.random.default_rng(seed)
    E = rng.uniform(-800, 800, N)
    ne = rng.uniform(1e8, 1e10, N + 1)
    ne[3] = 0.0                                  # an empty node
    nsrc = rng.uniform(0, 1e5, N + 1)
    nrg = ne * rng.uniform(1, 12, N + 1)
    return E, np.concatenate([ne, nsrc, nrg])


def _net():
    from boltzmann_amd.closure import ClosureNet
    net = ClosureNet(seed=4)
    net.norm = {"in_shift": np.array([1.5, 6.0, 16.0]), "in_scale": np.array([0.9, 0.12, 0.6]),
                "out_shift": np.array([-17.0, -16.0, 24.0]),
                "out_scale": np.array([2.0, 1.5, 0.4]),
                # a trained domain narrower than the test grid's features: clamped
                "in_lo": np.array([1.0, 2.0, 15.5]), "in_hi": np.array([2.0, 10.0, 17.5])}
    return net


def _want_src(net, E, mom, N, gas, dt):
    nn = N + 1
    en = np.empty(nn)
    en[0], en[N] = abs(E[0]), abs(E[N - 1])
    en[1:N] = 0.5 * (np.abs(E[:-1]) + np.abs(E[1:]))
    en_td = np.maximum(en * 100.0 / gas * 1e21, 1e-3)
    ne = mom[:nn]
    me = np.where(ne > 0, mom[2 * nn:] / np.where(ne > 0, ne, 1), 0.0)
    lne = np.log10(np.maximum(ne * 1e6, 1.0))
    x = np.stack([np.log10(en_td), me, lne], 1)
    lo, hi = net.norm.get("in_lo", np.zeros(3)), net.norm.get("in_hi", np.zeros(3))
    for k in range(3):
        if lo[k] < hi[k]:
            x[:, k] = np.clip(x[:, k], lo[k], hi[k])
    y = M.forward(M.bf16_value(net.bf16_bits()), net.norm, x)
    return 10.0 ** y[:, 0] * gas * ne * dt


def test_closure_source_matches_oracle_network(gpu):
    import ctypes as C
    import torch
    from boltzmann_amd import _lib
    net = _net()
    N, gas, dt = 200, 3.22e21, 1.8e-10
    E, mom = _grid(N)
    Ed = torch.from_numpy(E).cuda()
    md = torch.from_numpy(mom).cuda()
    src = torch.zeros(N + 1, dtype=torch.float64, device="cuda")
    acc = torch.full((N + 1,), 1.0, dtype=torch.float64, device="cuda")
    s = net.struct()
    p = lambda t: C.c_void_p(t.data_ptr())
    for _ in range(2):
        _lib.check(net.lib.bz_closure_source(net.ctx, C.byref(s), N, p(Ed), p(md), gas, dt,
                                             p(src), p(acc)))
    want = _want_src(net, E, mom, N, gas, dt)
    got = src.cpu().numpy()
    # FP32 network: the log-rate is good to ~1e-5 of its scale, 10^y amplifies by ln 10
    np.testing.assert_allclose(got, want, rtol=2e-4 * 2.0 * math.log(10), atol=0)
    np.testing.assert_allclose(acc.cpu().numpy(), 1.0 + 2 * got, rtol=1e-12)
    assert got[3] == 0.0
    net.close()


def test_field_substep_and_commit(gpu):
    import torch
    from boltzmann_amd import configs
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine
    import ctypes as C
    N = 128
    cfg = configs.c1().with_(n_cells=N, n_particles=1000)
    eng = KineticEngine(cfg, capacity=4096)
    E, mom = _grid(N, seed=1)
    src = np.random.default_rng(2).uniform(0, 1e6, N + 1)
    ni0 = np.random.default_rng(3).uniform(0.9e9, 1.1e9, N + 1)
    eng.moments.copy_(torch.from_numpy(mom))
    eng.E.copy_(torch.from_numpy(E))
    eng.ni.copy_(torch.from_numpy(ni0))
    eng.load_state(tuple(np.zeros(1) for _ in range(5)), t=2e-9)
    sd = torch.from_numpy(src).cuda()
    p = lambda t: C.c_void_p(t.data_ptr())
    from boltzmann_amd import _lib
    _lib.check(eng.lib.bz_field_substep(eng.ctx, C.byref(eng.fp), p(eng.moments), p(sd),
                                        p(eng.ni), p(eng.ns), p(eng.Te), p(eng.g2), p(eng.V),
                                        p(eng.E), p(eng.E_t), p(eng.chol)))
    fp = eng.fp
    m2 = mom.copy()
    m2[:N + 1] += src
    from scipy.linalg import cholesky_banded
    Ac = np.zeros((2, N - 1))
    Ac[0, 1:] = -1.0 / cfg.dx ** 2
    Ac[1, :] = 2.0 / cfg.dx ** 2
    c = cholesky_banded(Ac)
    chol = np.stack([c[1], np.concatenate([[0.0], c[0, 1:]])])
    want = O.field_step(N, fp.dx, fp.dt, fp.eps, fp.v0, fp.freq, fp.mu_i, fp.D_i, fp.D_s, 1, 1,
                        2e-9 + cfg.dt, np.concatenate([m2[:N + 1], src, m2[2 * N + 2:]]),
                        ni0, np.zeros(N + 1), np.zeros(N + 1), np.zeros(N + 1),
                        np.zeros(N - 1), E.copy(), chol, ion_scheme=fp.ion_scheme,
                        max_substeps=fp.max_substeps)
    for nm, w in zip(("ni", "V", "E"), (want[0], want[4], want[5])):
        g = getattr(eng, nm).cpu().numpy()
        np.testing.assert_allclose(g, w, rtol=1e-10, atol=1e-10 * np.abs(w).max(), err_msg=nm)
    np.testing.assert_allclose(eng.moments.cpu().numpy()[:N + 1], m2[:N + 1], rtol=1e-15)
    # commit: ni += nsrc - acc; tally sums
    acc = torch.from_numpy(src * 0.5).cuda()
    tally = torch.zeros(8, dtype=torch.float64, device="cuda")
    ni1 = eng.ni.cpu().numpy()
    _lib.check(eng.lib.bz_closure_commit(eng.ctx, N + 1, p(eng.moments), p(acc), p(eng.ni),
                                         p(eng.g2), p(eng.Te), p(tally)))
    nsrc = mom[N + 1:2 * N + 2]
    np.testing.assert_allclose(eng.ni.cpu().numpy(), ni1 + nsrc - 0.5 * src, rtol=1e-13)
    j = tally.cpu().numpy()
    assert j[0] == pytest.approx(nsrc.sum(), rel=1e-13)
    assert j[1] == pytest.approx(0.5 * src.sum(), rel=1e-13)
    assert j[2] == pytest.approx(abs(nsrc.sum() - 0.5 * src.sum()), rel=1e-12)
    assert j[3] == 1.0
    assert np.all(acc.cpu().numpy() == 0.0)
    eng.close()


def test_trained_closure_beats_untrained(gpu):
    """Ablation: a network trained on swarm data predicts held-out rates far
    better than the same network untrained."""
    from boltzmann_amd.closure_run import swarm_training_set, train_closure
    x, t = swarm_training_set(n_samples=60_000, per_instance=1500, window=10,
                              temps=(0.5, 10.0), n_bins=400)
    assert np.isfinite(x).all() and np.isfinite(t).all()
    net, rep = train_closure(x, t, epochs=25, batch=2048, lr=3e-3)
    a, b = rep["holdout_untrained"], rep["holdout_trained"]
    # the loss keeps the sampling noise of the small swarms (mobility above
    # all); the ionization rate, which the closure feeds to the fluid, gains most
    assert b["loss"] < 0.3 * a["loss"], rep
    assert b["k_ion_rel_err_median"] < 0.25 * a["k_ion_rel_err_median"], rep
    assert b["k_ion_rel_err_median"] < 0.2, rep
    net.close()


def test_closure_discharge_two_macro_steps(gpu):
    """The coupled loop of ClosureDischarge for two macro steps (a kinetic
    update, then 8 sub-steps, then the commit), each stage checked on the
    device's own input: every sub-step's learned source against the FP64
    oracle network on that sub-step's field and kinetic moments, the
    sub-step's ions and field against the FP64 field oracle, and the commit
    against ni + (kinetic - learned) ionization, with the tally counting the
    updates."""
    import torch
    from boltzmann_amd import _lib, configs
    from boltzmann_amd import physics as ph
    from boltzmann_amd.closure_run import ClosureDischarge
    net = _net()
    params = {"master": net.parameters(), "norm": dict(net.norm)}
    cfg = configs.c3().with_(n_particles=100_000, n_cells=64, xs_mode=_lib.BZ_XS_FLIGHT,
                             target_ppc=0)
    run = ClosureDischarge(cfg, params, n_sub=8)
    run.load(ph.initial_state(cfg))
    e = run.eng
    N, nn = cfg.n_cells, cfg.n_nodes
    fp = e.fp
    chol = e.chol.cpu().numpy()
    for macro in range(2):
        run.kinetic_update()
        mom_kin = e.moments.cpu().numpy().copy()
        ni_start = e.ni.cpu().numpy()
        acc_host = np.zeros(nn)
        for sub in range(run.n_sub):
            E0 = e.E.cpu().numpy()
            mom0 = e.moments.cpu().numpy()
            ni0, ns0 = e.ni.cpu().numpy(), e.ns.cpu().numpy()
            g20, V0 = e.g2.cpu().numpy(), e.V.cpu().numpy()
            t0 = e.clock()[2]
            run.sub_step()
            src = run.src.cpu().numpy()
            want = _want_src(run.net, E0, mom0, N, cfg.gas_density_m3, cfg.dt)
            np.testing.assert_allclose(src, want, rtol=2e-4 * 2.0 * math.log(10), atol=0,
                                       err_msg=f"source {macro}.{sub}")
            acc_host += src
            m2 = np.concatenate([mom0[:nn] + src, src, mom0[2 * nn:]])
            w_ = O.field_step(N, fp.dx, fp.dt, fp.eps, fp.v0, fp.freq, fp.mu_i, fp.D_i, fp.D_s,
                              fp.fluid_on, fp.poisson_on, t0 + cfg.dt, m2, ni0, ns0,
                              np.zeros(nn), g20, V0, E0, chol, ion_scheme=fp.ion_scheme,
                              max_substeps=fp.max_substeps)
            for nm, w in zip(("ni", "V", "E"), (w_[0], w_[4], w_[5])):
                g = getattr(e, nm).cpu().numpy()
                np.testing.assert_allclose(g, w, rtol=1e-10, atol=1e-10 * np.abs(w).max(),
                                           err_msg=f"{nm} {macro}.{sub}")
            assert e.clock()[2] == pytest.approx(t0 + cfg.dt, rel=1e-14)
        np.testing.assert_allclose(run.acc.cpu().numpy(), acc_host, rtol=1e-12, atol=0)
        ni_pre = e.ni.cpu().numpy()
        run.commit()
        nsrc = mom_kin[nn:2 * nn]
        np.testing.assert_allclose(e.ni.cpu().numpy(), ni_pre + nsrc - acc_host, rtol=1e-12,
                                   atol=1e-12 * np.abs(ni_pre).max())
        assert not run.acc.cpu().numpy().any()
        assert run.tally.cpu().numpy()[3] == macro + 1
        assert np.all(np.isfinite(ni_start))
    rep = run.report()
    assert rep["kinetic_updates"] == 2 and rep["ionization_kinetic"] >= 0.0
    run.close()
    net.close()
