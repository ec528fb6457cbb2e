"""Binning keys written by the 1-D push, through the C ABI.

With push keys on, the 1-D push also writes the bin-sort key of every slot
it leaves (holes key n_keys, secondaries in their appended slots), and the
next bin sort of that state sorts those keys instead of recomputing them.
The sort's counts, offsets and permutation must equal the ones recomputed
from the state and the oracle's, bit for bit; the direct and queued push
paths both write keys.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

KEV = 0.5 * O.M_E * O.J_EV


def _state(n, dtype, L=2.5, seed=0, dead_frac=0.02):
    rng = np.random.default_rng(seed)
    e = np.exp(rng.uniform(np.log(0.05), np.log(80.0), n))
    sp = np.sqrt(e / KEV)
    d = rng.normal(size=(3, n))
    d /= np.linalg.norm(d, axis=0)
    v = sp * d
    x = rng.uniform(1e-4 * L, L * (1 - 1e-4), n)
    w = rng.uniform(0.5, 2.0, n)
    w[rng.uniform(size=n) < dead_frac] = 0.0
    return tuple(np.ascontiguousarray(a, dtype=dtype) for a in (x, v[0], v[1], v[2], w))


def _engine(prec, xs, nb, n, push_keys, queue):
    from boltzmann_amd import _lib
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine
    gas = 3.2e24 if xs == _lib.BZ_XS_TABLE else 3.22e23
    cfg = ph.RunConfig(name="keys", dim=1, n_particles=n, precision=prec, xs_mode=xs, nb=nb,
                       n_cells=48, length_cm=2.5, dt=2e-11, gas_density_m3=gas, target_ppc=100,
                       bin_fractions=ph.bin_fractions(nb), bin_thresholds=ph.bin_thresholds(nb))
    # 32-bit keys: the bin sort (the key-order path) takes the push's keys
    eng = KineticEngine(cfg, capacity=3 * n + 64, push_keys=push_keys, push_keys16=False)
    _lib.check(eng.lib.bz_set_push_queue(eng.ctx, queue))
    if queue == 1:
        # one block per CU: every wave walks several tiles and its batches cross them
        _lib.check(eng.lib.bz_set_push_grid(eng.ctx, 1))
    return cfg, eng


def _sorted(eng):
    n = eng.count()
    return (eng.counts.cpu().numpy().copy(), eng.offsets.cpu().numpy().copy(),
            eng.perm[:n].cpu().numpy().copy())


CASES = [(4, 1, 3, 0), (8, 1, 5, 1), (4, 3, 3, 1), (8, 3, 1, 0), (4, 2, 5, 0)]


@pytest.mark.parametrize("prec,xs,nb,queue", CASES,
                         ids=[f"f{8 * c[0]}-xs{c[1]}-nb{c[2]}-q{c[3]}" for c in CASES])
def test_push_keys_sort_equals_recomputed(gpu, prec, xs, nb, queue):
    import torch
    from boltzmann_amd import physics as ph
    n = 1_000_000
    cfg, a = _engine(prec, xs, nb, n, True, queue)
    _, b = _engine(prec, xs, nb, n, False, queue)
    st = _state(n, cfg.np_dtype, seed=41 + prec + xs + nb)
    E = torch.from_numpy(np.random.default_rng(2).uniform(-400, 400, cfg.n_cells)
                         .astype(cfg.np_dtype))
    for eng in (a, b):
        eng.load_state(st)
        eng.E_t.copy_(E)
        eng.push()        # leaves holes and appends births
        eng.push()        # pushes a state with holes and earlier births
    n_a = a.count()
    assert n_a == b.count() and n_a > n
    assert a.counters()["births"] > 0 and a.counters()["absorbed"] > 0
    sa = a.fetch_slots()
    for u, v in zip(sa, b.fetch_slots()):
        np.testing.assert_array_equal(u, v)
    # the keys sorted are the push's: zeroing the weights after the push
    # (which the push keys do not see) leaves the sort unchanged
    a.state.arrays[4][:n_a].zero_()
    a.bin_sort()
    b.bin_sort()
    got, want = _sorted(a), _sorted(b)
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)
    s = O.Setup(dim=1, dt=cfg.dt, n_cells=cfg.n_cells, length_cm=cfg.length_cm)
    c, o, p = O.bin_sort(s, nb, ph.bin_thresholds(nb), sa)
    np.testing.assert_array_equal(got[0], c)
    np.testing.assert_array_equal(got[1], o)
    np.testing.assert_array_equal(got[2].view(np.uint32), p)
    a.close()
    b.close()


def test_push_keys_dropped_by_other_calls(gpu):
    """A state loaded after the push (clock set), or a second sort of the
    same state, recomputes the keys from the state."""
    from boltzmann_amd import physics as ph
    n = 200_000
    cfg, a = _engine(4, 1, 3, n, True, 0)
    st = _state(n, cfg.np_dtype, seed=5)
    st2 = _state(n, cfg.np_dtype, seed=6)
    s = O.Setup(dim=1, dt=cfg.dt, n_cells=cfg.n_cells, length_cm=cfg.length_cm)
    a.load_state(st)
    a.push()
    a.load_state(st2)           # same buffers, new contents
    a.bin_sort()
    c, o, p = O.bin_sort(s, 3, ph.bin_thresholds(3), st2)
    np.testing.assert_array_equal(a.counts.cpu().numpy(), c)
    np.testing.assert_array_equal(a.perm[:n].cpu().numpy().view(np.uint32), p)
    a.push()
    a.bin_sort()                # the push's keys
    a.state.arrays[4][:a.count()].zero_()
    a.bin_sort()                # the sort used them up: recomputed, every slot dead
    cnt = a.counts.cpu().numpy()
    assert cnt[:-1].sum() == 0 and cnt[-1] == a.count()
    a.close()


def test_push_keys_with_births_past_capacity(gpu):
    """Births that do not fit the capacity are dropped and counted; the keys
    the push wrote still cover exactly the slots kept."""
    from boltzmann_amd import _lib
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine
    n = 200_000
    cfg = ph.RunConfig(name="keys-cap", dim=1, n_particles=n, precision=4,
                       xs_mode=_lib.BZ_XS_TABLE, nb=3, n_cells=48, length_cm=2.5, dt=2e-11,
                       gas_density_m3=3.2e24, target_ppc=100, bin_fractions=ph.bin_fractions(3),
                       bin_thresholds=ph.bin_thresholds(3))
    cap = n + 64  # far fewer free slots than births
    a = KineticEngine(cfg, capacity=cap, push_keys=True, push_keys16=False)
    b = KineticEngine(cfg, capacity=cap, push_keys=False)
    st = _state(n, cfg.np_dtype, seed=9, dead_frac=0.0)
    for e in (a, b):
        e.load_state(st)
        e.push()
    assert a.count() == cap and a.counters()["overflow"] > 0
    a.bin_sort()
    b.bin_sort()
    for g, w in zip(_sorted(a), _sorted(b)):
        np.testing.assert_array_equal(g, w)
    a.close()
    b.close()


def test_push_keys_empty_state(gpu):
    from boltzmann_amd import physics as ph
    cfg, a = _engine(4, 1, 3, 1000, True, 0)
    a.load_state(tuple(np.zeros(0, dtype=cfg.np_dtype) for _ in range(5)))
    a.push()
    a.bin_sort()
    assert a.count() == 0
    assert int(a.counts.cpu().numpy().sum()) == 0
    a.close()
    _ = ph


def test_push_keys16_inplace_resampling_equals_32bit(gpu):
    """16-bit push keys (bz_set_push_keys16, the engine's choice for in-place
    resampling): a run that resamples every step in place, births appended
    with their keys by k_push_post, equals the run on 32-bit keys slot for
    slot (C4's layout scaled: 5 bins, 256 cells, compaction every 4th pass)."""
    import torch
    from boltzmann_amd import configs
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine
    cfg = configs.c4().with_(n_particles=300_000, n_cells=256, target_ppc=1200)
    st = ph.initial_state(cfg)
    runs = []
    for k16 in (True, False):
        e = KineticEngine(cfg, push_keys16=k16)
        assert e.push_keys and e.inplace_resample and e.push_keys16 == k16
        e.load_state(st)
        e.initialize_fields()
        for i in range(9):
            e.step(i)
        torch.cuda.synchronize()
        runs.append(e)
    a, b = runs
    ca, cb = a.counters(), b.counters()
    assert ca == cb and ca["merge"] > 0 and ca["split"] > 0 and ca["births"] > 0
    for u, v in zip(a.fetch_slots(), b.fetch_slots()):
        np.testing.assert_array_equal(u, v)
    np.testing.assert_array_equal(a.moments.cpu().numpy(), b.moments.cpu().numpy())
    a.close()
    b.close()


def test_push_keys16_rejects_keys_beyond_16_bits(gpu):
    """n_nodes * nb >= 65535 does not fit 16-bit keys: the push fails with
    an error and leaves the state as it was."""
    from boltzmann_amd import _lib
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine
    cfg = ph.RunConfig(name="keys16-wide", dim=1, n_particles=10_000, precision=4,
                       xs_mode=_lib.BZ_XS_TABLE, nb=5, n_cells=13106, length_cm=2.5, dt=2e-11,
                       gas_density_m3=3.2e24, target_ppc=10, bin_fractions=ph.bin_fractions(5),
                       bin_thresholds=ph.bin_thresholds(5))
    assert cfg.n_nodes * 5 == 65535
    e = KineticEngine(cfg, push_keys=True, push_keys16=False)
    st = _state(10_000, cfg.np_dtype, seed=3, dead_frac=0.0)
    e.load_state(st)
    _lib.check(e.lib.bz_set_push_keys16(e.ctx, 1))
    with pytest.raises(_lib.BzError, match="16-bit push keys"):
        e.push()
    for u, v in zip(e.fetch_slots(), st):
        np.testing.assert_array_equal(u, v)
    e.close()
