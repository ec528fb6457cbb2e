"""Compaction after the push (kernels/post_processing_1D.py:26-38) on the GPU.

bz_compact is a stable filter of the w > 0 slots into the other buffer;
bz_compact_inplace fills the holes below the live count L from the live slots
at or above it, k-th hole (slot order) from the k-th such slot, and leaves
live slots below L in place (oracle.fill_holes).  Both are integer/index
work: bit-exact against the CPU on every slot, for states smaller than one
block, spread over many blocks with several tiles each, with no holes, with
a hole at every slot and with the live count on a block boundary."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _engine(n, prec, fill):
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine
    cfg = ph.RunConfig(name="compact", dim=1, n_particles=max(n, 1), precision=prec, nb=1,
                       n_cells=16, length_cm=2.5, dt=1e-11, bin_fractions=ph.bin_fractions(1),
                       bin_thresholds=ph.bin_thresholds(1))
    return KineticEngine(cfg, capacity=n + 4096, inplace_resample=True, fill_holes=fill)


def _state(n, p_hole, dtype, seed):
    rng = np.random.default_rng(seed)
    w = np.where(rng.uniform(size=n) < p_hole, 0.0, rng.uniform(0.5, 2.0, n))
    return tuple(np.ascontiguousarray(a, dtype=dtype) for a in
                 (rng.uniform(0, 2.5, n), rng.normal(size=n), rng.normal(size=n),
                  rng.normal(size=n), w))


CASES = [(700, 0.3), (3_000_000, 0.02), (3_000_000, 0.5), (1_000_000, 0.0),
         (20_000, 1.0), (2_500_000, 0.001)]


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("fill", [True, False], ids=["fill_holes", "stable"])
@pytest.mark.parametrize("n,p_hole", CASES)
def test_compaction_matches_the_cpu(gpu, n, p_hole, prec, fill):
    dtype = np.float64 if prec == 8 else np.float32
    st = _state(n, p_hole, dtype, seed=n + int(100 * p_hole))
    eng = _engine(n, prec, fill)
    eng.load_state(st)
    eng.compact()
    got = eng.fetch_slots()
    if fill:
        want = O.fill_holes(st)
    else:
        live = st[4] > 0
        want = tuple(a[live] for a in st)
    assert got[0].size == want[0].size == int((st[4] > 0).sum())
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)
    eng.close()


def test_fill_holes_with_the_live_count_on_a_block_boundary(gpu):
    """L a multiple of a block's slot range: the boundary block has no holes
    below L and every live slot of it moves.  2^21 slots are 2048 tiles of
    1024; a grid of 256 to 2048 blocks gives ranges of 1, 2, 4 or 8 tiles,
    all of which divide L = 40 tiles."""
    n = 1 << 21
    eng = _engine(n, 4, True)
    L = 40 * 1024
    rng = np.random.default_rng(9)
    w = np.zeros(n)
    live = rng.permutation(n)[:L]
    w[live] = rng.uniform(0.5, 2.0, L)
    st = tuple(np.ascontiguousarray(a, dtype=np.float32) for a in
               (rng.uniform(0, 2.5, n), rng.normal(size=n), rng.normal(size=n),
                rng.normal(size=n), w))
    eng.load_state(st)
    eng.compact()
    got = eng.fetch_slots()
    assert got[0].size == L
    for g, w_ in zip(got, O.fill_holes(st)):
        np.testing.assert_array_equal(g, w_)
    eng.close()
