"""bz_reduction_static, the reduction entry point that allocates nothing,
against bz_reduction (the context entry point it is a variant of) and a
brute-force histogram: two calls in a row, the second smaller, bit for bit
equal to the context entry point; sizes beyond its static scratch are
rejected with an error."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(rng, RM, r, ptcl, bnf=4):
    M = bnf * RM
    dx = rng.uniform(0, 1, ptcl)
    dv = rng.normal(size=(ptcl, r))
    dv[rng.uniform(size=ptcl) < 0.1] = 0.0
    return M, dx, dv


def test_reduction_static_matches_the_context_entry(gpu):
    import torch

    from boltzmann_amd import _lib
    from boltzmann_amd.kernels import reduction as red
    lib = _lib.load()
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream(dev)
    rng = np.random.default_rng(3)
    for RM, r, ptcl in ((4097, 4, 400_000), (257, 3, 30_000)):
        M, dx, dv = _case(rng, RM, r, ptcl)
        tdx = torch.from_numpy(dx).to(dev)
        tdv = torch.from_numpy(dv).to(dev)
        want_ctx = torch.full((M, r), 5.0, dtype=torch.float64, device=dev)
        red.reduction(r, 0, M, RM, want_ctx, tdv, tdx, ptcl)
        got = torch.full((M, r), -9.0, dtype=torch.float64, device=dev)
        rc = lib.bz_reduction_static(r, M, RM, C.c_void_p(got.data_ptr()),
                                     C.c_void_p(tdv.data_ptr()), C.c_void_p(tdx.data_ptr()),
                                     ptcl, C.c_void_p(stream.cuda_stream))
        assert rc == 0, lib.bz_last_error()
        torch.cuda.synchronize(dev)
        g = got.cpu().numpy()
        np.testing.assert_array_equal(g, want_ctx.cpu().numpy())
        k = (dx * M).astype(int) // ((M + RM - 1) // RM)
        want = np.stack([np.bincount(k, weights=dv[:, j], minlength=RM) for j in range(r)], 1)
        np.testing.assert_allclose(g[:RM], want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
        assert np.all(g[RM:] == 0.0)


def test_reduction_static_rejects_sizes_beyond_its_scratch(gpu):
    from boltzmann_amd import _lib
    lib = _lib.load()
    RM, r = 200_000, 4  # 2 * RM * r + 64 doubles > 2^20
    rc = lib.bz_reduction_static(r, RM, RM, C.c_void_p(1), C.c_void_p(1), C.c_void_p(1), 0,
                                 None)
    assert rc != 0
    assert b"static array holds" in lib.bz_last_error()
