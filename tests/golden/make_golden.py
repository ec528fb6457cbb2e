"""Write the golden fixtures of the electron kinetic step.

Each fixture is a small .npz holding seeded inputs and the oracle's outputs
for one stage: the fused push (every collision model, both update orders,
both precisions, both layouts), binning + sort, ES-WPNC resampling, the FP64
grid step and the closure network.  The reference ships no test vectors, so
these freeze the oracle: tests/test_golden.py re-derives every output from
the inputs (a change to the oracle must show up as a fixture change), and
tests/test_golden_gpu.py runs the HIP path on the same inputs.

    python tests/golden/make_golden.py        (rewrites tests/golden/*.npz)
"""
from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

from oracle import mlp as M  # noqa: E402
from oracle import oracle as O  # noqa: E402

KEV = 0.5 * O.M_E * O.J_EV
SEED = 20261203

# name: (dim, xs_mode, dtype, push_order, rng "stream"|"philox", n_cells, gas, dt)
PUSH_CASES = {
    "push_table_f32_1d_leapfrog": (1, 1, np.float32, 1, "philox", 48, 3.2e24, 2e-11),
    "push_table_f64_1d_ref": (1, 1, np.float64, 0, "stream", 48, 3.2e24, 2e-11),
    "push_fit_f64_1d_ref": (1, 0, np.float64, 0, "stream", 48, 3.22e23, 2e-11),
    "push_table_f64_0d": (0, 1, np.float64, 0, "philox", 1, 3.2e24, 2e-11),
    "push_flight_f32_1d_leapfrog": (1, 2, np.float32, 1, "philox", 48, 3.2e23, 2e-11),
    "push_direct_f32_1d_leapfrog": (1, 3, np.float32, 1, "philox", 48, 3.22e23, 2e-11),
}
N_PUSH = 1000
L = 2.5


def xsec_table(gas, dtype):
    from boltzmann_amd import physics as ph
    t = ph.build_table(ph.XSecModel(), gas, dtype)
    return t


def push_state(n, dtype, dim, seed):
    rng = np.random.default_rng(seed)
    e = np.exp(rng.uniform(np.log(0.05), np.log(80.0), n))
    d = rng.normal(size=(3, n))
    d /= np.linalg.norm(d, axis=0)
    v = np.sqrt(e / KEV) * d
    x = rng.uniform(1e-4 * L, L * (1 - 1e-4), n) if dim == 1 else np.full(n, 0.5 * L)
    w = rng.uniform(0.5, 2.0, n)
    w[rng.uniform(size=n) < 0.02] = 0.0
    return tuple(np.ascontiguousarray(a, dtype=dtype) for a in (x, v[0], v[1], v[2], w))


def push_config(name):
    """The run configuration of a push fixture (the GPU test builds its
    engine from it)."""
    from boltzmann_amd import physics as ph
    dim, xs, dt_, order, rng_mode, n_cells, gas, dt = PUSH_CASES[name]
    prec = 8 if dt_ == np.float64 else 4
    return ph.RunConfig(name=name, dim=dim, n_particles=N_PUSH, precision=prec, xs_mode=xs,
                        n_cells=n_cells if dim == 1 else 1, length_cm=L, dt=dt,
                        gas_density_m3=gas, e_over_n_td=100.0, push_order=order, seed=SEED)


def push_inputs(name):
    dim, xs, dt_, order, rng_mode, n_cells, gas, dt = PUSH_CASES[name]
    cfg = push_config(name)
    seed = SEED + sum(map(ord, name))
    st = push_state(N_PUSH, dt_, dim, seed)
    rng = np.random.default_rng(seed + 1)
    E = rng.uniform(-400, 400, n_cells).astype(dt_)
    if rng_mode == "stream":
        R = rng.uniform(0, 1, size=(7, N_PUSH)).astype(dt_)
        G = rng.normal(0, 6.7e4, size=(3, N_PUSH)).astype(dt_) if xs == 0 else None
    else:
        R = O.collision_uniforms(SEED, 0, N_PUSH, dt_) if xs != 2 else None
        G = None
    tab = xsec_table(gas, dt_) if xs != 0 else None
    setup = O.Setup(dim=dim, dt=dt, xs_mode=xs, n_cells=n_cells if dim == 1 else 1,
                    length_cm=L, bnf=4, gas_density_m3=gas,
                    nu_max=tab.nu_max if tab is not None else 0.0,
                    e_field0=0.0 if dim == 1 else cfg.e_field0_vcm(), seed=SEED, step=0,
                    push_order=order)
    return dict(state=st, E=E if dim == 1 else None, R=R, G=G, tab=tab, setup=setup)


def push_expected(inp):
    t = inp["tab"]
    table = None if t is None else (t.energy, t.sig, t.slope)
    out, mom, cnt = O.push_step(inp["setup"], inp["state"], inp["E"], inp["R"], G=inp["G"],
                                table=table, capacity=2 * N_PUSH + 16)
    return out, mom, cnt


def resample_inputs(nb, dtype):
    from boltzmann_amd import physics as ph
    N = 16
    rng = np.random.default_rng(SEED + nb)
    counts = rng.integers(0, 120, N + 1)
    counts[3] = 0
    counts[5] = 400  # merge
    counts[7] = 12   # split
    dx = L / N
    xs, vs, ws = [], [], []
    for c, m in enumerate(counts):
        lo, hi = max((c - 0.5) * dx, 1e-6), min((c + 0.5) * dx, L - 1e-6)
        xs.append(rng.uniform(lo, hi, m))
        v = rng.normal(0, math.sqrt(12.0 / KEV / 3), size=(3, m))
        v[0] += 3e5
        vs.append(v)
        ws.append(rng.uniform(0.5, 2.0, m))
    x = np.concatenate(xs)
    v = np.concatenate(vs, axis=1)
    w = np.concatenate(ws)
    p = rng.permutation(x.size)
    st = tuple(np.ascontiguousarray(a[p], dtype=dtype) for a in (x, v[0], v[1], v[2], w))
    targets = np.array([int(f * 100) for f in ph.bin_fractions(nb)] * (N + 1), dtype=np.int32)
    setup = O.Setup(dim=1, dt=1e-11, n_cells=N, length_cm=L)
    pol = dict(split_below=0.5, merge_at=3.0, merge_keep=0.6, trigger_at=3.0, min_wt=0.1,
               split_to=2.0)
    return dict(state=st, targets=targets, setup=setup, pol=pol, nb=nb,
                thr=ph.bin_thresholds(nb))


def resample_expected(inp):
    s = inp["setup"]
    counts, offsets, perm = O.bin_sort(s, inp["nb"], inp["thr"], inp["state"])
    out, cnt = O.resample(s, O.policy(**inp["pol"]), inp["nb"], inp["thr"], inp["targets"],
                          SEED, 3, inp["state"], capacity=4 * inp["state"][0].size + 1024)
    return counts, offsets, perm, out, cnt


def resample_inplace_expected(inp):
    """The same pass in the reference's in-place layout (every slot, holes
    included)."""
    s = inp["setup"]
    out, cnt = O.resample_inplace(s, O.policy(**inp["pol"]), inp["nb"], inp["thr"],
                                  inp["targets"], SEED, 3, inp["state"],
                                  capacity=4 * inp["state"][0].size + 1024)
    return out, cnt


def field_inputs():
    from boltzmann_amd import physics as ph
    N = 64
    dx = L / N
    rng = np.random.default_rng(SEED + 7)
    ne = rng.uniform(0.8e9, 1.2e9, N + 1)
    mom = np.concatenate([ne, rng.uniform(0, 1e6, N + 1), ne * rng.uniform(2, 4, N + 1)])
    ni = rng.uniform(0.8e9, 1.2e9, N + 1)
    E = rng.uniform(-50, 50, N)
    return dict(N=N, dx=dx, dt=1.8e-10, eps=ph.EPS_E_VCM, v0=200.0, freq=13.56e6,
                mu_i=4.65e19 / 3.2e16, D_i=2.07e18 / 3.2e16, D_s=2.42e18 / 3.2e16, t=3.3e-9,
                mom=mom, ni=ni, E=E, chol=ph.poisson_cholesky(N, dx))


def field_expected(inp, ion_scheme):
    N = inp["N"]
    return O.field_step(N, inp["dx"], inp["dt"], inp["eps"], inp["v0"], inp["freq"],
                        inp["mu_i"], inp["D_i"], inp["D_s"], 1, 1, inp["t"], inp["mom"],
                        inp["ni"].copy(), np.zeros(N + 1), np.zeros(N + 1), np.zeros(N + 1),
                        np.zeros(N - 1), inp["E"].copy(), inp["chol"], ion_scheme=ion_scheme,
                        max_substeps=256)


def mlp_inputs():
    rng = np.random.default_rng(SEED + 11)
    theta = (rng.normal(0, 0.3, M.N_PARAM)).astype(np.float32)
    norm = {"in_shift": np.array([1.5, 6.0, 16.0]), "in_scale": np.array([0.9, 0.12, 0.6]),
            "out_shift": np.array([-17.0, -16.0, 24.0]), "out_scale": np.array([2.0, 1.5, 0.4])}
    n = 64
    x = np.stack([rng.uniform(0, 3, n), rng.uniform(0.3, 20, n), rng.uniform(13, 19, n)], 1)
    t = np.stack([-16 - 2 * np.cos(x[:, 0]), -15.5 - 0.05 * x[:, 1], 24 + 0.2 * x[:, 0]], 1)
    return dict(master=theta, norm=norm, x=x.astype(np.float32), t=t.astype(np.float32))


def mlp_expected(inp):
    th = M.bf16_value(M.bf16_bits(inp["master"]))
    y = M.forward(th, inp["norm"], inp["x"].astype(np.float64))
    loss, grad = M.loss_and_grad(th, inp["norm"], inp["x"].astype(np.float64),
                                 inp["t"].astype(np.float64))
    return M.bf16_bits(inp["master"]), y, loss, grad


def _opt(a):
    return np.zeros(0) if a is None else a


def write_all(out_dir: Path = HERE) -> list:
    written = []
    for name in PUSH_CASES:
        inp = push_inputs(name)
        out, mom, cnt = push_expected(inp)
        np.savez_compressed(out_dir / f"{name}.npz", x=inp["state"][0], vx=inp["state"][1],
                            vy=inp["state"][2], vz=inp["state"][3], w=inp["state"][4],
                            E=_opt(inp["E"]), R=_opt(inp["R"]), G=_opt(inp["G"]),
                            out=np.stack(out), moments=mom, counters=cnt)
        written.append(name)
    for nb in (3, 5):
        inp = resample_inputs(nb, np.float64)
        counts, offsets, perm, out, cnt = resample_expected(inp)
        ip, _ = resample_inplace_expected(inp)
        st = inp["state"]
        name = f"resample_{nb}bin_f64"
        np.savez_compressed(out_dir / f"{name}.npz", x=st[0], vx=st[1], vy=st[2], vz=st[3],
                            w=st[4], targets=inp["targets"], counts=counts, offsets=offsets,
                            perm=perm, out=np.stack(out), counters=cnt,
                            out_inplace=np.stack(ip))
        written.append(name)
    inp = field_inputs()
    for scheme in (0, 1):
        ni, ns, Te, g2, V, E = field_expected(inp, scheme)
        name = f"field_scheme{scheme}"
        np.savez_compressed(out_dir / f"{name}.npz", mom=inp["mom"], ni_in=inp["ni"],
                            E_in=inp["E"], chol=inp["chol"], ni=ni, ns=ns, Te=Te, g2=g2, V=V,
                            E=E)
        written.append(name)
    inp = mlp_inputs()
    bits, y, loss, grad = mlp_expected(inp)
    np.savez_compressed(out_dir / "mlp.npz", master=inp["master"], x=inp["x"], t=inp["t"],
                        bits=bits, y=y, loss=np.array(loss), grad=grad)
    written.append("mlp")
    return written


if __name__ == "__main__":
    for n in write_all():
        print(n, (HERE / f"{n}.npz").stat().st_size, "bytes")
