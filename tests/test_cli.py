"""CLI surface of the reference (1bin/main.py:22-34) plus the strategy flag."""
from boltzmann_amd import main as M


def test_reference_flags():
    a = M.parse(["-s", "5000", "-g", "1200", "-N", "500000", "-name", "test_", "-rsf", "1000",
                 "-dt", "100000", "-spc", "100000", "--bins", "3"])
    assert (a.steps, a.gridpts, a.num_particles, a.data_name) == (5000, 1200, 500000, "test_")
    assert (a.resample_freq, a.in_dt, a.in_spc, a.bins, a.bnf) == (1000, 1e5, 100000, 3, 4)


def test_defaults_match_reference():
    a = M.parse([])
    assert (a.steps, a.num_particles, a.gridpts, a.bnf, a.num_nodes, a.pykokkos,
            a.restart_cyclenum, a.data_name, a.resample_freq, a.in_dt, a.in_spc) == \
        (400, 200000, 256, 4, 1, False, 0, "", 1, 50000.0, 5000)


def test_reference_config_targets():
    from boltzmann_amd.boltzmann import reference_config
    c = reference_config(500000, 1200, 4, 1e5, nb=3)
    assert c.n_cells == 1199 and c.bin_thresholds == (10.0, 15.76)
    assert c.dt == (1.0 / 13.56e6) / 1e5
    assert c.gas_density_m3 == 1e6 * 3.22e16 * 0.1
    # 3bin targets int(0.7*Nc/N), int(0.2*Nc/N), int(0.1*Nc/N) = 291, 83, 41 per node
    assert c.target_ppc == 291 + 83 + 41
    assert c.n_particles == 415 * 1200  # 498,000 as recorded for the 3bin test run
