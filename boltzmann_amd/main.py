"""Command line of the reference (1bin/main.py:11-37) with the binning
strategy as a flag instead of a directory:

    python -m boltzmann_amd.main -s 5000 -g 1200 -N 500000 -name run_ -rsf 1000 \
        -dt 100000 -spc 100000 --bins 3
"""
from __future__ import annotations

import argparse

from .boltzmann import BoltzmannDSMC


def main(restart_cyclenum, num_steps, name, num_particles, gridpts, bnf, num_nodes, pyk,
         resample_freq, in_dt, in_spc, nbins=1, ref_rescale=False, fix_quirks=False):
    boltzmann = BoltzmannDSMC(num_steps, name, restart_cyclenum, num_particles, gridpts, bnf,
                              num_nodes, pyk, resample_freq, in_dt, in_spc, nbins=nbins,
                              ref_rescale=ref_rescale, fix_quirks=fix_quirks)
    boltzmann.run(num_steps, name, restart_cyclenum)
    return boltzmann


def run(restart_cyclenum, num_steps, name, num_particles, gridpts, bnf, num_nodes, pyk,
        resample_freq, in_dt, in_spc, nbins=1, ref_rescale=False, fix_quirks=False):
    return main(restart_cyclenum, num_steps, name, num_particles, gridpts, bnf, num_nodes, pyk,
                resample_freq, in_dt, in_spc, nbins, ref_rescale, fix_quirks)


def parse(argv=None):
    parser = argparse.ArgumentParser()
    parser.add_argument("-s", "--steps", type=int, default=400)
    parser.add_argument("-N", "--num_particles", type=int, default=200000)
    parser.add_argument("-g", "--gridpts", type=int, default=256)
    parser.add_argument("-bnf", type=int, default=4)
    parser.add_argument("-n", "--num_nodes", type=int, default=1)
    parser.add_argument("-p", "--pykokkos", action="store_true")
    parser.add_argument("-res", "--restart_cyclenum", type=int, default=0)
    parser.add_argument("-name", "--data_name", type=str, default="")
    parser.add_argument("-rsf", "--resample_freq", type=int, default=1)
    parser.add_argument("-dt", "--in_dt", type=float, default=50000.0)
    parser.add_argument("-spc", "--in_spc", type=int, default=5000)
    parser.add_argument("--bins", type=int, default=1, choices=(1, 3, 5),
                        help="energy bins per cell (the reference's 1bin/3bin/5bin directories)")
    parser.add_argument("--ref-rescale", action="store_true",
                        help="merge rescale from total energy as the reference does")
    parser.add_argument("--fix-quirks", action="store_true",
                        help="write the collision counters and the ns average, and include "
                             "the output step in the cycle averages")
    return parser.parse_args(argv)


if __name__ == "__main__":
    args = parse()
    run(args.restart_cyclenum, args.steps, args.data_name, args.num_particles, args.gridpts,
        args.bnf, args.num_nodes, args.pykokkos, args.resample_freq, args.in_dt, args.in_spc,
        args.bins, args.ref_rescale, args.fix_quirks)
