"""Stage functions with the reference's names and argument order
(1bin/boltzmann.py:758-795 binds them as driver attributes):

    electron_kernel       kernels/electron_kernel_numba.py:291
    bin_particles         kernels/bin_particles.py (1/3/5-bin signatures)
    post_processing       kernels/post_processing_1D.py:40
    reduction             kernels/reduction/__init__.py (init_reduction)
    EF_kernel             kernels/EF_kernel_1D.py:8
    heavies_kernel_fluid  kernels/heavies_kernel_1D.py:5

Device arrays are torch tensors on the GPU in the reference's AoS layout;
every stage runs in the HIP kernels of libbz.so.
"""
from .bin_particles import bin_particles
from .electron_kernel import electron_kernel, set_deposit_bins
from .field import EF_kernel, heavies_kernel_fluid
from .post_processing import post_processing
from . import reduction

__all__ = ["electron_kernel", "bin_particles", "post_processing", "reduction", "EF_kernel",
           "heavies_kernel_fluid", "set_deposit_bins"]
