"""Python side of the MI355X lattice-Boltzmann engine: ctypes binding of the
C ABI (include/badchimp_amd.h) and synthetic geometry helpers."""
from .ffi import (D2Q9, D3Q19, FP16S, FP32S, FP64, FORCE_FIXED, FORCE_MOMX, LATTICE_NAME, ND_OF,
                  PHI, Q_OF, RHO, VEL, VISC, Batch, Context, LBError, Topology, batch_footprint, bench_copy, bench_stream, lib,
                  periodic_mask, sphere_pack, VtuOutput)

__all__ = ["D2Q9", "D3Q19", "FP64", "FP32S", "FP16S", "FORCE_FIXED", "FORCE_MOMX", "RHO", "PHI", "VISC",
           "VEL", "Batch", "batch_footprint", "bench_copy", "bench_stream", "Q_OF", "ND_OF", "LATTICE_NAME", "Context", "Topology", "LBError", "lib",
           "periodic_mask", "sphere_pack", "VtuOutput"]
