"""general_code_amd -- MI355X-native Maximal Abelian gauge fixing.

The compute path is libmag_gpu.so (hand-written gfx950 HIP kernels behind the
C ABI of include/mag_gpu.h); this package is its Python front end.
"""
from .mag import (  # noqa: F401
    F64, FORM_LINK, FORM_SPIN, MAX_COPIES, MIXED, SU2, SU3, MAG_OK, MAG_ERR_NONFINITE,
    MAG_ERR_NOT_CONVERGED, MAG_ERR_CANCELLED, MAG_ERR_STATE, MagBatch, MagError, MagJob, MagParams, MagServer, NotConverged,
    SPIN_HEATBATH, SPIN_OMEGA, SPIN_OR, SPIN_RELAX, SpinSession, device_count,
    SHAPES, TUNE_AUTO, TUNE_DEFAULT, TUNE_FIXED, functional_spin, gauge_transform_spins,
    tune_clear, tune_set, generate_spins_uniform, hbm_copy_bandwidth, lib, plaquette, su2_functional, su3_functional,
    su3_monopole_currents, u1_monopole_currents, current_clusters, CLUSTER_INTS,
)
