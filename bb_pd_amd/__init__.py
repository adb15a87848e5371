"""bb_pd_amd -- MI355X-native explicit bond-based peridynamics step (mark-hobbs/BB_PD hot path).

Modules
  _lib     ctypes binding of libbbpd_hip.so (include/bbpd.h)
  build    hipcc build of the library for gfx950
  beam     three-point-bend beam inputs in the reference data model
  mex      the reference MEX entry points on device arrays
  solver   the fused device-resident step
"""
from ._lib import (ADR, EULER_CROMER, FP64, LAW_BILINEAR, LAW_EXPONENTIAL, LAW_LINEAR, LAW_TRILINEAR, MIXED,
                   BBPDError)

__all__ = ["ADR", "EULER_CROMER", "FP64", "MIXED", "LAW_LINEAR", "LAW_BILINEAR", "LAW_TRILINEAR", "LAW_EXPONENTIAL",
           "BBPDError"]
