"""MI355X-native HOPS propagation engine behind the MesoHOPS trajectory API.

The propagation path (hierarchy super-operator build, RK4 derivative stages,
noise filter) runs in hand-written HIP kernels in libhops_mi355x.so, called
through a C ABI (include/hops_mi355x.h) with ctypes.
"""
from .errors import AuxError, EngineError, LockedException, TrajectoryError, UnsupportedRequest
from .params import bcf_convert_dl_to_exp, bcf_exp, hbar, kB, precision
from .trajectory import HopsBatch, HopsStorage, HopsTrajectory

__all__ = ["HopsTrajectory", "HopsBatch", "HopsStorage", "bcf_exp", "bcf_convert_dl_to_exp", "hbar", "kB",
           "precision", "UnsupportedRequest", "LockedException", "TrajectoryError", "AuxError", "EngineError"]
