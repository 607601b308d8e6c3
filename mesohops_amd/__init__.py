"""MI355X-native HOPS propagation engine behind the MesoHOPS trajectory API.

The propagation path (hierarchy super-operator build, RK4 derivative stages,
noise filter) runs in hand-written HIP kernels in libhops_mi355x.so, called
through a C ABI (include/hops_mi355x.h) with ctypes.
"""
from .errors import AuxError, EngineError, LockedException, TrajectoryError, UnsupportedRequest
from .params import bcf_convert_dl_to_exp, bcf_exp, hbar, ishizaki_decomposition_bcf_dl, kB, precision
from .trajectory import HopsBatch, HopsStorage, HopsTrajectory
from .dyadic import (DyadicSpectra, DyadicTrajectory, prepare_chromophore_input_dict,
                     prepare_convergence_parameter_dict, prepare_spectroscopy_input_dict, zeropadded_fft)

__all__ = ["HopsTrajectory", "HopsBatch", "HopsStorage", "DyadicTrajectory", "DyadicSpectra",
           "prepare_spectroscopy_input_dict", "prepare_chromophore_input_dict", "prepare_convergence_parameter_dict",
           "zeropadded_fft", "bcf_exp", "bcf_convert_dl_to_exp", "ishizaki_decomposition_bcf_dl", "hbar", "kB",
           "precision", "UnsupportedRequest", "LockedException", "TrajectoryError", "AuxError", "EngineError"]
