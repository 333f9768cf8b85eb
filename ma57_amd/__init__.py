"""ma57_amd -- MI355X-native numeric phase of symla's multifrontal LDL^T solver.

The compute path is libsymla_hip.so (HIP kernels for gfx950 behind the C-ABI
in include/symla_hip.h); this package holds its build script, a ctypes
binding, a Python mirror of the ``symla::SymLDLT<double>`` facade and the
synthetic problem generators used by the tests and the benchmark.
"""
from .solver import (SymLDLT, Mode, OrderingType, Inertia, SymbolicFactorOptions, RefinementOptions,
                     InvalidArgument, LogicError, RuntimeFailure, DeviceFailure, dense_factor, analyze)

__all__ = ["SymLDLT", "Mode", "OrderingType", "Inertia", "SymbolicFactorOptions", "RefinementOptions",
           "InvalidArgument", "LogicError", "RuntimeFailure", "DeviceFailure", "dense_factor", "analyze"]
