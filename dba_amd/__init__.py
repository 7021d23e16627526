"""dba_amd - MI355X-native dynamic-block-activation time step.

The product is libdba_hip.so (dba_amd/csrc, C ABI in include/dba/dba.h);
this package binds it (abi, runtime) and defines the workloads (configs).
"""
from . import abi, configs  # noqa: F401
from .runtime import Context, Instance  # noqa: F401
