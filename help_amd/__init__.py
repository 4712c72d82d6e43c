"""help_amd - MI355X-native normal-equation backend for the HELP HSD LP solver.

The product is libhelp_mi355x.so (HIP/CDNA4 kernels + C ABI declared in
include/help_mi355x.h); this package holds its Python mirror of HELP's
backend interface (backend.LinearEquation), the instance generators and the
HSD driver binding.
"""
from . import _lib
from .backend import HelpError, Instance, LinearEquation, device_count

__all__ = ["HelpError", "Instance", "LinearEquation", "device_count", "_lib"]
