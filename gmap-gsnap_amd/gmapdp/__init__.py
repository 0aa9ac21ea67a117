"""Python front end of the MI355X banded affine-gap DP core (libgmapdp.so).

The compute path is the HIP library; this package only describes problems,
moves buffers (through torch's device allocator) and calls the C ABI
declared in include/gmapdp.h.
"""
from .lib import (GDP_FULL, GDP_UPPER, GDP_LOWER, HIGHQ, MEDQ, LOWQ, ENDQ, Problem, TbRequest, Pair,
                  load, library_path)

__all__ = ["GDP_FULL", "GDP_UPPER", "GDP_LOWER", "HIGHQ", "MEDQ", "LOWQ", "ENDQ", "Problem", "TbRequest",
           "Pair", "load", "library_path"]
