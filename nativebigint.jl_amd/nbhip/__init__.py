"""nbhip -- MI355X-native FP64 two-prime NTT big-integer multiplication.

Host-side Python mirror of NativeBigInt.jl's engine interface over the C ABI
of libnbh.so (include/nbh.h).  `engine` holds the host-buffer drop-ins
(mul_fpntt2/sqr_fpntt2/mul/sqr/mul_bal), `device` the HBM-resident and batched
entry points, `workloads` the seeded synthetic operands.
"""
from . import build as _build  # noqa: F401
from ._lib import ArgumentError, NbhError  # noqa: F401
