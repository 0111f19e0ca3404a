"""MI355X (gfx950) implementation of AMDAT's reciprocal-space and
time-correlation path: rho_k(t), S(k), F(k,t), Fs(k,t), MSD and the
non-Gaussian parameter over AMDAT's blocked-exponential time scheme.

The product is the C ABI library amdat_amd/lib/libamdat_hip.so
(include/amdat_hip.h) and the C++ Analysis-class layer in amdat_amd/host;
this package holds the Python binding of the C ABI (_abi) and the seeded
synthetic trajectory generators (synth).
"""
from . import _abi  # noqa: F401

__all__ = ["_abi"]
