"""zkpoly_amd: MI355X-native BN254 prover hot path behind the reference's C ABI.

The compute lives in the HIP shared libraries under <package>/lib (libntt.so,
libmsm_bn254_*.so, generated fused-kernel libraries); this package is the
host-side mirror of the reference runtime's operator interface plus build
helpers.  There is no CPU fallback: operators raise if a library is missing.
"""
from . import field, hip, runtime  # noqa: F401

__all__ = ["field", "hip", "runtime"]
