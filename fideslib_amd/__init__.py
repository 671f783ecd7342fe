"""fideslib_amd: MI355X (gfx950) CKKS hybrid key-switching core.

Hand-written HIP kernels behind the C ABI in include/fideslib_amd.h
(libfideslib_amd.so, loaded through fideslib_amd.capi); host-side helpers
for parameter selection live in fideslib_amd.params.
"""
from . import params  # noqa: F401
