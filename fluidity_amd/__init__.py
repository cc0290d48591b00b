"""MI355X-native (gfx950) finite-volume Euler step of fluidity.

The HIP kernels and the C ABI live in libfluidity_amd.so (sources in csrc/,
declarations in include/fluidity_amd.h).  ``fluidity_amd.euler`` mirrors the
reference's split-solver API over that ABI; importing it requires the built
library and never falls back to another implementation.
"""
__all__ = ["build", "workloads"]
