"""MI355X-native PBD cloth substep with Morton/LBVH self-collision.

The compute path is the HIP library ``libclothmi.so`` (hand-written gfx950
kernels behind the C ABI in include/clothmi.h); this package holds its ctypes
binding, the host-side mirror of the reference's simulation-object interface and
the synthetic scene builders.
"""
