"""bearing_amd: the contact pipeline of the Bearing physics engine on an AMD Instinct MI355X.

The product is the HIP library built from ``csrc/`` into ``_build/libbearing_mi355x.so``
(C ABI in ``include/bearing_mi355x.h``). ``native`` binds it with ctypes; ``scenes``
builds the seeded workloads.
"""
