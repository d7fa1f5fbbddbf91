"""MI355X-native HiG-Flow projection time step (HIP, gfx950).

``higflow_amd.native``  ctypes binding of libhigflow_hip.so (C ABI in
                        include/higflow_hip.h): the projection step and the
                        HiG-Tree solver-plugin engine.
``higflow_amd.cases``   synthetic case definitions (domain/boundary .amr
                        texts, boundary values, parameters).
"""
__version__ = "0.1.0"
