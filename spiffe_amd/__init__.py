"""MI355X-native time step of spiffe's 2.5D axisymmetric EM-PIC gun simulation.

The device library (libspiffe_amd.so, HIP for gfx950) holds fields and
particles in HBM and runs the FDTD, push, deposition and Poisson kernels
behind the C ABI of include/spiffe_amd.h; this package is its host side.
"""
__version__ = "0.1.0"
