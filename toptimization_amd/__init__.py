"""toptimization_amd: SIMP topology optimization on one AMD Instinct MI355X.

Drop-in for the Taichi kernels of ``tomohiron907/toptimization``: the same
YAML-problem / solver / filter / optimizer Python API, with the 2-D Q4 hot
path (matvec, PCG, filter, OC) in hand-written HIP kernels for gfx950 behind a
C ABI (``libtopt.so``) loaded with ctypes.
"""

__version__ = "0.1.0"
