"""FEM layer: device fields, matrix-free Q4 kernels, PCG solver, preconditioners."""
