"""Optimizer layer: SIMP loop, density/sensitivity filter, OC update."""
