"""``toptimization``: the reference's import name for the MI355X drop-in.

Code written against tomohiron907/toptimization (``from toptimization.fem
import solver``, ``from toptimization.optimizer.simp import run``,
``python -m toptimization run problem.yaml``) runs unchanged on the HIP
engine: every public module of the reference's path is the corresponding
module of ``toptimization_amd``, registered under this package's name, so
both names refer to the same module objects and share one set of device
fields (``toptimization.fem.fields`` is ``toptimization_amd.fem.fields``).
"""

from __future__ import annotations

import importlib
import sys

# modules of the reference package (src/toptimization/) served by the drop-in
_MODULES = (
    "problem", "mesh", "material", "cli", "viz",
    "fem", "fem.fields", "fem.kernels", "fem.fused_kernels", "fem.preconditioner", "fem.solver",
    "optimizer", "optimizer.filter", "optimizer.oc", "optimizer.simp",
)

for _name in _MODULES:
    _mod = importlib.import_module(f"toptimization_amd.{_name}")
    sys.modules[f"{__name__}.{_name}"] = _mod
    if "." not in _name:
        globals()[_name] = _mod

__version__ = importlib.import_module("toptimization_amd").__version__
