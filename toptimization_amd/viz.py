"""Result images (density map and convergence plot) for ``simp.run``.

Same files as the reference ``viz/plot2d.py:9-84``; matplotlib is optional
and imported lazily (Agg backend).
"""

from __future__ import annotations

from pathlib import Path

import numpy as np


def save_density_png(rho_2d: np.ndarray, path: Path, title: str = "") -> None:
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    fig, ax = plt.subplots(figsize=(8, 8 * rho_2d.shape[1] / max(rho_2d.shape[0], 1) + 0.6))
    ax.imshow(1.0 - rho_2d.T, cmap="gray", vmin=0.0, vmax=1.0, origin="lower", aspect="equal")
    ax.set_title(title)
    ax.axis("off")
    fig.savefig(path, dpi=100, bbox_inches="tight")
    plt.close(fig)


def save_convergence_png(compliance, volume, path: Path) -> None:
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    fig, (a, b) = plt.subplots(1, 2, figsize=(10, 4))
    a.semilogy(compliance)
    a.set_xlabel("Iteration")
    a.set_ylabel("Compliance")
    b.plot(volume)
    b.set_xlabel("Iteration")
    b.set_ylabel("Volume fraction")
    fig.tight_layout()
    fig.savefig(path, dpi=100)
    plt.close(fig)
