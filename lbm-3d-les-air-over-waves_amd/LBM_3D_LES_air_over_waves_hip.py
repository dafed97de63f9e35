"""Drop-in solver module: loads like LBM_3D_LES_air_over_waves_CODEX.py.

The reference solver is a module whose import parses the LBM_* environment
and allocates the fields (CODEX:35-529, 991-1021), and whose ``main()`` runs
the init sequence and the step loop (CODEX:3757-3959).  Its harness loads it
from a file path (run_ab_case.py:261-279), sets ``steps``, ``plot_freq``,
``plot_step_start`` and ``visualize`` on the module, calls ``main()`` and,
from ``visualize``, reads fields with ``.to_numpy()``, counters with
``[None]`` and the configuration constants as module attributes.

Importing this module does the same against liblbm.so: it parses the
environment, creates the case on the GPU and exposes every attribute of the
``Solver`` facade at module level.  ``main()`` hands the module-level
``steps`` / ``plot_freq`` / ``plot_step_start`` / ``visualize`` to the solver
and runs it.  There is no CPU path: without liblbm.so the import fails.
"""
from __future__ import annotations

import os
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))
if _HERE not in sys.path:
    sys.path.insert(0, _HERE)

from lbm_solver import make_solver  # noqa: E402

_solver = make_solver(os.environ)

for _name in dir(_solver):
    if not _name.startswith("_") and _name not in ("main", "visualize"):
        globals()[_name] = getattr(_solver, _name)

steps = _solver.steps
plot_freq = _solver.plot_freq
plot_step_start = _solver.plot_step_start
visualize = _solver.visualize


def main():
    """main() (CODEX:3757-3959) with the harness's module-level settings."""
    _solver.steps = steps
    _solver.plot_freq = plot_freq
    _solver.plot_step_start = plot_step_start
    _solver.visualize = visualize
    _solver.main()


if __name__ == "__main__":
    main()
