"""MI355X-native spectral-element tendency and Rosenbrock-W step of CGDycore.jl.

The compute path is ``csrc/libcgdycore_hip.so`` (hand-written gfx950 HIP
kernels behind the C ABI declared in ``include/cgdycore_hip.h``).  This
package holds the host-side mirror of the reference interface
(``FcnPrepareGPU!``, ``FcnGPU!``, ``JacSchurGPU!``, ``SchurSolveGPU!``,
``RosenbrockSchur!``, ``TimeStepper!``) in :mod:`.host` and the one-time set
up of the grid, metric and initial state in :mod:`.grid`,
:mod:`.discretization` and :mod:`.state`.

The package directory name contains a dot, so it is loaded by path; see
``load_package`` in ``__graft_entry__.py``.
"""
from . import discretization, grid, state  # noqa: F401

__all__ = ["grid", "discretization", "state", "host", "output", "lib_path"]


def lib_path():
    import os

    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "libcgdycore_hip.so")


def __getattr__(name):
    if name in ("host", "output"):
        import importlib

        return importlib.import_module(__name__ + "." + name)
    raise AttributeError(name)
