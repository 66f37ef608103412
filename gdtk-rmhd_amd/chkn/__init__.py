"""chkn -- MI355X (gfx950) explicit finite-volume step of the chicken solver.

Python front-end around libchkn.so: job preparation (prep), the array form of a job (job),
the C-ABI binding (lib), post-processing readers (post) and synthetic workloads (cases).
"""
from .job import JobData  # noqa: F401

__all__ = ["JobData"]
