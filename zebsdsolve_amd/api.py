"""The plain-function API: one call per job (python/zebsdsolve/api.py:1-132).

    import zebsdsolve_amd as ze
    res = ze.index(patterns, ["Ni_master_20kV.npz"], pc=(0.5, 0.5, 0.55))

`scan` is an array of patterns (rows, cols, h, w) or (n, h, w) (scan files are outside this
build). Keyword arguments naming a geometry field set the geometry; every other keyword is a
setting.
"""
from dataclasses import dataclass, field

import numpy as np

GEOMETRY_FIELDS = ("pc", "rot", "sample_tilt_deg", "detector_tilt_deg", "px_um", "step_x_um", "step_y_um")


@dataclass
class Result:
    """Indexing results per map cell, with the geometry they were indexed under (api.py:22-47)."""

    quaternions: np.ndarray
    euler: np.ndarray
    ncc: np.ndarray
    nmi: np.ndarray
    phase: np.ndarray
    z: np.ndarray
    sweep: np.ndarray
    geometry: dict
    phases: list
    settings: dict
    fit: dict = field(default_factory=dict)


def _open(solver, scan, phases, kwargs):
    geo = {k: kwargs.pop(k) for k in list(kwargs) if k in GEOMETRY_FIELDS}
    solver.configure(**kwargs)
    if isinstance(scan, (str, bytes)) or hasattr(scan, "__fspath__"):
        from . import Error
        raise Error("scan files are not read by ze.index here: pass the patterns as an array (Solver.open reads kikuchipy's h5ebsd)")
    solver.set_patterns(scan)
    if geo:
        solver.set_geometry(**geo)
    for p in [phases] if isinstance(phases, str) or hasattr(phases, "__fspath__") else phases:
        solver.add_phase(p)


def index(scan, phases, fit=None, progress=None, **kwargs):
    """Index a scan against candidate phases. fit: None only (the geometry fits are not
    implemented here). progress(status) is called after each step; returning False stops early."""
    from . import Error, Solver
    if fit is not None:
        raise Error("fit: the pattern-centre and plane fits are not implemented here")
    with Solver() as s:
        _open(s, scan, phases, kwargs)
        r = s.index(progress)
        return Result(geometry=s.geometry, phases=s.phases, settings=s.settings, fit={}, **r)
