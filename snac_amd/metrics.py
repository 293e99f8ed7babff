"""Algorithmic HBM bytes of the explicit step, for roofline reporting.

Per element-step, compulsory persistent-state traffic (each array read once
and written once per step; structured connectivity costs nothing; r = N/E):
  elastic  960 (tet stress 10x6x8, read + write) + 96 r (coords + velocity, r + w)
  plastic  + 160 (plastic strain 10x8, r + w)
  maxwell  + 8 (per-element viscosity read) + 48 r (previous coords read)
The fused element kernel alone must move (its own compulsory bytes):
  960 + 48 r (coords + velocity read) [+ 160 plastic] [+ 8 + 24 r maxwell]
"""
from . import _abi

HBM_PEAK_GBPS = 8000.0   # MI355X HBM3E, vendor figure


def _r(problem):
    return problem.n_nodes / problem.n_elements


def _law(problem):
    rh = 0
    for m in problem.materials:
        rh |= m.rheology
    if rh & (_abi.RHEOLOGY_PLASTIC | _abi.RHEOLOGY_VISCOPLASTIC):
        return "plastic"
    if rh & _abi.RHEOLOGY_MAXWELL:
        return "maxwell"
    return "elastic"


def step_bytes_per_element(problem) -> float:
    r = _r(problem)
    law = _law(problem)
    b = 960.0 + 96.0 * r
    if law == "plastic":
        b += 160.0
    elif law == "maxwell":
        b += 8.0 + 48.0 * r
    return b


def elem_kernel_bytes_per_element(problem) -> float:
    r = _r(problem)
    law = _law(problem)
    b = 960.0 + 48.0 * r
    if law == "plastic":
        b += 160.0
    elif law == "maxwell":
        b += 8.0 + 24.0 * r
    return b


def node_kernel_bytes_per_element(problem) -> float:
    """Gather kernel: stress read once (480) + coords/velocity r+w (96 r)."""
    return 480.0 + 96.0 * _r(problem)
