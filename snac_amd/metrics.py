"""Algorithmic HBM bytes of the explicit step, for roofline reporting.

Per element-step, compulsory persistent-state traffic (each array read once
and written once per step; structured connectivity costs nothing; r = N/E):
  elastic  960 (tet stress 10x6x8, read + write) + 96 r (coords + velocity, r + w)
  plastic  + 160 (plastic strain 10x8, r + w)
  maxwell  + 8 (per-element viscosity read) + 48 r (previous coords read)
  temperature plugin + 160 (tet density 10x8, written by the element pass,
           read by the next force and heat passes) + 48 r (nodal temperature,
           previous temperature, pre-BC temperature, r + w)
The fused element kernel alone must move (its own compulsory bytes):
  960 + 48 r (coords + velocity read) [+ 160 plastic] [+ 8 + 24 r maxwell]
The heat-flux kernel moves 240 (flux write) + 24 r + 16 r (coords, T, T0
read) [+ 80 thermal stress write]; the heat gather 240 (flux read) + 80
(density read) + 24 r (coords) + 32 r (T read, T, T0, pre-BC T write).
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
    if problem.thermal is not None:
        b += 160.0 + 48.0 * r
    return b


def heat_elem_kernel_bytes_per_element(problem) -> float:
    b = 240.0 + 40.0 * _r(problem)
    if problem.thermal is not None and problem.thermal.computeThermalStress:
        b += 80.0
    return b


def heat_node_kernel_bytes_per_element(problem) -> float:
    return 320.0 + 56.0 * _r(problem)


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
