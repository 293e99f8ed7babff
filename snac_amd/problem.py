"""Problem descriptions for the explicit step: materials, velocity BCs, mesh,
and the named workloads the benchmarks and parity tests run.

Defaults follow the reference's dictionary defaults
(Snac/libSnac/src/Context.c:752-766,816-914), which are float literals; the
workloads override them with the parity settings (alpha = beta = 0,
topo_kappa = 0, demf = 0.8 as a double, lambda = mu = 3e10 as exact doubles).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _abi

SECONDS_PER_YEAR = 365.25 * 86400.0
CM_PER_YEAR = 0.01 / SECONDS_PER_YEAR          # 3.16881e-10 m/s


def f32(x: float) -> float:
    """The double a C float literal becomes."""
    return float(np.float32(x))


@dataclass
class MaterialSpec:
    """One phase (Snac_Material).  Defaults are the reference's dictionary
    defaults (Context.c:826-914)."""
    rheology: int = _abi.RHEOLOGY_ELASTIC
    density: float = f32(2700.0)
    alpha: float = f32(3.0e-05)
    beta: float = f32(0.0)
    lam: float = f32(3.0e+10)
    mu: float = f32(3.0e+10)
    plstrain: Sequence[float] = (0.0, 0.0, 0.0)
    frictionAngle: Sequence[float] = (0.0, 0.0, 0.0)
    dilationAngle: Sequence[float] = (0.0, 0.0, 0.0)
    cohesion: Sequence[float] = (0.0, 0.0, 0.0)
    ten_off: float = 0.0
    vis_min: float = 1.0e18
    vis_max: float = 3.0e27
    refvisc: float = 1.0e19
    refsrate: float = 1.0e-15
    reftemp: float = 1400.0
    activationE: float = 5.4e03
    srexponent: float = 1.0
    thermal_conduct: float = 2.0
    heatCapacity: float = 1000.0

    def to_c(self) -> _abi.Material:
        m = _abi.Material()
        m.rheology = self.rheology
        m.phsDensity = self.density
        m.alpha = self.alpha
        m.beta = self.beta
        m.lambda_ = self.lam
        m.mu = self.mu
        nseg = len(self.plstrain) - 1
        if not (1 <= nseg <= _abi.SNH_MAX_SEGMENTS):
            raise ValueError("softening table needs 2..9 points")
        for arr in (self.frictionAngle, self.dilationAngle, self.cohesion):
            if len(arr) != nseg + 1:
                raise ValueError("softening tables must have equal length")
        m.nsegments = nseg
        for i in range(nseg + 1):
            m.plstrain[i] = self.plstrain[i]
            m.frictionAngle[i] = self.frictionAngle[i]
            m.dilationAngle[i] = self.dilationAngle[i]
            m.cohesion[i] = self.cohesion[i]
        m.ten_off = self.ten_off
        m.vis_min = self.vis_min
        m.vis_max = self.vis_max
        m.refvisc = self.refvisc
        m.refsrate = self.refsrate
        m.reftemp = self.reftemp
        m.activationE = self.activationE
        m.srexponent = self.srexponent
        m.thermal_conduct = self.thermal_conduct
        m.heatCapacity = self.heatCapacity
        return m


@dataclass
class ThermalSpec:
    """The SnacTemperature plugin: initial nodal temperature (nodeICs),
    temperatureBCs (wall, value; later entries win) and computeThermalStress.
    top2bottom(problem, top, bottom) gives the reference's
    _SnacTemperature_Top2BottomSweep initial field (temperature/
    VariableConditions.c:55-71): T = bottom + (top - bottom) / (J - 1) * j."""
    temperature: np.ndarray                      # [N]
    bcs: List[Tuple[str, float]] = field(default_factory=list)
    computeThermalStress: bool = False

    def to_c(self):
        T = np.ascontiguousarray(self.temperature, dtype=np.float64)
        bcs = (_abi.TemperatureBC * max(1, len(self.bcs)))()
        for idx, (wall, val) in enumerate(self.bcs):
            bcs[idx].wall = _abi.WALL_NAMES[wall] if isinstance(wall, str) else wall
            bcs[idx].value = val
        th = _abi.Thermal()
        th.computeThermalStress = int(self.computeThermalStress)
        th.nbcs = len(self.bcs)
        th.bcs = C.cast(bcs, C.POINTER(_abi.TemperatureBC))
        th.temperature = T.ctypes.data_as(C.POINTER(C.c_double))
        return th, (T, bcs)


def top2bottom(nodes, top: float = 0.0, bottom: float = 1300.0) -> np.ndarray:
    I, J, K = nodes
    n = np.arange(I * J * K)
    j = (n // I) % J
    return bottom + (top - bottom) / (J - 1) * j


@dataclass
class Problem:
    """Mesh + constants + BCs + initial state of one subdomain.

    ``elements`` is (I, J, K) hex elements along x, y (vertical), z.
    """
    elements: Tuple[int, int, int]
    extent: Tuple[Tuple[float, float], Tuple[float, float], Tuple[float, float]]
    materials: List[MaterialSpec]
    bcs: List[Tuple[str, int, Tuple[float, float, float]]] = field(default_factory=list)
    dtType: int = _abi.DT_DYNAMIC
    timeStep: float = 0.0
    gravity: float = f32(9.81)
    demf: float = f32(0.8)
    strain_inert: float = f32(1.0e-5)
    topo_kappa: float = 0.0
    topoGradCriterion: float = 1.0e-16
    precision: int = _abi.PRECISION_FP64
    element_material: Optional[np.ndarray] = None        # uint8 [E]
    fields: Dict[int, np.ndarray] = field(default_factory=dict)   # SNH_FIELD_* -> [E]
    plastic_strain0: Optional[np.ndarray] = None          # [E,10]
    hydrostatic: bool = False      # SnacHydroStaticIC initial stress
    thermal: Optional["ThermalSpec"] = None   # SnacTemperature plugin
    name: str = "custom"

    @property
    def nodes(self) -> Tuple[int, int, int]:
        return tuple(n + 1 for n in self.elements)

    @property
    def n_elements(self) -> int:
        i, j, k = self.elements
        return i * j * k

    @property
    def n_nodes(self) -> int:
        i, j, k = self.nodes
        return i * j * k

    def to_c(self):
        """Return (Config, keepalive) -- keep the second alive while the
        first is in use."""
        mats = (_abi.Material * len(self.materials))(*[m.to_c() for m in self.materials])
        bcs = (_abi.VelocityBC * max(1, len(self.bcs)))()
        for idx, (wall, mask, val) in enumerate(self.bcs):
            bcs[idx].wall = _abi.WALL_NAMES[wall] if isinstance(wall, str) else wall
            bcs[idx].mask = mask
            for d in range(3):
                bcs[idx].value[d] = val[d]
        cfg = _abi.Config()
        ni, nj, nk = self.nodes
        cfg.meshSizeI, cfg.meshSizeJ, cfg.meshSizeK = ni, nj, nk
        (cfg.minX, cfg.maxX), (cfg.minY, cfg.maxY), (cfg.minZ, cfg.maxZ) = self.extent
        cfg.dtType = self.dtType
        cfg.timeStep = self.timeStep
        cfg.gravity = self.gravity
        cfg.demf = self.demf
        cfg.strain_inert = self.strain_inert
        cfg.topo_kappa = self.topo_kappa
        cfg.topoGradCriterion = self.topoGradCriterion
        cfg.nmaterials = len(self.materials)
        cfg.materials = C.cast(mats, C.POINTER(_abi.Material))
        cfg.nbcs = len(self.bcs)
        cfg.bcs = C.cast(bcs, C.POINTER(_abi.VelocityBC))
        cfg.precision = self.precision
        return cfg, (mats, bcs)


# ---------------------------------------------------------------------------
# named workloads
# ---------------------------------------------------------------------------

def _parity_material(**kw) -> MaterialSpec:
    """Material with the parity settings: alpha = beta = 0, lambda = mu = 3e10
    exactly, density 2700."""
    base = dict(alpha=0.0, beta=0.0, lam=3.0e10, mu=3.0e10, density=2700.0)
    base.update(kw)
    return MaterialSpec(**base)


def _parity_problem(**kw) -> Problem:
    base = dict(dtType=_abi.DT_DYNAMIC, timeStep=0.0, gravity=10.0, demf=0.8,
                strain_inert=f32(1.0e-5), topo_kappa=0.0)
    base.update(kw)
    return Problem(**base)


def splitmix64(seed: int):
    """SplitMix64 stream (Steele, Lea & Flood 2014): deterministic seeds."""
    mask = (1 << 64) - 1
    state = seed & mask
    while True:
        state = (state + 0x9E3779B97F4A7C15) & mask
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        yield z ^ (z >> 31)


def _box_bcs(v_left: float):
    """x-faces moved at +/-v_left (inward for v_left > 0), free-slip base and
    z-faces, free top.  With gravity and dynamic dt the z-faces must be
    constrained: a laterally free box spreads and the mass-scaled pseudo
    velocity then grows without bound."""
    return [("left", _abi.VX, (v_left, 0.0, 0.0)),
            ("right", _abi.VX, (-v_left, 0.0, 0.0)),
            ("front", _abi.VZ, (0.0, 0.0, 0.0)),
            ("back", _abi.VZ, (0.0, 0.0, 0.0)),
            ("bottom", _abi.VY, (0.0, 0.0, 0.0))]


def elastic_box_compression(elements=(16, 8, 16), extent_km=(100.0, 50.0, 100.0)) -> Problem:
    """Elastic box compression: x-faces pushed inward at 1 cm/yr, free top,
    free-slip base, gravity 10, rho 2700, lambda = mu = 3e10."""
    lx, ly, lz = (1e3 * e for e in extent_km)
    v = CM_PER_YEAR
    return _parity_problem(
        name="elastic_box_compression",
        elements=tuple(elements),
        extent=((0.0, lx), (-ly, 0.0), (0.0, lz)),
        materials=[_parity_material(rheology=_abi.RHEOLOGY_ELASTIC)],
        bcs=_box_bcs(+v),
        hydrostatic=True,
    )


def mohr_coulomb_extension(elements=(256, 64, 128), extent_km=(400.0, 100.0, 200.0), seed=1,
                           seed_size=2) -> Problem:
    """Mohr-Coulomb extension: cohesion softening 44 -> 4 MPa over plastic
    strain 0..0.5, friction 30 deg, dilation 0, a weak seed of seed_size^3
    elements with plastic strain 0.5 at a SplitMix64-drawn position in the
    lower half, x-faces pulled apart at 1 cm/yr."""
    lx, ly, lz = (1e3 * e for e in extent_km)
    v = CM_PER_YEAR
    ni, nj, nk = elements
    mat = _parity_material(
        rheology=_abi.RHEOLOGY_PLASTIC,
        plstrain=(0.0, 0.5, 1.0e3),
        frictionAngle=(30.0, 30.0, 30.0),
        dilationAngle=(0.0, 0.0, 0.0),
        cohesion=(44.0e6, 4.0e6, 4.0e6),
        ten_off=0.0)
    rng = splitmix64(seed)
    si = next(rng) % max(1, ni - seed_size + 1)
    sj = next(rng) % max(1, nj // 2 - seed_size + 1)
    sk = next(rng) % max(1, nk - seed_size + 1)
    ps = np.zeros((nk, nj, ni, 10))
    ps[sk:sk + seed_size, sj:sj + seed_size, si:si + seed_size, :] = 0.5
    p = _parity_problem(
        name="mohr_coulomb_extension",
        elements=tuple(elements),
        extent=((0.0, lx), (-ly, 0.0), (0.0, lz)),
        materials=[mat],
        bcs=_box_bcs(-v),
        hydrostatic=True,
        plastic_strain0=ps.reshape(-1, 10),
    )
    p.seed_position = (int(si), int(sj), int(sk))
    return p


def lognormal_field(elements, extent, mean=1.0e21, sigma_ln=1.0, corr_len=20e3, seed=3) -> np.ndarray:
    """Per-element log-normal field exp(N(ln mean, sigma_ln)) with an
    exponential correlation of length corr_len, made by a separable first-order
    autoregressive filter of white noise along each axis (seeded)."""
    ni, nj, nk = elements
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((nk, nj, ni))
    spacing = [(extent[d][1] - extent[d][0]) / elements[d] for d in range(3)]
    for axis, d in ((2, 0), (1, 1), (0, 2)):
        rho = float(np.exp(-spacing[d] / corr_len))
        s = np.sqrt(1.0 - rho * rho)
        g = np.moveaxis(g, axis, 0).copy()
        for i in range(1, g.shape[0]):
            g[i] = rho * g[i - 1] + s * g[i]
        g = np.moveaxis(g, 0, axis)
    return np.exp(np.log(mean) + sigma_ln * g).reshape(-1)


def maxwell_viscoelastic(elements=(192, 48, 192), extent_km=(300.0, 75.0, 300.0), seed=3) -> Problem:
    """Maxwell visco-elastic box compression with a log-normal viscosity field
    (mean 1e21 Pa s, sigma 1 in ln, correlation 20 km), n = 1, H = 0."""
    lx, ly, lz = (1e3 * e for e in extent_km)
    v = CM_PER_YEAR
    extent = ((0.0, lx), (-ly, 0.0), (0.0, lz))
    mat = _parity_material(rheology=_abi.RHEOLOGY_MAXWELL, activationE=0.0, srexponent=1.0,
                           refvisc=1.0e21)
    p = _parity_problem(
        name="maxwell_viscoelastic",
        elements=tuple(elements),
        extent=extent,
        materials=[mat],
        bcs=_box_bcs(+v),
        hydrostatic=True,
    )
    p.fields[_abi.FIELD_REFVISC] = lognormal_field(elements, extent, seed=seed)
    return p


def mohr_coulomb_shear(elements=(64, 32, 64), extent_km=(64.0, 32.0, 64.0), seed=1) -> Problem:
    """Simple shear of a Mohr-Coulomb box without gravity: the top moves in x
    at 1 cm/yr over a fixed base, z-faces free-slip.  Cohesion softens from
    1 MPa to 0.2 MPa over plastic strain 0..0.05 (friction 30 deg, dilation
    0), strain_inert 1e-4: the box yields within a few steps, with distinct
    principal stresses (the return is then a smooth function of the trial
    stress).  A seeded 2x2x2 weak seed sits in the lower half."""
    p = mohr_coulomb_extension(elements=elements, extent_km=extent_km, seed=seed)
    m = p.materials[0]
    m.cohesion = (1.0e6, 2.0e5, 2.0e5)
    m.plstrain = (0.0, 0.05, 1.0e3)
    p.gravity = 0.0
    p.hydrostatic = False
    p.strain_inert = 1.0e-4
    v = CM_PER_YEAR
    p.bcs = [("top", _abi.VX, (v, 0.0, 0.0)),
             ("bottom", _abi.VX | _abi.VY | _abi.VZ, (0.0, 0.0, 0.0)),
             ("front", _abi.VZ, (0.0, 0.0, 0.0)),
             ("back", _abi.VZ, (0.0, 0.0, 0.0))]
    p.name = "mohr_coulomb_shear"
    return p


def mohr_coulomb_weak_crust(elements=(64, 32, 64), extent_km=(400.0, 100.0, 200.0), seed=1) -> Problem:
    """The Mohr-Coulomb extension with a weak crust, for the cohesion
    adjoint: cohesion 0.1 MPa softening to 0.05 MPa over plastic strain
    0..0.5 (friction 30 deg, dilation 0) and strain_inert 1e-3, so that the
    top layers yield within a few hundred steps.  Gravity 10, hydrostatic
    initial stress, x-faces pulled apart at 1 cm/yr."""
    p = mohr_coulomb_extension(elements=elements, extent_km=extent_km, seed=seed)
    m = p.materials[0]
    m.cohesion = (1.0e5, 5.0e4, 5.0e4)
    m.plstrain = (0.0, 0.5, 1.0e3)
    p.strain_inert = 1.0e-3
    p.name = "mohr_coulomb_weak_crust"
    return p


def maxwell_thermal(elements=(64, 32, 64), extent_km=(100.0, 50.0, 100.0), seed=5, activationE=0.0,
                    thermal_stress=True) -> Problem:
    """Maxwell visco-elastic box compression with the SnacTemperature plugin:
    initial temperature the reference's top-to-bottom sweep (0 to 1300 C)
    plus a seeded Gaussian anomaly of 200 C (radius 10 km) in the lower
    half, top and bottom held at 0 and 1300 C, thermal expansion
    alpha = 3e-5 (density and, with computeThermalStress, thermal stress),
    conductivity 2 W/m/K, heat capacity 1000 J/kg/K.  activationE = 0 keeps
    the viscosity linear (exp(0) = 1); a positive value makes it Arrhenius
    in the tet-average temperature."""
    lx, ly, lz = (1e3 * e for e in extent_km)
    v = CM_PER_YEAR
    extent = ((0.0, lx), (-ly, 0.0), (0.0, lz))
    mat = _parity_material(rheology=_abi.RHEOLOGY_MAXWELL, activationE=activationE, srexponent=1.0,
                           refvisc=1.0e21, alpha=f32(3.0e-5), thermal_conduct=2.0, heatCapacity=1000.0)
    p = _parity_problem(
        name="maxwell_thermal",
        elements=tuple(elements),
        extent=extent,
        materials=[mat],
        bcs=_box_bcs(+v),
        hydrostatic=True,
    )
    p.fields[_abi.FIELD_REFVISC] = lognormal_field(elements, extent, seed=seed)
    I, J, K = p.nodes
    T = top2bottom(p.nodes, 0.0, 1300.0)
    rng = splitmix64(seed)
    cx, cz = (next(rng) % 1000) / 1000.0 * lx, (next(rng) % 1000) / 1000.0 * lz
    cy = -0.7 * ly
    n = np.arange(I * J * K)
    x = (n % I) / (I - 1) * lx
    y = -ly + ((n // I) % J) / (J - 1) * ly
    z = (n // (I * J)) / (K - 1) * lz
    T = T + 200.0 * np.exp(-((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) / (1.0e4) ** 2)
    p.thermal = ThermalSpec(temperature=T, bcs=[("top", 0.0), ("bottom", 1300.0)],
                            computeThermalStress=thermal_stress)
    return p


WORKLOADS = {
    "elastic_box_compression": elastic_box_compression,
    "mohr_coulomb_extension": mohr_coulomb_extension,
    "maxwell_viscoelastic": maxwell_viscoelastic,
    "mohr_coulomb_shear": mohr_coulomb_shear,
    "mohr_coulomb_weak_crust": mohr_coulomb_weak_crust,
    "maxwell_thermal": maxwell_thermal,
}


def basic_hydrostatic() -> Problem:
    """Snac/tests/basic-hydroStatic.xml: 13^3 nodes over 7.5 km, density 2700,
    gravity 9.81, demf 0.8, alpha 0, topo_kappa 0, no velocity conditions,
    the SnacHydroStaticIC plugin; the other values are the dictionary
    defaults (lambda = mu = 3e10f, dynamic dt, strain_inert 1e-5f)."""
    mat = MaterialSpec(alpha=0.0, density=2700.0)
    p = Problem(elements=(12, 12, 12), extent=((0.0, 7500.0), (-7500.0, 0.0), (0.0, 7500.0)), materials=[mat],
                gravity=9.81, demf=0.8, topo_kappa=0.0, hydrostatic=True)
    p.name = "basic_hydrostatic"
    return p


def basic_temperature() -> Problem:
    """Snac/tests/basic-temp.xml: 13^3 nodes over 7.5 km, density 2700,
    gravity 0, alpha 0, thermal_conduct 1.6, heatCapacity 0.2, the
    SnacTemperature plugin with T = 0 and temperature conditions 100 on the
    top wall, 0 on the bottom; the other values are the dictionary
    defaults."""
    mat = MaterialSpec(alpha=0.0, density=2700.0, thermal_conduct=1.6, heatCapacity=0.2)
    p = Problem(elements=(12, 12, 12), extent=((0.0, 7500.0), (-7500.0, 0.0), (0.0, 7500.0)), materials=[mat],
                gravity=0.0, demf=0.8, topo_kappa=0.0,
                thermal=ThermalSpec(temperature=np.zeros(13 ** 3), bcs=[("top", 100.0), ("bottom", 0.0)]))
    p.name = "basic_temperature"
    return p
