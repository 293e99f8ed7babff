"""The reference's output files, written from a GPU run.

Per rank, in the output directory (Snac/libSnac/src/Context.c:425-460,
1311-1390, 1444-1531; StGermain/Base/Automation/src/VariableDumpStream.c:
135-184; plastic_BI/ConstructExtensions.c:59, Output.c:60-80;
maxwell/ConstructExtensions.c:79, Output.c:62-85):

  sim.<rank>             text: global element counts, process counts, local counts
  timeStep.<rank>        text "%16u %16g %16g\\n": timeStep, currentTime, dt, per dump
  minLengthScale.0       text "%e": the initial minimum length scale
  coord.<rank>           float32 x, y, z per node, appended per dump
  vel.<rank>             float32 vx, vy, vz per node
  force.<rank>           float32 fx, fy, fz per node (damped force of the step)
  strainRate.<rank>      float32 per element (strain-rate invariant)
  stress.<rank>          float32 per element (stress invariant)
  hydroPressure.<rank>   float32 per element
  phaseIndex.<rank>      int32 per element (material index)
  stressTensor.<rank>    float32 x 6 per element: volume-weighted tet average,
                         accumulated in float as the reference does
  plStrain.<rank>        float32 per element (plastic laws): volume-averaged plastic strain
  viscosity.<rank>       float32 per element (Maxwell): log10 of the tet-average viscosity
  temperature.<rank>     float32 per node (temperature plugin, temperature/Output.c:47-75),
                         the temperature after the step's energy update

Checkpoint files (Context.c:1294-1298, 1393-1434, 1471-1520;
plastic_BI/Output.c:83-106; maxwell/Output.c:88-119), written when
isTimeToCheckpoint (Context.c:1545-1561) holds:

  checkpointTimeStep.<rank>  text, as timeStep.<rank>
  coordCP.<rank>, velCP.<rank>, phaseIndexCP.<rank>   as the dumps
  stressTensorCP.<rank>  float32 x 6 per tet (10 per element), for restarting
  plStrainCP.<rank>      float32 per tet (plastic laws)
  viscosityCP.<rank>     float32 per element, as the viscosity dump
  temperatureCP.<rank>   float32 per node (temperature/Output.c:78-102)

A dump happens after step n when n == 1 or n % dumpEvery == 0
(isTimeToDump, Context.c:1536-1543); the dump holds the state after the step
(coordinates, velocities, rotated stresses, volumes of the new mesh) and the
element scalars computed during it.  These files are what the reference's
post-processors (snac2vtk, snac2restart) read.
"""
from __future__ import annotations

import os

import numpy as np

from . import _abi
from .context import Context


def is_time_to_dump(time_step: int, dump_every: int) -> bool:
    return time_step == 1 or (time_step > 1 and dump_every > 0 and time_step % dump_every == 0)


def is_time_to_checkpoint(time_step: int, checkpoint_every: int, max_time_steps: int | None) -> bool:
    if max_time_steps is not None and time_step == max_time_steps:
        return True
    if time_step == 0 or checkpoint_every == 0:
        return False
    return time_step % checkpoint_every == 0


def log10_mean_viscosity(viscosity: np.ndarray) -> np.ndarray:
    """maxwell/Output.c:62-80: float sum of viscosity_t / 10 in tet order,
    then log10 (in double, rounded to float) when positive."""
    E = viscosity.shape[0]
    v = np.zeros(E, dtype=np.float32)
    for t in range(10):
        v = (v.astype(np.float64) + viscosity[:, t] / 10.0).astype(np.float32)
    pos = v > 0.0
    out = v.copy()
    out[pos] = np.log10(v[pos].astype(np.float64)).astype(np.float32)
    return out


def stress_tensor_float(stress: np.ndarray, volume: np.ndarray) -> np.ndarray:
    """_Snac_Context_DumpStressTensor (Context.c:1482-1505): in float,
    totalVolume += volume_t, then stressVector[c] += stress_t[c] * volume_t /
    totalVolume, each sum rounded to float after every term."""
    E = stress.shape[0]
    tot = np.zeros(E, dtype=np.float32)
    for t in range(10):
        tot = (tot.astype(np.float64) + volume[:, t]).astype(np.float32)
    out = np.zeros((E, 6), dtype=np.float32)
    for t in range(10):
        term = stress[:, t, :] * volume[:, t, None] / tot.astype(np.float64)[:, None]
        out = (out.astype(np.float64) + term).astype(np.float32)
    return out


class DumpWriter:
    """Appends one rank's dump files, in the reference's formats."""

    def __init__(self, outdir: str, ctx: Context, rank: int = 0, dump_every: int = 1,
                 checkpoint_every: int = 0, max_time_steps: int | None = None):
        self.dir = outdir
        self.ctx = ctx
        self.rank = rank
        self.dump_every = dump_every
        self.checkpoint_every = checkpoint_every
        self.max_time_steps = max_time_steps
        os.makedirs(outdir, exist_ok=True)
        p = ctx.problem
        ei, ej, ek = p.elements
        with open(self._path("sim"), "w") as f:
            f.write(f"{ei} {ej} {ek} 1 1 1 {ei} {ej} {ek}\n"
                    "(global element numbers)[3]=(proc numbers)[3]*(local element numbers)[3]\n")
        self._time_step_file = open(self._path("timeStep"), "w")
        self.names = ["strainRate", "stress", "hydroPressure", "stressTensor", "coord", "vel", "force",
                      "phaseIndex"]
        rh = 0
        for m in p.materials:
            rh |= m.rheology
        self.plastic = bool(rh & (_abi.RHEOLOGY_PLASTIC | _abi.RHEOLOGY_VISCOPLASTIC))
        self.maxwell = bool(rh & _abi.RHEOLOGY_MAXWELL)
        if self.plastic:
            self.names.append("plStrain")
        if self.maxwell:
            self.names.append("viscosity")
        self.thermal = p.thermal is not None
        if self.thermal:
            self.names.append("temperature")
        self.files = {n: open(self._path(n), "wb") for n in self.names}
        self._cp_time_step_file = open(self._path("checkpointTimeStep"), "w")
        self.cp_names = ["phaseIndexCP", "coordCP", "velCP", "stressTensorCP"]
        if self.plastic:
            self.cp_names.append("plStrainCP")
        if self.maxwell:
            self.cp_names.append("viscosityCP")
        if self.thermal:
            self.cp_names.append("temperatureCP")
        self.cp_files = {n: open(self._path(n), "wb") for n in self.cp_names}
        sc = ctx.scalars()
        with open(os.path.join(outdir, "minLengthScale.0"), "w") as f:
            f.write("%e" % sc["minLengthScale"])
        self.current_time = 0.0

    def _path(self, name):
        return os.path.join(self.dir, f"{name}.{self.rank}")

    def close(self):
        self._time_step_file.close()
        self._cp_time_step_file.close()
        for f in list(self.files.values()) + list(self.cp_files.values()):
            f.close()

    def step(self, nsteps: int):
        """Run nsteps, dumping after each step that is time to dump."""
        c = self.ctx
        for _ in range(nsteps):
            ts = c.scalars()["timeStep"] + 1
            dt = c.scalars()["dt"]
            self.current_time += dt
            dump = is_time_to_dump(ts, self.dump_every)
            cp = is_time_to_checkpoint(ts, self.checkpoint_every, self.max_time_steps)
            if dump:
                self._time_step_file.write("%16u %16g %16g\n" % (ts, self.current_time, dt))
                self._time_step_file.flush()
                c.set_diagnostics(True)
            if cp:
                self._cp_time_step_file.write("%16u %16g %16g\n" % (ts, self.current_time, dt))
                self._cp_time_step_file.flush()
            c.step(1)
            if dump:
                c.set_diagnostics(False)
                self.dump()
            if cp:
                self.checkpoint()
        return self

    def _phase_index(self):
        p = self.ctx.problem
        return (np.zeros(p.n_elements, dtype=np.int32) if p.element_material is None
                else np.asarray(p.element_material, dtype=np.int32))

    def checkpoint(self):
        fields = ("coord", "velocity", "stress", "plastic_strain", "viscosity")
        st = self.ctx.state(fields=fields + (("temperature",) if self.thermal else ()))
        w = self.cp_files
        w["stressTensorCP"].write(st["stress"].astype(np.float32).tobytes())
        w["coordCP"].write(st["coord"].astype(np.float32).tobytes())
        w["velCP"].write(st["velocity"].astype(np.float32).tobytes())
        w["phaseIndexCP"].write(self._phase_index().tobytes())
        if self.plastic:
            w["plStrainCP"].write(st["plastic_strain"].astype(np.float32).tobytes())
        if self.maxwell:
            w["viscosityCP"].write(log10_mean_viscosity(st["viscosity"]).tobytes())
        if self.thermal:
            w["temperatureCP"].write(st["temperature"].astype(np.float32).tobytes())
        for f in w.values():
            f.flush()

    def dump(self):
        c = self.ctx
        p = c.problem
        fields = ("coord", "velocity", "stress", "plastic_strain", "viscosity", "volume")
        st = c.state(fields=fields + (("temperature",) if self.thermal else ()))
        dg = c.diagnostics()
        w = self.files
        w["strainRate"].write(dg["strain_rate"].astype(np.float32).tobytes())
        w["stress"].write(dg["stress"].astype(np.float32).tobytes())
        w["hydroPressure"].write(dg["hydro_pressure"].astype(np.float32).tobytes())
        w["stressTensor"].write(stress_tensor_float(st["stress"], st["volume"]).tobytes())
        w["coord"].write(st["coord"].astype(np.float32).tobytes())
        w["vel"].write(st["velocity"].astype(np.float32).tobytes())
        w["force"].write(dg["force"].astype(np.float32).tobytes())
        w["phaseIndex"].write(self._phase_index().tobytes())
        if self.plastic:
            w["plStrain"].write(dg["plastic_strain"].astype(np.float32).tobytes())
        if self.maxwell:
            w["viscosity"].write(log10_mean_viscosity(st["viscosity"]).tobytes())
        if self.thermal:
            w["temperature"].write(st["temperature"].astype(np.float32).tobytes())
        for f in w.values():
            f.flush()


def read_checkpoint(outdir: str, problem, rank: int = 0) -> list:
    """The checkpoints in outdir, one dict per record: time_step, time, dt
    and the float32 arrays (coord [N,3], velocity [N,3], stress [E,10,6],
    phase_index [E], plastic_strain [E,10], log10_viscosity [E],
    temperature [N] when present), e.g. to upload a restart state with Context.set_state."""
    N, E = problem.n_nodes, problem.n_elements

    def path(name):
        return os.path.join(outdir, f"{name}.{rank}")

    recs = []
    with open(path("checkpointTimeStep")) as f:
        for line in f:
            ts, t, dt = line.split()
            recs.append({"time_step": int(ts), "time": float(t), "dt": float(dt)})
    shapes = {"coordCP": ("coord", np.float32, (N, 3)), "velCP": ("velocity", np.float32, (N, 3)),
              "stressTensorCP": ("stress", np.float32, (E, 10, 6)), "phaseIndexCP": ("phase_index", np.int32, (E,)),
              "plStrainCP": ("plastic_strain", np.float32, (E, 10)), "viscosityCP": ("log10_viscosity", np.float32, (E,)),
              "temperatureCP": ("temperature", np.float32, (N,))}
    for name, (key, dt, shape) in shapes.items():
        if not os.path.exists(path(name)):
            continue
        a = np.fromfile(path(name), dtype=dt)
        n = int(np.prod(shape))
        for k, r in enumerate(recs):
            r[key] = a[k * n:(k + 1) * n].reshape(shape)
    return recs
