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

A dump happens after step n when n == 1 or n % dumpEvery == 0
(isTimeToDump, Context.c:1536-1543); the dump holds the state after the step
(coordinates, velocities, rotated stresses, volumes of the new mesh) and the
element scalars computed during it.  These files are what the reference's
post-processors (snac2vtk) read.
"""
from __future__ import annotations

import os

import numpy as np

from . import _abi
from .context import Context


def is_time_to_dump(time_step: int, dump_every: int) -> bool:
    return time_step == 1 or (time_step > 1 and dump_every > 0 and time_step % dump_every == 0)


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

    def __init__(self, outdir: str, ctx: Context, rank: int = 0, dump_every: int = 1):
        self.dir = outdir
        self.ctx = ctx
        self.rank = rank
        self.dump_every = dump_every
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
        self.files = {n: open(self._path(n), "wb") for n in self.names}
        sc = ctx.scalars()
        with open(os.path.join(outdir, "minLengthScale.0"), "w") as f:
            f.write("%e" % sc["minLengthScale"])
        self.current_time = 0.0

    def _path(self, name):
        return os.path.join(self.dir, f"{name}.{self.rank}")

    def close(self):
        self._time_step_file.close()
        for f in self.files.values():
            f.close()

    def step(self, nsteps: int):
        """Run nsteps, dumping after each step that is time to dump."""
        c = self.ctx
        for _ in range(nsteps):
            ts = c.scalars()["timeStep"] + 1
            dt = c.scalars()["dt"]
            self.current_time += dt
            dump = is_time_to_dump(ts, self.dump_every)
            if dump:
                self._time_step_file.write("%16u %16g %16g\n" % (ts, self.current_time, dt))
                self._time_step_file.flush()
                c.set_diagnostics(True)
            c.step(1)
            if dump:
                c.set_diagnostics(False)
                self.dump()
        return self

    def dump(self):
        c = self.ctx
        p = c.problem
        st = c.state(fields=("coord", "velocity", "stress", "plastic_strain", "viscosity", "volume"))
        dg = c.diagnostics()
        w = self.files
        w["strainRate"].write(dg["strain_rate"].astype(np.float32).tobytes())
        w["stress"].write(dg["stress"].astype(np.float32).tobytes())
        w["hydroPressure"].write(dg["hydro_pressure"].astype(np.float32).tobytes())
        w["stressTensor"].write(stress_tensor_float(st["stress"], st["volume"]).tobytes())
        w["coord"].write(st["coord"].astype(np.float32).tobytes())
        w["vel"].write(st["velocity"].astype(np.float32).tobytes())
        w["force"].write(dg["force"].astype(np.float32).tobytes())
        mat = (np.zeros(p.n_elements, dtype=np.int32) if p.element_material is None
               else np.asarray(p.element_material, dtype=np.int32))
        w["phaseIndex"].write(mat.tobytes())
        if self.plastic:
            w["plStrain"].write(dg["plastic_strain"].astype(np.float32).tobytes())
        if self.maxwell:
            v = np.zeros(p.n_elements, dtype=np.float32)
            for t in range(10):
                v = (v.astype(np.float64) + st["viscosity"][:, t] / 10.0).astype(np.float32)
            v = np.where(v > 0.0, np.log10(v.astype(np.float64)).astype(np.float32), v)
            w["viscosity"].write(v.tobytes())
        for f in w.values():
            f.flush()
