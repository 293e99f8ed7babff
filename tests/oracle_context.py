"""The Context interface the dump writer uses, over the CPU oracle (test
infrastructure)."""
import ctypes as C

import numpy as np

from oracle.oracle import Oracle, lib as orc_lib


class OracleContext:
    """The interface DumpWriter uses, over the CPU oracle."""

    def __init__(self, problem):
        self.problem = problem
        self.o = Oracle(problem, threads=4)
        if problem.hydrostatic:
            self.o.hydrostatic_ic()
        self.o.timestep_zero()
        self.E, self.N = self.o.E, self.o.N

    def scalars(self):
        return self.o.scalars()

    def set_diagnostics(self, on):
        return self

    def step(self, n):
        self.o.step(n)

    def state(self, fields=None):
        return self.o.state()

    def diagnostics(self):
        out = {k: np.zeros(self.E) for k in ("strain_rate", "stress", "hydro_pressure", "plastic_strain")}
        for e in range(self.E):
            v = self.o.element(e)
            out["strain_rate"][e] = v["strainRate"]
            out["stress"][e] = v["stress"]
            out["hydro_pressure"][e] = v["hydroPressure"]
            out["plastic_strain"][e] = v["aps"]
        f = np.zeros((self.N, 3))
        m = C.c_double()
        for n in range(self.N):
            fn = (C.c_double * 3)()
            orc_lib().orc_get_node_force(self.o.h, n, fn, C.byref(m))
            f[n] = list(fn)
        out["force"] = f
        return out
