"""chimera_amd: CHIMERA's quasi-cylindrical PSATD particle-in-cell step on the
AMD Instinct MI355X (gfx950).

* chimera_amd.fimera     drop-in for the reference's f2py module `fimera`
                         (host arrays in, host arrays out, GPU in between)
* Solver, Specie,        the reference's driver API with all state resident
  ChimeraRun             in device memory
* chimera_amd.workloads  the benchmark configurations

All computation goes through libchimera_hip.so (include/chimera_hip.h); there is
no CPU fallback.
"""
from .chimera_main import ChimeraRun
from .solvers import Solver
from .species import Specie

__all__ = ["Solver", "Specie", "ChimeraRun"]
