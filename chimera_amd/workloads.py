"""Named workloads of the PIC step, written as the reference's configuration
dictionaries (the shape of doc/tests/lpa-testrun.py), and a builder that
instantiates them with any module exposing Solver / Specie / ChimeraRun.

  lpa_cfg0   laser in uniform plasma, Nx=512 x Nkr=64, modes m=0,1,
             FixedCell 2x2x4 (515,088 electrons) + still ions, a0=1, waist 5,
             duration 4, dt = dx = 0.05, fixed window, 100 steps
  lpa_cfg1   production drop-in: Nx=4096 x Nkr=256, modes m=0..2, 16.7M
             electrons, a0=2, waist 8, duration 6, moving window (plasma
             injected on the right, absorbing layer on the left), 200 steps
  lpa_small  the cfg0 physics on a reduced grid, for quick parity runs

Plasma density n0 = 0.005 n_c (as doc/tests/lpa-testrun.py:27) and thermal
spread 1e-3 m_e c.  Random draws come from NumPy's legacy generator, seeded
once before the species are generated (electrons first, then ions).
"""
from __future__ import annotations

import copy

import numpy as np


def lpa(Nx, Nkr, nko, dx=0.05, dr=0.3, a0=1.0, waist=5.0, duration=4.0, x0_from_right=12.8, dens=0.005,
        steps=100, moving=False, absorb=0, fixed_cell=(2, 2, 4)):
    xmin = -Nx * dx
    rmax = Nkr * dr
    solver_in = {"Grid": (xmin, 0.0, rmax, dx, dr), "TimeStep": dx, "MaxAzimuthMode": nko,
                 "Features": ("SpaceCharge", "StillAsBackground")}
    laser_in = {"a0": a0, "k0": 1.0, "x0": -x0_from_right, "x_foc": -x0_from_right, "Lx": duration, "LR": waist}
    electrons_in = {"Grid": (xmin, 0.0, rmax, dx, dr), "TimeStep": dx, "Density": dens, "FixedCell": fixed_cell,
                    "MomentaSpreads": (1e-3, 1e-3, 1e-3)}
    ions_in = copy.deepcopy(electrons_in)
    ions_in["Charge"] = 1.0
    ions_in["Mass"] = 1886.0
    ions_in["Features"] = ("Still",)
    ions_in["MomentaSpreads"] = (0.0, 0.0, 0.0)
    if moving:
        frames = ({"TimeStep": dx, "Steps": 10, "Velocity": 1.0, "AbsorbLayer": absorb,
                   "AddPlasma": _uniform, "Features": ("IonsOnTop",)},)
    else:
        frames = ()
    return {"solver": solver_in, "laser": laser_in, "electrons": electrons_in, "ions": ions_in,
            "frames": frames, "features": ("PreDepositBackground",), "steps": steps}


def _uniform(x):
    return 1.0 + 0 * x


def lpa_cfg0():
    return lpa(512, 64, 1, a0=1.0, waist=5.0, duration=4.0, x0_from_right=12.8, steps=100)


def lpa_cfg1():
    return lpa(4096, 256, 2, a0=2.0, waist=8.0, duration=6.0, x0_from_right=30.0, steps=200, moving=True,
               absorb=256)


def lpa_small(Nx=128, Nkr=32, nko=1, moving=False, steps=10):
    return lpa(Nx, Nkr, nko, a0=1.0, waist=2.0, duration=1.5, x0_from_right=2.5, steps=steps, moving=moving,
               absorb=16 if moving else 0)


WORKLOADS = {"lpa_cfg0": lpa_cfg0, "lpa_cfg1": lpa_cfg1, "lpa_small": lpa_small}


def with_precision(spec, precision):
    """A copy of spec whose solver and species run in `precision` ('fp64' or
    'mixed'; the oracle ignores the key and always runs FP64)."""
    spec = copy.deepcopy(spec)
    for key in ("solver", "electrons", "ions"):
        spec[key]["Precision"] = precision
    return spec


def build(mod, spec, seed=12345, **run_options):
    """Instantiate a workload with mod.Solver / mod.Specie / mod.ChimeraRun;
    run_options are extra ChimeraRun keys (e.g. SortEvery)."""
    spec = copy.deepcopy(spec)
    solver = mod.Solver(spec["solver"])
    solver.add_gauss_beam(spec["laser"])
    electrons = mod.Specie(spec["electrons"])
    ions = mod.Specie(spec["ions"])
    np.random.seed(seed)
    electrons.add_particles(*electrons.gen_parts())
    ions.add_particles(*ions.gen_parts())
    comps = {"Solvers": (solver,), "Particles": (electrons, ions), "MovingFrames": spec["frames"],
             "Features": spec["features"]}
    comps.update(run_options)
    run = mod.ChimeraRun(comps)
    return run
