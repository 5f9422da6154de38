"""Named workloads of the PIC step, written as the reference's configuration
dictionaries (the shape of doc/tests/lpa-testrun.py), and a builder that
instantiates them with any module exposing Solver / Specie / ChimeraRun.

  lpa_cfg0   laser in uniform plasma, Nx=512 x Nkr=64, modes m=0,1,
             FixedCell 2x2x4 (515,088 electrons) + still ions, a0=1, waist 5,
             duration 4, dt = dx = 0.05, fixed window, 100 steps
  lpa_cfg1   production drop-in: Nx=4096 x Nkr=256, modes m=0..2, 16.7M
             electrons, a0=2, waist 8, duration 6, moving window (plasma
             injected on the right, absorbing layer on the left), 200 steps
  lpa_cfg2   the cfg1 physics at the precision-form size: Nx=8192 x Nkr=512,
             modes m=0..3, 67.0M electrons, moving window, 500 steps
  lpa_small  the cfg0 physics on a reduced grid, for quick parity runs
  lpa_testrun  the reference's own CI scenario (doc/tests/lpa-testrun.py):
             Nx=528 on 4 x-chunks, modes m=0,1, density ramp injected
  beam       one relativistic bunch with space charge in an undulator, in a
             window moving with it; beam_scan draws the 64 bunches of a scan
             (Nx 256-2048 x Nkr 32-256, modes m=0,1, 1e4-2e6 particles, 300 steps)
  fel_testrun  the reference's envelope CI scenario (doc/tests/fel-testrun.py):
             KxShift at the FEL resonance, Nx=120 on 4 x-chunks x Nr=120 cut
             at Rcut (85 nodes), mode m=0 of the envelope, RandCell 50 beam
             (doubled by denoise) in an analytic undulator, seed laser, window
             moving at the beam velocity, 300 steps

Plasma density n0 = 0.005 n_c (as doc/tests/lpa-testrun.py:27) and thermal
spread 1e-3 m_e c.  Random draws come from NumPy's legacy generator, seeded
once before the species are generated (electrons first, then ions).
"""
from __future__ import annotations

import copy

import numpy as np


def lpa(Nx, Nkr, nko, dx=0.05, dr=0.3, a0=1.0, waist=5.0, duration=4.0, x0_from_right=12.8, dens=0.005,
        steps=100, moving=False, absorb=0, fixed_cell=(2, 2, 4), xchunked=None):
    """xchunked = (T, n): the reference's x-chunked layout (solver and species
    'Xchunked'): T chunks deposited by one OpenMP thread each, particles binned
    into chunks every n + 1 steps; Nx stays a multiple of 2T."""
    xmin = -Nx * dx
    rmax = Nkr * dr
    solver_in = {"Grid": (xmin, 0.0, rmax, dx, dr), "TimeStep": dx, "MaxAzimuthMode": nko,
                 "Features": ("SpaceCharge", "StillAsBackground")}
    laser_in = {"a0": a0, "k0": 1.0, "x0": -x0_from_right, "x_foc": -x0_from_right, "Lx": duration, "LR": waist}
    electrons_in = {"Grid": (xmin, 0.0, rmax, dx, dr), "TimeStep": dx, "Density": dens, "FixedCell": fixed_cell,
                    "MomentaSpreads": (1e-3, 1e-3, 1e-3)}
    if xchunked is not None:
        solver_in["Xchunked"] = tuple(xchunked)
        electrons_in["Xchunked"] = tuple(xchunked)
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


def lpa_cfg0(**kw):
    args = dict(a0=1.0, waist=5.0, duration=4.0, x0_from_right=12.8, steps=100)
    args.update(kw)
    return lpa(512, 64, 1, **args)


def lpa_cfg1(**kw):
    args = dict(a0=2.0, waist=8.0, duration=6.0, x0_from_right=30.0, steps=200, moving=True, absorb=256)
    args.update(kw)
    return lpa(4096, 256, 2, **args)


def lpa_cfg2(**kw):
    args = dict(a0=2.0, waist=8.0, duration=6.0, x0_from_right=30.0, steps=500, moving=True, absorb=512)
    args.update(kw)
    return lpa(8192, 512, 3, **args)


def lpa_small(Nx=128, Nkr=32, nko=1, moving=False, steps=10, **kw):
    return lpa(Nx, Nkr, nko, a0=1.0, waist=2.0, duration=1.5, x0_from_right=2.5, steps=steps, moving=moving,
               absorb=16 if moving else 0, **kw)


def lpa_testrun(steps=550):
    """The reference's own LPA test scenario (doc/tests/lpa-testrun.py:11-63):
    x in [-20, 1] with dx = 0.04 on 4 x-chunks (Nx = 528, not a power of two),
    R = 16 with dr = 0.25, modes m = 0, 1; a0 = 3 laser focused at x = 45;
    FixedCell 2x2x4 electrons (NoSorting) and still ions at n = 0.005 n_c; a
    window moving every 10 steps with a 175-point absorbing layer and a density
    ramp injected on the right."""
    xgmin, xgmax, Rg, dx, dr = -20.0, 1.0, 16.0, 0.04, 0.25
    solver_in = {"Grid": (xgmin, xgmax, Rg, dx, dr), "TimeStep": dx, "MaxAzimuthMode": 1, "Xchunked": (4, 10),
                 "Features": ("SpaceCharge", "StillAsBackground")}
    laser_in = {"a0": 3.0, "k0": 1.0, "x0": -14.0, "x_foc": 45.0, "Lx": 4.0, "LR": 4.0}
    electrons_in = {"Grid": (xgmin, xgmax, Rg, dx, dr), "TimeStep": dx, "Density": 0.005, "FixedCell": (2, 2, 4),
                    "Xchunked": (4, 10), "Features": ("NoSorting",)}
    ions_in = copy.deepcopy(electrons_in)
    ions_in["Charge"] = 1
    ions_in["Mass"] = 1886
    ions_in["Features"] += ("Still",)
    frames = ({"TimeStep": dx, "Steps": 10, "AbsorbLayer": 175, "AddPlasma": _density_ramp,
               "Features": ("IonsOnTop",)},)
    return {"solver": solver_in, "laser": laser_in, "electrons": electrons_in, "ions": ions_in,
            "frames": frames, "features": (), "steps": steps}


def _density_ramp(x):
    return np.interp(x, [0.0, 40.0, 50.0, 300], [0.0, 1.0, 0.5, 0.5])


WORKLOADS = {"lpa_cfg0": lpa_cfg0, "lpa_cfg1": lpa_cfg1, "lpa_cfg2": lpa_cfg2, "lpa_small": lpa_small,
             "lpa_testrun": lpa_testrun}


def beam(Nx, Nkr, gamma, sig_x, sig_r, emittance, n_target, K, lambda_u, dens=1e-3, steps=300):
    """A relativistic electron bunch with space charge in an analytic
    undulator, in a window that follows it (the layout of the reference's
    space-charge demo, doc/space-charge-demo(vs_ocelot).ipynb, with the
    undulator device of doc/tests/fel-testrun.py).  The Gaussian bunch (rms
    length sig_x, rms radius sig_r) fills the box +-3.5 sig_x, 0..3.5 sig_r on
    a FixedCell lattice sized for about n_target particles; the transverse
    momentum spread is emittance / sig_r, the longitudinal one 1e-3 of the mean.
    Modes m = 0, 1; dt = dx; the undulator amplitude is K k_u (k_u = 2 pi / lambda_u)."""
    ncx = max(8, min(int(0.7 * Nx), int(round(np.sqrt(n_target / 4.0 * Nx / Nkr)))))
    ncr = max(4, min(int(0.7 * Nkr), int(round(n_target / 4.0 / ncx))))
    per = max(1.0, n_target / (4.0 * ncx * ncr))
    fx = max(1, int(round(np.sqrt(per))))
    fr = max(1, int(round(per / fx)))
    dx = 7.0 * sig_x / ncx
    dr = 3.5 * sig_r / ncr
    xmin, xmax, rmax = -0.5 * Nx * dx, 0.5 * Nx * dx, Nkr * dr
    solver_in = {"Grid": (xmin, xmax, rmax, dx, dr), "TimeStep": dx, "MaxAzimuthMode": 1,
                 "Features": ("SpaceCharge",)}
    pz0 = float(np.sqrt(gamma ** 2 - 1.0))
    pt = emittance / sig_r
    ku = 2.0 * np.pi / lambda_u
    undulator = np.array([K * ku, lambda_u, xmin - 2.0 * lambda_u, 1.0e12])
    beam_in = {"Grid": (xmin, xmax, rmax, dx, dr), "TimeStep": dx, "Charge": -1.0, "Mass": 1.0, "Density": dens,
               "FixedCell": (fx, fr, 4), "MomentaMeans": (pz0, 0.0, 0.0),
               "MomentaSpreads": (1e-3 * pz0, pt, pt), "Features": ("NoSorting",),
               "Devices": (("undul_analytic", undulator),)}
    frame = {"TimeStep": dx, "Steps": 1, "Velocity": 1.0, "Features": ("NoSorting", "Staged")}
    return {"kind": "beam", "solver": solver_in, "beam": beam_in, "frames": (frame,), "steps": steps,
            "domain": (-3.5 * sig_x, 3.5 * sig_x, 0.0, 3.5 * sig_r), "sigmas": (sig_x, sig_r),
            "params": {"Nx": Nx, "Nkr": Nkr, "gamma": gamma, "sig_x": sig_x, "sig_r": sig_r,
                       "emittance": emittance, "n_target": n_target, "K": K, "lambda_u": lambda_u}}


def beam_scan(n=64, seed=2026, steps=300):
    """n bunches with the parameters of a scan drawn from one seed: gamma
    uniform 50-500, sig_x uniform 1-10, sig_r uniform 0.5-5, normalised
    emittance log-uniform 0.1-2, particle count log-uniform 1e4-2e6, Nx a power
    of two in 256-2048, Nkr uniform 32-256, K uniform 1-3, undulator period
    uniform 100-1000."""
    rng = np.random.RandomState(seed)
    out = []
    for _ in range(n):
        out.append(beam(Nx=int(2 ** rng.randint(8, 12)), Nkr=int(rng.randint(32, 257)),
                        gamma=float(rng.uniform(50, 500)), sig_x=float(rng.uniform(1, 10)),
                        sig_r=float(rng.uniform(0.5, 5)), emittance=float(np.exp(rng.uniform(np.log(0.1), np.log(2)))),
                        n_target=float(np.exp(rng.uniform(np.log(1e4), np.log(2e6)))), K=float(rng.uniform(1, 3)),
                        lambda_u=float(rng.uniform(100, 1000)), steps=steps))
    return out


def fel_testrun(Nx=120, Nr=120, steps=None, periods=10):
    """doc/tests/fel-testrun.py:11-63 as a configuration: units of the
    undulator period (lam0 = 2.8 cm), K0 = 1.95, a 200 MeV beam of 20 pC in a
    80 um x 80 um box, the envelope solver shifted to the resonant wave number
    2 gamma_z^2 and co-propagating with the beam, no Poisson correction, dt =
    1/30.  steps defaults to the reference's periods / dt = 300."""
    K0, lam0 = 1.95, 2.8
    g0 = 200 / 0.511
    lbx = lbr = 80e-4 / lam0
    dens = 20e-12 / 1.6022e-19 / (np.pi * 80e-4 ** 3) / (1.1e21 / 2.8e4 ** 2)
    gg = g0 / (1.0 + K0 ** 2 / 2) ** 0.5
    k_res = 2 * gg ** 2
    vb = (1.0 - gg ** -2) ** 0.5
    Lgx, Rg, Rg_cut = 200e-4 / lam0, 1000e-4 / lam0, 700e-4 / lam0
    dt = 1.0 / 30
    solver_in = {"Grid": (-0.5 * Lgx, 0.5 * Lgx, Rg, Lgx / Nx, Rg / Nr), "TimeStep": dt, "MaxAzimuthMode": 0,
                 "KxShift": k_res, "Rcut": Rg_cut, "CoPropagative": vb, "Xchunked": (4, 6),
                 "Features": {"NoPoissonCorrection": True}}
    seed_in = {"a0": 0.15, "k0": k_res, "x0": -30e-4 / lam0, "x_foc": 70.0 / lam0, "Lx": 15e-4 / lam0,
               "LR": 180e-4 / lam0}
    beam_in = {"Grid": (-0.5 * Lgx, 0.5 * Lgx, Rg, Lgx / Nx, lbr / Nr), "TimeStep": dt, "Charge": -1.0, "Mass": 1.0,
               "Density": dens, "RandCell": 50, "MomentaMeans": (g0, 0.0, 0.0),
               "MomentaSpreads": (1e-4 * g0, 2e-5 * g0, 2e-5 * g0), "Xchunked": (4, 6),
               "Devices": (("undul_analytic", np.array([K0, 1.0, 1.0, float(periods)])),)}
    frame = {"TimeStep": dt, "Steps": 1, "Velocity": vb, "Features": ("Staged", "NoSorting")}
    return {"kind": "fel", "solver": solver_in, "laser": seed_in, "beam": beam_in, "frames": (frame,),
            "domain": (-0.5 * lbx, 0.5 * lbx, 0.0, lbr), "k_res": k_res,
            "steps": int(periods / dt) if steps is None else steps}


def build_fel(mod, spec, seed=12345, own_rng=False):
    """Instantiate fel_testrun with mod.Solver / mod.Specie / mod.ChimeraRun in
    the reference script's order: seed laser, beam generation, denoise at the
    resonant wave number, run."""
    spec = copy.deepcopy(spec)
    solver = mod.Solver(spec["solver"])
    solver.add_gauss_beam(spec["laser"])
    bunch = mod.Specie(spec["beam"])
    _seed((bunch,), seed, own_rng)
    bunch.add_particles(*bunch.gen_parts(Domain=spec["domain"]))
    bunch.denoise((spec["k_res"],))
    return mod.ChimeraRun({"Solvers": (solver,), "Particles": (bunch,), "MovingFrames": spec["frames"]})


def _seed(species, seed, own_rng):
    """Seed the draws of gen_parts: NumPy's global stream (the reference's), or
    with own_rng one RandomState shared by the run's species in the same order,
    which gives the same draws without touching the global stream."""
    if own_rng:
        rng = np.random.RandomState(seed)
        for sp in species:
            sp.rng = rng
    else:
        np.random.seed(seed)


def build_beam(mod, spec, seed=12345, own_rng=False):
    """Instantiate a beam workload with mod.Solver / mod.Specie / mod.ChimeraRun."""
    spec = copy.deepcopy(spec)
    solver = mod.Solver(spec["solver"])
    bunch = mod.Specie(spec["beam"])
    sx, sr = spec["sigmas"]

    def profile(x, y, z):
        return np.exp(-0.5 * x ** 2 / sx ** 2 - 0.5 * (y ** 2 + z ** 2) / sr ** 2)

    _seed((bunch,), seed, own_rng)
    bunch.add_particles(*bunch.gen_parts(Domain=spec["domain"], ProfileFunc=profile))
    return mod.ChimeraRun({"Solvers": (solver,), "Particles": (bunch,), "MovingFrames": spec["frames"]})


def x_slab(spec, parts):
    """The workload on 1/parts of its x extent (the window's right edge kept):
    the same radial grid, modes, plasma and laser, parts times fewer x nodes
    and particles."""
    spec = copy.deepcopy(spec)
    for key in ("solver", "electrons", "ions"):
        if key in spec:
            xmin, xmax, rmax, dx, dr = spec[key]["Grid"]
            spec[key]["Grid"] = (xmax - (xmax - xmin) / parts, xmax, rmax, dx, dr)
    return spec


def with_precision(spec, precision):
    """A copy of spec whose solver and species run in `precision` ('fp64' or
    'mixed'; the oracle ignores the key and always runs FP64)."""
    spec = copy.deepcopy(spec)
    for key in ("solver", "electrons", "ions", "beam"):
        if key in spec:
            spec[key]["Precision"] = precision
    return spec


def build(mod, spec, seed=12345, own_rng=False, **run_options):
    """Instantiate a workload with mod.Solver / mod.Specie / mod.ChimeraRun;
    run_options are extra ChimeraRun keys (e.g. SortEvery).  With own_rng the
    species draw from their own seeded stream (also for the plasma injected by
    the moving window), not NumPy's global one."""
    spec = copy.deepcopy(spec)
    solver = mod.Solver(spec["solver"])
    solver.add_gauss_beam(spec["laser"])
    electrons = mod.Specie(spec["electrons"])
    ions = mod.Specie(spec["ions"])
    _seed((electrons, ions), seed, own_rng)
    electrons.add_particles(*electrons.gen_parts())
    ions.add_particles(*ions.gen_parts())
    comps = {"Solvers": (solver,), "Particles": (electrons, ions), "MovingFrames": spec["frames"],
             "Features": spec["features"]}
    comps.update(run_options)
    run = mod.ChimeraRun(comps)
    return run
