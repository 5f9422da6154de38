"""Time the steps of one bunch simulation (workloads.beam: a relativistic
electron bunch with space charge in an analytic undulator, window moving with
the bunch) on the GPU, eagerly, with the FB transforms and operators of one
step timed one by one.

    python -m chimera_amd.bunch_step --nx 1024 --nkr 128 --precision mixed

prints one line per stage (us per call, averaged over --reps calls) and the
step time over --steps steps.
"""
from __future__ import annotations

import argparse
import time


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--nkr", type=int, default=128)
    ap.add_argument("--particles", type=float, default=3e5)
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="mixed")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--reps", type=int, default=50)
    ap.add_argument("--seed", type=int, default=5)
    a = ap.parse_args(argv)

    import chimera_amd as G
    from chimera_amd import workloads as W
    from chimera_amd._lib import lib
    L = lib()
    spec = W.with_precision(W.beam(Nx=a.nx, Nkr=a.nkr, gamma=200.0, sig_x=4.0, sig_r=2.0, emittance=0.5,
                                   n_target=a.particles, K=2.0, lambda_u=300.0), a.precision)
    run = W.build_beam(G, spec, a.seed)
    for i in (1, 2, 3):
        run.make_step(i)
    L.chimera_synchronize()
    s = run.Solvers[0]
    print(f"bunch {a.nx} x {a.nkr}, modes 0-1, {run.Particles[0].np} particles, {a.precision}", flush=True)

    def timed(name, fn):
        fn()
        L.chimera_synchronize()
        t0 = time.perf_counter()
        for _ in range(a.reps):
            fn()
        L.chimera_synchronize()
        print(f"{name:24s} {(time.perf_counter() - t0) / a.reps * 1e6:8.1f} us", flush=True)

    timed("fb_curr_in", s.fb_curr_in)
    timed("fb_dens_in", s.fb_dens_in)
    timed("FBGradDens", s.FBGradDens)
    timed("poiss_corr (3 iter)", s.poiss_corr)
    timed("G2B_FBRot", s.G2B_FBRot)
    timed("fb_fld_out", s.fb_fld_out)
    t0 = time.perf_counter()
    for i in range(4, 4 + a.steps):
        run.make_step(i)
    L.chimera_synchronize()
    print(f"{'step':24s} {(time.perf_counter() - t0) / a.steps * 1e6:8.1f} us", flush=True)


if __name__ == "__main__":
    main()
