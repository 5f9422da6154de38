"""Whole PIC step on the GPU (chimera_amd resident driver) against the CPU
oracle driver (oracle/driver.py), from identical seeded inputs.

Bars (FP64): fields after the initial half step and after 1 step agree to
relative L2 <= 1e-12; after more steps to <= 1e-9, as do the particle moments
(total charge, weighted mean and RMS momentum).  Particles keep the same order
on both sides (compaction is stable), so positions and momenta are also
compared particle by particle.
"""
import sys

import numpy as np
import pytest

from tests._cases import rel_l2

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def moments(u, w):
    """Total charge, weighted mean and RMS of each momentum component.  The
    arrays are made C-contiguous first: NumPy sums a Fortran-ordered (3, n)
    array along axis 1 in another order (not pairwise), and for a beam whose
    mean momentum is 1e4 times its spread that alone moves the mean by 3e-10
    of the spread."""
    u, w = np.ascontiguousarray(u), np.ascontiguousarray(w)
    W = w.sum()
    mean = (u * w).sum(1) / W
    rms = np.sqrt(((u - mean[:, None]) ** 2 * w).sum(1) / W)
    return np.concatenate([[W], mean, rms])


def moments_err(mg, mc):
    """Largest moment difference: total charge and RMS momenta relative to
    themselves, mean momenta relative to the RMS spread of their component
    (the mean of a thermal plasma is ~1e-6 of its spread, so its own relative
    error only measures round-off)."""
    rms = np.maximum(np.abs(mc[4:7]), 1e-300)
    e_w = abs(mg[0] - mc[0]) / abs(mc[0])
    e_mean = np.abs(mg[1:4] - mc[1:4]) / rms
    e_rms = np.abs(mg[4:7] - mc[4:7]) / rms
    return float(max(e_w, e_mean.max(), e_rms.max()))


def snapshot_gpu(run):
    s, e = run.Solvers[0], run.Particles[0]
    snap = {k: s.Data[k].numpy() for k in ("EG_fb", "J_fb", "EB")}
    snap.update(np=e.np, momenta=e.host("momenta"), coords=e.host("coords"), weights=e.host("weights"),
                np_ions=run.Particles[1].np if len(run.Particles) > 1 else 0)
    return snap


def snapshot_cpu(run):
    s, e = run.Solvers[0], run.Particles[0]
    snap = {k: s.Data[k].copy() for k in ("EG_fb", "J_fb", "EB")}
    snap.update(np=e.Data["coords"].shape[1], momenta=e.Data["momenta"].copy(), coords=e.Data["coords"].copy(),
                weights=e.Data["weights"].copy(),
                np_ions=run.Particles[1].Data["coords"].shape[1] if len(run.Particles) > 1 else 0)
    return snap


def compare_snaps(g, c, tol_fields, tol_parts, what, floor=None):
    """floor: the oracle's rounding floor per field key (rounding_floor); a
    field's bar is then max(tol_fields, 4 x its floor).  Only J_fb ever needs
    it: its Poisson correction adds (grad rho_n+1 - grad rho_n)/dt / w^2, a
    difference of two nearly equal terms that amplifies rounding."""
    out = {k: rel_l2(g[k], c[k]) for k in ("EG_fb", "J_fb", "EB")}
    assert g["np"] == c["np"] and g["np_ions"] == c["np_ions"], (what, g["np"], c["np"])
    out["momenta"] = rel_l2(g["momenta"], c["momenta"])
    out["coords"] = rel_l2(g["coords"], c["coords"])
    out["moments"] = moments_err(moments(g["momenta"], g["weights"]), moments(c["momenta"], c["weights"]))
    say(f"{what}: " + " ".join(f"{k}={v:.2e}" for k, v in out.items()))
    for key in ("EG_fb", "J_fb", "EB"):
        bar = max(tol_fields, 4 * floor[key]) if floor else tol_fields
        assert out[key] <= bar, (what, key, out[key], bar)
    for key in ("momenta", "coords", "moments"):
        assert out[key] <= tol_parts, (what, key, out[key])


def noise_floor(c, cs, what):
    """Print the oracle's own noise floor beside a comparison: the oracle run
    c against the same run with its particles shuffled (cs), either runs or
    their snapshots; fields and moments only, since the particle order
    differs."""
    a = c if isinstance(c, dict) else snapshot_cpu(c)
    b = cs if isinstance(cs, dict) else snapshot_cpu(cs)
    out = {k: rel_l2(b[k], a[k]) for k in ("EG_fb", "J_fb", "EB")}
    out["moments"] = moments_err(moments(b["momenta"], b["weights"]), moments(a["momenta"], a["weights"]))
    say(f"{what} oracle noise floor (shuffled particle order): " + " ".join(f"{k}={v:.2e}" for k, v in out.items()))
    return out


def rounding_floor(c, cf, what):
    """The oracle run c against the same run on the oracle compiled with FMA
    contraction (cf, built and stepped under driver.rounding_variant()): the
    spread that rounding alone gives, per field.  Printed; returned for the
    bars of compare_snaps."""
    a = c if isinstance(c, dict) else snapshot_cpu(c)
    b = cf if isinstance(cf, dict) else snapshot_cpu(cf)
    out = {k: rel_l2(b[k], a[k]) for k in ("EG_fb", "J_fb", "EB")}
    say(f"{what} oracle rounding floor (FMA contraction): " + " ".join(f"{k}={v:.2e}" for k, v in out.items()))
    return out


def compare(gpu_run, cpu_run, tol_fields, tol_parts, what):
    compare_snaps(snapshot_gpu(gpu_run), snapshot_cpu(cpu_run), tol_fields, tol_parts, what)


def _runs(spec, seed=7):
    import chimera_amd as G
    from chimera_amd import workloads as W
    from oracle import driver as O
    g = W.build(G, spec, seed)
    c = W.build(O, spec, seed)
    return g, c


@pytest.mark.parametrize("nko", [1, 2])
def test_small_fixed_window(nko):
    from chimera_amd import workloads as W
    say(f"case: lpa_small nko={nko} fixed window")
    spec = W.lpa_small(Nx=128, Nkr=32, nko=nko)
    g, c = _runs(spec)
    compare(g, c, 1e-12, 1e-12, "half step")
    g.make_step(1)
    c.make_step(1)
    compare(g, c, 1e-12, 1e-12, "step 1")
    for i in range(2, 11):
        g.make_step(i)
        c.make_step(i)
    compare(g, c, 1e-9, 1e-9, "step 10")


def test_small_moving_window():
    """Moving window every 10 steps: absorbing layer, plasma injection with
    ions on top, particle absorption and background re-deposit.  Each side
    runs to the end on its own (the injected plasma comes from the global
    NumPy stream, which both sides must see in the same state)."""
    import chimera_amd as G
    from chimera_amd import workloads as W
    from oracle import driver as O
    say("case: lpa_small moving window, 21 steps")
    spec = W.lpa_small(Nx=128, Nkr=32, nko=1, moving=True)
    checks = (1, 10, 11, 21)
    snaps = {}
    g = W.build(G, spec, 11)
    for i in range(1, 22):
        g.make_step(i)
        if i in checks:
            snaps[i] = snapshot_gpu(g)
    c = W.build(O, spec, 11)
    for i in range(1, 22):
        c.make_step(i)
        if i in checks:
            tol = 1e-12 if i == 1 else 1e-9
            compare_snaps(snaps[i], snapshot_cpu(c), tol, tol, f"step {i}")
    assert snaps[21]["np_ions"] != snaps[1]["np_ions"]      # the window injected and absorbed ions


def test_cfg0_one_step():
    """The full lpa_cfg0 configuration (Nx=512, Nkr=64, m=0,1, 515,088
    electrons): half step and one step against the oracle."""
    from chimera_amd import workloads as W
    say("case: lpa_cfg0 one step")
    spec = W.lpa_cfg0()
    g, c = _runs(spec, seed=2024)
    assert g.Particles[0].np == 515088
    compare(g, c, 1e-12, 1e-12, "cfg0 half step")
    from oracle import driver as O
    cs = W.build(O, spec, seed=2024)
    cs.shuffle_particles(1)
    g.make_step(1)
    c.make_step(1)
    cs.make_step(1)
    compare(g, c, 1e-12, 1e-12, "cfg0 step 1")
    noise_floor(c, cs, "cfg0 step 1")


@pytest.mark.timeout(600)
def test_cfg0_100_steps():
    """lpa_cfg0 for its full 100 steps: the FP64 bar after many steps is
    relative L2 <= 1e-9 for the fields and for the particle moments (total
    charge, weighted mean and RMS momentum).  Atomic deposition sums in another
    order than the oracle's serial loop, so the difference grows from 1e-16
    per step; positions and momenta particle by particle get a looser sanity
    bar (1e-7) that still catches any wrong step."""
    import chimera_amd as G
    from chimera_amd import workloads as W
    from oracle import driver as O
    say("case: lpa_cfg0 100 steps")
    spec = W.lpa_cfg0()
    checks = (1, 10, 50, 100)
    snaps = {}
    g = W.build(G, spec, 99)
    for i in range(1, 101):
        g.make_step(i)
        if i in checks:
            snaps[i] = snapshot_gpu(g)
    say("  GPU run done; oracle run and the same run with shuffled particles (about 0.2 s per step each)")
    c = W.build(O, spec, 99)
    c2 = W.build(O, spec, 99)
    c2.shuffle_particles(3)
    for i in range(1, 101):
        c.make_step(i)
        c2.make_step(i)
        if i in checks:
            noise_floor(c, c2, f"  step {i}")
            cs = snapshot_cpu(c)
            tf = 1e-12 if i == 1 else 1e-9
            out = {k: rel_l2(snaps[i][k], cs[k]) for k in ("EG_fb", "J_fb", "EB")}
            out["moments"] = moments_err(moments(snaps[i]["momenta"], snaps[i]["weights"]),
                                         moments(cs["momenta"], cs["weights"]))
            out["momenta"] = rel_l2(snaps[i]["momenta"], cs["momenta"])
            out["coords"] = rel_l2(snaps[i]["coords"], cs["coords"])
            say(f"  step {i}: " + " ".join(f"{k}={v:.2e}" for k, v in out.items()))
            assert snaps[i]["np"] == cs["np"]
            for key in ("EG_fb", "J_fb", "EB"):
                assert out[key] <= tf, (i, key, out[key])
            assert out["moments"] <= 1e-9, (i, out["moments"])
            assert out["momenta"] <= 1e-7 and out["coords"] <= 1e-7, (i, out["momenta"], out["coords"])
