"""External-field devices of devices.f90 inside the fused gather + push of the
resident step (chimera_soa_gather_push_dev), against the oracle driver whose
species apply the same Devices entries through the C port (bit-identical to
the reference Fortran for these routines).

Each case is the small space-charge bunch of the batch tests with one device
(or two, in order) in its Devices tuple, the reference's entry form
(device, *args): 6 steps eager on the GPU and on the oracle.  Bars (FP64):
fields and particle moments <= 1e-9, momenta and positions particle by
particle <= 1e-7.  The time-dependent devices (plane wave, Gaussian beam)
also make the run refuse step recording.  Mixed precision: the same runs
against the FP64 GPU path at <= 1e-5.
"""
import sys

import numpy as np
import pytest

from tests._cases import rel_l2

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _spec(devices, precision="fp64"):
    from chimera_amd import workloads as W
    spec = W.beam(Nx=128, Nkr=32, gamma=100.0, sig_x=2.0, sig_r=1.0, emittance=0.5, n_target=2e4, K=2.0,
                  lambda_u=150.0, dens=1e-3, steps=6)
    spec["beam"]["Devices"] = devices
    return W.with_precision(spec, precision)


def _devices(kind, spec):
    """Device entries acting on the bunch (x around 0, r up to ~3.5) over 6 steps."""
    xmin = spec["solver"]["Grid"][0]
    lu = 150.0
    ku = 2 * np.pi / lu
    und = np.array([2.0 * ku, lu, xmin - 2 * lu, 1.0e12])
    nmap = 400
    xs = xmin - 50.0 + 0.5 * np.arange(nmap)
    a0 = np.asfortranarray(np.vstack([2.0 * ku * np.exp(-(xs / 80.0) ** 2), 1.5 * ku * np.cos(xs / 40.0)]))
    if kind == "undul_analytic":
        return (("undul_analytic", und),)
    if kind == "undul_analytic_taper":
        return (("undul_analytic_taper", np.array([2.0 * ku, lu, xmin - 2 * lu, 3000.0, 0.15])),)
    if kind == "undul_mapped":
        return (("undul_mapped", a0, np.array([lu, xs[0], 0.5])),)
    if kind == "undul_mapped_tap":
        return (("undul_mapped_tap", a0, np.array([lu, xs[0], 0.5, 150.0, 0.2])),)
    if kind == "planewave":
        return (("planewave", np.array([0.05, 2.0, xmin - 10.0, 200.0, 20.0, 0.3, 0.7])),)
    if kind == "gaussbeam":
        return (("gaussbeam", 0.05, np.array([2.0, -1.0, 5.0, 0.3, -0.2, 6.0, 4.0, 4.0])),)
    # two devices in order
    return (("undul_analytic", und), ("planewave", np.array([0.05, 2.0, xmin - 10.0, 200.0, 20.0, -0.4, 0.1])))


def _run(mod, spec, steps=6):
    from chimera_amd import workloads as W
    run = W.build_beam(mod, spec, 7)
    for i in range(1, steps + 1):
        run.make_step(i)
    return run


def _snap_gpu(run):
    s, e = run.Solvers[0], run.Particles[0]
    return {"EB": s.Data["EB"].numpy().astype(np.complex128), "EG_fb": s.Data["EG_fb"].numpy().astype(np.complex128),
            "momenta": e.host("momenta").astype(np.float64), "coords": e.host("coords"),
            "weights": e.host("weights").astype(np.float64)}


def _moments(u, w):
    W = w.sum()
    mean = (u * w).sum(1) / W
    rms = np.sqrt(((u - mean[:, None]) ** 2 * w).sum(1) / W)
    return np.concatenate([[W], mean, rms])


KINDS = ["undul_analytic", "undul_analytic_taper", "undul_mapped", "undul_mapped_tap", "planewave", "gaussbeam",
         "two"]


@pytest.mark.parametrize("kind", KINDS)
def test_device_in_fused_push_matches_oracle(kind):
    import chimera_amd as G
    from oracle import driver as O
    spec = _spec(None)
    spec = _spec(_devices(kind, spec))
    say(f"case: beam 128x32 with device {kind}, 6 steps vs oracle")
    g = _run(G, spec)
    if kind in ("planewave", "gaussbeam", "two"):
        assert any("time" in b for b in g.graph_blockers())
    a = _snap_gpu(g)
    c = _run(O, spec)
    cs, ce = c.Solvers[0], c.Particles[0]
    b = {"EB": cs.Data["EB"], "EG_fb": cs.Data["EG_fb"], "momenta": ce.Data["momenta"], "coords": ce.Data["coords"],
         "weights": ce.Data["weights"]}
    out = {k: rel_l2(a[k], b[k]) for k in ("EB", "EG_fb", "momenta", "coords")}
    ma, mb = _moments(a["momenta"], a["weights"]), _moments(b["momenta"], b["weights"])
    out["moments"] = float(np.max(np.abs(ma - mb) / np.maximum(np.abs(mb), 1e-300)))
    # the device must have acted: the run without it differs
    ref0 = _snap_gpu(_run(G, _spec(())))
    out["device_effect"] = rel_l2(ref0["momenta"], a["momenta"])
    say("  " + " ".join(f"{k}={v:.2e}" for k, v in out.items()))
    assert out["device_effect"] > 1e-6
    for k in ("EB", "EG_fb", "moments"):
        assert out[k] <= 1e-9, (k, out[k])
    assert out["momenta"] <= 1e-7 and out["coords"] <= 1e-7


@pytest.mark.parametrize("kind", ["undul_mapped_tap", "gaussbeam"])
def test_device_in_fused_push_mixed(kind):
    import chimera_amd as G
    spec = _spec(None)
    devs = _devices(kind, spec)
    say(f"case: beam 128x32 with device {kind}, mixed vs fp64, 6 steps")
    a = _snap_gpu(_run(G, _spec(devs, "mixed")))
    b = _snap_gpu(_run(G, _spec(devs, "fp64")))
    e = {k: rel_l2(a[k], b[k]) for k in ("EB", "EG_fb", "momenta")}
    say("  " + " ".join(f"{k}={v:.2e}" for k, v in e.items()))
    for k, v in e.items():
        assert v <= 1e-5, (k, v)


def test_device_descriptors_checks():
    """Wrong device names, short parameter arrays and maps of the wrong shape
    are refused on the host before any launch."""
    import chimera_amd as G
    from chimera_amd import workloads as W
    spec = _spec(None)
    run = W.build_beam(G, _spec(()), 7)
    sp = run.Particles[0]
    for devs, exc in ((("no_such_device", np.zeros(4)),), NotImplementedError), \
                     ((("undul_analytic", np.zeros(3)),), ValueError), \
                     ((("undul_mapped", np.zeros((3, 10)), np.zeros(3)),), ValueError):
        sp.Devices = devs
        with pytest.raises(exc):
            sp.device_descriptors(1)
    sp.Devices = _devices("two", spec)
    recs, t = sp.device_descriptors(3)
    assert len(recs) == 2 and recs[0].kind == 0 and recs[1].kind == 4 and t == 3 * sp.Args["TimeStep"]
