"""Host-side logic that needs no GPU: the device window frame's recurrence,
which steps can be recorded and replayed, the bunch scan and request stream,
the memory estimate used for admission, and the per-grid table cache.

The frame test pins the one-thread kernel k_frame_step (frame[0] = frame[1] +
shift1, frame[1] = frame[0] + shift2) against the host's move_frame
(chimera_main.py:286-291: Xgrid + shiftX, leftX = Xgrid[0]) by replaying the
same FP64 additions: the edges must agree exactly, step after step.
"""
import numpy as np
import pytest

from chimera_amd import operators as ops
from chimera_amd import serve as S
from chimera_amd import workloads as W
from chimera_amd.chimera_main import ChimeraRun


class _Comp:
    def __init__(self, xgrid, features=()):
        self.Args = {"Xgrid": np.array(xgrid, dtype=np.float64), "Features": features}
        self.Args["leftX"], self.Args["rightX"] = self.Args["Xgrid"][0], self.Args["Xgrid"][-1]
        self.Data = {"gradRho_fb_prv": "prv", "gradRho_fb_nxt": "nxt"}
        self.frame = None
        self.env = False


def _run(frames, features=("SpaceCharge",), xgrid=None):
    run = object.__new__(ChimeraRun)
    run.MovingFrames = frames
    run.Solvers = (_Comp(np.linspace(-20.0, 0.0, 401) if xgrid is None else xgrid, features),)
    run.Particles = ()
    run.SortEvery = 0
    run.init_Moving_Frames()
    return run


@pytest.mark.parametrize("staged", [True, False])
def test_device_frame_follows_move_frame_exactly(staged):
    dt = 0.0173
    feats = ("Staged", "NoSorting") if staged else ("NoSorting",)
    run = _run(({"TimeStep": dt, "Steps": 1, "Velocity": 0.9993, "Features": feats},))
    A = run.Solvers[0].Args
    f = [A["leftX"], A["leftX"]]
    for i in range(1, 2001):
        s1, s2 = run._frame_shifts(i)
        a = f[1] + s1                       # k_frame_step, in FP64
        f = [a, a + s2]
        # host: the stage-1 move, then the stage-2 move (staged windows only)
        run.move_frame(run.MovingFrames[0])
        assert A["leftX"] == f[0], (i, A["leftX"], f[0])
        if staged:
            run.move_frame(run.MovingFrames[0])
        assert A["leftX"] == f[1], (i, A["leftX"], f[1])
    assert (f[1] - run.Solvers[0].Args["Xgrid"][0]) == 0.0


def test_host_step_moves_and_swaps_like_make_step():
    run = _run(({"TimeStep": 0.05, "Steps": 1, "Features": ("Staged", "NoSorting")},))
    s = run.Solvers[0]
    x0 = s.Args["leftX"]
    run.host_step(1)
    assert s.Data["gradRho_fb_prv"] == "nxt" and s.Data["gradRho_fb_nxt"] == "prv"
    sh = run.MovingFrames[0]["shiftX"]
    assert s.Args["leftX"] == (x0 + sh) + sh
    run.host_step(2)
    assert s.Data["gradRho_fb_prv"] == "prv"


def test_graph_blockers():
    ok = _run(({"TimeStep": 0.05, "Steps": 1, "Features": ("Staged", "NoSorting")},))
    assert ok.graph_blockers() == []
    assert _run(()).graph_blockers() == []
    every10 = _run(({"TimeStep": 0.05, "Steps": 10, "Features": ()},))
    assert any("every step" in b for b in every10.graph_blockers())
    inject = _run(({"TimeStep": 0.05, "Steps": 1, "AddPlasma": lambda x: 1 + 0 * x, "Features": ()},))
    assert any("injection" in b for b in inject.graph_blockers())
    bg = _run(({"TimeStep": 0.05, "Steps": 1, "Features": ()},), features=("SpaceCharge", "StillAsBackground"))
    assert any("background" in b for b in bg.graph_blockers())


def test_beam_scan_is_seeded_and_in_range():
    a, b = W.beam_scan(64, seed=7), W.beam_scan(64, seed=7)
    assert [x["params"] for x in a] == [x["params"] for x in b]
    assert [x["params"] for x in a] != [x["params"] for x in W.beam_scan(64, seed=8)]
    for sp in a:
        p = sp["params"]
        assert p["Nx"] in (256, 512, 1024, 2048) and 32 <= p["Nkr"] <= 256
        assert 50 <= p["gamma"] <= 500 and 1 <= p["sig_x"] <= 10 and 0.5 <= p["sig_r"] <= 5
        assert 0.1 <= p["emittance"] <= 2 and 1e4 <= p["n_target"] <= 2e6
        assert 1 <= p["K"] <= 3 and 100 <= p["lambda_u"] <= 1000
        xmin, xmax, rmax, dx, dr = sp["solver"]["Grid"]
        assert int(round((xmax - xmin) / dx)) == p["Nx"] and int(round(rmax / dr)) == p["Nkr"]
        # the bunch box fits the grid
        lo, hi, _, rtop = sp["domain"]
        assert xmin < lo and hi < xmax and rtop < rmax


def test_poisson_requests():
    r1 = S.poisson_requests(20.0, 30.0, seed=3)
    r2 = S.poisson_requests(20.0, 30.0, seed=3)
    assert [(r.name, r.t_arrival, r.seed) for r in r1] == [(r.name, r.t_arrival, r.seed) for r in r2]
    t = np.array([r.t_arrival for r in r1])
    assert np.all(np.diff(t) > 0) and t[-1] <= 30.0
    assert abs(len(r1) - 600) < 4 * np.sqrt(600)          # Poisson count at 20/s over 30 s
    kinds = {r.name.split("#")[0] for r in r1}
    assert kinds == set(S.MIX)
    for r in r1:
        assert r.bytes > 0 and r.steps in (100, 200, 300)


def test_memory_estimate_scales_with_size():
    small, big = W.lpa_cfg0(), W.lpa_cfg1()
    e0, e1 = S.estimate_bytes(small), S.estimate_bytes(big)
    assert 0 < e0 < e1
    # cfg1 FP64: 33.4 M particles of 80 B and ~1.3 GB of fields and tables; the estimate bounds both
    assert e1 > 2 * 16_707_600 * 80
    assert S.estimate_bytes(big, "mixed") < e1


def test_table_cache_returns_read_only_shared_tables():
    kx = 2 * np.pi * np.fft.fftfreq(64, 0.1)
    R = 0.3 * (np.arange(17) - 0.5)
    a = ops.spectral_operators(64, 16, 1, R[-1] + 0.3, kx, R, ext=0)
    b = ops.spectral_operators(64, 16, 1, R[-1] + 0.3, kx.copy(), R.copy(), ext=0)
    assert a is b
    with pytest.raises(ValueError):
        a["In"][0, 0, 0] = 1.0
    c = ops.spectral_operators(64, 16, 1, R[-1] + 0.3, kx, R, ext=1)
    assert c is not a and c["DpS2S"].shape[2] == 3
    E, G = ops.psatd_tables(a["w"], 0.1, True)
    assert ops.psatd_tables(a["w"], 0.1, True)[0] is E
    assert ops.psatd_tables(a["w"], 0.2, True)[0] is not E
