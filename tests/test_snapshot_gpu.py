"""Snapshots of a resident run (drop_snap / read_snap, chimera_main.py:323-410):
a run stopped after a few steps, its snapshot loaded into a fresh run of the
same configuration, and both continued, give the same state.

* The snapshot round trip is exact: a snapshot of the restored run holds the
  same names and the same bits.
* After the resumed steps the two runs agree to 1e-12 (FP64) and 1e-4 (mixed
  precision) -- the bars of recorded against eager steps, since the only
  difference is the order of the global atomic adds of deposit tiles that do
  not fit a slot.  The moving window resumes from the restored edge.
Workloads without plasma injection (a bunch with its window moving every step,
lpa_cfg0): the injected plasma of a moving window is drawn from a random
stream that a snapshot does not hold, in the reference as here.
"""
import sys

import numpy as np
import pytest

from tests._cases import rel_l2

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "mixed": 1e-4}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _build(kind, precision):
    import chimera_amd as G
    from chimera_amd import workloads as W
    if kind == "beam":
        spec = W.beam(Nx=256, Nkr=48, gamma=150.0, sig_x=2.0, sig_r=1.0, emittance=0.5, n_target=5e4, K=2.0,
                      lambda_u=150.0, steps=8)
        return W.build_beam(G, W.with_precision(spec, precision), 3)
    return W.build(G, W.with_precision(W.lpa_cfg0(), precision), 5)


@pytest.mark.parametrize("kind", ["beam", "lpa_cfg0"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_snapshot_resume(kind, precision, tmp_path):
    say(f"case: snapshot {kind} {precision}: 4 steps, snapshot, 4 more; fresh run from the snapshot, 4 steps")
    a = _build(kind, precision)
    for i in range(1, 5):
        a.make_step(i)
    fn = a.drop_snap(str(tmp_path / "first_"))
    for i in range(5, 9):
        a.make_step(i)
    b = _build(kind, precision)
    b.read_snap(fn)
    fn2 = b.drop_snap(str(tmp_path / "again_"))
    with np.load(fn) as f1, np.load(fn2) as f2:
        assert set(f1.files) == set(f2.files)
        for k in f1.files:
            assert f1[k].dtype == f2[k].dtype and np.array_equal(f1[k], f2[k], equal_nan=f1[k].dtype.kind in "fc"), k
    for i in range(5, 9):
        b.make_step(i)
    sa, sb = a.Solvers[0], b.Solvers[0]
    assert float(sa.Args["leftX"]) == float(sb.Args["leftX"])
    for key in ("EB", "EG_fb"):
        e = rel_l2(sb.Data[key].numpy().astype(np.complex128), sa.Data[key].numpy().astype(np.complex128))
        say(f"  {key}: {e:.2e}")
        assert e <= TOL[precision], (key, e)
    for pa, pb in zip(a.Particles, b.Particles):
        assert pa.np == pb.np
        for key in ("coords", "momenta"):
            e = rel_l2(pb.host(key).astype(np.float64), pa.host(key).astype(np.float64))
            assert e <= TOL[precision], (key, e)
