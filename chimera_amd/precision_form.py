"""Error statement of the mixed-precision PIC step against the FP64 step.

    python -m chimera_amd.precision_form [--workload lpa_cfg2] [--checkpoints 1,100,500] [--seed S]
                                         [--markdown FILE]

Runs the workload twice from the same seed, first in FP64 and then in mixed
precision (float32 momenta and weights, complex64 grids, split-BF16 MFMA
transforms with FP32 accumulation; FP64 positions, tables and reductions), and
at every checkpoint compares:

* each component of the gathered field EB (Ex, Ey, Ez, Bx, By, Bz), relative L2
  over the whole grid and per azimuthal mode, and against the norm of all six
  components (a component that is still near zero has a meaningless relative
  error of its own);
* the spectral fields EG_fb (E and G = rot B in Fourier-Bessel space);
* the particle moments, reduced in FP64 on the host: total charge (sum of
  weights), mean momentum per component (difference over the RMS spread), RMS
  momentum per component, and the kinetic energy sum w (gamma - 1).

The runs are sequential because the moving window injects plasma drawn from
NumPy's global random stream; each run starts from the same seed and makes the
same draws.  The FP64 run is the reference: it is the path whose whole steps
agree with the CPU restatement of the reference to 1e-12 after one step and
1e-9 after 100.  One JSON line goes to stdout; progress to stderr.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

COMPS = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")


def log(msg):
    print(f"[precision_form] {msg}", file=sys.stderr, flush=True)


def moments(u, w):
    """FP64 reductions of momenta u (3, n) and weights w (n)."""
    u = u.astype(np.float64)
    w = w.astype(np.float64)
    W = w.sum()
    mean = (u * w).sum(1) / W
    rms = np.sqrt(((u - mean[:, None]) ** 2 * w).sum(1) / W)
    kin = float((w * (np.sqrt(1.0 + (u * u).sum(0)) - 1.0)).sum())
    return {"charge": float(W), "mean": mean, "rms": rms, "kinetic": kin}


def snapshot(run):
    s, e = run.Solvers[0], run.Particles[0]
    return {"EB": s.Data["EB"].numpy().astype(np.complex128),
            "EG_fb": s.Data["EG_fb"].numpy().astype(np.complex128),
            "moments": moments(e.host("momenta"), e.host("weights")),
            "np": [sp.np for sp in run.Particles]}


def compare(mx, fp):
    """Error record of snapshot mx against snapshot fp."""
    out = {"fields": {}, "fields_per_mode": {}, "moments": {}}
    EBm, EBd = mx["EB"], fp["EB"]
    allnorm = np.linalg.norm(EBd.ravel())
    for l, name in enumerate(COMPS):
        d = EBm[..., l] - EBd[..., l]
        ref = np.linalg.norm(EBd[..., l].ravel())
        dn = np.linalg.norm(d.ravel())
        out["fields"][name] = {"rel_l2": float(dn / max(ref, 1e-300)), "of_all": float(dn / max(allnorm, 1e-300))}
        per = []
        for m in range(EBd.shape[2]):
            rm = np.linalg.norm(EBd[:, :, m, l].ravel())
            per.append(float(np.linalg.norm(d[:, :, m].ravel()) / max(rm, 1e-300)))
        out["fields_per_mode"][name] = per
    d = mx["EG_fb"] - fp["EG_fb"]
    out["fields"]["EG_fb"] = {"rel_l2": float(np.linalg.norm(d.ravel()) / np.linalg.norm(fp["EG_fb"].ravel()))}
    a, b = mx["moments"], fp["moments"]
    out["moments"]["charge"] = abs(a["charge"] - b["charge"]) / abs(b["charge"])
    out["moments"]["mean_over_rms"] = [float(v) for v in np.abs(a["mean"] - b["mean"]) / b["rms"]]
    out["moments"]["rms"] = [float(v) for v in np.abs(a["rms"] - b["rms"]) / b["rms"]]
    out["moments"]["kinetic"] = abs(a["kinetic"] - b["kinetic"]) / abs(b["kinetic"])
    out["np_equal"] = mx["np"] == fp["np"]
    return out


def run_precision(workload, precision, seed, checkpoints, on_snapshot):
    import chimera_amd as G
    from chimera_amd import workloads as W
    from chimera_amd._lib import lib

    L = lib()
    t0 = time.time()
    run = W.build(G, W.with_precision(W.WORKLOADS[workload](), precision), seed)
    L.chimera_synchronize()
    A = run.Solvers[0].Args
    log(f"{precision}: built {workload} in {time.time() - t0:.1f}s (Nx={A['Nx']} Nkr={A['Nkr']} "
        f"modes={A['Mtot']} electrons={run.Particles[0].np})")
    stepping, psteps = 0.0, 0
    for i in range(1, max(checkpoints) + 1):
        t1 = time.perf_counter()
        run.make_step(i)
        psteps += run.Particles[0].np
        if i in checkpoints:
            L.chimera_synchronize()
            stepping += time.perf_counter() - t1
            on_snapshot(i, snapshot(run))
            log(f"{precision}: step {i}")
        else:
            stepping += time.perf_counter() - t1
    L.chimera_synchronize()
    return {"particle_steps_per_s": psteps / stepping, "ms_per_step": stepping / max(checkpoints) * 1e3,
            "electrons": run.Particles[0].np, "Nx": A["Nx"], "Nkr": A["Nkr"], "modes": A["Mtot"]}


def markdown(rec):
    lines = [f"| step | " + " | ".join(COMPS) + " | EG_fb | charge | mean u / rms u (x, y, z) | rms u (x, y, z) | kinetic |",
             "|---" * (len(COMPS) + 6) + "|"]
    for step, e in rec["errors"].items():
        f, m = e["fields"], e["moments"]
        cells = [f"{f[c]['rel_l2']:.1e} ({f[c]['of_all']:.0e})" for c in COMPS]
        lines.append(f"| {step} | " + " | ".join(cells) + f" | {f['EG_fb']['rel_l2']:.1e} | {m['charge']:.1e} | "
                     + ", ".join(f"{v:.1e}" for v in m["mean_over_rms"]) + " | "
                     + ", ".join(f"{v:.1e}" for v in m["rms"]) + f" | {m['kinetic']:.1e} |")
    lines.append("")
    lines.append("Per azimuthal mode (relative L2 of each component within the mode), last checkpoint:")
    lines.append("")
    last = list(rec["errors"].values())[-1]["fields_per_mode"]
    nm = len(next(iter(last.values())))
    lines.append("| component | " + " | ".join(f"m={m}" for m in range(nm)) + " |")
    lines.append("|---" * (nm + 1) + "|")
    for c in COMPS:
        lines.append(f"| {c} | " + " | ".join(f"{v:.1e}" for v in last[c]) + " |")
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="lpa_cfg2")
    ap.add_argument("--checkpoints", default="1,100,500")
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--markdown", default=None, help="also write the error tables to this file")
    args = ap.parse_args()
    checkpoints = sorted({int(c) for c in args.checkpoints.split(",")})
    ref = {}
    fp = run_precision(args.workload, "fp64", args.seed, checkpoints, lambda i, s: ref.__setitem__(i, s))
    errors = {}

    def cmp(i, s):
        errors[i] = compare(s, ref.pop(i))
        e = errors[i]
        log(f"step {i}: " + " ".join(f"{c}={e['fields'][c]['rel_l2']:.2e}" for c in COMPS)
            + f" EG_fb={e['fields']['EG_fb']['rel_l2']:.2e} charge={e['moments']['charge']:.1e}")

    mx = run_precision(args.workload, "mixed", args.seed, checkpoints, cmp)
    rec = {"workload": args.workload, "seed": args.seed, "checkpoints": checkpoints,
           "fp64": fp, "mixed": mx, "speedup": mx["particle_steps_per_s"] / fp["particle_steps_per_s"],
           "errors": {str(k): v for k, v in errors.items()}}
    print(json.dumps(rec), flush=True)
    if args.markdown:
        with open(args.markdown, "w") as f:
            f.write(markdown(rec) + "\n")


if __name__ == "__main__":
    main()
