"""TEST INFRASTRUCTURE — study of mixed-precision linear solves on the oracle's Newton systems.

For seeded samples of the C3 (net86) and C4 (net87) distributions, one hydro step is integrated
with the oracle's FP64 solve (the reference's eigen::solve) and with models of mixed solves:
  LU-IR     FP32 unpivoted LU in the reference's elimination order, FP64 residual, `sweeps`
            refinement sweeps (the device's mixed path),
  GMRES-IR  FP64 GMRES on the Ruiz-equilibrated system, left-preconditioned by its FP32 LU
            (Carson & Higham, SIAM J. Sci. Comput. 40(2), 2018), `sweeps` outer iterations of at
            most `inner` GMRES steps,
each converged when every correction |d_i| <= tol (|x_i| + s_i), s = (T, Y); a particle with a
solve that does not converge is integrated again in FP64 (the device's redo pass).  Reported per
model: the share of Newton solves and of particles the mixed solve carries, the GMRES iterations
per solve, and, over the carried particles, the final-state error against the FP64 run (mass
fractions, temperature, specific energy release).

    python oracle/mixed_study.py [out.json]
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path[:0] = [HERE, os.path.join(ROOT, "sph-exa-nnet_amd")]

import workloads  # noqa: E402
from oracle import Oracle  # noqa: E402

TABLE = os.path.join(ROOT, "sph-exa-nnet_amd", "data", "helm_table.bin")
MODELS = [("LU-IR", 1, tol, sw, 0) for tol in (1e-11, 1e-10, 1e-9, 1e-8) for sw in (1, 3)] + \
         [("GMRES-IR", 2, tol, 3, 8) for tol in (1e-11, 1e-9)]


def run(o, cfg, model, max_iter):
    name, mode, tol, sweeps, inner = model
    o.set_solver_model(mode, tol, sweeps, inner)
    Y = np.ascontiguousarray(cfg["Y"].copy())
    T, dt = cfg["temp"].copy(), cfg["dt"].copy()
    t0 = time.time()
    st, sub, new = o.batch(cfg["net"], cfg["rho"].copy(), T, Y, dt, cfg["hydro_dt"], eos_kind=cfg["eos"], mu=cfg["mu"],
                           Z=cfg["Z"], max_iter=max_iter)
    el = time.time() - t0
    counts = o.solver_counts()
    redone = o.last_redone.copy()
    o.set_solver_model(0)
    return dict(Y=Y, T=T, st=st, new=new, redone=redone, counts=counts, seconds=el)


def errors(cfg, ref, mix):
    A = workloads.mass_numbers(cfg["net"])
    BE = workloads.binding_energies(cfg["net"])
    carried = (mix["redone"] == 0) & (ref["st"] != 1) & (ref["st"] == mix["st"]) & (ref["st"] != 4)
    m = int(carried.sum())
    if m == 0:
        return dict(carried_particles=0)
    Xr, Xm = ref["Y"][:, carried] * A[:, None], mix["Y"][:, carried] * A[:, None]
    dx = np.abs(Xm - Xr)
    relx = (dx / np.maximum(np.abs(Xr), 1e-14)).max(0)
    relT = np.abs(mix["T"][carried] - ref["T"][carried]) / ref["T"][carried]
    er = BE @ (ref["Y"][:, carried] - cfg["Y"][:, carried])
    em = BE @ (mix["Y"][:, carried] - cfg["Y"][:, carried])
    nz = np.abs(er) > 0
    rele = np.abs(em - er)[nz] / np.abs(er[nz])
    ok = np.all(dx <= 1e-6 * np.abs(Xr) + 1e-14, axis=0) & (relT <= 1e-8)
    return dict(carried_particles=m, X_max_rel=float(relx.max()), X_median_rel=float(np.median(relx)),
                T_max_rel=float(relT.max()), eps_max_rel=float(rele.max()) if rele.size else 0.0,
                within_1e6_X_1e8_T=float(ok.mean()))


def main():
    o = Oracle()
    o.helm_load(TABLE)
    out = []
    for cid, n, max_iter in ((3, 4096, 10000), (4, 256, 10000)):
        cfg = workloads.config(cid, n=n, seed=7000 + cid)
        ref = run(o, cfg, ("FP64", 0, 0, 0, 0), max_iter)
        live = int((ref["st"] != 1).sum())
        print(f"{cfg['label']}: {n} particles ({live} integrated), FP64 {int(ref['new'].sum())} Newton solves, "
              f"{int((ref['st'] == 4).sum())} at the cap, {ref['seconds']:.1f} s", flush=True)
        for model in MODELS:
            mix = run(o, cfg, model, max_iter)
            c = mix["counts"]
            rec = dict(config=cfg["label"], model=model[0], tol=model[2], sweeps=model[3], inner=model[4],
                       solves_converged=c["converged"] / max(c["solves"], 1),
                       particles_carried=float(((mix["redone"] == 0) & (ref["st"] != 1)).sum()) / max(live, 1),
                       gmres_iterations_per_solve=c["gmres_iterations"] / max(c["solves"], 1),
                       seconds=mix["seconds"])
            rec.update(errors(cfg, ref, mix))
            out.append(rec)
            print("  " + json.dumps({k: (round(v, 4) if isinstance(v, float) and abs(v) > 1e-3 else v)
                                     for k, v in rec.items() if k != "config"}), flush=True)
    if len(sys.argv) > 1:
        json.dump(out, open(sys.argv[1], "w"), indent=1)


if __name__ == "__main__":
    main()
