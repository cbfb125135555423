"""Validation run (not collected by pytest; needs a GPU and ~2 minutes of host CPU):
the benchmark configurations 1, 2 and 3 to their final times on the HIP path and on
the oracle, from the same inputs.

    python tests/run_full_length.py [c1 c2 c3]

Prints per configuration the step, halving, MOOD and failure counts of both, the
recovery-flag flips, and the largest final T and u differences over the cells both
recovered. Configurations 1 and 2 are also GPU tests (test_gpu_fullsize.py);
configuration 3 takes ~110 s on the oracle."""
import ctypes as C
import os
import sys
import time

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "fivohydro.jl_amd"))

import numpy as np  # noqa: E402

import bench  # noqa: E402
import fivo2d as F  # noqa: E402

TAUF = {"c1": 3.0, "c2": 8.0, "c3": 10.0}


def main(names):
    hip = F.load_hip()
    orc = F.Lib(os.path.join(REPO, "oracle", "liboracle2d.so"), "orc_")
    orc.dll.orc_set_threads.restype = C.c_int
    print("oracle threads", orc.dll.orc_set_threads(16), flush=True)
    for name in names:
        cfg = bench.CONFIGS[name]
        out = {}
        for lib, tag in ((hip, "hip"), (orc, "oracle")):
            g, m, prims = bench.make_inputs(cfg)
            s = F.Solver(g, m, lib=lib)
            s.init_from_prims(cfg["tau0"], **prims)
            t0 = time.perf_counter()
            r = s.run_sim_2d(tau0=cfg["tau0"], taufinal=TAUF[name], CFL=cfg["cfl"], CFLtau=cfg["cfltau"],
                             integrator=cfg["integrator"])
            dt = time.perf_counter() - t0
            out[tag] = (r, s.get_work(("yT", "ux", "uy", "ok")))
            print(f"{name} {tag}: ok {r['ok']} tau {r['tau']:.4f} steps {r['nsteps']} halvings {r['nhalvings']} "
                  f"mood {r['nmood']} primfail {r['nprimfail']} in {dt:.1f} s", flush=True)
            s.close()
        (rh, wh), (ro, wo) = out["hip"], out["oracle"]
        sx, sy = g.interior()
        grab = lambda a: a.reshape(g.Nxtot, g.Nytot)[sx, sy]
        okh, oko = grab(wh["ok"]).astype(bool), grab(wo["ok"]).astype(bool)
        both = okh & oko
        Th, To = np.exp(grab(wh["yT"])), np.exp(grab(wo["yT"]))
        dT = np.abs(Th - To) / To
        du = np.maximum(np.abs(grab(wh["ux"]) - grab(wo["ux"])), np.abs(grab(wh["uy"]) - grab(wo["uy"])))
        print(f"{name}: flips {int((okh != oko).sum())}, max dT/T {dT[both].max():.2e}, max du {du[both].max():.2e}, "
              f"cells with dT/T > 1e-9: {int((dT[both] > 1e-9).sum())} of {int(both.sum())}", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or ["c1", "c2", "c3"])
