"""Writes tests/golden/config2_trace.json: the FP64 oracle's diagnostics
trace (L1, mass, free energy every 50 steps) for the 256^3 perturbed-interface
configuration, the reference's own logging cadence (smaiso-wcomp.cpp:2202-2263).

Configuration: 256^3, dx = dy = dz = pi/4, dt = 0.01, rho0 = 0.8, rho_m = 1,
divvSwitch = 0; interface height h = Lz/2 + 0.5 cos(2 pi x/Lx) cos(2 pi y/Ly),
width 2, psi amplitude 1, rho 1.0 below / 0.8 above, 1e-3 splitmix64 noise
with seed 2 (the bench's workload state).

The oracle is the checker, run here on the CPU (hours for 2000 steps); the
fixture is rewritten after every 50 steps, so a partial run leaves a valid
shorter trace.  Usage:

    python tests/golden/make_config2_trace.py [--steps 2000] [--threads 6]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "config2_trace.json")

CONFIG = dict(n=256, dt=0.01, rho0=0.8, rho_m=1.0, divv_switch=0,
              ic=dict(amplitude=1.0, width=2.0, rho_iso=0.8, rho_sm=1.0, undulation=0.5,
                      noise=1e-3, seed=2),
              cadence=50)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--threads", type=int, default=6)
    ap.add_argument("--n", type=int, default=CONFIG["n"])
    ap.add_argument("--out", default=OUT)
    args = ap.parse_args()
    os.environ["SMAISO_ORACLE_THREADS"] = str(args.threads)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import smaiso_oracle as O
    O._WORKERS = args.threads

    n = args.n
    p = O.Params(nx=n, ny=n, nz=n, dt=CONFIG["dt"], rho0=CONFIG["rho0"], rho_m=CONFIG["rho_m"],
                 divv_switch=CONFIG["divv_switch"])
    psi, rho = O.make_interface_ic(p, O.InterfaceIC(**CONFIG["ic"]))
    o = O.Oracle(p)
    o.set_state(psi, rho)
    del psi, rho
    cad = CONFIG["cadence"]
    trace = []
    t0 = time.time()
    done = 0
    while done < args.steps:
        o.step(cad)
        done += cad
        l1, mass, en = o.diagnostics()
        trace.append(dict(step=done, l1=l1, mass=mass, energy=en))
        doc = dict(config=dict(CONFIG, n=n), oracle="oracle/smaiso_oracle.py (FP64, numpy + scipy.fft)",
                   steps_done=done, trace=trace)
        tmp = args.out + ".tmp"
        with open(tmp, "w") as f:
            json.dump(doc, f, indent=1)
        os.replace(tmp, args.out)
        print(f"step {done}: L1 {l1:.17g} mass {mass:.17g} energy {en:.17g} "
              f"({time.time() - t0:.0f} s)", flush=True)


if __name__ == "__main__":
    main()
