"""Benchmark of the MI355X pseudo-spectral stepper (one GPU).

Workload: the 256^3 FP64 plain stepper on the perturbed-interface initial
state (interface height Lz/2 + 0.5 cos(2 pi x/Lx) cos(2 pi y/Ly), density
1.0/0.8, 1e-3 seeded noise, dt = 0.01), with the diagnostics (L1, mass,
free energy) every 50 steps as the reference's loop logs them and at the
end of the timed region, so every timed run holds at least one.  The FP32-mixed stepper on the same state is
timed beside it.  A "step" is one iteration of the reference's loop body
(smaiso-wcomp.cpp:1081-2198).

Prints one JSON line (rank 0, N=1).  The timed region is bracketed by a
device-wide synchronisation of the HIP runtime libsmaiso.so uses (torch's
ROCm wheel carries a separate copy of the runtime that cannot see this
library's stream, so torch.cuda.synchronize() would not be a barrier here).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "smaiso-wcomp_amd"))

METRIC = "steps/s at 256^3: FP64 plain vs FP32-mixed @ ≤1e-6 rel L2; % HBM peak"
HBM_PEAK_GBS = 8000.0
# Algorithmic field passes per step on the canonical dataflow (30 transforms
# counted as 1 read + 1 write each, fused single-operand multipliers free):
# 133 passes; FP64 moves 8 B per pass, FP32-mixed 126 passes at 4 B + 7 at 8 B.
BYTES_PER_POINT = {"f64": 133 * 8, "mixed": 126 * 4 + 7 * 8}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def make_params(S, n, precision):
    return S.Params(nx=n, ny=n, nz=n, dt=0.01, rho0=0.8, rho_m=1.0, divv_switch=0,
                    precision=precision)


def time_stepper(S, params, psi, rho, steps, warmup, cadence=50, graph=True, tune=True):
    st = S.Stepper(params)
    st.set_graph(graph)
    tuned = None
    if tune:
        ms, chosen = st.autotune()
        tuned = dict(chosen=chosen.tolist(), ms=[[[round(float(v), 4) for v in ax] for ax in dt] for dt in ms])
        log(f"[bench] autotuned tile variants (fp64, fp32) x (x, y, z): {chosen.tolist()}")
    st.set_state(psi, rho)
    st.step(warmup)
    S.device_synchronize()
    t0 = time.perf_counter()
    done = 0
    diag = None
    ndiag = 0
    while done < steps:
        chunk = min(cadence, steps - done)
        st.step(chunk)
        done += chunk
        # every 50 steps (the reference's stepL1) and after the last chunk
        diag = st.diagnostics(allow_nonfinite=True)
        ndiag += 1
    S.device_synchronize()
    t1 = time.perf_counter()
    st.tuned = tuned
    st.ndiag = ndiag
    return st, (t1 - t0), diag


def kernel_timing(st, nsteps):
    """Per-kernel-class HIP-event timing over nsteps (events on the stepper's
    stream).  The y and x passes form one class, r2r_yx: they are the same
    kernels on the two strided axes."""
    st.timing(True)
    st.step(nsteps)
    res = {}
    for kc in ("r2r_z", "r2r_y", "r2r_x", "r2r_axis", "spectral", "pointwise"):
        ms, n, by = st.timing_get(kc)
        res[kc] = dict(total_ms=ms, launches=n, bytes=by)
    st.timing(False)
    y, x = res.pop("r2r_y"), res.pop("r2r_x")
    res["r2r_yx"] = {k: y[k] + x[k] for k in y}
    return res


def class_table(kt, nsteps):
    """avg launch, launches per step, achieved GB/s and fraction of the HBM
    peak per kernel class."""
    out = {}
    for k, v in kt.items():
        avg = v["total_ms"] / max(v["launches"], 1)
        gbs = v["bytes"] / max(v["total_ms"], 1e-12) / 1e6
        out[k] = dict(avg_ms=avg, launches_per_step=v["launches"] / nsteps, ms_per_step=v["total_ms"] / nsteps,
                      gbs=gbs, frac=gbs / HBM_PEAK_GBS)
    return out


def copy_peak(nbytes=1 << 30, reps=10):
    """The HBM rate of a plain copy kernel (smaiso_hbm_copy_rate: four 16-byte
    elements per lane over 1 GiB, 10 launches timed with HIP events),
    reported beside the 8 TB/s specification.  The buffers are well above the
    256 MB memory-side cache, so the figure is DRAM's."""
    import ctypes as C
    import smaiso_amd as S
    L = S.lib()
    L.smaiso_hbm_copy_rate.argtypes = [C.c_int64, C.c_int32, C.POINTER(C.c_double)]
    g = C.c_double()
    if L.smaiso_hbm_copy_rate(nbytes, reps, C.byref(g)) != 0:
        return None
    return dict(gbs=g.value, frac_of_spec=g.value / HBM_PEAK_GBS, bytes_per_copy=2 * nbytes,
                method="copy kernel (four 16-byte elements per lane) over 1 GiB, 10 launches timed with HIP events")

def mixed_accuracy(S, n, psi, rho, checkpoints=(100, 500, 2000), params=None, label=""):
    """FP32-mixed vs FP64 on the same state: relative L2 errors of psi, rho and
    v (all three components against the FP64 velocity's norm) at checkpoints.
    The FP64 GPU stepper is the reference here; its own parity with the FP64
    oracle is checked by the test suite.  params(precision) builds the
    parameters (default: make_params at n^3)."""
    import numpy as np
    mk = params or (lambda prec: make_params(S, n, prec))
    a = S.Stepper(mk(S.FP64))
    b = S.Stepper(mk(S.FP32_MIXED))
    a.set_state(psi, rho)
    b.set_state(psi, rho)
    out, done = {}, 0
    for c in checkpoints:
        a.step(c - done)
        b.step(c - done)
        done = c
        res = {}
        for name in ("psi", "rho"):
            fa, fb = a.field(name), b.field(name)
            res[name] = float(np.linalg.norm(fb - fa) / np.linalg.norm(fa))
        va = [a.field(k) for k in ("vx", "vy", "vz")]
        vb = [b.field(k) for k in ("vx", "vy", "vz")]
        num = math.sqrt(sum(float(np.linalg.norm(y - x)) ** 2 for x, y in zip(va, vb)))
        den = math.sqrt(sum(float(np.linalg.norm(x)) ** 2 for x in va))
        res["v"] = num / den
        out[str(c)] = res
        log(f"[bench] mixed vs fp64{label} at step {c}: {res}")
    a.close()
    b.close()
    return out


def flat_interface_accuracy(S, shapes, checkpoints=(100, 500, 2000)):
    """FP32-mixed vs FP64 on the flat-interface initial state (layers
    cos(q0 z) below z = Lz/2 times 1/2 (1 - tanh((z - Lz/2)/2)), density 1.0 in
    the smectic and 0.9 in the isotropic phase, 1e-3 seeded noise; rho0 = 0.9,
    dt = 0.01, rho_m = 1) on each grid of `shapes`, errors at the checkpoints."""
    out = {}
    for (nx, ny, nz) in shapes:
        def params(prec, nx=nx, ny=ny, nz=nz):
            return S.Params(nx=nx, ny=ny, nz=nz, dt=0.01, rho0=0.9, rho_m=1.0, divv_switch=0,
                            precision=prec)
        ic = S.InterfaceIC(amplitude=1.0, width=2.0, rho_iso=0.9, rho_sm=1.0, undulation=0.0,
                           noise=1e-3, seed=1)
        psi, rho = S.make_interface_ic(params(S.FP64), ic)
        key = f"{nx}x{ny}x{nz}"
        out[key] = mixed_accuracy(S, None, psi, rho, checkpoints, params=params, label=f" (flat, {key})")
        del psi, rho
    out["workload"] = ("flat interface at Lz/2, rho 1.0/0.9, 1e-3 noise (seed 1), rho0=0.9, dt=0.01, rho_m=1; "
                       "relative L2 of FP32-mixed vs FP64 (v: all components against the FP64 velocity norm)")
    return out


ENSEMBLE_SHAPES = ((64, 64, 64), (96, 96, 96), (128, 128, 64))
ENSEMBLE_DT = 0.0025


def ensemble_members(S, count, seed=4):
    """Batched-ensemble workload: `count` simulations split evenly over 64^3,
    96^3 and 128x128x64 grids, FP32-mixed, with seeded draws of the density
    ratio (rho_iso = rho0 in [0.7, 0.95]), viscosity nu in [0.1, 2], coupling
    zeta in [0.5, 2] and interface undulation in [0, 1].  dt = 0.0025: the
    reference scheme (and the oracle) diverges within 75-300 steps at
    dt = 0.01 for the draws with nu below about 0.35, and stays bounded for
    every draw at dt = 0.0025."""
    import numpy as np
    rng = np.random.default_rng(seed)
    members, ics = [], []
    for m in range(count):
        shape = ENSEMBLE_SHAPES[m * len(ENSEMBLE_SHAPES) // count]
        ratio = float(rng.uniform(0.7, 0.95))
        p = S.Params(nx=shape[0], ny=shape[1], nz=shape[2], dt=ENSEMBLE_DT, rho0=ratio, rho_m=1.0,
                     divv_switch=0, nu=float(rng.uniform(0.1, 2.0)), zeta=float(rng.uniform(0.5, 2.0)),
                     precision=S.FP32_MIXED)
        ic = S.InterfaceIC(amplitude=1.0, width=2.0, rho_iso=ratio, rho_sm=1.0,
                           undulation=float(rng.uniform(0.0, 1.0)), noise=1e-3,
                           seed=int(rng.integers(1 << 62)))
        members.append(p)
        ics.append(ic)
    return members, ics


def time_ensemble(S, count, steps, warmup=4):
    members, ics = ensemble_members(S, count)
    points = sum(p.nx * p.ny * p.nz for p in members)
    en = S.Ensemble(members)
    states = [S.make_interface_ic(p, ic) for p, ic in zip(members, ics)]

    def timed(k):
        for m, (psi, rho) in enumerate(states):
            en.set_state(m, psi, rho)
        en.step(warmup)
        S.device_synchronize()
        t0 = time.perf_counter()
        en.step(k)
        S.device_synchronize()
        return time.perf_counter() - t0

    k_default = min(steps, 100)
    sec_default = timed(k_default)
    en.autotune()
    sec = timed(steps)
    finite = sum(all(math.isfinite(v) for v in en.diagnostics(m, allow_nonfinite=True)) for m in range(count))
    eps = steps / sec
    gbs = BYTES_PER_POINT["mixed"] * points * eps / 1e9
    res = dict(members=count, groups=en.groups, points=points, steps=steps,
               ensemble_steps_per_s=eps, member_steps_per_s=eps * count,
               point_steps_per_s=eps * points, algorithmic_gbs=gbs, hbm_frac=gbs / HBM_PEAK_GBS,
               member_steps_per_s_default_tiles=k_default * count / sec_default,
               autotuned=True, graph=True, members_finite=finite,
               workload="16 x 64^3 + 16 x 96^3 + 16 x 128x128x64 (per 48 members), FP32-mixed, "
                        f"dt={ENSEMBLE_DT}, seeded rho0/nu/zeta/undulation draws, {steps} timed steps")
    en.close()
    return res


def time_adjoint(S, n=128, nsteps=500, every=50):
    """Adjoint workload: 128^3, loss ||psi(T) - target||^2 after 500 steps
    (target from a perturbed interface height), checkpoint every 50; the whole
    smaiso_adjoint call (forward sweep, segment recompute, transposed steps)
    is timed, and a plain forward run of the same steps beside it."""
    import dataclasses
    import numpy as np
    out = {}
    for name, prec in (("fp64", S.FP64), ("fp32_mixed", S.FP32_MIXED)):
        p = S.Params(nx=n, ny=n, nz=n, dt=0.01, rho0=0.8, rho_m=1.0, divv_switch=0, nu=0.6, zeta=1.4,
                     precision=prec)
        x = np.arange(n) * p.dx
        h = n * p.dz / 2 + 0.5 * np.cos(2 * np.pi * x / (n * p.dx))[:, None] * np.cos(2 * np.pi * x / (n * p.dy))[None, :]
        ic = S.InterfaceIC(amplitude=1.0, width=2.0, rho_iso=0.8, rho_sm=1.0, noise=1e-3, seed=4)
        psi0, rho0 = S.make_interface_ic(p, ic, height=h)
        target = S.make_interface_ic(p, dataclasses.replace(ic, seed=11), height=h + 0.6)[0]
        st = S.Stepper(p)
        st.adjoint(psi0, rho0, target, 4, 2)  # warm-up (allocations, tables)
        S.device_synchronize()
        t0 = time.perf_counter()
        g = st.adjoint(psi0, rho0, target, nsteps, every)
        S.device_synchronize()
        sec = time.perf_counter() - t0
        st.set_state(psi0, rho0)
        S.device_synchronize()
        t0 = time.perf_counter()
        st.step(nsteps)
        S.device_synchronize()
        fsec = time.perf_counter() - t0
        st.close()
        out[name] = dict(seconds=sec, adjoint_steps_per_s=nsteps / sec, forward_steps_per_s=nsteps / fsec,
                         adjoint_over_forward=sec / fsec, loss=g.loss, dzeta=g.zeta, dnu=g.nu)
        log(f"[bench] adjoint {name}: {nsteps} steps in {sec:.3f} s ({sec / fsec:.2f}x the forward run)")
    out["workload"] = f"{n}^3, {nsteps} steps, checkpoint every {every}, loss ||psi(T)-target||^2"
    return out


def time_outputs(S, st, reps=5):
    """The output writers' device work on the 256^3 state: the surface tracker
    (one wavefront per column), one x-z slice, and a full-state checkpoint
    written to and read back from a temporary file."""
    import tempfile
    import numpy as np
    p = st.params
    prev = np.zeros(p.nx * p.ny)
    st.surface_height(0.85, prev)
    t0 = time.perf_counter()
    for _ in range(reps):
        st.surface_height(0.85, prev)
    surf_ms = (time.perf_counter() - t0) / reps * 1e3
    st.slice("psi")
    t0 = time.perf_counter()
    for _ in range(reps):
        st.slice("psi")
    slice_ms = (time.perf_counter() - t0) / reps * 1e3
    nbytes = p.nx * p.ny * p.nz * (4 * 8 + 2 * (8 if p.precision == S.FP64 else 4))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "state.ck")
        t0 = time.perf_counter()
        st.save_checkpoint(path)
        save_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        st.load_checkpoint(path)
        load_s = time.perf_counter() - t0
    res = dict(surface_ms=surf_ms, slice_ms=slice_ms, checkpoint_bytes=nbytes,
               checkpoint_save_s=save_s, checkpoint_load_s=load_s,
               note="host-visible times of the synchronous calls; the checkpoint includes disk I/O")
    log(f"[bench] outputs: surface {surf_ms:.2f} ms, slice {slice_ms:.2f} ms, checkpoint save {save_s:.2f} s "
        f"load {load_s:.2f} s ({nbytes / 2**20:.0f} MiB)")
    return res


def cpu_baseline(n, psi, rho, seconds_budget=20.0):
    """The FP64 oracle (numpy + scipy.fft) on the host on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    threads = min(16, os.cpu_count() or 1)
    os.environ["SMAISO_ORACLE_THREADS"] = str(threads)
    import smaiso_oracle as O
    O._WORKERS = threads
    op = O.Params(nx=n, ny=n, nz=n, dt=0.01, rho0=0.8, rho_m=1.0, divv_switch=0)
    o = O.Oracle(op)
    o.set_state(psi, rho)
    t0 = time.perf_counter()
    k = 0
    while True:
        o.step(1)
        k += 1
        if time.perf_counter() - t0 > seconds_budget * 0.6 or k >= 20:
            break
    dt = time.perf_counter() - t0
    return dict(value=k / dt, unit="steps/s", cores=threads, kind="port",
                sample=f"{k} FP64 steps of the numpy/scipy.fft oracle at {n}^3 from the same initial state "
                       f"(scipy.fft workers={threads}, numpy elementwise single-threaded; "
                       f"host {cpu_model()}, nproc={os.cpu_count()})")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--accuracy-steps", type=str, default="100,500,2000",
                    help="checkpoints of the FP32-mixed vs FP64 comparison ('' to skip)")
    ap.add_argument("--flat-grids", type=str, default="256x256x256,512x512x256",
                    help="grids of the flat-interface FP32-mixed vs FP64 comparison ('' to skip)")
    ap.add_argument("--no-graph", action="store_true", help="launch every step eagerly")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default r2r tiles")
    ap.add_argument("--ensemble", type=int, default=48, help="members of the ensemble leg (0 to skip)")
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--adjoint-steps", type=int, default=500, help="0 to skip the adjoint leg")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    import smaiso_amd as S

    n = args.n
    log(f"[bench] building IC {n}^3")
    p64 = make_params(S, n, S.FP64)
    ic = S.InterfaceIC(amplitude=1.0, width=2.0, rho_iso=0.8, rho_sm=1.0, undulation=0.5,
                       noise=1e-3, seed=2)
    psi, rho = S.make_interface_ic(p64, ic)

    log(f"[bench] FP64: warmup {args.warmup}, timed {args.steps}")
    st64, sec64, diag64 = time_stepper(S, p64, psi, rho, args.steps, args.warmup, graph=not args.no_graph,
                                       tune=not args.no_autotune)
    tuned64 = st64.tuned
    st64_ndiag = st64.ndiag
    sps64 = args.steps / sec64
    log(f"[bench] FP64 {sps64:.2f} steps/s, diag {diag64}")
    kt64 = kernel_timing(st64, args.profile_steps)
    outputs = time_outputs(S, st64)
    del st64

    mixed = None
    if not args.no_mixed:
        pm = make_params(S, n, S.FP32_MIXED)
        log(f"[bench] FP32-mixed: warmup {args.warmup}, timed {args.steps}")
        stm, secm, diagm = time_stepper(S, pm, psi, rho, args.steps, args.warmup, graph=not args.no_graph,
                                        tune=not args.no_autotune)
        spsm = args.steps / secm
        ktm = kernel_timing(stm, args.profile_steps)
        del stm
        log(f"[bench] mixed {spsm:.2f} steps/s, diag {diagm}")
        mixed = dict(value=spsm, ms_per_step=1e3 / spsm, speedup_vs_fp64=spsm / sps64,
                     step_roofline_frac=BYTES_PER_POINT["mixed"] * n ** 3 * spsm / (HBM_PEAK_GBS * 1e9),
                     fp32_stages=S.fp32_stages_default(),
                     kernels=class_table(ktm, args.profile_steps))

    if mixed is not None and args.accuracy_steps:
        cps = tuple(int(x) for x in args.accuracy_steps.split(","))
        mixed["rel_l2_vs_fp64"] = mixed_accuracy(S, n, psi, rho, cps)
        if args.flat_grids:
            shapes = [tuple(int(v) for v in g.split("x")) for g in args.flat_grids.split(",")]
            log(f"[bench] flat-interface accuracy on {shapes}")
            mixed["rel_l2_vs_fp64_flat"] = flat_interface_accuracy(S, shapes, cps)

    # the dominant kernel class of the FP64 stepper (most time per step) and
    # the one-axis operator passes beside it
    tr = None
    tpath = os.path.join(ROOT, "profiles", "pmc_traffic_256cubed_fp64.json")
    if n == 256 and os.path.exists(tpath):
        # HBM bytes per launch of each kernel class from the committed
        # rocprofv3 FETCH_SIZE/WRITE_SIZE runs of this bench (copy-calibrated)
        with open(tpath) as f:
            tr = json.load(f)
    copy = copy_peak()

    def class_roofline(kc):
        d = kt64[kc]
        avg_ms = d["total_ms"] / max(d["launches"], 1)
        bytes_per_launch = d["bytes"] / max(d["launches"], 1)
        achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
        cls = tr.get(kc) if tr else None
        return dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=achieved / HBM_PEAK_GBS,
                    traffic=(cls["read_bytes"] + cls["write_bytes"]) if cls else None, kernel=kc,
                    avg_launch_ms=avg_ms, launches_per_step=d["launches"] / args.profile_steps,
                    algorithmic_bytes_per_launch=bytes_per_launch,
                    frac_of_copy_peak=achieved / copy["gbs"] if copy else None,
                    traffic_source=tr.get("source") if (tr and cls) else None)

    dom = max(kt64, key=lambda k: kt64[k]["total_ms"])
    roofline = class_roofline(dom)
    roofline_axis = class_roofline("r2r_axis") if dom != "r2r_axis" else None
    step_bytes = BYTES_PER_POINT["f64"] * n ** 3
    step_roof = dict(algorithmic_bytes_per_step=step_bytes,
                     achieved=step_bytes * sps64 / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                     frac=step_bytes * sps64 / (HBM_PEAK_GBS * 1e9),
                     frac_of_copy_peak=step_bytes * sps64 / 1e9 / copy["gbs"] if copy else None)

    ens = None
    if args.ensemble > 0:
        log(f"[bench] ensemble: {args.ensemble} members, {args.ensemble_steps} steps")
        ens = time_ensemble(S, args.ensemble, args.ensemble_steps)
        log(f"[bench] ensemble {ens['member_steps_per_s']:.1f} member-steps/s, "
            f"{ens['algorithmic_gbs']:.0f} GB/s")

    adj = None
    if args.adjoint_steps > 0:
        adj = time_adjoint(S, nsteps=args.adjoint_steps)

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle)")
        cpu = cpu_baseline(n, psi, rho)
        log(f"[bench] CPU {cpu['value']:.4f} steps/s")

    out = {
        "metric": METRIC,
        "value": sps64,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 / sps64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded perturbed-interface initial state built on the host",
        "config": {"workload": f"plain FP64 stepper, {n}^3 grid, interface h=Lz/2+0.5cos(2pi x/Lx)cos(2pi y/Ly), "
                               f"rho 1.0/0.8, dt=0.01, rho_m=1, diagnostics every 50 steps and at the end of "
                               f"the timed region ({st64_ndiag} in it)",
                   "grid": [n, n, n], "precision": "fp64", "dt": 0.01, "rho_m": 1.0,
                   "hipgraph": not args.no_graph, "autotuned_tiles": tuned64},
        "roofline": roofline,
        "roofline_r2r_axis": roofline_axis,
        "hbm_copy_peak": copy,
        "step_roofline": step_roof,
        "kernels_fp64": class_table(kt64, args.profile_steps),
        "fp32_mixed": mixed,
        "ensemble": ens,
        "adjoint": adj,
        "outputs": outputs,
        "cpu_baseline": cpu,
        "diagnostics_fp64": diag64,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
