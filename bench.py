"""Benchmark: FastS NS residual + LU-SGS iteration, Taylor-Green vortex 256^3, FP64, one GPU.

One step = one implicit time step with one Newton sub-iteration: viscosity and LU
coefficients, fused inviscid (AUSM+ o3) + viscous face fluxes with the BDF residual, LU-SGS
lower and upper sweeps with the Newton update, and the ghost-cell fill (periodic faces and
corners).  Residual-only verification sub-iterations are not part of a step.

Prints one JSON line.  Timing: W untimed steps, then R repeats of exactly K steps, each between
two device synchronisations (wall clock) and cross-checked with HIP events on the solver stream;
the line reports the median repeat (value, ms_per_step) and the spread of the R repeats.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic (compulsory) HBM bytes per interior cell on a Cartesian zone, FP64, 5 equations,
# for the implicit BDF iteration with one Newton sub-iteration:
#   residual: read W^n 40 + W^{n-1} 40, write R 40                       = 120
#   LU-SGS:   lower (read R 40 + W 40, write dW* 40) + upper and Newton
#             update (read dW* 40 + W 40, write W^{n+1} 40)               = 240
#   iteration (320 of the LU-SGS iteration + 80 for the two BDF levels)  = 400
#   coefficients (read W 40, write mu 8 + coe 40), counted separately     =  88
BYTES_PER_CELL = {"coe": 88.0, "residual": 120.0, "lu": 240.0, "bc": 0.0, "iteration": 400.0}
KERNELS = {"coe": "k_coe", "bc": "k_bc_face",
           "residual": {"cells": "k_residual", "faces": "k_residual_faces", "kcol": "k_residual_kcol",
                        "auto": "k_residual (six faces per cell, the default on Cartesian zones)"},
           "lu": {"blocks": "k_lu_block (lower + upper sweeps, one launch per block diagonal)",
                  "pipeline": "k_lu_pipe (lower + upper sweeps)",
                  "dataflow": "k_lu_flow (lower + upper sweeps, one launch each)",
                  "stream": "LU-SGS phase: k_lu_pack (skewed copy of the inputs), k_lu_stream lower + "
                            "upper, k_lu_unpack (one launch each)"}}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--scheme", default="ausmpred")
    ap.add_argument("--repeats", type=int, default=5, help="timed repeats of K steps (median reported)")
    ap.add_argument("--cpu-iters", type=int, default=4)
    ap.add_argument("--cpu-procs", type=int, default=0,
                    help="processes of the parallel CPU leg (0: the host's cores, at most 16)")
    ap.add_argument("--no-fasts-loop", action="store_true",
                    help="skip the FastS-style PyTree._compute loop timing")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/primal timing")
    ap.add_argument("--phase-reps", type=int, default=5)
    ap.add_argument("--lu", choices=["pipeline", "blocks", "dataflow", "stream"], default="stream")
    ap.add_argument("--residual", choices=["auto", "cells", "faces", "kcol"], default="auto")
    ap.add_argument("--lu-precision", choices=["fp64", "fp32"], default="fp64",
                    help="fp32: mixed precision, LU-SGS sweep arithmetic in FP32 (state, residual FP64)")
    ap.add_argument("--workload", choices=["tgv256", "channel", "zones64"], default="tgv256",
                    help="tgv256: the headline line; channel: 257x129x33 curvilinear channel with a "
                         "bump (3dfull, local time step, CFL 10); zones64: 64 abutting blocks "
                         "(4x4x4 tiling, sizes from default_rng(1) in [16, 96]) joined by transfers")
    ap.add_argument("--autotune", action="store_true",
                    help="time the interchangeable schedules first and keep the fastest (fast_amd.tune)")
    ap.add_argument("--graph", type=int, default=1, choices=[0, 1],
                    help="1 replays captured HIP graphs of the time step (FM_OPT_GRAPH)")
    ap.add_argument("--zone-batch", type=int, default=1, choices=[0, 1],
                    help="zones64: 1 batches the LU-SGS of all zones per block diagonal, 0 loops")
    ap.add_argument("--mixed-report", action="store_true",
                    help="FP64 against FP32 LU-SGS sweeps: Newton convergence of one BDF step to 1e-10 on "
                         "TGV 256^3 and the 513x257x65 channel (time, sub-iterations, state and KE errors)")
    ap.add_argument("--mixed-nmax", type=int, default=200, help="sub-iteration limit of --mixed-report")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy a MIB MiB device buffer 5 times (HBM counter calibration)")
    return ap.parse_args()


def copy_calibration(mib):
    """A plain device-to-device copy of known size, for scaling the FETCH_SIZE / WRITE_SIZE
    counters of the solver kernels."""
    import torch
    x = torch.ones(mib * (1 << 20) // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": mib * (1 << 20), "copies": 5}))


def make_zone(n, scheme):
    from fast_amd import cases, params as P
    num = P.Numerics(scheme=scheme, slope="o3", temporal_scheme="implicit", ss_iteration=1,
                     time_step_nature="global", time_step=1.0)
    z = cases.tgv_zone(n, num)
    z.param_real[P.DTC] = cases.tgv_dt(z, cfl=2.0)
    return z


def traffic_from_profiles(phase):
    """HBM bytes per launch of a phase's kernels ("lu/blocks", "residual/cells", ...) from the
    committed rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE, scaled by a plain copy)."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_tgv256.json")
    if not os.path.exists(path):
        return None
    try:
        d = json.load(open(path))
        return d.get(phase, {}).get("hbm_bytes_per_launch")
    except Exception:
        return None


def _reference_backend():
    import oracle
    if oracle.Reference.available():
        return oracle.Reference(), "reference"
    return oracle.Oracle(), "port"


def _cpu_run(dims, scheme, iters, barrier=None):
    """Seconds for `iters` implicit iterations of the TGV case of size dims on one host thread
    (and the wall-clock start and end of the timed part)."""
    from fast_amd import cases, params as P
    be, kind = _reference_backend()
    num = P.Numerics(scheme=scheme, slope="o3", temporal_scheme="implicit", ss_iteration=1,
                     time_step_nature="global", time_step=1.0)
    z = cases.tgv_zone(dims, num)
    z.param_real[P.DTC] = cases.tgv_dt(z, cfl=2.0)
    be.init_metric(z)
    be.apply_bc(z, z.ro)
    z.ro_m1[:] = z.ro
    z.ro_p1[:] = z.ro
    if barrier is not None:
        barrier.wait()
    w0 = time.time()
    t0 = time.perf_counter()
    for s in range(iters):
        be.substep(z, 1, first_it=0 if s == 0 else 1)
        m1 = z.ro_m1.copy(); z.ro_m1[:] = z.ro; z.ro[:] = z.ro_p1; z.ro_p1[:] = m1
    return time.perf_counter() - t0, kind, w0, time.time()


_BARRIER = None


def _cpu_init(barrier):
    global _BARRIER
    _BARRIER = barrier


def _cpu_slab(args):
    dims, scheme, iters = args
    return _cpu_run(dims, scheme, iters, _BARRIER)


def _host_description():
    """Core count and model of the host the baseline ran on."""
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "processor": platform.processor() or platform.machine()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core"):
                info[k.strip()] = v.strip()
    except Exception:
        pass
    return info


def cpu_baseline(n, scheme, iters, procs):
    """The reference FastS Fortran kernels (compiled from source into oracle/_ref) on the host
    (the oracle's C restatement if the reference library is not present), on the same workload:
    one thread on the whole 256^3 zone, and `procs` processes at once, each one thread on a slab
    of n / procs planes in k.  The reference parallelises one zone over OpenMP threads that each
    run the kernels on their own sub-box (block-Jacobi LU-SGS across the sub-boxes); independent
    slabs run the same kernels with no synchronisation between them, so the parallel leg bounds
    the OpenMP reference's throughput from above."""
    import multiprocessing as mp
    host = _host_description()
    dt, kind, _, _ = _cpu_run(n, scheme, iters)
    out = {"value": round(n ** 3 * iters / dt / 1e6, 3), "unit": "Mcell-iter/s", "cores": 1, "kind": kind,
           "sample": "%d implicit iterations of the same TGV %d^3 %s case, one thread" % (iters, n, scheme),
           "host": host}
    if procs <= 0:
        procs = min(16, host["affinity"])
    if procs > 1:
        nk = max(4, n // procs)
        mpc = mp.get_context("spawn")
        barrier = mpc.Barrier(procs)
        with mpc.Pool(procs, initializer=_cpu_init, initargs=(barrier,)) as pool:
            res = pool.map(_cpu_slab, [((n, n, nk), scheme, iters)] * procs, chunksize=1)
        # all slabs start together (barrier); the leg lasts until the last one ends
        wall = max(r[3] for r in res) - min(r[2] for r in res)
        cells = n * n * nk * procs
        out["parallel"] = {
            "value": round(cells * iters / wall / 1e6, 3), "unit": "Mcell-iter/s", "cores": procs,
            "kind": res[0][1],
            "sample": "%d processes x 1 thread, each %d implicit iterations on a %dx%dx%d slab of the "
                      "TGV case (independent slabs: block-Jacobi LU-SGS across them, no "
                      "synchronisation; an upper bound of the OpenMP reference)" % (procs, iters, n, n, nk),
            "slowest_process_s": round(max(r[0] for r in res), 3)}
    return out


# algorithmic bytes per interior cell for the other workloads (3dfull adds the 80 B of face
# normals and volume to the residual and to each LU sweep; local time stepping reads no BDF levels)
BYTES_FULL_LOCAL = {"coe": 168.0, "residual": 160.0, "lu": 400.0, "bc": 0.0, "iteration": 560.0}
BYTES_CART_LOCAL = {"coe": 88.0, "residual": 80.0, "lu": 240.0, "bc": 0.0, "iteration": 320.0}


def adjoint_ratio(n, scheme, reps):
    """Time of the transposed residual product (fm_residual_vjp) over the primal residual, both
    on the Euler variant of the workload (the transposed product covers the AUSM+(P) o3 Euler
    face balance)."""
    from fast_amd import _lib, cases, params as P
    num = P.Numerics(scheme=scheme, slope="o3", temporal_scheme="implicit", ss_iteration=1,
                     time_step_nature="global", time_step=1.0)
    z = cases.tgv_zone(n, num, flow="euler")
    z.param_real[P.DTC] = cases.tgv_dt(z, cfl=2.0)
    ctx = _lib.Context(0)
    zid = ctx.add_zone(z)
    del z
    ctx.apply_bc(zid, 0)
    res = ctx.time_phase(zid, _lib.PHASE_RESIDUAL, reps=reps)
    vjp = ctx.time_phase(zid, _lib.PHASE_VJP, reps=reps)
    ctx.close()
    return {"workload": "TGV %d^3 Euler, %s o3: residual vs transposed product (dR/dW)^T lam" % (n, scheme),
            "residual_ms": round(res, 4), "vjp_ms": round(vjp, 4), "ratio": round(vjp / res, 3),
            "vjp_gbs_algorithmic": round(480.0 * n ** 3 / (vjp * 1e-3) / 1e9, 1),
            "vjp_traffic": traffic_from_profiles("adjoint/vjp")}


def fasts_loop(a, resident_ms):
    """The same workload driven the way a FastS script drives it: a tree, warmup, then
    `for it: _compute(t, metrics, it)` (Fast/FastS/FastS/PyTree.py _compute), one time step per
    call, the state kept on the GPU between calls (no verification steps, as in the resident
    loop).  Returns its ms per step and the ratio to the resident fm_compute loop."""
    from fast_amd import Internal, PyTree as FP, params as P
    z = make_zone(a.n, a.scheme)
    num = P.Numerics(scheme=a.scheme, slope="o3", temporal_scheme="implicit", ss_iteration=1,
                     time_step_nature="global", time_step=float(z.param_real[P.DTC]), modulo_verif=0)
    t = Internal.newCGNSTree()
    b = Internal.newCGNSBase("Base", parent=t)
    FP.newZoneFromHost(z, b)
    FP._setNum2Zones(t, dict(num))
    FP._setNum2Base(t, {"modulo_verif": 0})
    FP._addState(t, z.state, model="NSLaminar")
    del z
    t, tc, metrics = FP.warmup(t, None)
    ctx = FP.HOOK["ctx"]
    for it in range(a.warmup):
        FP._compute(t, metrics, it)
    ctx.sync()
    t0 = time.perf_counter()
    for it in range(a.warmup, a.warmup + a.steps):
        FP._compute(t, metrics, it)
    ctx.sync()
    ms = (time.perf_counter() - t0) / a.steps * 1e3
    dens = Internal.getNodeFromName(t, "Density")[1]   # reading the tree pulls the state once
    ok = bool(np.isfinite(dens).all())
    ctx.close()
    FP.HOOK.clear()
    return {"ms_per_step": round(ms, 4), "resident_ms_per_step": round(resident_ms, 4),
            "ratio": round(ms / resident_ms, 4), "state_finite": ok,
            "loop": "for it: PyTree._compute(t, metrics, it), %d steps after %d warm-up calls"
                    % (a.steps, a.warmup)}


def zones64_case():
    """64 abutting blocks: a 4 x 4 x 4 tiling whose per-direction sizes are drawn from
    default_rng(1) uniformly in [16, 96] (so that neighbouring blocks share whole faces),
    periodic, laminar AUSM+ o3, implicit LU-SGS with local time stepping at CFL 5."""
    from fast_amd import cases, connect, params as P
    d = np.random.default_rng(1).integers(16, 97, size=(3, 4))
    num = P.Numerics(scheme="ausmpred", slope="o3", temporal_scheme="implicit", ss_iteration=1,
                     time_step_nature="local", cfl=5.0)
    zones, pairs, _ = cases.tiled_box(tuple(int(v) for v in d[0]), tuple(int(v) for v in d[1]),
                                      tuple(int(v) for v in d[2]), num, flow="lamin")
    return zones, connect.build(zones, pairs), d


def run_other(a):
    """channel / zones64: the same iteration on another mesh; one JSON line."""
    from fast_amd import _lib, cases
    ctx = _lib.Context(0)
    ctx.set_option(_lib.OPT_ZONE_BATCH, a.zone_batch)
    ctx.set_option(_lib.OPT_GRAPH, a.graph)
    ctx.set_option(_lib.OPT_LU_PRECISION, _lib.LU_FP32 if a.lu_precision == "fp32" else _lib.LU_FP64)
    ctx.set_option(_lib.OPT_LU_SCHEDULE, {"blocks": _lib.LU_BLOCKS, "pipeline": _lib.LU_PIPELINE,
                                          "dataflow": _lib.LU_DATAFLOW, "stream": _lib.LU_STREAM}[a.lu])
    if a.workload == "channel":
        zones, racs = [cases.channel_zone(257, 129, 33)], []
        desc = ("channel with a bump 257x129x33 (3dfull, jittered), laminar AUSM+ o3, implicit LU-SGS, "
                "local time step CFL 10, FP64")
        bpc = BYTES_FULL_LOCAL
    else:
        zones, racs, dims = zones64_case()
        desc = ("64 zones %s x %s x %s (4x4x4 abutting blocks, default_rng(1) in [16, 96]) joined by "
                "1-to-1 transfers, laminar AUSM+ o3, implicit LU-SGS, local time step CFL 5, FP64"
                % tuple(list(map(int, r)) for r in dims))
        bpc = BYTES_CART_LOCAL
    ids = [ctx.add_zone(z) for z in zones]
    ctx.set_transfers(racs)
    ctx.transfer(0)
    for i, z in zip(ids, zones):
        ctx.apply_bc(i, 0)
        buf = np.zeros_like(z.ro)
        ctx.get(i, _lib.FIELD_RO, buf)
        ctx.set(i, _lib.FIELD_RO_M1, buf)
        ctx.set(i, _lib.FIELD_RO_P1, buf)
    ncell = sum(int(np.prod(z.dims)) for z in zones)
    del zones
    ctx.compute(0, 1, 1, nit_c=a.warmup, modulo_verif=0, first_it=0)
    ctx.sync()
    t0 = time.perf_counter()
    ms_ev = ctx.compute_timed(a.warmup, 1, 1, nit_c=a.steps, modulo_verif=0, first_it=1)
    ctx.sync()
    wall = time.perf_counter() - t0
    value = ncell * a.steps / wall / 1e6
    tke = ctx.tke(ids[0], 0)
    if not np.isfinite(tke).all():
        raise RuntimeError("non-finite state after the benchmark")
    out = {"metric": "Mcell-iter/s NS residual+LU-SGS", "value": round(value, 2), "unit": "Mcell-iter/s",
           "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(wall / a.steps * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64" if a.lu_precision == "fp64" else "f64 (LU-SGS sweeps f32)",
           "data": "synthetic", "config": {"workload": desc, "cells": ncell, "zones": len(ids),
                                           "zone_batch": a.zone_batch, "graph": a.graph, "lu_schedule": a.lu},
           "event_ms_per_step": round(ms_ev / a.steps, 3),
           "iteration_hbm_frac": round(value * 1e6 * bpc["iteration"] / 1e9 / HBM_PEAK_GBS, 4)}
    if len(ids) == 1:
        phases = {}
        for name, ph in (("coe", _lib.PHASE_COE), ("residual", _lib.PHASE_RESIDUAL),
                         ("lu", _lib.PHASE_LU), ("bc", _lib.PHASE_BC)):
            phases[name] = ctx.time_phase(ids[0], ph, reps=a.phase_reps)
        out["phase_ms"] = {k: round(v, 4) for k, v in phases.items()}
        ach = bpc["lu"] * ncell / (phases["lu"] * 1e-3) / 1e9
        kname = KERNELS["lu"][a.lu]
        out["roofline"] = {"bound": "hbm", "kernel": kname, "achieved": round(ach, 1),
                           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                           "traffic": None, "bytes_per_cell": bpc["lu"]}
    ctx.close()
    print(json.dumps(out))


def _newton_to(ctx, zid, nmax, tol):
    """Sub-iterations of one BDF step until the L2 norm of the Newton residual is tol times its
    first value: (sub-iterations, seconds of GPU work including one norm read per sub-iteration,
    L2 history relative to the first)."""
    hist, first, secs = [], None, 0.
    for it in range(1, nmax + 1):
        t0 = time.perf_counter()
        ctx.substep(zid, it, first_it=0, lssiter_verif=1)
        r = ctx.residual_norms(zid)[:5]
        secs += time.perf_counter() - t0
        l2 = float(np.sqrt(np.sum(r ** 2)))
        first = l2 if first is None else first
        hist.append(l2 / first)
        if l2 <= tol * first:
            return it, secs, hist
    return None, secs, hist


def run_mixed(a):
    """FP64 sweeps against FP32 sweeps (state, Newton residual and update FP64) on one BDF step
    driven to a 1e-10 relative Newton residual: time, sub-iterations, and the converged state's
    error.  The FP64 GPU path equals the oracle bit for bit (the parity tests), so the error
    against the FP64 run is the error against the oracle."""
    from fast_amd import _lib, cases, params as P
    tol = 1e-10
    out = {"metric": "time to a 1e-10 Newton residual, FP64 vs FP32 LU-SGS sweeps", "configs": []}
    cfgs = [("TGV 256^3, ausmpred o3 laminar, BDF, global time step CFL 2", "tgv"),
            ("channel with a bump 513x257x65 (3dfull), ausmpred o3 laminar, BDF, local time step CFL 10",
             "channel")]
    runs = (("fp64 (default schedule)", None, _lib.LU_FP64), ("fp64 (dataflow)", _lib.LU_DATAFLOW, _lib.LU_FP64),
            ("fp32 sweeps (dataflow)", _lib.LU_DATAFLOW, _lib.LU_FP32))
    for desc, kind in cfgs:
        res = {"workload": desc, "tolerance": tol, "runs": {}}
        states = {}
        for name, sched, prec in runs:
            if kind == "tgv":
                num = P.Numerics(scheme="ausmpred", slope="o3", temporal_scheme="implicit",
                                 ss_iteration=a.mixed_nmax, time_step_nature="global", time_step=1.0)
                z = cases.tgv_zone((256, 256, 256), num)
                z.param_real[P.DTC] = cases.tgv_dt(z, cfl=2.0)
            else:
                num = P.Numerics(scheme="ausmpred", slope="o3", temporal_scheme="implicit",
                                 ss_iteration=a.mixed_nmax, time_step_nature="local", cfl=10.0)
                z = cases.channel_zone(513, 257, 65, numerics=num)
            ctx = _lib.Context(0)
            ctx.set_option(_lib.OPT_GRAPH, 0)
            if sched is not None:
                ctx.set_option(_lib.OPT_LU_SCHEDULE, sched)
            ctx.set_option(_lib.OPT_LU_PRECISION, prec)
            zid = ctx.add_zone(z)
            ctx.apply_bc(zid, 0)
            buf = np.zeros_like(z.ro)
            ctx.get(zid, _lib.FIELD_RO, buf)
            ctx.set(zid, _lib.FIELD_RO_M1, buf)
            ctx.set(zid, _lib.FIELD_RO_P1, buf)
            # one untimed warm-up sub-iteration sequence on a copy of the start (first-use
            # allocations), then the measured one from the same start
            _newton_to(ctx, zid, 2, 0.)
            ctx.set(zid, _lib.FIELD_RO_P1, buf)
            n, secs, hist = _newton_to(ctx, zid, a.mixed_nmax, tol)
            w = np.zeros_like(z.ro)
            ctx.get(zid, _lib.FIELD_RO_P1, w)
            states[name] = (z.interior(w), ctx.tke(zid, 2)[0])
            ctx.close()
            del z
            print("mixed-report: %s, %s: %s sub-iterations, %.3f s" % (kind, name, n, secs), file=sys.stderr, flush=True)
            res["runs"][name] = {"subiterations": n, "seconds": round(secs, 4),
                                 "ms_per_subiteration": round(secs / len(hist) * 1e3, 3),
                                 "final_relative_l2": hist[-1],
                                 "subiterations_to": {"%g" % t: next((q + 1 for q, h in enumerate(hist) if h <= t), None)
                                                      for t in (1e-4, 1e-6, 1e-8, 1e-9, 1e-10)}}
        ref, ke_ref = states["fp64 (default schedule)"]
        for name, (w, ke) in states.items():
            d = w - ref
            scale = np.abs(ref).reshape(5, -1).max(axis=1)
            res["runs"][name]["state_error_vs_fp64"] = {
                "l2_rel": [float(np.sqrt(np.mean(d[e] ** 2)) / scale[e]) for e in range(5)],
                "linf_rel": [float(np.abs(d[e]).max() / scale[e]) for e in range(5)],
                "ke_rel": float(abs(ke - ke_ref) / abs(ke_ref))}
        f64, f32 = res["runs"]["fp64 (default schedule)"], res["runs"]["fp32 sweeps (dataflow)"]
        if f64["subiterations"] and f32["subiterations"]:
            res["speedup_fp32_vs_fp64_default"] = round(f64["seconds"] / f32["seconds"], 3)
            res["speedup_fp32_vs_fp64_dataflow"] = round(res["runs"]["fp64 (dataflow)"]["seconds"] / f32["seconds"], 3)
        out["configs"].append(res)
    print(json.dumps(out))


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py runs on one GPU only (this build has no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    if a.mixed_report:
        run_mixed(a)
        return
    if a.workload != "tgv256":
        run_other(a)
        return
    # the host baseline first, before this process touches the GPU (its parallel leg starts
    # worker processes)
    cpu = None if a.no_cpu_baseline else cpu_baseline(a.n, a.scheme, a.cpu_iters, a.cpu_procs)
    from fast_amd import _lib
    z = make_zone(a.n, a.scheme)
    ctx = _lib.Context(0)
    ctx.set_option(_lib.OPT_LU_SCHEDULE, {"pipeline": _lib.LU_PIPELINE, "blocks": _lib.LU_BLOCKS,
                                          "dataflow": _lib.LU_DATAFLOW, "stream": _lib.LU_STREAM}[a.lu])
    ctx.set_option(_lib.OPT_RESIDUAL_SCHEDULE,
                   {"cells": _lib.RES_CELLS, "faces": _lib.RES_FACES, "kcol": _lib.RES_KCOL,
                    "auto": _lib.RES_AUTO}[a.residual])
    ctx.set_option(_lib.OPT_GRAPH, a.graph)
    ctx.set_option(_lib.OPT_LU_PRECISION, _lib.LU_FP32 if a.lu_precision == "fp32" else _lib.LU_FP64)
    zid = ctx.add_zone(z)
    ctx.apply_bc(zid, 0)
    buf = np.zeros_like(z.ro)
    ctx.get(zid, _lib.FIELD_RO, buf)
    ctx.set(zid, _lib.FIELD_RO_M1, buf)
    ctx.set(zid, _lib.FIELD_RO_P1, buf)
    del buf
    ncell = a.n ** 3
    tuned = None
    if a.autotune:
        from fast_amd import tune
        tuned = tune.autotune(ctx, zid)
        a.lu, a.residual = tuned["lu"][0], tuned["residual"][0]
    # warm-up (first step BDF1, then BDF2)
    ctx.compute(0, 1, 1, nit_c=a.warmup, modulo_verif=0, first_it=0)
    ctx.sync()
    walls, evs = [], []
    it = a.warmup
    for r in range(max(1, a.repeats)):
        t0 = time.perf_counter()
        evs.append(ctx.compute_timed(it, 1, 1, nit_c=a.steps, modulo_verif=0, first_it=1))
        ctx.sync()
        walls.append(time.perf_counter() - t0)
        it += a.steps
    order = np.argsort(walls)
    med = int(order[len(walls) // 2])
    wall, ms_ev = walls[med], evs[med]
    value = ncell * a.steps / wall / 1e6
    # health check of the state after the timed steps
    tke = ctx.tke(zid, 0)
    if not np.isfinite(tke).all():
        raise RuntimeError("non-finite state after the benchmark")
    # per-phase timings (HIP events on the solver stream)
    phases = {}
    for name, ph in (("coe", _lib.PHASE_COE), ("residual", _lib.PHASE_RESIDUAL),
                     ("lu", _lib.PHASE_LU), ("bc", _lib.PHASE_BC)):
        phases[name] = ctx.time_phase(zid, ph, reps=a.phase_reps)
    ctx.close()
    step_ms = wall / a.steps * 1e3
    rep_ms = [w / a.steps * 1e3 for w in walls]
    dom = max(("residual", "lu"), key=lambda k: phases[k])
    dom_ms = phases[dom]
    achieved = BYTES_PER_CELL[dom] * ncell / (dom_ms * 1e-3) / 1e9
    kname = KERNELS[dom] if isinstance(KERNELS[dom], str) else KERNELS[dom][a.lu if dom == "lu" else a.residual]
    roofline = {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic_from_profiles(dom + "/" + (a.lu if dom == "lu" else
                                                              "cells" if a.residual == "auto" else a.residual)),
                "bytes_per_cell": BYTES_PER_CELL[dom],
                "launch_ms": round(dom_ms, 4),
                "phase_gbs": {k: round(BYTES_PER_CELL[k] * ncell / (v * 1e-3) / 1e9, 1)
                              for k, v in phases.items() if BYTES_PER_CELL[k] > 0}}
    it_gbs = value * 1e6 * BYTES_PER_CELL["iteration"] / 1e9
    out = {"metric": "Mcell-iter/s NS residual+LU-SGS at 256^3", "value": round(value, 2),
           "unit": "Mcell-iter/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": round(step_ms, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64" if a.lu_precision == "fp64" else "f64 (LU-SGS sweeps f32)",
           "data": "synthetic (Taylor-Green vortex initial field)",
           "config": {"workload": "TGV %d^3 implicit BDF residual (%s o3 + laminar viscous) + LU-SGS, FP64"
                      % (a.n, a.scheme), "cells": ncell, "zones": 1, "mesh": "3dcart periodic",
                      "lu_schedule": a.lu, "residual_schedule": a.residual, "graph": a.graph,
                      "lu_precision": a.lu_precision,
                      "autotuned": None if tuned is None else {k: {n: round(t, 4) for n, t in v[1].items()}
                                                               for k, v in tuned.items()}},
           "iteration_hbm_frac": round(it_gbs / HBM_PEAK_GBS, 4),
           "phase_ms": {k: round(v, 4) for k, v in phases.items()},
           "event_ms_per_step": round(ms_ev / a.steps, 3),
           "protocol": {"warmup_steps": a.warmup, "repeats": len(walls), "steps_per_repeat": a.steps,
                        "ms_per_step_repeats": [round(v, 4) for v in rep_ms],
                        "ms_per_step_median": round(float(np.median(rep_ms)), 4),
                        "ms_per_step_stdev": round(float(np.std(rep_ms, ddof=1)) if len(rep_ms) > 1 else 0.0, 4),
                        "event_ms_per_step_repeats": [round(v / a.steps, 4) for v in evs]},
           "roofline": roofline}
    if not a.no_fasts_loop:
        out["fasts_loop"] = fasts_loop(a, step_ms)
    if not a.no_adjoint and a.scheme == "ausmpred":
        out["adjoint"] = adjoint_ratio(a.n, a.scheme, a.phase_reps)
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out))


if __name__ == "__main__":
    main()
