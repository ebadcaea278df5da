"""Benchmark: FastS NS residual + LU-SGS iteration, Taylor-Green vortex 256^3, FP64, one GPU.

One step = one implicit time step with one Newton sub-iteration: viscosity and LU
coefficients, fused inviscid (AUSM+ o3) + viscous face fluxes with the BDF residual, LU-SGS
lower and upper sweeps with the Newton update, and the ghost-cell fill (periodic faces and
corners).  Residual-only verification sub-iterations are not part of a step.

Prints one JSON line.  Timing: W untimed steps, then exactly K steps between two device
synchronisations (wall clock), cross-checked with HIP events on the solver stream.
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
           "residual": {"cells": "k_residual", "faces": "k_residual_faces", "kcol": "k_residual_kcol"},
           "lu": {"blocks": "k_lu_block (lower + upper sweeps, one launch per block diagonal)",
                  "pipeline": "k_lu_pipe (lower + upper sweeps)",
                  "dataflow": "k_lu_flow (lower + upper sweeps, one launch each)",
                  "stream": "k_lu_stream (lower + upper sweeps, one launch each)"}}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--scheme", default="ausmpred")
    ap.add_argument("--cpu-iters", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/primal timing")
    ap.add_argument("--phase-reps", type=int, default=5)
    ap.add_argument("--lu", choices=["pipeline", "blocks", "dataflow", "stream"], default="dataflow")
    ap.add_argument("--residual", choices=["cells", "faces", "kcol"], default="kcol")
    ap.add_argument("--lu-precision", choices=["fp64", "fp32"], default="fp64",
                    help="fp32: mixed precision, LU-SGS sweep arithmetic in FP32 (state, residual FP64)")
    ap.add_argument("--workload", choices=["tgv256", "channel", "zones64"], default="tgv256",
                    help="tgv256: the headline line; channel: 257x129x33 curvilinear channel with a "
                         "bump (3dfull, local time step, CFL 10); zones64: 64 abutting blocks "
                         "(4x4x4 tiling, sizes from default_rng(1) in [16, 96]) joined by transfers")
    ap.add_argument("--autotune", action="store_true",
                    help="time the interchangeable schedules first and keep the fastest (fast_amd.tune)")
    ap.add_argument("--graph", type=int, default=0, choices=[0, 1],
                    help="1 replays captured HIP graphs of the time step (FM_OPT_GRAPH)")
    ap.add_argument("--zone-batch", type=int, default=1, choices=[0, 1],
                    help="zones64: 1 batches the LU-SGS of all zones per block diagonal, 0 loops")
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


def cpu_baseline(n, scheme, iters):
    """The reference FastS Fortran kernels (compiled from source into oracle/_ref) on the host,
    one thread, on the same workload; falls back to the oracle's C restatement if the
    reference library is not present."""
    import oracle
    from fast_amd import params as P
    kind = "reference"
    if oracle.Reference.available():
        be = oracle.Reference()
    else:
        be = oracle.Oracle()
        kind = "port"
    z = make_zone(n, scheme)
    be.init_metric(z)
    be.apply_bc(z, z.ro)
    z.ro_m1[:] = z.ro
    z.ro_p1[:] = z.ro
    t0 = time.perf_counter()
    for s in range(iters):
        be.substep(z, 1, first_it=0 if s == 0 else 1)
        m1 = z.ro_m1.copy(); z.ro_m1[:] = z.ro; z.ro[:] = z.ro_p1; z.ro_p1[:] = m1
    dt = time.perf_counter() - t0
    ncell = n ** 3
    return {"value": ncell * iters / dt / 1e6, "unit": "Mcell-iter/s", "cores": 1, "kind": kind,
            "sample": "%d implicit iterations of the same TGV %d^3 %s case, single thread, %s"
                      % (iters, n, scheme, platform.processor() or platform.machine())}


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
                                           "zone_batch": a.zone_batch, "graph": a.graph},
           "event_ms_per_step": round(ms_ev / a.steps, 3),
           "iteration_hbm_frac": round(value * 1e6 * bpc["iteration"] / 1e9 / HBM_PEAK_GBS, 4)}
    if len(ids) == 1:
        phases = {}
        for name, ph in (("coe", _lib.PHASE_COE), ("residual", _lib.PHASE_RESIDUAL),
                         ("lu", _lib.PHASE_LU), ("bc", _lib.PHASE_BC)):
            phases[name] = ctx.time_phase(ids[0], ph, reps=a.phase_reps)
        out["phase_ms"] = {k: round(v, 4) for k, v in phases.items()}
        ach = bpc["lu"] * ncell / (phases["lu"] * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "kernel": "k_lu_block (lower + upper sweeps)", "achieved": round(ach, 1),
                           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                           "traffic": None, "bytes_per_cell": bpc["lu"]}
    ctx.close()
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
    if a.workload != "tgv256":
        run_other(a)
        return
    from fast_amd import _lib
    z = make_zone(a.n, a.scheme)
    ctx = _lib.Context(0)
    ctx.set_option(_lib.OPT_LU_SCHEDULE, {"pipeline": _lib.LU_PIPELINE, "blocks": _lib.LU_BLOCKS,
                                          "dataflow": _lib.LU_DATAFLOW, "stream": _lib.LU_STREAM}[a.lu])
    ctx.set_option(_lib.OPT_RESIDUAL_SCHEDULE,
                   {"cells": _lib.RES_CELLS, "faces": _lib.RES_FACES, "kcol": _lib.RES_KCOL}[a.residual])
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
    t0 = time.perf_counter()
    ms_ev = ctx.compute_timed(a.warmup, 1, 1, nit_c=a.steps, modulo_verif=0, first_it=1)
    ctx.sync()
    t1 = time.perf_counter()
    wall = t1 - t0
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
    dom = max(("residual", "lu"), key=lambda k: phases[k])
    dom_ms = phases[dom]
    achieved = BYTES_PER_CELL[dom] * ncell / (dom_ms * 1e-3) / 1e9
    kname = KERNELS[dom] if isinstance(KERNELS[dom], str) else KERNELS[dom][a.lu if dom == "lu" else a.residual]
    roofline = {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic_from_profiles(dom + "/" + (a.lu if dom == "lu" else a.residual)),
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
           "roofline": roofline}
    if not a.no_adjoint and a.scheme == "ausmpred":
        out["adjoint"] = adjoint_ratio(a.n, a.scheme, a.phase_reps)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(a.n, a.scheme, a.cpu_iters)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
