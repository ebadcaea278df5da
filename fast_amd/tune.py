"""Schedule and tile-shape autotuner: times the interchangeable kernel schedules of a context on
one of its zones and keeps the fastest, then the tile shapes of the schedules it kept.  Every
choice gives the same results bit for bit, so it only affects speed.

Knobs (fm_ctx_set_option): the residual schedule (six faces per cell / every face once /
columns of cells up k) and the LU-SGS schedule (block diagonals / persistent pipelined tiles /
dataflow blocks / streaming pencils); tile shapes: the k-column length of the column residual
and the number of workgroups of a dataflow sweep."""
from . import _lib

KNOBS = (
    ("residual", _lib.OPT_RESIDUAL_SCHEDULE, _lib.PHASE_RESIDUAL,
     (("cells", _lib.RES_CELLS), ("faces", _lib.RES_FACES), ("kcol", _lib.RES_KCOL))),
    ("lu", _lib.OPT_LU_SCHEDULE, _lib.PHASE_LU,
     (("dataflow", _lib.LU_DATAFLOW), ("blocks", _lib.LU_BLOCKS), ("pipeline", _lib.LU_PIPELINE),
      ("stream", _lib.LU_STREAM))),
)

# (tile knob, schedule knob, schedule it shapes, option, phase, values)
TILES = (
    ("kcol_n", "residual", "kcol", _lib.OPT_RES_KCOL_N, _lib.PHASE_RESIDUAL,
     (("2", 2), ("4", 4), ("8", 8))),
    ("flow_grid", "lu", "dataflow", _lib.OPT_LU_FLOW_GRID, _lib.PHASE_LU,
     (("resident", 0), ("256", 256), ("384", 384))),
)


def _pick(ctx, zone, opt, phase, values, reps):
    times = {}
    for vname, v in values:
        ctx.set_option(opt, v)
        ctx.time_phase(zone, phase, reps=1)
        times[vname] = ctx.time_phase(zone, phase, reps=reps)
    best = min(times, key=times.get)
    ctx.set_option(opt, dict(values)[best])
    return best, times


def autotune(ctx, zone, reps=3, knobs=KNOBS, tiles=TILES):
    """Time every value of every knob on ``zone`` (HIP events on the context stream, one
    untimed run first) and set the fastest; then the tile shapes of the chosen schedules.
    Returns {knob: (chosen name, {name: ms})}."""
    out = {}
    for name, opt, phase, values in knobs:
        out[name] = _pick(ctx, zone, opt, phase, values, reps)
    for name, knob, sched, opt, phase, values in tiles:
        if knob in out and out[knob][0] == sched:
            out[name] = _pick(ctx, zone, opt, phase, values, reps)
    return out
