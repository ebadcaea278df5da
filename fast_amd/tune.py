"""Schedule autotuner: times the interchangeable kernel schedules of a context on one of its
zones and keeps the fastest.  Every schedule gives the same results bit for bit, so the choice
only affects speed.

Knobs (fm_ctx_set_option): the residual schedule (six faces per cell / every face once /
columns of cells up k) and
the LU-SGS schedule (block diagonals / persistent pipelined tiles)."""
from . import _lib

KNOBS = (
    ("residual", _lib.OPT_RESIDUAL_SCHEDULE, _lib.PHASE_RESIDUAL,
     (("cells", _lib.RES_CELLS), ("faces", _lib.RES_FACES), ("kcol", _lib.RES_KCOL))),
    ("lu", _lib.OPT_LU_SCHEDULE, _lib.PHASE_LU,
     (("dataflow", _lib.LU_DATAFLOW), ("blocks", _lib.LU_BLOCKS), ("pipeline", _lib.LU_PIPELINE))),
)


def autotune(ctx, zone, reps=3, knobs=KNOBS):
    """Time every value of every knob on ``zone`` (HIP events on the context stream, one
    untimed run first) and set the fastest.  Returns {knob: (chosen name, {name: ms})}."""
    out = {}
    for name, opt, phase, values in knobs:
        times = {}
        for vname, v in values:
            ctx.set_option(opt, v)
            ctx.time_phase(zone, phase, reps=1)
            times[vname] = ctx.time_phase(zone, phase, reps=reps)
        best = min(times, key=times.get)
        ctx.set_option(opt, dict(values)[best])
        out[name] = (best, times)
    return out
