"""The C ABI library loads (no GPU needed to load it) and exports every entry point declared
in include/fastmi.h; the Python binding declares the same set."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fastmi.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fm_\w+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    for n in ("fm_ctx_create", "fm_zone_add", "fm_compute", "fm_zone_substep", "fm_apply_bc",
              "fm_residual_norms"):
        assert n in names


def test_library_exports_all_declared_symbols():
    from fast_amd import _lib
    if not os.path.exists(_lib.LIBPATH):
        pytest.fail("libfastmi.so not built: run __graft_entry__.build()")
    lib = _lib.load()
    for n in declared():
        assert hasattr(lib, n), n
    assert set(declared()) == set(_lib.SIGNATURES)


def test_missing_library_fails_loudly(tmp_path):
    from fast_amd import _lib
    with pytest.raises(_lib.FastMIError):
        _lib.load(str(tmp_path / "libfastmi.so"))


def test_autotune_picks_the_fastest_schedule():
    """The autotuner sets, for every knob, the value with the smallest measured time."""
    from fast_amd import tune, _lib

    class FakeCtx:
        def __init__(self):
            self.opts = {}
            self.cost = {(_lib.OPT_RESIDUAL_SCHEDULE, _lib.RES_CELLS): 2.0,
                         (_lib.OPT_RESIDUAL_SCHEDULE, _lib.RES_FACES): 1.5,
                         (_lib.OPT_RESIDUAL_SCHEDULE, _lib.RES_KCOL): 1.7,
                         (_lib.OPT_LU_SCHEDULE, _lib.LU_BLOCKS): 3.0,
                         (_lib.OPT_LU_SCHEDULE, _lib.LU_PIPELINE): 4.0,
                         (_lib.OPT_LU_SCHEDULE, _lib.LU_DATAFLOW): 3.5,
                         (_lib.OPT_LU_SCHEDULE, _lib.LU_STREAM): 3.2}

        def set_option(self, opt, v):
            self.opts[opt] = v

        def time_phase(self, zone, phase, reps=1):
            opt = _lib.OPT_RESIDUAL_SCHEDULE if phase == _lib.PHASE_RESIDUAL else _lib.OPT_LU_SCHEDULE
            return self.cost[(opt, self.opts[opt])]

    c = FakeCtx()
    out = tune.autotune(c, 0)
    assert out["residual"][0] == "faces" and out["lu"][0] == "blocks"
    assert c.opts[_lib.OPT_RESIDUAL_SCHEDULE] == _lib.RES_FACES
    assert c.opts[_lib.OPT_LU_SCHEDULE] == _lib.LU_BLOCKS


def test_autotune_shapes_the_schedule_it_kept():
    """Tile shapes are tuned only for the schedule chosen: the k-column length when the column
    residual wins, the dataflow grid when the dataflow LU-SGS wins."""
    from fast_amd import tune, _lib

    class FakeCtx:
        def __init__(self):
            self.opts = {_lib.OPT_RES_KCOL_N: 0, _lib.OPT_LU_FLOW_GRID: 0}
            self.cost = {(_lib.OPT_RESIDUAL_SCHEDULE, _lib.RES_CELLS): 2.0,
                         (_lib.OPT_RESIDUAL_SCHEDULE, _lib.RES_FACES): 2.5,
                         (_lib.OPT_RESIDUAL_SCHEDULE, _lib.RES_KCOL): 1.7,
                         (_lib.OPT_LU_SCHEDULE, _lib.LU_BLOCKS): 3.0,
                         (_lib.OPT_LU_SCHEDULE, _lib.LU_PIPELINE): 4.0,
                         (_lib.OPT_LU_SCHEDULE, _lib.LU_DATAFLOW): 2.5,
                         (_lib.OPT_LU_SCHEDULE, _lib.LU_STREAM): 3.2}
            self.calls = []

        def set_option(self, opt, v):
            self.opts[opt] = v

        def time_phase(self, zone, phase, reps=1):
            self.calls.append(phase)
            if phase == _lib.PHASE_RESIDUAL:
                base = self.cost[(_lib.OPT_RESIDUAL_SCHEDULE, self.opts[_lib.OPT_RESIDUAL_SCHEDULE])]
                return base - {0: 0., 2: 0.1, 4: 0., 8: 0.2}[self.opts[_lib.OPT_RES_KCOL_N]]
            base = self.cost[(_lib.OPT_LU_SCHEDULE, self.opts[_lib.OPT_LU_SCHEDULE])]
            return base - {0: 0., 256: 0.3, 384: 0.1}[self.opts[_lib.OPT_LU_FLOW_GRID]]

    c = FakeCtx()
    out = tune.autotune(c, 0)
    assert out["residual"][0] == "kcol" and out["kcol_n"][0] == "8"
    assert out["lu"][0] == "dataflow" and out["flow_grid"][0] == "256"
    assert c.opts[_lib.OPT_RES_KCOL_N] == 8 and c.opts[_lib.OPT_LU_FLOW_GRID] == 256
