"""Names of the reference module FastS.FastS (Fast/FastS/FastS/FastS.py).

``metric(xyz, numerics)`` there forwards to ``fasts.metric``, which the extension module does not
register (Fast/FastS/FastS/fastS.cpp:25-67), so it cannot run in the reference.  The name is kept
with the behaviour its docstring announces ("Compute Metric"): the numerics, when given, are set
on every zone of the tree, then the metric of every zone is allocated and computed on the GPU
(``allocate_metric`` + ``_init_metric`` of fast_amd.PyTree).
"""
from . import PyTree


def metric(xyz, numerics=None):
    """Metric (face vectors and volumes) of every zone of the tree ``xyz``; returns the list of
    per-zone metric arrays, as PyTree.allocate_metric does."""
    if numerics is not None:
        PyTree._setNum2Zones(xyz, dict(numerics))
    metrics = PyTree.allocate_metric(xyz)
    PyTree._init_metric(xyz, metrics)
    return metrics
