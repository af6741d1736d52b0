"""Oracle (test infrastructure only) for earthquake events built from unit-source
gauge series: a numpy restatement of make_tsunami_event_from_unit_sources
(R/examples/austptha_template/SOURCE_ZONES/TEMPLATE/TSUNAMI_EVENTS/
sum_tsunami_unit_sources.R:555-711).

Per event, in the reference's order:
  * msl is subtracted from every unit source's stage (:606-609);
  * the template starts at zero and each of the event's unit sources is added with
    an axpy, weight 1 for uniform slip (:658-669) or the source's slip for
    stochastic slip (:672-687);
  * uniform slip then scales the sum by the event slip (:669);
  * msl is added back to the stage (:690);
  * an optional summary reduces the event series (:694-699); the summary kept here
    is the maximum stage over time at each gauge.

Series are arrays [sample][3][gauge] (variables stage, uh, vh), the layout the
batched solver records. Pinned by tests/test_superpose_oracle.py against the
linearity of the truly-linear solver (a superposed event equals the solver run from
the superposed initial condition, with and without msl).
"""
import numpy as np


def event_series(series, nsamples, src, weights=None, slip=None, msl=0.0):
    """series: list of [sample][3][gauge] arrays (one per unit source, at least as
    long as nsamples[k]); src: the event's unit-source indices; weights: per-source
    slips (stochastic) or None (uniform, then scaled by slip). Returns the event
    series over the samples common to its sources."""
    n = min(nsamples[k] for k in src) if len(src) else min(nsamples)
    flow = [np.array(series[k][:n], dtype=np.float64) for k in src]
    for f in flow:
        f[:, 0, :] = f[:, 0, :] - msl
    template = np.zeros_like(series[src[0]][:n] if len(src) else series[0][:n], dtype=np.float64)
    for q, f in enumerate(flow):
        a = 1.0 if weights is None else float(weights[q])
        template = template + a * f
    if weights is None and slip is not None:
        template = template * float(slip)
    template[:, 0, :] = template[:, 0, :] + msl
    return template


def event_max_stage(series, nsamples, src, weights=None, slip=None, msl=0.0):
    """Maximum over time of the event's stage at each gauge."""
    return event_series(series, nsamples, src, weights, slip, msl)[:, 0, :].max(axis=0)
