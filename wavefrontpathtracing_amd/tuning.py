"""Traversal kernel configuration and the autotuner over it.

Axes: workgroup size {64, 128, 256} x stack placement {LDS 4/8/12/16 entries
per lane with global spill, private (scratch) stack} x popped-interior skip
x node layout {2-wide pair records, 4-wide records} x refill threshold
(finished lanes of a wave before it claims new rays) x cost-ordered queue of
caller rays (on, off).
"""
from __future__ import annotations

import ctypes as C
import itertools

from . import _lib as L

# the library's default configuration (the autotuner's choice on the incoherent
# 256-instance workload); a candidate replaces it only when measurably faster
DEFAULT = dict(block=256, lds_stack=16, skip_popped=1, bvh_layout=0, refill_min=16, tlas_in_lds=0, interior_visits=4,
               cost_order=1)


def set_extend_config(renderer, block: int = DEFAULT["block"], lds_stack: int = DEFAULT["lds_stack"],
                      skip_popped: int = None, bvh_layout: int = DEFAULT["bvh_layout"], stack_limit: int = 0,
                      refill_min: int = DEFAULT["refill_min"], tlas_in_lds: int = DEFAULT["tlas_in_lds"],
                      interior_visits: int = DEFAULT["interior_visits"], big_claims: int = 0,
                      claim_segments: int = 0, cost_order: int = DEFAULT["cost_order"]) -> None:
    """skip_popped defaults to the library default, or to 1 with the 4-wide
    layout (which runs with the popped-interior skip only)."""
    if skip_popped is None:
        skip_popped = 1 if bvh_layout == 1 else DEFAULT["skip_popped"]
    c = L.TraversalConfig(block, lds_stack, skip_popped, bvh_layout, stack_limit, refill_min, tlas_in_lds,
                          interior_visits, big_claims, claim_segments, cost_order)
    L.check(L.lib().wf_renderer_set_traversal_config(renderer.handle, C.byref(c)), renderer.handle, "set_traversal_config")


def get_extend_config(renderer) -> dict:
    c = L.TraversalConfig()
    L.check(L.lib().wf_renderer_get_traversal_config(renderer.handle, C.byref(c)), renderer.handle, "get_traversal_config")
    return dict(block=c.workgroup_size, lds_stack=c.lds_stack_entries, skip_popped=c.skip_popped_interior,
                bvh_layout=c.bvh_layout, stack_limit=c.stack_limit, refill_min=c.refill_min,
                tlas_in_lds=c.tlas_in_lds, interior_visits=c.interior_visits, big_claims=c.big_claims,
                claim_segments=c.claim_segments, cost_order=c.cost_order)


def candidates():
    """Every workgroup size x stack placement x skip x layout the kernels
    support (the 4-wide layout runs with the popped-interior skip only)."""
    for block, lds, skip, layout in itertools.product((64, 128, 256), (0, 4, 8, 12, 16), (1, 0), (0, 1)):
        if layout == 1 and not skip:
            continue
        yield dict(block=block, lds_stack=lds, skip_popped=skip, bvh_layout=layout)


def label(cfg: dict) -> str:
    """Short name of a configuration: block/lds/skip/layout/visits/refill/cost-order (1 on, 0 off)."""
    c = {**DEFAULT, **cfg}
    order = 0 if c["cost_order"] == 2 else 1
    return (f"{c['block']}/{c['lds_stack']}/{c['skip_popped']}/{c['bvh_layout']}/{c['interior_visits']}/"
            f"{c['refill_min']}/{order}")


def autotune(renderer, time_fn, configs=None, repeats: int = 3, margin: float = 0.01):
    """Time every configuration with ``time_fn()`` (returns seconds for one
    measured pass, run after the config is set) and keep the fastest, unless
    it beats the first configuration in the list that equals DEFAULT by less
    than `margin` (then the default stays, so run-to-run noise between
    near-equal configurations does not flip the choice).
    Returns (chosen_config, [(config, seconds)] fastest first)."""
    results = []
    for cfg in (configs or list(candidates())):
        set_extend_config(renderer, **cfg)
        time_fn()   # warm-up
        best = min(time_fn() for _ in range(repeats))
        results.append((cfg, best))
    results.sort(key=lambda x: x[1])
    chosen = results[0][0]
    for cfg, t in results:
        if {**DEFAULT, **cfg} == DEFAULT and results[0][1] > t * (1.0 - margin):
            chosen = cfg
            break
    set_extend_config(renderer, **chosen)
    return chosen, results
