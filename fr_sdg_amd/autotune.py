"""Launch-shape autotuner for the batched sweep.

Knobs (frsdg_batch_set_tuning): the number of persistent workgroups of the
stage kernels (or one workgroup per element tile), the element tile shape
(the default tile of each order, or that tile with its rows halved: more,
shorter work items, better balance across uneven instances), and the launch
mode (one kernel per stage over all orders; one specialised kernel per order
and mesh kind; or the same cut into cache-sized runs of instances that take
their three stages back to back, so stages 2 and 3 find their inputs in the
last-level cache).  Each candidate is timed over
a few graph-free batched steps on a scratch copy of the state; the winner is
cached in a JSON file keyed by the batch signature (per-instance mesh size,
order, gas model) and the device name.
"""
import hashlib
import json
import os

import torch

CACHE = os.environ.get("FRSDG_TUNING_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "fr_sdg_amd", "tuning.json"))


def signature(batch):
    items = [(c["nx"], c["ny"], c["p"], c.get("gas_kind", 0), c.get("geometry", 0), c.get("apply_filter", 0))
             for c in batch.cfgs]
    dev = torch.cuda.get_device_name(batch.device)
    return hashlib.sha1(json.dumps([dev, items]).encode()).hexdigest()


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(db):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        with open(CACHE, "w") as f:
            json.dump(db, f, indent=1)
    except OSError:
        pass


def candidates():
    """(grid, epb_div, mode); grid 0 = one workgroup per element tile."""
    ncu = torch.cuda.get_device_properties(0).multi_processor_count
    return [(g * ncu, d, m) for m in (1, 2, 0) for g in (0, 1, 2, 4, 8) for d in (1, 2)]


def tune_batch(batch, u, steps=3, use_cache=True):
    """Pick (grid, epb_div, mode) for `batch` by timing; returns (best, table)."""
    key = signature(batch)
    db = _load() if use_cache else {}
    if use_cache and key in db:
        g, d, m = db[key]["best"]
        batch.set_tuning(g, d, m)
        return (g, d, m), db[key]["table"]
    scratch = u.clone()
    table = []
    s = torch.cuda.current_stream()
    for g, d, m in candidates():
        batch.set_tuning(g, d, m)
        scratch.copy_(u)
        batch.ssprk_step(scratch)  # warm
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(s)
        for _ in range(steps):
            batch.ssprk_step(scratch)
        b.record(s)
        b.synchronize()
        table.append([g, d, m, a.elapsed_time(b) / steps])
    best = min(table, key=lambda r: r[3])
    batch.set_tuning(best[0], best[1], best[2])
    db[key] = {"best": best[:3], "table": table}
    if use_cache:
        _store(db)
    return (best[0], best[1], best[2]), table
