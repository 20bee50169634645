"""Launch-shape autotuner for the batched sweep.

Knobs: the number of persistent workgroups of the stage kernels (or one
workgroup per element tile), the launch mode (one kernel per stage over all
orders; one specialised kernel per order and mesh kind; or the same cut into
cache-sized runs of instances that take their three stages back to back, so
stages 2 and 3 find their inputs in the last-level cache), both through
frsdg_batch_set_tuning, and the element tile of each order
(frsdg_batch_set_tile_shape: the default tile of the order, that tile halved,
or the other tile with the fewest rim faces per element).  The search runs in
two passes, each candidate timed over a few graph-free batched steps on a
scratch copy of the state: (1) grid x {default, halved tiles} x mode over the
whole batch; (2) from the winner, a coordinate pass over the tile shape of one
order at a time.  The result is cached in a JSON file keyed by the batch
signature (per-instance mesh size, order, gas model) and the device name.
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


def _time(batch, u, scratch, steps):
    s = torch.cuda.current_stream()
    scratch.copy_(u)
    batch.ssprk_step(scratch)  # warm
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(s)
    for _ in range(steps):
        batch.ssprk_step(scratch)
    b.record(s)
    b.synchronize()
    return a.elapsed_time(b) / steps


def _apply(batch, best):
    g, d, m, shapes = best
    batch.set_tuning(g, d, m)
    for n, sh in shapes.items():
        batch.set_tile_shape(int(n), int(sh))


def tune_batch(batch, u, steps=3, use_cache=True):
    """Pick (grid, epb_div, mode, {order: tile shape}) for `batch` by timing;
    returns (best, table): table rows [grid, epb_div, mode, shapes, ms]."""
    key = signature(batch)
    db = _load() if use_cache else {}
    if use_cache and key in db:
        best = tuple(db[key]["best"])
        _apply(batch, best)
        return best, db[key]["table"]
    scratch = u.clone()
    table = []
    for g, d, m in candidates():
        batch.set_tuning(g, d, m)
        table.append([g, d, m, {}, _time(batch, u, scratch, steps)])
    g, d, m, _, t_best = min(table, key=lambda r: r[4])
    # pass 2: tile shape per order, one order at a time from the pass-1 winner
    shapes = {n: d - 1 for n in batch.orders()}
    batch.set_tuning(g, d, m)
    for n in batch.orders():
        for sh in (0, 1, 2):
            if sh == shapes[n]:
                continue
            batch.set_tile_shape(n, sh)
            trial = dict(shapes)
            trial[n] = sh
            t = _time(batch, u, scratch, steps)
            table.append([g, d, m, trial, t])
            if t < t_best:
                t_best, shapes = t, trial
            else:
                batch.set_tile_shape(n, shapes[n])
    best = (g, d, m, shapes)
    _apply(batch, best)
    db[key] = {"best": list(best), "table": table}
    if use_cache:
        _store(db)
    return best, table
