"""A file cache of er_autotune's launch shapes.

The shapes depend on the system (atoms, molecules, PME grid and order,
interaction type, box shape, run type), on the solutes' site counts and
insertions per frame, and on the GPU, so a cache file belongs to one kind of
GPU.  A key found in the file is applied with er_set_tuning; otherwise the
context is tuned on its current frame and insertions and the result is
added to the file (written to a temporary name, then renamed).
"""
from __future__ import annotations

import json
import os


def tuning_key(topo, params: dict, nsites, nins: int) -> str:
    """The cache key of a context: everything the launch shapes depend on
    apart from the GPU."""
    ms = params.get("ms", (0, 0, 0))
    parts = [f"atoms={topo.natom}", f"molecules={topo.nmol}", "grid=" + "x".join(str(int(m)) for m in ms),
             f"order={params.get('splodr', 0)}", f"cltype={params.get('cltype', 2)}",
             f"boxshp={params.get('boxshp', 1)}", f"slttype={params.get('slttype', 3)}",
             "sites=" + ",".join(str(int(n)) for n in nsites), f"insertions={int(nins)}"]
    return " ".join(parts)


def load(path: str) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        data = json.load(fh)
    if not isinstance(data, dict):
        raise ValueError(f"{path}: not a tuning cache")
    return data


def tune_cached(ctx, key: str, path: str, reps: int = 0):
    """(shapes, found): the cached shapes applied to ctx, or ctx tuned now
    and the cache file extended."""
    cache = load(path)
    if key in cache:
        ctx.set_tuning(cache[key])
        return cache[key], True
    shapes = ctx.autotune(reps)
    cache[key] = shapes
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(cache, fh, indent=1, sort_keys=True)
        fh.write("\n")
    os.replace(tmp, path)
    return shapes, False
