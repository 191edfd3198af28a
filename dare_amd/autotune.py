"""Autotuner cache for batched solves.

dare_batch_autotune (C-ABI) times every candidate configuration — lanes per tile along k
(16/32/64 or by each system's nz), planes per tile, persistent grid size and the order of the
work list — on the batch itself and keeps the fastest. This module keys the result by the
batch's shape (system sizes, components) and the device, and stores it in a JSON file so a
later run with the same batch skips the search.
"""
import hashlib
import json
import os

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "dare_amd", "batch_autotune.json")


def batch_key(dims, ncomp, device_name):
    """dims: list of (n0, n1, n2) per system, in batch order."""
    h = hashlib.sha1(json.dumps([list(map(int, d)) for d in dims]).encode()).hexdigest()[:16]
    return f"{device_name}|N{ncomp}|{len(dims)}sys|{h}"


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def tune(batch, dims, ncomp, device_name, cache_path=None, reps=5, force=False):
    """Configuration for `batch` (a dare_amd.Batch): from the cache, else measured and stored.
    Returns (config dict, microseconds per operator sweep or None when cached)."""
    path = cache_path or os.environ.get("DARE_AUTOTUNE_CACHE", DEFAULT_CACHE)
    key = batch_key(dims, ncomp, device_name)
    cache = _load(path)
    if not force and key in cache:
        cfg = cache[key]["config"]
        batch.set_config(**cfg)
        return cfg, None
    cfg, us = batch.autotune(reps)
    cache[key] = {"config": cfg, "us_per_sweep": us}
    try:
        os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass
    return cfg, us
