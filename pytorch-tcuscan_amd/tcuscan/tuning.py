"""Persisted autotuner cache of the batched plans.

BatchPlan.tune() times candidate tilings of every scan op on the device and
keeps the fastest per (op kind, dtype, ceil(log2 n)) for the process.  These
helpers write that cache to a JSON file and read it back in a later
process, so a plan built there skips the timing runs.  The file records the
device it was measured on (architecture and compute-unit count); entries from
another device are not loaded.  No counterpart in the reference, whose
tiling comes from fixed host heuristics (src/tiling/heuristics/).
"""
from __future__ import annotations

import json
import os

import torch

import tcuscan_ops

_FIELDS = ("kind", "dtype", "log2n", "S", "tiles_per_wg", "method", "window", "ms")


def default_path() -> str:
    """$TCUSCAN_TUNE_CACHE, else ~/.cache/tcuscan/batch_tuning.json."""
    return os.environ.get("TCUSCAN_TUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "tcuscan", "batch_tuning.json")


def device_key(device: int = 0) -> dict:
    """What a cache file must match to be loaded: the GPU architecture
    (gcnArchName without its feature suffixes) and compute-unit count."""
    p = torch.cuda.get_device_properties(device)
    arch = getattr(p, "gcnArchName", "") or ""
    return {"arch": arch.split(":")[0], "compute_units": int(p.multi_processor_count)}


def save(path: str | None = None, device: int = 0) -> str:
    """Writes this process's tuner cache to `path` (atomically); returns the path."""
    path = path or default_path()
    entries = [dict(zip(_FIELDS, e)) for e in tcuscan_ops._tuned_entries()]
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = f"{path}.tmp{os.getpid()}"
    with open(tmp, "w") as f:
        json.dump({"version": 1, "device": device_key(device), "entries": entries}, f, indent=1)
    os.replace(tmp, path)
    return path


def load(path: str | None = None, device: int = 0) -> int:
    """Loads the entries of `path` measured on this device into the tuner's
    cache; returns how many were loaded (0 when the file is missing or was
    written for another device)."""
    path = path or default_path()
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return 0
    if d.get("version") != 1 or d.get("device") != device_key(device):
        return 0
    n = 0
    for e in d.get("entries", []):
        tcuscan_ops._set_tuned(*(e[k] for k in _FIELDS))
        n += 1
    return n


def tune(plan, reps: int = 3, path: str | None = None, force: bool = False, device: int = 0) -> None:
    """plan.tune() with the persisted cache: loads `path` first, tunes what
    it lacks, and saves the result."""
    load(path, device)
    plan.tune(reps, force)
    save(path, device)
