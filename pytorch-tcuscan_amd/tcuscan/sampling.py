"""Top-p (nucleus) filtering and sampling on the scan and sort operators.

The reference composes top-p in Python from its operators
(scripts/profiling/profile_tcuscan_ops.py:335-392): sort the probabilities,
scan them, zero every entry whose preceding mass already exceeds p,
renormalise, draw with torch.multinomial and map the draw back through the
sort's indices.  Here the sort is run_radix_sort_ex (descending, stable: the
order of sample_top_p at :372-392; the reference's tcuscan_top_p sorts
ascending) and the scan run_scan_multi_core (fp16 in, float32 out, each
prefix rounded once).  The mask, renormalisation and draw are elementwise
torch ops on the device.
"""
from __future__ import annotations

import torch

import tcuscan_ops


def top_p_filter(probs: torch.Tensor, p: float, S: int = 128):
    """(kept, order): probabilities sorted largest first with every entry
    whose preceding mass exceeds p set to 0 and the rest renormalised to sum
    1 (float32), and the int32 input positions of that order.

    probs: 1-D float16 on the GPU.
    """
    if probs.dim() != 1 or probs.dtype != torch.float16:
        raise RuntimeError("top_p_filter: probs must be a 1-D float16 tensor")
    probs_sort, order = tcuscan_ops.run_radix_sort_ex(probs, True)
    mass = tcuscan_ops.run_scan_multi_core(probs_sort, S)
    kept = probs_sort.float()
    kept = torch.where(mass - kept > p, torch.zeros_like(kept), kept)
    kept = kept / kept.sum()
    return kept, order


def sample_top_p(probs: torch.Tensor, p: float, S: int = 128, generator: torch.Generator | None = None):
    """One token index (int64 tensor of shape [1]) drawn from the top-p
    nucleus of `probs`."""
    kept, order = top_p_filter(probs, p, S)
    pick = torch.multinomial(kept, num_samples=1, generator=generator)
    return order.long().gather(0, pick)
