"""Top-p (nucleus) filtering and sampling on the scan and sort operators.

The reference composes top-p in Python from its operators
(scripts/profiling/profile_tcuscan_ops.py:335-392): sort the probabilities,
scan them, zero every entry whose preceding mass already exceeds p,
renormalise, draw with torch.multinomial and map the draw back through the
sort's indices.  Here the sort is run_radix_sort_ex (descending, stable: the
order of sample_top_p at :372-392; the reference's tcuscan_top_p sorts
ascending), the scan run_scan_multi_core (fp16 in, float32 out, each prefix
rounded once) and the mask one fused kernel (run_top_p_keep).

sample_top_p draws by inverting the kept weights' cumulative distribution:
their inclusive scan (run_scan_single_core, float32 rounded once per prefix)
and one uniform variate located by binary search.  That is the same
distribution as renormalising and calling torch.multinomial, with no
renormalisation pass, but not the same random stream.
"""
from __future__ import annotations

import torch

import tcuscan_ops


def _kept(probs: torch.Tensor, p: float, S: int):
    if probs.dim() != 1 or probs.dtype != torch.float16:
        raise RuntimeError("top_p: probs must be a 1-D float16 tensor")
    probs_sort, order = tcuscan_ops.run_radix_sort_ex(probs, True)
    mass = tcuscan_ops.run_scan_multi_core(probs_sort, S)
    return tcuscan_ops.run_top_p_keep(probs_sort, mass, p), order


def top_p_filter(probs: torch.Tensor, p: float, S: int = 128):
    """(kept, order): probabilities sorted largest first with every entry
    whose preceding mass exceeds p set to 0 and the rest renormalised to sum
    1 (float32), and the int32 input positions of that order.

    probs: 1-D float16 on the GPU.
    """
    kept, order = _kept(probs, p, S)
    return kept / kept.sum(), order


def sample_top_p(probs: torch.Tensor, p: float, S: int = 128, generator: torch.Generator | None = None):
    """One token index (int64 tensor of shape [1]) drawn from the top-p
    nucleus of `probs`."""
    kept, order = _kept(probs, p, S)
    cdf = tcuscan_ops.run_scan_single_core(kept, S, 0.0)
    total = cdf[-1:]
    u = torch.rand(1, generator=generator, device=probs.device) * total
    pick = torch.searchsorted(cdf, u, right=True)
    # u rounds up to the total at most: never past the last kept entry.
    last = torch.searchsorted(cdf, total)
    return order.gather(0, torch.minimum(pick, last)).long()
