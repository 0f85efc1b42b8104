"""tcuscan for AMD Instinct MI355X.

Re-exports the `tcuscan_ops` extension module, as the reference package does
(python/tcuscan/__init__.py:3-5 of huawei-csl/pytorch-tcuscan).  Importing
fails loudly when the HIP extension has not been built: there is no CPU
fallback for the GPU operators.
"""
from __future__ import annotations

import os
import sys

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if _PKG_ROOT not in sys.path:
    sys.path.insert(0, _PKG_ROOT)

import torch  # noqa: E402,F401  (loads the HIP runtime the extension shares)

try:
    import tcuscan_ops  # noqa: E402
except ImportError as exc:  # pragma: no cover - exercised only without a build
    raise ImportError(
        "tcuscan_ops is not built: run `python pytorch-tcuscan_amd/build.py` "
        "(or __graft_entry__.build()) first"
    ) from exc

from tcuscan_ops import *  # noqa: E402,F401,F403

from . import tuning  # noqa: E402
from .sampling import sample_top_p, top_p_filter  # noqa: E402

__all__ = [n for n in dir(tcuscan_ops) if n.startswith("run_")] + ["BatchPlan", "Server", "sample_top_p",
                                                                   "top_p_filter", "tuning"]
