"""Helpers shared by the tests (tensor <-> numpy, seeded inputs)."""
import numpy as np
import torch


def to_np(t: torch.Tensor) -> np.ndarray:
    t = t.detach().cpu()
    if t.dtype == torch.bfloat16:
        return t.float().numpy()
    if t.dtype == torch.uint32:
        return t.view(torch.int32).numpy().view(np.uint32)
    return t.numpy()


def gen(seed: int) -> torch.Generator:
    return torch.Generator().manual_seed(seed)


def uniform_pm1(n: int, dtype: torch.dtype, seed: int) -> torch.Tensor:
    """U[-1, 1] rounded to `dtype` (the large-scan workload's input)."""
    return (torch.rand(n, generator=gen(seed), dtype=torch.float32) * 2 - 1).to(dtype)


def assert_exact(actual: np.ndarray, expected: np.ndarray, what: str = ""):
    """Bit-for-bit equality of values (+0.0 and -0.0 compare equal)."""
    assert actual.shape == expected.shape, f"{what}: shape {actual.shape} != {expected.shape}"
    assert actual.dtype == expected.dtype, f"{what}: dtype {actual.dtype} != {expected.dtype}"
    bad = np.flatnonzero(~((actual == expected) | (np.isnan(actual) & np.isnan(expected)))
                         if actual.dtype.kind == "f" else actual != expected)
    assert bad.size == 0, (
        f"{what}: {bad.size} mismatches, first at {bad[:5].tolist()}: "
        f"{actual[bad[:5]].tolist()} vs {expected[bad[:5]].tolist()}")


def sparse_random(nrow: int, ncol: int, density: float, seed: int):
    """CSR matrix with round(density * nrow * ncol) distinct uniformly random
    positions and values uniform in [-1, 1) (scipy.sparse.random's shape and
    density; its legacy sampler permutes all nrow * ncol positions, which
    takes seconds at 16K x 16K)."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    k = int(round(density * nrow * ncol))
    pos = rng.choice(nrow * ncol, size=k, replace=False)
    data = 2 * rng.random(k) - 1
    m = sp.csr_matrix((data, (pos // ncol, pos % ncol)), shape=(nrow, ncol))
    m.sort_indices()
    return m
