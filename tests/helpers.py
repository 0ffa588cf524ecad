"""Shared helpers of the test suite."""
import sys

import numpy as np
import torch

from oracle import hcipy_oracle as o


def og(grid):
    """Oracle grid with the same coordinates and weights as a product grid."""
    if grid.is_regular:
        g = o.Grid(np.array(grid.delta), tuple(int(d) for d in grid.dims), np.array(grid.zero))
    else:
        g = o.Grid(separated=grid.separated_coords)
    g._weights = grid._weights
    return g


def dev(a, dtype=None):
    t = torch.as_tensor(np.asarray(a), device='cuda')
    return t if dtype is None else t.to(dtype)


def host(t):
    if isinstance(t, torch.Tensor):
        return t.as_subclass(torch.Tensor).detach().cpu().numpy()
    return np.asarray(t)


def peak_rel_field(a, ref):
    return np.abs(a - ref).max() / np.abs(ref).max()


def peak_rel_intensity(a, ref, peak=None):
    Ia = np.abs(a) ** 2
    Ir = np.abs(ref) ** 2
    return np.abs(Ia - Ir).max() / (Ir.max() if peak is None else peak)


def say(*args):
    print(*args, file=sys.stderr, flush=True)
