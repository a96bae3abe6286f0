"""Window functions (reference: windows.py).  Only `cosine` is wired into a
decomposition (decompositions.py:199), and only it is accelerated."""

import math

import torch


def cosine(xmin, xmax, x):
    """windows.py:25-35 for a SINGLE point x of shape (xd,)."""
    mu, sd = (xmin + xmax) / 2, (xmax - xmin) / 2
    ws = ((1 + torch.cos(math.pi * (x - mu) / sd)) / 2) ** 2
    ws = (x >= xmin).to(x.dtype) * (xmax >= x).to(x.dtype) * ws
    return torch.prod(ws, dim=0, keepdim=True)
