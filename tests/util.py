"""Shared helpers for the parity tests."""
import numpy as np


def rel_err(a, b):
    """max over cells of |a-b|/|b| (0 where both are 0; |a-b| where b=0)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    d = np.abs(a - b)
    d[both_nan] = 0.0
    den = np.abs(b)
    r = np.where(den > 0, d / np.where(den > 0, den, 1.0), d)
    return float(np.max(r)) if r.size else 0.0


def to_dev(x, device):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


def to_host(t):
    return t.detach().cpu().numpy()
