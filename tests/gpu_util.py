"""Helpers for the GPU tests: device buffers via torch, copies to numpy."""
import numpy as np


def dev_empty(torch, n):
    return torch.empty(n * 4, dtype=torch.int64, device="cuda")


def to_dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int64).reshape(-1)).to("cuda")


def to_host(t):
    return t.cpu().numpy().view(np.uint64).reshape(-1, 4).copy()


def ptr(t):
    return t.data_ptr()
