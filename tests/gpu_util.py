"""Helpers shared by the GPU tests: device buffers through torch, and the oracle side of a check."""
import sys

import numpy as np


def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        import pytest
        pytest.skip("no GPU")
    return torch


def dev(a, dtype=None):
    torch = torch_cuda()
    a = np.ascontiguousarray(a)
    t = torch.from_numpy(a.copy())
    return t.to("cuda:0")


def host(t):
    return t.detach().cpu().numpy()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and a.view(np.uint64).tolist() == b.view(np.uint64).tolist()


def first_diff(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    d = np.nonzero(a.view(np.uint64) != b.view(np.uint64))[0]
    if len(d) == 0:
        return None
    i = d[0]
    return dict(index=int(i), count=int(len(d)), got=a[i].hex(), want=b[i].hex())
