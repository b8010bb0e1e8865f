"""Helpers shared by the GPU parity tests: numpy <-> device tensors, comparisons."""
from __future__ import annotations

import numpy as np

import solver_amd as sa


def dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def host(t):
    return t.detach().cpu().numpy()


def zeros_like_np(a):
    import torch

    return torch.zeros(a.shape, dtype=torch.float64, device="cuda")


def zeros(g, loc):
    import torch

    return torch.zeros(sa.storage_shape(g.nx, g.ny, g.nz, loc), dtype=torch.float64, device="cuda")


def zeros_vel(g):
    return [zeros(g, sa.FACE_X), zeros(g, sa.FACE_Y), zeros(g, sa.FACE_Z)]


def rel_max(a, b):
    scale = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)))) / scale


def rel_l2_active(a, b, mean_removed=False):
    """Relative L2 over the active interior (reference helper tests/test_runtime.cpp:304-345)."""
    x = np.asarray(a)[1:-1, 1:-1, 1:-1].astype(np.float64)
    y = np.asarray(b)[1:-1, 1:-1, 1:-1].astype(np.float64)
    if mean_removed:
        x = x - x.mean()
        y = y - y.mean()
    den = np.sqrt(np.mean(y * y))
    return float(np.sqrt(np.mean((x - y) ** 2)) / max(den, 1e-300))


def grid_of(p):
    return sa.grid(p["nx"], p["ny"], p["nz"], p["dx"], p["dy"], p["dz"])


def bcs_of(p):
    vel = np.array(p.get("bcvel", [0.0] * 18)).reshape(6, 3)
    return sa.bc_set(p["bc"], vel, p.get("bcp", [0.0] * 6))


def pbcs_of(p):
    return sa.pressure_bc_set(p["pbc"], p.get("pbcval"), p.get("pbcgrad"))
