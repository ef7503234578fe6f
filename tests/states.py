"""Synthetic states with known answers, shared by the oracle and GPU tests."""
import numpy as np


def hydrostatic_state(p, grad, along_y=False):
    """Fluid at rest with p = 1 + grad * x (or y) at cell centres, SoA (5, ny, nx)."""
    nx, ny = p.nx, p.ny
    x = (np.arange(nx) + 0.5) * p.dx
    y = (np.arange(ny) + 0.5) * p.dy
    X, Y = np.meshgrid(x, y)
    pres = 1.0 + grad * (Y if along_y else X)
    e = pres / (p.gamma - 1.0)
    return np.stack([np.ones_like(pres), np.zeros_like(pres), np.zeros_like(pres), e, e]).astype(np.float64)
