"""TEST INFRASTRUCTURE - NOT PRODUCT CODE.

Interface interpolation weights of a mask formula, restated with numpy for the
oracle (the oracle's C++ march takes the weights as an input array).  Only
tests/ and bench.py's cpu_baseline leg use this module, as the checker.

Follows UtilitiesLib/VectorField.cpp:64-113 of the reference: for every
column (ix, iy) and slab iz (planes iz, iz+1, centred mesh coordinates), when
the formula changes sign (or vanishes) between the two planes, the crossing
is bracketed by bisection until the bracket is narrower than 1e-3 dz, and
the weights of the slab's two planes become ((z - z_iz) / dz, 1 - that),
with z the last midpoint evaluated (or the midpoint where the formula is
exactly zero).  A crossing between two planes of equal mask value means the
formula disagrees with the mask; the reference then prints a notice and
keeps every weight at 0.5 (:107-111).

The bisection runs on every slab of every column at once: each bracket
halves per iteration, so all of them finish after the same number of steps,
except those that hit an exact zero, which stop there.
"""
import numpy as np

MISMATCH_NOTICE = "Mask values do not corresponds to the given formula, will ignore interpolation correction"


def interface_weights(formula_fn, dims, spacings, mask):
    """formula_fn(x, y, z) -> numpy array (broadcasting); dims (Nx, Ny, Nz);
    mask (Nz, Ny, Nx) float64 (>= 0 is LC).  Returns (weights, mismatch):
    weights of shape (Nz - 1, Ny, Nx, 2), all 0.5 on a mismatch."""
    Nx, Ny, Nz = dims
    dx, dy, dz = spacings
    x = ((np.arange(Nx) - 0.5 * (Nx - 1)) * dx)[None, None, :]
    y = ((np.arange(Ny) - 0.5 * (Ny - 1)) * dy)[None, :, None]
    iz = np.arange(Nz - 1)[:, None, None]
    zi = (iz - 0.5 * (Nz - 1)) * dz
    zm = np.broadcast_to(zi, (Nz - 1, Ny, Nx)).astype(np.float64).copy()
    zp = np.broadcast_to((iz + 1 - 0.5 * (Nz - 1)) * dz, (Nz - 1, Ny, Nx)).astype(np.float64).copy()
    xb = np.broadcast_to(x, zm.shape)
    yb = np.broadcast_to(y, zm.shape)
    f = lambda z: np.asarray(formula_fn(xb, yb, z), dtype=np.float64)  # noqa: E731
    cross = f(zm) * f(zp) <= 0
    m = np.asarray(mask) >= 0
    mismatch = bool(np.any(cross & (m[:-1] == m[1:])))
    z = zm.copy()
    active = cross.copy()
    while True:
        active &= np.abs(zp - zm) >= 1e-3 * dz
        if not active.any():
            break
        fm = f(zm)
        mid = 0.5 * (zp + zm)
        fz = f(mid)
        z = np.where(active, mid, z)
        zero = active & (fz == 0)
        move_p = active & ~zero & (fz * fm < 0)
        move_m = active & ~zero & ~(fz * fm < 0)
        zp = np.where(move_p, mid, zp)
        zm = np.where(move_m, mid, zm)
        active &= ~zero
    w = np.full((Nz - 1, Ny, Nx, 2), 0.5)
    if mismatch:
        return w, True
    w0 = (z - zi) / dz
    w[..., 0] = np.where(cross, w0, 0.5)
    w[..., 1] = np.where(cross, 1 - w0, 0.5)
    return w, False
