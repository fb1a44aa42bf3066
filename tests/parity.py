"""Parity metric shared by the GPU tests.

Scalars: max over cells of |gpu - ref| / max(|ref|, floor), floor = 1e-8 * max|ref| over
the field.  Vector components (velocity, momentum): max over cells of
|gpu_i - ref_i| / max(|ref vector|, floor) -- judged against the magnitude of the vector
at that cell, the rotation-invariant reading of "relative error per step".  The reference
rotates cell velocities into each face frame and back in place (onedinterp.d:761-770,
983-986), which perturbs every component by rounding of the vector's magnitude, so a
small component cannot be compared against its own size.  The FP64 bar is 1e-12.
"""
import numpy as np

TOL = 1.0e-12
VECTORS = {"vx": ("vx", "vy", "vz"), "vy": ("vx", "vy", "vz"), "vz": ("vx", "vy", "vz")}


def rel_err(a, b, scale=None):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if not b.size:
        return 0.0
    ref = np.abs(b) if scale is None else np.asarray(scale, dtype=np.float64)
    floor = max(1.0e-8 * float(np.max(ref)), 1.0e-300)
    return float(np.max(np.abs(a - b) / np.maximum(ref, floor)))


def vec_mag(get, names):
    return np.sqrt(sum(get(n) ** 2 for n in names))


def field_errors(gpu_get, ref_get, names, U_gpu=None, U_ref=None, dims=3):
    """Errors of flow-state fields (names) and, if given, conserved quantities U (ncell x ncq)."""
    out = {}
    vmag = vec_mag(ref_get, ("vx", "vy", "vz"))
    for n in names:
        scale = vmag if n in VECTORS else None
        out[n] = rel_err(gpu_get(n), ref_get(n), scale)
    if U_gpu is not None:
        nmom = 3 if dims == 3 else 2
        mmag = np.sqrt(np.sum(U_ref[:, 1:1 + nmom] ** 2, axis=1))
        for j in range(U_ref.shape[1]):
            scale = mmag if 1 <= j <= nmom else None
            out[f"U{j}"] = rel_err(U_gpu[:, j], U_ref[:, j], scale)
    return out
