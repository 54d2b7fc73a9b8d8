"""Error metrics shared by the parity tests."""
import numpy as np

# FP32 output vs the FP64 oracle of the same lattice: per-element relative
# error with a floor of 1e-3 of the largest magnitude (pixels whose
# homogeneous weight is tiny would otherwise amplify rounding without bound).
FP32_REL_TOL = 1e-5
FLOOR = 1e-3


def rel_err(a, b, floor=FLOOR):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if b.size == 0:
        return 0.0
    den = np.maximum(np.abs(b), floor * np.abs(b).max())
    return float((np.abs(a - b) / den).max())


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def digest(arr) -> str:
    """SHA-256 (first 16 hex digits) of an array's bytes: fixture fingerprints."""
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest()[:16]
