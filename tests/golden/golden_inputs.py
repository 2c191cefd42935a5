"""Closed-form inputs shared by the golden-vector writers and the tests that
read their fixtures.  Imports nothing from the reference, so the GPU tests may
use it where the reference is absent."""
import numpy as np


def golden_sysmat(detector: np.ndarray, coor: np.ndarray) -> np.ndarray:
    """Positive smooth stand-in system matrix [D, P] (float32), a closed form of the inputs."""
    d = detector[:, None, :3].astype(np.float64) - coor[None, :, :3].astype(np.float64)
    dist = np.sqrt((d ** 2).sum(-1))
    k = np.arange(detector.shape[0])[:, None]
    return (np.exp(-dist / 150.0) * (1.25 + np.sin(0.37 * k))).astype(np.float32)
