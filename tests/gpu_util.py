"""Helpers shared by the GPU parity tests: device transfers and the parity
measure the task states for the 8-bit output (max |diff| <= 1 with >= 99.9 %
of pixels identical; the integer blend itself is expected to be exact)."""
import sys

import numpy as np


def note(msg: str):
    print(msg, flush=True)
    sys.stdout.flush()


def to_dev(a: np.ndarray):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def from_dev(t) -> np.ndarray:
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy()


def parity(got_bgr: np.ndarray, ref_bgr: np.ndarray):
    d = np.abs(got_bgr.astype(np.int32) - ref_bgr.astype(np.int32))
    px_equal = (d.max(axis=-1) == 0).mean()
    return int(d.max()), float(px_equal)


def assert_parity(got_bgr, ref_bgr, exact=True, label=""):
    mx, eq = parity(got_bgr, ref_bgr)
    note(f"parity {label}: max|diff|={mx} identical={eq * 100:.4f}%")
    if exact:
        if mx != 0:
            bad = np.argwhere(np.abs(got_bgr.astype(int) - ref_bgr.astype(int)).max(-1) > 0)
            note(f"first mismatches (y, x): {bad[:8].tolist()}")
        assert mx == 0, f"{label}: max diff {mx}, identical {eq:.6f}"
    else:
        assert mx <= 1 and eq >= 0.999, f"{label}: max diff {mx}, identical {eq:.6f}"
