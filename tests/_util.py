"""Helpers shared by the GPU parity tests."""
import numpy as np
import torch


def to_dev(a: np.ndarray, device="cuda:0") -> torch.Tensor:
    """numpy (any layout) -> device tensor with Julia column-major memory."""
    a = np.asarray(a)
    perm = tuple(range(a.ndim - 1, -1, -1))
    return torch.from_numpy(np.ascontiguousarray(a.transpose(perm))).to(device).permute(*perm)


def to_host(t: torch.Tensor) -> np.ndarray:
    return t.detach().cpu().numpy()


def rel_l2(x, ref) -> float:
    x = np.asarray(x)
    ref = np.asarray(ref)
    den = np.linalg.norm(ref.ravel())
    return float(np.linalg.norm((x - ref).ravel()) / (den if den > 0 else 1.0))
