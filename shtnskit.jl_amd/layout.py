"""Packed SHTns LM order (src/layout.jl:143-162, 282-339): m-major blocks,
``index = (im*(2*lmax+2-(im+1)*mres))/2 + l`` with ``im = m/mres``."""
from __future__ import annotations

import torch


def LM_index(lmax: int, mres: int, l: int, m: int) -> int:
    if m < 0:
        raise ValueError("m must be ≥ 0 for packed layout")
    if m % mres:
        raise ValueError("m must be a multiple of mres")
    if not (m <= l <= lmax):
        raise ValueError("require m ≤ l ≤ lmax")
    im = m // mres
    return (im * (2 * lmax + 2 - (im + 1) * mres)) // 2 + l


def _index_pairs(cfg):
    ls, ms = [], []
    for m in range(0, cfg.mmax + 1, cfg.mres):
        for l in range(m, cfg.lmax + 1):
            ls.append(l)
            ms.append(m)
    return torch.tensor(ls), torch.tensor(ms)


def pack_lm(cfg, A: torch.Tensor) -> torch.Tensor:
    """Dense (lmax+1, mmax+1) -> packed vector of length cfg.nlm."""
    if tuple(A.shape[:2]) != (cfg.lmax + 1, cfg.mmax + 1):
        raise ValueError(f"A must have shape ({cfg.lmax + 1}, {cfg.mmax + 1})")
    ls, ms = _index_pairs(cfg)
    return A[ls.to(A.device), ms.to(A.device)]


def unpack_lm(cfg, q: torch.Tensor) -> torch.Tensor:
    """Packed vector -> dense (lmax+1, mmax+1); entries without a packed slot are zero."""
    if q.shape[0] != cfg.nlm:
        raise ValueError(f"Qlm must have length {cfg.nlm}")
    ls, ms = _index_pairs(cfg)
    A = torch.zeros((cfg.lmax + 1, cfg.mmax + 1) + tuple(q.shape[1:]), dtype=q.dtype, device=q.device)
    A[ls.to(q.device), ms.to(q.device)] = q
    return A
