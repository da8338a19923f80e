"""Signal-edge extensions along the last axis for stft(boundary=...)
(reference: src/mlx_signal_processing/_arraytools.py:16-50)."""

from __future__ import annotations

import torch


def _check_ext_len(x: torch.Tensor, n: int) -> None:
    if n > x.shape[-1] - 1:
        raise ValueError(
            f"The extension length n ({n}) is too big. It must not exceed "
            f"x.shape[-1]-1, which is {x.shape[-1] - 1}."
        )


def even_ext(x: torch.Tensor, n: int) -> torch.Tensor:
    """[x[n]..x[1], x, x[-2]..x[-n-1]] (mirror, edge not repeated)."""
    if n < 1:
        return x
    _check_ext_len(x, n)
    left = torch.flip(x[..., 1 : n + 1], dims=(-1,))
    right = torch.flip(x[..., -(n + 1) : -1], dims=(-1,))
    return torch.cat([left, x, right], dim=-1)


def odd_ext(x: torch.Tensor, n: int) -> torch.Tensor:
    """Antisymmetric extension about the edge values."""
    if n < 1:
        return x
    _check_ext_len(x, n)
    left = 2 * x[..., :1] - torch.flip(x[..., 1 : n + 1], dims=(-1,))
    right = 2 * x[..., -1:] - torch.flip(x[..., -(n + 1) : -1], dims=(-1,))
    return torch.cat([left, x, right], dim=-1)


def const_ext(x: torch.Tensor, n: int) -> torch.Tensor:
    """Repeat the edge values n times."""
    if n < 1:
        return x
    left = x[..., :1].expand(*x.shape[:-1], n)
    right = x[..., -1:].expand(*x.shape[:-1], n)
    return torch.cat([left, x, right], dim=-1)


def zero_ext(x: torch.Tensor, n: int) -> torch.Tensor:
    """Zero-pad n samples on both sides."""
    if n < 1:
        return x
    z = torch.zeros(*x.shape[:-1], n, dtype=x.dtype, device=x.device)
    return torch.cat([z, x, z], dim=-1)
