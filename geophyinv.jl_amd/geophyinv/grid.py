"""Uniform 1-D grids, the host counterpart of Julia's ``StepRangeLen{Float64}``.

The reference stores every medium, field and time grid as a Julia range
(src/media/media.jl:39-62, src/fields.jl:92-671).  ``Grid`` keeps (first, step, length)
and evaluates points as ``first + k*step`` in float64.
"""
from __future__ import annotations

import numpy as np


class Grid:
    __slots__ = ("first", "step", "n")

    def __init__(self, first: float, step: float, n: int):
        if n < 1:
            raise ValueError("grid length must be positive")
        self.first = float(first)
        self.step = float(step)
        self.n = int(n)

    @classmethod
    def range(cls, start: float, stop: float | None = None, *, length: int | None = None,
              step: float | None = None) -> "Grid":
        """``range(start, stop=..., length=...)`` or ``range(start, step=..., length=...)``."""
        if length is None:
            raise ValueError("length is required")
        if step is None:
            if stop is None:
                raise ValueError("need stop or step")
            step = (float(stop) - float(start)) / (length - 1) if length > 1 else 0.0
        return cls(start, step, length)

    def __len__(self) -> int:
        return self.n

    def __getitem__(self, k: int) -> float:
        if k < 0:
            k += self.n
        if not 0 <= k < self.n:
            raise IndexError(k)
        return self.first + k * self.step

    @property
    def last(self) -> float:
        return self.first + (self.n - 1) * self.step

    def values(self) -> np.ndarray:
        return self.first + np.arange(self.n, dtype=np.float64) * self.step

    def __repr__(self) -> str:
        return f"Grid({self.first!r}, step={self.step!r}, n={self.n})"

    def __eq__(self, other) -> bool:
        return isinstance(other, Grid) and (self.first, self.step, self.n) == (other.first, other.step, other.n)
