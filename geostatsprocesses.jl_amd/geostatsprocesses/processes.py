"""GaussianProcess (src/processes/field/gaussian.jl:22-48) and traversal paths
(Meshes `traverse`: LinearPath, RandomPath)."""
from __future__ import annotations

import numpy as np


class FieldProcess:
    pass


class GaussianProcess(FieldProcess):
    """GaussianProcess(func, mean=0.0)."""

    def __init__(self, func, mean=None):
        nf = func.nvariables()
        if mean is None:
            mean = 0.0 if nf == 1 else np.zeros(nf)
        nm = 1 if np.ndim(mean) == 0 else len(mean)
        assert nm == nf, f"mean must have {nf} components, received {nm}"
        self.func = func
        self.mean = mean

    def iscontinuous(self):
        return True

    def isanalytical(self):
        return True

    def defaultnames(self):
        nf = self.func.nvariables()
        return ("field",) if nf == 1 else tuple(f"field{i + 1}" for i in range(nf))


class IndicatorProcess(FieldProcess):
    """IndicatorProcess(func, prob) or IndicatorProcess(transiogram)
    (src/processes/field/indicator.jl:29-47):
    categorical field with multivariate function `func` (one variable per
    category) and a priori probabilities `prob` (rescaled to sum one)."""

    def __init__(self, func, prob=None):
        if prob is None:
            # IndicatorProcess(transiogram) takes the transiogram's proportions
            # (indicator.jl:46)
            if not hasattr(func, "proportions"):
                raise TypeError("IndicatorProcess needs probabilities or a transiogram")
            prob = func.proportions()
        prob = np.asarray(prob, dtype=np.float64)
        prob = prob / prob.sum()
        nf = func.nvariables()
        if not nf > 1:
            raise AssertionError("indicator process requires multivariate function")
        if len(prob) != nf:
            raise AssertionError(f"probabilities must have {nf} components, received {len(prob)}")
        if not np.all((prob >= 0) & (prob <= 1)):
            raise AssertionError("probabilities must be in [0, 1]")
        if not np.isclose(prob.sum(), 1.0):
            raise AssertionError("probabilities must sum up to one")
        self.func = func
        self.prob = prob

    def iscontinuous(self):
        return False

    def isanalytical(self):
        return False

    def defaultnames(self):
        return ("field",)


class LinearPath:
    """Meshes LinearPath: elements in linear-index order."""

    def traverse(self, domain) -> np.ndarray:
        return np.arange(domain.nelements(), dtype=np.int64)


class RandomPath:
    """Meshes RandomPath: a permutation drawn from the path's own generator."""

    def __init__(self, rng=None, seed=None):
        if rng is None:
            rng = np.random.default_rng(seed)
        self.rng = rng

    def traverse(self, domain) -> np.ndarray:
        return self.rng.permutation(domain.nelements()).astype(np.int64)


class SourcePath:
    """Meshes SourcePath-like order: the given sources first, then the rest by
    linear index (a permutation of all elements)."""

    def __init__(self, sources):
        self.sources = np.asarray(sources, dtype=np.int64)

    def traverse(self, domain) -> np.ndarray:
        n = domain.nelements()
        seen = np.zeros(n, dtype=bool)
        src = []
        for s in self.sources:
            if not seen[s]:
                seen[s] = True
                src.append(int(s))
        rest = np.flatnonzero(~seen)
        return np.concatenate([np.asarray(src, dtype=np.int64), rest.astype(np.int64)])


class FixedPath:
    """An explicit traversal order (a permutation of 0..N-1)."""

    def __init__(self, order):
        self.order = np.asarray(order, dtype=np.int64)

    def traverse(self, domain) -> np.ndarray:
        return self.order
