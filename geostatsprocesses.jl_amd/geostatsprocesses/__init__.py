"""geostatsprocesses — MI355X backend for GeoStatsProcesses' GaussianProcess
simulation by SEQSIM and LUSIM, behind the reference's
`rand(process, domain, nreals; method=...)` API (src/simulation/field.jl:47-124).

The arithmetic runs in libgsp_hip.so (HIP, gfx950) through its C ABI
(include/gsp.h); this package mirrors the reference's Julia interface so the
same calls and the same plugin contract (preprocess / randsingle) apply.
The Gaussian posterior mean (src/expectation/field/gaussian.jl:21-25) runs the
same neighbour search and kriging kernels over the data only.
"""
from . import _lib
from ._lib import (ArgumentError, Context, GspError, PosDefException, UnsupportedError,
                   default_context)
from .ensembles import Ensemble
from .expectation import KrigingPlan, expectedvalue, mean, posteriormean, priormean
from .functions import (CoregionalizedFunction, ExponentialCovariance, ExponentialTransiogram,
                        ExponentialVariogram, GaussianCovariance, GaussianTransiogram,
                        GaussianVariogram, SphericalCovariance, SphericalTransiogram,
                        SphericalVariogram, Transiogram)
from .geometry import CartesianGrid, GeoTable, PointSet, georef
from .initialization import ExplicitInit, NearestInit, initialize
from .processes import (FixedPath, GaussianProcess, IndicatorProcess, LinearPath, RandomPath,
                        SourcePath)
from .simulation import (FFTSIM, LUSIM, SEQSIM, MetricBall, defaultsimulation, preprocess, rand,
                         randsingle)

KNN_THEN_MASK = _lib.KNN_THEN_MASK
MASKED_KNN = _lib.MASKED_KNN
DATA_ONLY = _lib.DATA_ONLY
FP64 = _lib.FP64
FP32_REFINE = _lib.FP32_REFINE
LU_FP64 = _lib.LU_FP64
LU_MIXED = _lib.LU_MIXED

__all__ = [
    "ArgumentError", "Context", "GspError", "PosDefException", "UnsupportedError",
    "default_context", "Ensemble", "ExponentialCovariance", "ExponentialVariogram",
    "GaussianCovariance", "GaussianVariogram", "SphericalCovariance", "SphericalVariogram",
    "CoregionalizedFunction", "IndicatorProcess", "Transiogram", "SphericalTransiogram",
    "ExponentialTransiogram", "GaussianTransiogram",
    "CartesianGrid", "GeoTable", "PointSet", "georef", "ExplicitInit", "NearestInit",
    "initialize", "FixedPath", "GaussianProcess", "LinearPath", "RandomPath", "SourcePath",
    "FFTSIM", "LUSIM", "SEQSIM", "MetricBall", "defaultsimulation", "preprocess", "rand",
    "randsingle", "KrigingPlan", "expectedvalue", "mean", "posteriormean", "priormean",
    "KNN_THEN_MASK", "MASKED_KNN", "DATA_ONLY", "FP64", "FP32_REFINE", "LU_FP64", "LU_MIXED",
]
