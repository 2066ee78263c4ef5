"""xerus_amd — MI355X-native Tensor-Train engine behind xerus's TT API.

The product is libxtt.so (HIP/C++ for gfx950, C ABI in include/xtt.h); this package is its Python face.
"""
from ._lib import XT_BF16, XT_F32, XT_F64, XttError, lib
from .tt import (ADF, ALS, ALS_SPD, DMRG, DMRG_SPD, EPSILON, ADFVariant, ALSVariant, FileFormat, PerformanceData, RankOneMeasurementSet,
                 SinglePointMeasurementSet, TTOperator, TTTensor, adf_gradient, als_round_batch, frob_norm, load_from_file, normal_draws, residual, save_to_file,
                 seed, solve, uniform_draws)

__all__ = [
    "ADF", "ALS", "ALS_SPD", "DMRG", "DMRG_SPD", "EPSILON", "ADFVariant", "ALSVariant", "FileFormat", "PerformanceData", "RankOneMeasurementSet",
    "SinglePointMeasurementSet",
    "TTOperator", "TTTensor", "adf_gradient", "als_round_batch", "frob_norm", "load_from_file", "normal_draws", "residual",
    "save_to_file", "seed", "solve", "uniform_draws",
    "XT_BF16", "XT_F32", "XT_F64", "XttError", "lib",
]
