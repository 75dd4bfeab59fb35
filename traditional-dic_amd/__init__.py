"""MI355X (gfx950) subset-correlation hot path of Traditional-DIC.

libdic_hip.so (csrc/, C-ABI in include/dic_hip.h) holds the B-spline
prefilter, the reference-gradient kernels and the IC-GN kernels; this
package is the Python host layer over that C-ABI.
"""
from . import _lib
from ._lib import (Context, DicError, INITIAL_DTYPE, RESULT_DTYPE, IcgnParams, load_library, device_count,
                   PRECISION_FP64, PRECISION_MIXED, SUCCESS, NOT_CONVERGED, INVALID_INPUT, NUMERICAL_FAILURE,
                   GRADIENT_CENTRAL, GRADIENT_BSPLINE_HALF, Batch, BatchItem, BatchTuning, BATCH_GRID,
                   BATCH_PERSISTENT, BATCH_TILE, BATCH_ORDER_NATURAL, BATCH_ORDER_COST, BATCH_ORDER_INTERLEAVE,
                   BATCH_ORDER_XCD, Server, ServedRequest, SEED_DTYPE, ESTIMATE_DTYPE, SEED_EVAL_DTYPE)
from .subset import (ICGNSolver, SubsetConfig, BSplinePrecomputeConfig, config_from_dict, compute_grid,
                     initial_guesses, make_params, results_to_dict, PreparedPair, ForwardGaussNewtonSolver,
                     make_solver, propagate, SeedIntegerSearchConfig, SeedSubpixelRefinementConfig,
                     SeedInitializationConfig, SeedSelectionConfig)
from .seeding import (IntegerSearchInitializer, SeedSelector, SeedSelectionResult, SubsetDIC, subset,
                      recommended_subset_padding, seed_params, search_params)
from . import workflow, run_contract

__all__ = [
    "Context", "DicError", "INITIAL_DTYPE", "RESULT_DTYPE", "IcgnParams", "load_library", "device_count",
    "PRECISION_FP64", "PRECISION_MIXED", "SUCCESS", "NOT_CONVERGED", "INVALID_INPUT", "NUMERICAL_FAILURE",
    "GRADIENT_CENTRAL", "GRADIENT_BSPLINE_HALF", "ICGNSolver", "SubsetConfig", "BSplinePrecomputeConfig",
    "config_from_dict", "compute_grid", "initial_guesses", "make_params", "results_to_dict", "PreparedPair",
    "ForwardGaussNewtonSolver", "make_solver", "propagate", "SEED_DTYPE",
    "Batch", "BatchItem", "BatchTuning", "BATCH_GRID", "BATCH_PERSISTENT", "BATCH_TILE", "BATCH_ORDER_NATURAL",
    "BATCH_ORDER_COST", "BATCH_ORDER_INTERLEAVE", "BATCH_ORDER_XCD", "Server", "ServedRequest",
    "ESTIMATE_DTYPE", "SEED_EVAL_DTYPE", "SeedIntegerSearchConfig", "SeedSubpixelRefinementConfig",
    "SeedInitializationConfig", "SeedSelectionConfig", "IntegerSearchInitializer", "SeedSelector",
    "SeedSelectionResult", "SubsetDIC", "subset", "recommended_subset_padding", "seed_params", "search_params",
    "workflow", "run_contract",
]
