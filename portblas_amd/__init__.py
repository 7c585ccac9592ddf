"""portblas_amd — MI355X-native implementation of portBLAS's hot path.

GEMM (plain, batched, strided-batched, split-precision, grouped), GEMV and the
BLAS-1 / extension reductions run as hand-written gfx950 HIP kernels in
``libportblas_amd.so`` behind the C ABI of ``include/portblas_amd.h``.  This
package is the Python mirror of the reference's API (``blas.py``) plus the
autotuner and learned tile selector (``tuning.py``).
"""
from .blas import (  # noqa: F401
    AbsoluteAddOperator,
    AddOperator,
    MaxOperator,
    MeanOperator,
    MinOperator,
    ProductOperator,
    SB_Handle,
    _asum,
    _axpy,
    _copy,
    _dot,
    _gemm,
    _gemm_batched,
    _gemm_bf16,
    _gemm_ex,
    _gemm_f16,
    _gemm_strided_batched,
    _gemv,
    _iamax,
    _iamin,
    _nrm2,
    _scal,
    _sdsdot,
    _swap,
    _symm,
    _trsm,
    check,
    extension,
    gemm_batch_type_t,
    gemm_precision_t,
    reduction_dim_t,
)
from ._lib import LIB_PATH, declared_symbols, load  # noqa: F401
