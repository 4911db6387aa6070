"""tridsolver_amd: batched scalar tridiagonal solves on AMD Instinct MI355X.

The product is the C ABI library ``lib/libtridhip.so`` (headers in
``include/``), a drop-in for the reference's ``trid{S,D}mtsvStridedBatch[Inc]``.
This package builds it (``tridsolver_amd.build``), binds it with ctypes
(``tridsolver_amd._lib``) and mirrors the C API on torch tensors
(``tridsolver_amd.api``).
"""
from ._lib import TridError, TridParams, lib  # noqa: F401

__all__ = ["TridError", "TridParams", "lib"]
