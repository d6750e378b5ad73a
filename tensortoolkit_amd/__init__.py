"""MI355X (gfx950) block-sparse contraction: the hot path of
QuantumLiquids/TensorToolkit's ``qlten::Contract``, rebuilt on HIP.

``capi`` binds the C ABI of ``libqlten_amd.so`` (include/qlten_amd.h);
``blocksparse`` holds block-sparse tensor metadata and the contraction plan;
``workloads`` defines the synthetic benchmark tensors.
"""
from . import capi  # noqa: F401

__all__ = ["capi", "blocksparse", "workloads"]
