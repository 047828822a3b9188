"""htool_amd: MI355X-native H-matrix product and partial-pivot ACA.

The hot path (kernel-entry generation, batched ACA, leaf products) runs in
libhtool_amd.so, built from htool_amd/csrc for gfx950. The cluster and block
trees are host C++ inside the same library.
"""
from . import geometry
from .capi import HtgError
from .hmatrix import Generator, HMatrix
from .lrmat import Block, aca_batch
from .predictor import RankPredictor

__all__ = ["geometry", "Generator", "HMatrix", "HtgError", "Block", "aca_batch", "RankPredictor"]
