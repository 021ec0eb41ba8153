"""MI355X-native bucket-elimination engine for QAOA lightcone energies.

The package is a drop-in bucket backend for QTensor's ``ProcessingFrameworks``
interface (``process_bucket`` / ``get_sliced_buckets`` / ``get_result_data``)
plus the host-side callers that drive it (composers, greedy ordering,
simulators) and a batched, plan-compiled executor for whole graphs.
"""
from .model import Var, Tensor
from .composer import (QtreeQAOAComposer, ZZQtreeQAOAComposer, SimpZZQtreeComposer,
                       DefaultQAOAComposer, get_edge_subgraph)
from .ordering import OrderingOptimizer, SlicesOptimizer, TamakiOptimizer, TreeTrimSplitter, DefaultOptimizer
from .tensornet import QtreeTensorNet
from .simulator import QtreeSimulator, QAOAQtreeSimulator, FeynmanSimulator

__all__ = [
    "Var", "Tensor", "QtreeQAOAComposer", "ZZQtreeQAOAComposer", "SimpZZQtreeComposer",
    "DefaultQAOAComposer", "get_edge_subgraph", "OrderingOptimizer", "SlicesOptimizer", "TreeTrimSplitter",
    "TamakiOptimizer", "DefaultOptimizer", "QtreeTensorNet", "QtreeSimulator",
    "QAOAQtreeSimulator", "FeynmanSimulator",
]
