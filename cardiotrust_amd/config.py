"""Hot-path fields of the reference's `Algorithm` config.

Mirrors `config/algorithm.rs:25-58` (field names and meaning) with the defaults
of `config/algorithm.rs:59-85`. Only the fields that the epoch reads are kept;
GUI/scheduler fields (epochs, snapshots, learning-rate schedule) belong to the
caller.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from enum import IntEnum


class Optimizer(IntEnum):
    """`refinement.rs:7-12`."""

    Sgd = 0
    Adam = 1


class APDerivative(IntEnum):
    """`config/algorithm.rs:16-21`."""

    Textbook = 0
    Simple = 1


class CAlgorithm(ctypes.Structure):
    """C layout of `ct_algorithm` in include/cardiotrust_hip.h."""

    _fields_ = [
        ("learning_rate", ctypes.c_float),
        ("batch_size", ctypes.c_int32),
        ("optimizer", ctypes.c_int32),
        ("mse_strength", ctypes.c_float),
        ("maximum_regularization_strength", ctypes.c_float),
        ("maximum_regularization_threshold", ctypes.c_float),
        ("difference_regularization_strength", ctypes.c_float),
        ("smoothness_regularization_strength", ctypes.c_float),
        ("freeze_gains", ctypes.c_int32),
        ("freeze_delays", ctypes.c_int32),
        ("ap_derivative", ctypes.c_int32),
    ]


@dataclass
class Algorithm:
    """Epoch configuration (`config/algorithm.rs:25-85`)."""

    learning_rate: float = 200.0
    batch_size: int = 0
    optimizer: Optimizer = Optimizer.Sgd
    mse_strength: float = 1.0
    maximum_regularization_strength: float = 1.0
    maximum_regularization_threshold: float = 1.01
    difference_regularization_strength: float = 0.0
    smoothness_regularization_strength: float = 0.0
    freeze_gains: bool = False
    freeze_delays: bool = True
    ap_derivative: APDerivative = APDerivative.Textbook

    def to_c(self) -> CAlgorithm:
        return CAlgorithm(
            float(self.learning_rate),
            int(self.batch_size),
            int(self.optimizer),
            float(self.mse_strength),
            float(self.maximum_regularization_strength),
            float(self.maximum_regularization_threshold),
            float(self.difference_regularization_strength),
            float(self.smoothness_regularization_strength),
            int(bool(self.freeze_gains)),
            int(bool(self.freeze_delays)),
            int(self.ap_derivative),
        )
