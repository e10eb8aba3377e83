"""Cell states and types (cells/model.py:14-30 of the reference): the integer codes the
kernels and the deposit layers use."""

from enum import IntEnum


class CellState(IntEnum):
    PROLIFERATIVE = 0
    QUIESCENT = 1
    HYPOXIC = 2
    DEAD = 3


class CellType(IntEnum):
    TUMOR = 0


NUM_STATES = len(CellState)
