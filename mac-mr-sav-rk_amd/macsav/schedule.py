"""Step-schedule arithmetic (solver/schedule.py:7-29)."""
from __future__ import annotations

import numpy as np

from .mac.grid import Array


def schedule_nodes(initial_time: float, sequence: Array) -> Array:
    """Node times t_0 + cumulative steps with Kahan-compensated accumulation."""
    nodes = np.empty(sequence.size + 1, dtype=float)
    nodes[0] = initial_time
    total, carry = float(initial_time), 0.0
    for k, step in enumerate(sequence, start=1):
        y = float(step) - carry
        t = total + y
        carry = (t - total) - y
        total = t
        nodes[k] = total
    return nodes


def nearest_node_indices(nodes: Array, requests: list[float]) -> list[int]:
    """Index of the node nearest each request; ties go to the earlier node."""
    out = []
    scale = float(np.max(np.abs(nodes)))
    for t in requests:
        gap = np.abs(nodes - t)
        slack = 8 * np.finfo(float).eps * max(abs(t), scale, np.finfo(float).tiny)
        out.append(int(np.flatnonzero(gap <= gap.min() + slack)[0]))
    return out
