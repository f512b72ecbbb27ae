"""Collision draw stream of the drop-in entry points.

The reference draws its collision uniforms with np.random.rand inside the
jitted collision steps (collision/all_or_nothing.py:134, 217, 314, 429, 545,
677) and seeds with np.random.seed (integration.py:2055,
collision/box_model.py:150).  Here every draw is a counter-based
Philox-4x32-10 value keyed by (seed, run, collision step, cell, pair), so the
GPU can produce any draw in any order; the consumption structure (one uniform
per unordered pair i < j, row-major pair order, cells in x-major order, none
for cells with fewer than two SIPs) is the reference's.

This module holds the process-wide key and the collision-step counter that
the drop-in functions use: ``seed(s)`` takes np.random.seed's place and
restarts the counter; every collision step a drop-in function runs consumes
one counter value.  An Engine created directly carries its own key and
counter instead.
"""

_state = {"seed": 0, "step": 0}


def seed(s, step=0):
    """Set the key of the collision draws and restart (or position) the step counter."""
    _state["seed"] = int(s) & 0xFFFFFFFFFFFFFFFF
    _state["step"] = int(step) & 0xFFFFFFFF


def get_seed():
    return _state["seed"]


def get_step():
    return _state["step"]


def set_step(step):
    _state["step"] = int(step) & 0xFFFFFFFF


def take_steps(n=1):
    """Reserve n consecutive collision-step counter values; returns the first."""
    s = _state["step"]
    _state["step"] = (s + int(n)) & 0xFFFFFFFF
    return s
