"""Iterative solvers (``warp.optim`` drop-in for the CG path)."""

from . import linear
from . import refine
