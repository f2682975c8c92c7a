"""Drivers of Monte Carlo runs (src/isle/drivers/__init__.py): ``hmc``."""
from . import hmc

__all__ = ["hmc"]
