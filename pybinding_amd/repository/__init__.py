"""Ready-made lattices and modifiers (pybinding/repository)."""
from . import graphene, group6_tmd  # noqa: F401
