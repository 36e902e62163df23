"""Offline problem generator writing the reference's problem file set."""
from . import btio, tree, build, problem, configs  # noqa: F401
