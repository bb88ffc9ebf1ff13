"""Shared helpers of the tests."""
from leela_amd.positions import synthetic_planes  # noqa: F401
