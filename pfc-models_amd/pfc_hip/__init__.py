"""pfc_hip: MI355X-native (gfx950) pseudo-spectral ETD step for the standard
and log phase-field-crystal models, behind the reference catalogue's model API.

Importing the package does not touch the GPU; the device path loads
libpfc_hip.so on first use and raises if it is absent (no CPU fallback).
"""
__all__ = ["fields", "discretization", "plan"]
