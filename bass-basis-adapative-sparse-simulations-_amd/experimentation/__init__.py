"""Experiment harness (src/experimentation of the reference) on the device drivers."""
