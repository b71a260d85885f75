"""Propagators of the light-propagation path (parallel-beam projector, curved-ray tracer)."""
