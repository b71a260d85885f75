"""CPU oracle for the VAMToolbox light-propagation hot path.

TEST INFRASTRUCTURE ONLY.  Nothing under ``oracle/`` is part of the product:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker (or the timed CPU port).
The product path (``vamtoolbox_amd``) never imports this package and fails
loudly when its HIP library is missing.

Every function restates, in float64 NumPy, the algorithm of the reference file
and line range cited in its docstring (paths relative to the VAMToolbox
repository root).  How each function is pinned (golden vectors from the
reference's own third-party routines, brute-force matrices, closed forms,
invariants) is listed in ``PINNING`` below; a function listed there as
"parity unpinned" has no independent check.
"""

PINNING = {
    "d1.clip_to_circle": "exact restatement; invariant tests (mask symmetry, count)",
    "d1.forward": "scikit-image 0.18.3 warp golden vectors; brute-force matrix; "
                  "90-degree rotation and integer-offset invariants; mass conservation",
    "d1.backward": "numpy.interp golden vectors; brute-force matrix; constant-sinogram closed form",
    "d1.forward_adjoint": "dot-product test against d1.forward; brute-force matrix transpose",
    "d1.backward_adjoint": "dot-product test against d1.backward; brute-force matrix transpose",
    "d1 absorption": "uniform-cylinder closed form (astra3Dabsorption.getExpabs); mu=0 reduction",
    "d1 inclination": "phi=0 reduction; z-linear invariance",
    "filters.filter_sinogram": "closed form of the Ram-Lak kernel DFT; scipy.fft reference",
    "filters.filter_target_osmo": "direct 2-D DFT on tiny arrays",
    "response": "closed form, analytic derivative vs finite differences, map_inv(map(f)) = f",
    "optimizers": "tiny-problem iterates recomputed with the explicit projector matrix",
    "raytrace": "straight-ray limit equals line integral; W/W^T dot-product test; "
                "step-halving order test for RK4",
}
