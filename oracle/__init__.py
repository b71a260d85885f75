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
    # --- D1 projector (d1.py; Projector3DParallel.py:57-216, util/data.py:14-40) ---
    "d1.outside_circle_mask / clip_to_circle / inside_circle_mask":
        "tests/test_oracle_d1.py::test_clip_to_circle_counts (symmetry, complement count, points on the "
        "circle kept at odd n)",
    "d1.warp_matrix / sample_coords / bilinear_zero_pad / clip_warp_output / forward":
        "scikit-image 0.18.3 warp golden vectors (test_golden_skimage_and_interp, test_warp_clip_matters); "
        "scalar brute-force loop (test_forward_brute_force_scalar_loop); 90-degree rotation and integer-offset "
        "invariants; mass conservation",
    "d1._interp_taps / backward": "numpy.interp golden vectors (test_golden_skimage_and_interp); "
                                  "constant-sinogram closed form (test_constant_sinogram_backprojection)",
    "d1.forward_adjoint / backward_adjoint": "brute-force matrix transposes (test_adjoints_are_transposes)",
    "d1.absorption_weight / default_radius": "uniform-cylinder closed form and entry-side orientation "
                                             "(test_absorption_closed_form_uniform_cylinder); mu = 0 reduction",
    "d1._z_taps (inclination)": "phi = 0 reduction (test_mu_zero_and_phi_zero_reduce_exactly); brute-force "
                                "transposes with inclination (test_adjoints_are_transposes)",
    "d1.occlusion_sinogram / forward(occ=) / backward(occ=)":
        "scikit-image / numpy.interp golden vectors in the reference's loop structure "
        "(test_occlusion_sinogram_matches_golden, test_occluded_forward_backward_match_golden); first-hit brute "
        "force; neutral without an occluder",
    # --- filters and response (optim.py; util/data.py:43-113, 302-438; response.py:12-291) ---
    "optim.fourier_filter": "ideal-ramp limit and symmetry (test_ramlak_filter_closed_form)",
    "optim.filter_sinogram": "direct zero-padded linear convolution with the spatial Ram-Lak kernel "
                             "(test_filter_sinogram_matches_direct_convolution)",
    "optim.filter_target_osmo": "closed form on single real Fourier modes (test_filter_target_osmo_single_mode)",
    "optim.Response": "closed form, derivative vs central differences for nu in {1, 0.5, 2.5}, "
                      "map_inv(map(f)) = f (test_response_closed_forms)",
    # --- optimisers (optim.py; OSMO.py, CAL.py, BCLP.py, FBP.py, PM.py, metrics.py) ---
    "optim.osmo": "scale invariance, fixed point, void-then-gel step order in (d_l, d_h), iteration count "
                  "(tests/test_oracle_pin_loops.py)",
    "optim.cal": "lr = 0 keeps the initial iterate, scale invariance, update direction and iteration count, "
                 "first momentum step (tests/test_oracle_pin_loops.py)",
    "optim.cal_threshold": "separable classes and brute-force counting (test_cal_threshold_separated_classes)",
    "optim.bclp_state / bclp": "identity-operator closed forms for p = q = 2 and p = 2, q = 1, band and bounds "
                               "(tests/test_oracle_pin_loops.py); gradient vs central differences "
                               "(test_bclp_gradient_finite_differences_oracle)",
    "optim.bclp_init": "filtered back projection recovers map_inv(T) (check_bclp_init_fbp_scale)",
    "optim.ver / circle_inds": "known answer and circle geometry (check_ver_known_answer, check_circle_inds)",
    "optim.fbp": "disk reconstruction and the offset variant (check_fbp_reconstructs)",
    # --- image-sequence export (imageseq.py; imagesequence.py:85-399) ---
    "imageseq.place": "hand-computed windows for one, three and two arrayed copies and the out-of-screen "
        "errors (check_place_known_answer)",
    "imageseq.rotate": "is scipy.ndimage.rotate, the reference's own call; quarter turn = index permutation "
        "(check_rotate_is_scipy)",
    "imageseq.resize_linear": "closed forms: same size = copy, constants, 2x ramp upscale at half-pixel centres "
        "with edge clamps, 2x downscale = pair means (check_resize_closed_forms); OpenCV itself is absent, so "
        "its exact float rounding is unpinned",
    "imageseq.frames / to_unsigned / percentile": "hand-computed bytes for truncation, clipping, 10-bit "
        "wrap modulo 256, 100th-percentile normalisation, flips (check_frames_known_answer); percentile is "
        "np.percentile",
    "optim.zero_rays and the zero_sino branches of osmo / cal / fbp / bclp": "every back-projected and every "
        "returned sinogram is zero on the masked rays; an empty mask changes nothing (check_zero_dose_rays)",
    "optim.discretize": "hand-computed levels incl. the index -1 wrap, brute force over 64 levels "
        "(check_discretize_known_answer)",
    "optim.osmo / cal / fbp / bclp with bit_depth": "one-bit two-level projections inside OSMO, returned "
        "sinograms on the [0, max] levels, quantisation never raises a value, 20 bits reproduce the "
        "unquantised loops (check_osmo_bit_depth, check_cal_fbp_bclp_bit_depth)",
    "optim.pm_loss_grad / pm": "hand-computed known answer (check_pm_known_answer); gradient vs central "
                               "differences (test_pm_gradient_finite_differences)",
    # --- ray tracer (raytrace.py; pyTorchRayTrace.py, medium.py) ---
    "raytrace.coord_vec": "spacing 1 / rate, centred, 2-D (check_coord_vec_and_attenuation_models)",
    "raytrace.GridMedium.sample / _index": "torch.nn.functional.grid_sample called as the reference calls it, "
                                          "border and zeros padding, 3-D and 2-D (check_grid_sample_parity)",
    "raytrace.central_difference": "exact on quadratic fields, zero faces, 2-D (check_central_difference)",
    "raytrace.luneburg_n": "analytic profile and radial gradient (check_luneburg_profile); focus on the "
                           "antipodal surface point (test_luneburg_lens_focus)",
    "raytrace.LuneburgMedium.sample": "closed-form lens: gradient vs central differences for p = 2 and 3, "
                                      "centre and rim values, grid values, focus on the antipodal surface point "
                                      "(check_luneburg_analytic)",
    "raytrace.att_alpha": "cylinder / ball point tests; grid = grid_sample with zeros padding "
                          "(check_coord_vec_and_attenuation_models, check_grid_sample_parity)",
    "raytrace.setup_rays": "orthonormal ray frame, distance D, lateral coordinates, energies "
                           "(check_setup_rays_frame)",
    "raytrace.Tracer._rhs / _dv / _dx / advance": "canonical, physical and optical-path parametrisations trace "
                                                  "the same geometric ray with RK4 and symplectic Euler "
                                                  "(check_parametrisations_same_geometry); RK4 and Euler orders "
                                                  "(test_rk4_fourth_order_and_euler_first_order)",
    "raytrace.Tracer.run / _taps / _alpha_att": "straight-ray closed form of integrate and deposit with "
                                                "cylinder attenuation at the step starts, h / dx weights and "
                                                "trilinear hats (check_straight_ray_attenuation_closed_form); "
                                                "W / W^T transposition (test_deposit_is_transpose_of_integrate)",
    "raytrace.Tracer._exited / exit schedule": "closed-form exit step counts, every 10 and every 3 steps, and "
                                               "the step cap (check_exit_step_count)",
    "raytrace_torch.TorchCPUTracer (CPU baseline only)": "equals the float64 oracle on rays off the voxel "
                                                         "centres, integrate and deposit "
                                                         "(test_torch_cpu_port_matches_oracle)",
    # --- fan-beam rebinning (rebin.py; geometry.py:736-944) ---
    "rebin.sampling / rebin_fan_beam": "telecentric limit, Bouguer's invariant, Fresnel factor at normal "
                                       "incidence (test_fan_beam_rebin_telecentric_limit, "
                                       "test_fan_beam_snell_and_centre_ray)",
}
"""Every entry's check fails under a planted one-line mistake in the oracle:
``tests/test_oracle_mutations.py`` plants 76 such mutations and asserts that
the named check catches each one."""
