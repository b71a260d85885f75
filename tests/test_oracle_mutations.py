"""Each oracle pinning check catches a planted one-line mistake (CPU).

For every entry below, one line of an oracle function is rewritten (a dropped
term, a swapped operand, a wrong sign, index or bound), the mutated function
replaces the original for the duration of the test, and the named check must
then fail.  The mutations live only in this process's copy of the oracle.
"""

import importlib
import importlib.util
import inspect
import os
import textwrap

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _check(module_file, name):
    spec = importlib.util.spec_from_file_location("_pin_" + module_file, os.path.join(HERE, module_file + ".py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return getattr(mod, name)


def _mutate(monkeypatch, module, qualname, old, new):
    mod = importlib.import_module(module)
    owner = mod
    parts = qualname.split(".")
    for p in parts[:-1]:
        owner = getattr(owner, p)
    fn = getattr(owner, parts[-1])
    src = textwrap.dedent(inspect.getsource(fn))
    assert src.count(old) == 1, "mutation site must be unique: %r in %s" % (old, qualname)
    ns = {}
    exec(compile(src.replace(old, new), "<mutant %s>" % qualname, "exec"), mod.__dict__, ns)
    monkeypatch.setattr(owner, parts[-1], ns[parts[-1]])


LOOPS = "test_oracle_pin_loops"
RAYS = "test_oracle_pin_raytrace"
D1 = "test_oracle_d1"
OPT = "test_oracle_optim"
IMG = "test_oracle_imageseq"

MUTATIONS = [
    # --- OSMO (OSMO.py:36-162) ---
    ("oracle.optim", "osmo", "b = b / np.amax(b)", "b = b", LOOPS, "check_osmo_scale_invariance"),
    ("oracle.optim", "osmo", "xu = xu / np.amax(xu)", "xu = xu", LOOPS, "check_osmo_scale_invariance"),
    ("oracle.optim", "osmo", "vd[vd < 0] = 0", "pass", LOOPS, "check_osmo_fixed_point"),
    ("oracle.optim", "osmo", "gd[gd < 0] = 0", "pass", LOOPS, "check_osmo_fixed_point"),
    ("oracle.optim", "osmo", "vd = x[void] - d_l", "vd = x[void] - d_h", LOOPS, "check_osmo_step_order_and_band"),
    ("oracle.optim", "osmo", "gd = d_h - x[gel]", "gd = d_l - x[gel]", LOOPS, "check_osmo_step_order_and_band"),
    ("oracle.optim", "osmo", "xm[void] = xm[void] - vd", "xm[void] = xm[void] + vd", LOOPS,
     "check_osmo_step_order_and_band"),
    ("oracle.optim", "osmo", "for it in range(1, n_iter):", "for it in range(0, n_iter):", LOOPS,
     "check_osmo_iteration_count"),
    # --- CAL (CAL.py:116-218) ---
    ("oracle.optim", "cal", "        x = x / np.amax(x)\n", "        x = x\n", LOOPS, "check_cal_scale_invariance"),
    ("oracle.optim", "cal", "grad = db * (1 - momentum)", "grad = db", LOOPS, "check_cal_momentum_first_step"),
    ("oracle.optim", "cal", "b = np.clip(b - lr * grad, 0, None)", "b = np.clip(b + lr * grad, 0, None)", LOOPS,
     "check_cal_iteration_count_and_direction"),
    ("oracle.optim", "cal", "for it in range(1, n_iter):", "for it in range(0, n_iter):", LOOPS,
     "check_cal_iteration_count_and_direction"),
    ("oracle.optim", "cal", "b_opt = b_opt / np.amax(b_opt)", "b_opt = b_opt", LOOPS, "check_cal_lr_zero"),
    # --- BCLP (BCLP.py:244-386, 477-526) ---
    ("oracle.optim", "bclp", "G = (q * loss ** ((q - p) / q)) * G", "G = (loss ** ((q - p) / q)) * G", LOOPS,
     "check_bclp_closed_form_p2_q2"),
    ("oracle.optim", "bclp", "G = (q * loss ** ((q - p) / q)) * G", "G = (q * loss ** ((q - p) / p)) * G", LOOPS,
     "check_bclp_closed_form_p2_q1"),
    ("oracle.optim", "bclp", "g = constrain(g - lr * G)", "g = constrain(g + lr * G)", LOOPS,
     "check_bclp_closed_form_p2_q2"),
    ("oracle.optim", "bclp_state", "** (q / p)", "** (1 / p)", LOOPS, "check_bclp_closed_form_p2_q2"),
    ("oracle.optim", "bclp_state", "* np.sign(eT)", "", LOOPS, "check_bclp_closed_form_p2_q1"),
    ("oracle.optim", "bclp_state", "eb = np.abs(eT) - eps", "eb = eT - eps", LOOPS, "check_bclp_band_and_bounds"),
    ("oracle.optim", "bclp_init", "* (np.pi / (2 * n_angles))", "* (np.pi / n_angles)", LOOPS,
     "check_bclp_init_fbp_scale"),
    # --- metrics, FBP, PM ---
    ("oracle.optim", "ver", "x[void] >= np.min(x[gel])", "x[void] > np.min(x[gel])", LOOPS,
     "check_ver_known_answer"),
    ("oracle.optim", "circle_inds", "cx ** 2 + cy ** 2 <= 1", "cx ** 2 + cy ** 2 < 1", LOOPS, "check_circle_inds"),
    ("oracle.optim", "fbp", "b = np.clip(b, 0, None)", "pass", LOOPS, "check_fbp_reconstructs"),
    ("oracle.optim", "pm_loss_grad", "dose = bwd(b) * np.pi / 180", "dose = bwd(b)", LOOPS, "check_pm_known_answer"),
    # --- media (medium.py) ---
    ("oracle.raytrace", "GridMedium.sample", "idx = np.clip(idx, 0.0, shp - 1.0)", "pass", RAYS,
     "check_grid_sample_parity"),
    ("oracle.raytrace", "GridMedium._index", "* (sizes - 1.0)", "* sizes", RAYS, "check_grid_sample_parity"),
    ("oracle.raytrace", "central_difference", "(2 * voxel[0])", "voxel[0]", RAYS, "check_central_difference"),
    ("oracle.raytrace", "central_difference", "(n[:, :, 2:] - n[:, :, :-2])", "(n[:, :, 2:] - n[:, :, 1:-1])", RAYS,
     "check_central_difference"),
    ("oracle.raytrace", "luneburg_n", "fac = -c *", "fac = c *", RAYS, "check_luneburg_profile"),
    ("oracle.raytrace", "LuneburgMedium.sample", "(n[inside] ** (1 - self.p))", "(n[inside] ** (self.p - 1))", RAYS,
     "check_luneburg_analytic"),
    ("oracle.raytrace", "LuneburgMedium.sample", "** (1 / self.p)", "** (1 / 2)", RAYS, "check_luneburg_analytic"),
    ("oracle.raytrace", "att_alpha", "np.linalg.norm(x, axis=1) < radius", "np.linalg.norm(x[:, :2], axis=1) < radius",
     RAYS, "check_coord_vec_and_attenuation_models"),
    # --- ray geometry and stepping (pyTorchRayTrace.py) ---
    ("oracle.raytrace", "setup_rays", "D = np.sqrt(n0max ** 2 + zv.max() ** 2)", "D = n0max", RAYS,
     "check_setup_rays_frame"),
    ("oracle.raytrace", "setup_rays", "- G2 * np.cos(G1) * np.sin(phi)", "+ G2 * np.cos(G1) * np.sin(phi)", RAYS,
     "check_setup_rays_frame"),
    ("oracle.raytrace", "Tracer._rhs", "return v * n ** (-2), gn / n", "return v / n, gn / n", RAYS,
     "check_parametrisations_same_geometry"),
    ("oracle.raytrace", "Tracer._rhs", "return v, n * gn", "return v, gn", RAYS,
     "check_parametrisations_same_geometry"),
    ("oracle.raytrace", "Tracer._dx", "return v * n ** (-2)", "return v / n", RAYS,
     "check_parametrisations_same_geometry"),
    ("oracle.raytrace", "Tracer._dv", "return gn / n, n", "return gn, n", RAYS,
     "check_parametrisations_same_geometry"),
    ("oracle.raytrace", "Tracer.run", "A[ia] += self._alpha_att(xo) * ds", "A[ia] += self._alpha_att(xn) * ds",
     RAYS, "check_straight_ray_attenuation_closed_form"),
    ("oracle.raytrace", "Tracer.run", "E = E0[ia] * np.exp(-A[ia])", "E = np.exp(-A[ia])", RAYS,
     "check_straight_ray_attenuation_closed_form"),
    ("oracle.raytrace", "Tracer.run", "hw = self.h / m.voxel[0]", "hw = self.h / m.voxel[0] / 2", RAYS,
     "check_straight_ray_attenuation_closed_form"),
    ("oracle.raytrace", "Tracer._taps", "idx = (x - (-m.span / 2)) * (1.0 / m.voxel)",
     "idx = (x - (-m.span / 2)) * (1.0 / m.voxel) + 0.5", RAYS, "check_straight_ray_attenuation_closed_form"),
    ("oracle.raytrace", "Tracer.run", "if step % self.exit_every == 0:",
     "if step % self.exit_every == self.exit_every - 1:", RAYS, "check_exit_step_count"),
    ("oracle.raytrace", "Tracer._exited", "pos = (x > m.vmax + m.voxel) & (v > 0.0)", "pos = (x > m.vmax) & (v > 0.0)",
     RAYS, "check_exit_step_count"),
    # --- D1 projector (Projector3DParallel.py:113-207) ---
    ("oracle.d1", "sample_coords", "col = M[0, 0] * kcol + M[0, 1] * rrow", "col = M[0, 0] * kcol - M[0, 1] * rrow",
     D1, "test_golden_skimage_and_interp"),
    ("oracle.d1", "warp_matrix", "-center * (cos_a - sin_a - 1)", "-center * (cos_a + sin_a - 1)", D1,
     "test_golden_skimage_and_interp"),
    ("oracle.d1", "clip_warp_output", "o[zero] = 0.0", "pass", D1, "test_golden_skimage_and_interp"),
    ("oracle.d1", "_interp_taps", "valid = (t >= xp0) & (t <= xp_last)", "valid = (t >= xp0) & (t < xp_last)", D1,
     "test_golden_skimage_and_interp"),
    ("oracle.d1", "backward", "t = x * cos_a - y * sin_a", "t = x * cos_a + y * sin_a", D1,
     "test_golden_skimage_and_interp"),
    ("oracle.d1", "outside_circle_mask", "circle_y ** 2 > 1", "circle_y ** 2 >= 1", D1, "test_clip_to_circle_counts"),
    ("oracle.d1", "absorption_weight", "half - s", "half + s", D1, "test_absorption_closed_form_uniform_cylinder"),
    # --- filters and response (util/data.py, response.py) ---
    ("oracle.optim", "fourier_filter", "f[0] = 0.25", "f[0] = 0.5", OPT, "test_ramlak_filter_closed_form"),
    ("oracle.optim", "filter_sinogram", "img[:pX] = s[:, :, z]", "img[P - pX:] = s[:, :, z]", OPT,
     "test_filter_sinogram_matches_direct_convolution"),
    ("oracle.optim", "filter_target_osmo", "X = fftshift(fft2(x, axes=(0, 1)), axes=(0, 1))",
     "X = fft2(x, axes=(0, 1))", OPT, "test_filter_target_osmo_single_mode"),
    ("oracle.optim", "filter_target_osmo", "f = np.sqrt(fx ** 2 + fy ** 2)", "f = fx ** 2 + fy ** 2", OPT,
     "test_filter_target_osmo_single_mode"),
    ("oracle.optim", "Response._dmap", "c2 = self.B / self.nu", "c2 = self.B", OPT, "test_response_closed_forms"),
    # --- zero-dose rays (Projector3DParallelCUDA.py:57-58) ---
    ("oracle.optim", "osmo", "        b = zero_rays(b, zero_sino)\n", "", LOOPS, "check_zero_dose_rays"),
    ("oracle.optim", "osmo", "b = zero_rays(np.clip(fwd(xm), 0, None), zero_sino)", "b = np.clip(fwd(xm), 0, None)",
     LOOPS, "check_zero_dose_rays"),
    ("oracle.optim", "cal", "        b = zero_rays(b, zero_sino)\n", "", LOOPS, "check_zero_dose_rays"),
    ("oracle.optim", "fbp", "b = zero_rays(b, zero_sino)", "pass", LOOPS, "check_zero_dose_rays"),
    ("oracle.optim", "bclp", "        g = zero_rays(g, zero_sino)\n        dose, m, loss, operand",
     "        dose, m, loss, operand", LOOPS, "check_zero_dose_rays"),
    ("oracle.optim", "bclp", "    g = zero_rays(g, zero_sino)\n    dose, m, loss, _", "    dose, m, loss, _", LOOPS,
     "check_zero_dose_rays"),
    ("oracle.optim", "zero_rays", "np.where(zero_sino, 0.0, b)", "np.where(zero_sino, b, 0.0)", LOOPS,
     "check_zero_dose_rays"),
    # --- image-sequence export (imagesequence.py:85-399) ---
    ("oracle.imageseq", "place", "extra = int(array_offset * d + v_offset)", "extra = int(array_offset * d)", IMG,
     "check_place_known_answer"),
    ("oracle.imageseq", "place", "d = int(np.ceil((k + 1) / 2) * (-1) ** k)", "d = int(np.ceil((k + 1) / 2))", IMG,
     "check_place_known_answer"),
    ("oracle.imageseq", "resize_linear", "f = (np.arange(n_dst) + 0.5) * (n_src / n_dst) - 0.5",
     "f = np.arange(n_dst) * (n_src / n_dst)", IMG, "check_resize_closed_forms"),
    ("oracle.imageseq", "frames", "m = to_unsigned(np.minimum(m, maxv), bit_depth)", "m = to_unsigned(m, bit_depth)",
     IMG, "check_frames_known_answer"),
    ("oracle.imageseq", "frames", "m = np.flip(m, axis=0)", "m = np.flip(m, axis=2)", IMG,
     "check_frames_known_answer"),
    ("oracle.imageseq", "frames", "img = m[:, j, :].T", "img = m[:, j, :]", IMG, "check_frames_known_answer"),
    # --- bit depth (util/data.py:495-523 and its call sites) ---
    ("oracle.optim", "discretize", "np.digitize(x, bins=levels) - 1", "np.digitize(x, bins=levels, right=True) - 1",
     LOOPS, "check_discretize_known_answer"),
    ("oracle.optim", "discretize", "2 ** bit_depth", "2 ** bit_depth - 1", LOOPS, "check_discretize_known_answer"),
    ("oracle.optim", "osmo", "b = discretize(b, bit_depth, np.min(b), np.max(b))",
     "b = discretize(b, bit_depth, 0.0, np.max(b))", LOOPS, "check_osmo_bit_depth"),
    ("oracle.optim", "osmo", "b = discretize(b, bit_depth, 0.0, np.max(b))", "b = b", LOOPS, "check_osmo_bit_depth"),
    ("oracle.optim", "cal", "b = discretize(b, bit_depth, 0.0, np.max(b))", "b = b", LOOPS,
     "check_cal_fbp_bclp_bit_depth"),
    ("oracle.optim", "fbp", "b = discretize(b, bit_depth, 0.0, np.max(b))", "b = b", LOOPS,
     "check_cal_fbp_bclp_bit_depth"),
    ("oracle.optim", "bclp", "g = discretize(g, bit_depth, 0.0, np.amax(g))", "g = g", LOOPS,
     "check_cal_fbp_bclp_bit_depth"),
]


@pytest.mark.parametrize("module,qualname,old,new,check_file,check_name", MUTATIONS,
                         ids=["%s:%s:%d" % (m[0].split(".")[-1], m[1], i) for i, m in enumerate(MUTATIONS)])
def test_mutation_is_caught(monkeypatch, module, qualname, old, new, check_file, check_name):
    check = _check(check_file, check_name)
    check()                       # passes on the unmutated oracle
    _mutate(monkeypatch, module, qualname, old, new)
    with pytest.raises(AssertionError):
        check()
