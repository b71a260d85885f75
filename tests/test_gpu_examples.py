"""The reference's example scripts (examples/*.py), at reduced sizes, through
this package's API unchanged: the calls a VAMToolbox user makes must work as
written (resources are replaced by generated images and arrays; STL
voxelisation is outside this package)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _png(tmp_path, n, seed=0):
    from PIL import Image
    yy, xx = np.mgrid[:n, :n] - (n - 1) / 2
    rng = np.random.default_rng(seed)
    img = np.zeros((n, n))
    for _ in range(4):
        cx, cy, r = rng.uniform(-0.2, 0.2) * n, rng.uniform(-0.2, 0.2) * n, rng.uniform(0.08, 0.2) * n
        img = np.maximum(img, np.clip(1.2 - np.hypot(xx - cx, yy - cy) / r, 0, 1))
    path = tmp_path / "target.png"
    Image.fromarray((255 * img).astype(np.uint8)).save(path)
    return str(path)


def test_2d_optimization_and_fbp_offset(gpu, tmp_path):
    """examples/2Doptimization.py (CAL, hamming filter) and FBPoffset.py
    (shepp-logan, with and without offset)."""
    import matplotlib
    matplotlib.use("Agg")
    import vamtoolbox_amd as vam
    target_geo = vam.geometry.TargetGeometry(imagefilename=_png(tmp_path, 96), pixels=64)
    target_geo.show()
    angles = np.linspace(0, 360 - 360 / 90, 90)
    proj_geo = vam.geometry.ProjectionGeometry(angles, ray_type="parallel", CUDA=True)
    opts = vam.optimize.Options(method="CAL", n_iter=10, d_h=0.85, d_l=0.6, filter="hamming", verbose=None)
    opt_sino, opt_recon, error = vam.optimize.optimize(target_geo, proj_geo, opts)
    assert opt_sino.array.shape == (64, 90, 1) and opt_recon.array.shape == (64, 64, 1)
    assert np.all(np.isfinite(error)) and error[-1] < error[0]
    opt_recon.show(savepath=str(tmp_path / "recon.png"))
    for offset in (False, True):
        opts = vam.optimize.Options(method="FBP", filter="shepp-logan", offset=offset, verbose=None)
        opt_sino, opt_recon, error = vam.optimize.optimize(target_geo, proj_geo, opts)
        assert opt_sino.array.min() >= 0 and np.isclose(opt_sino.array.max(), 1.0)


def test_3d_optimization_with_absorption(gpu):
    """examples/3Doptimizationabsorption.py: OSMO without and with the radial
    absorption mask (units "dose"; the reference ignores n_rot and rot_vel)."""
    import vamtoolbox_amd as vam
    n, nz = 48, 8
    yy, xx = np.mgrid[:n, :n] - (n - 1) / 2
    target = np.repeat(((xx / 14) ** 2 + (yy / 9) ** 2 <= 1)[:, :, None], nz, axis=2).astype(float)
    target_geo = vam.geometry.TargetGeometry(target=target)
    angles = np.linspace(0, 360 - 360 / 120, 120)
    sinos = []
    for kw, units in (({}, "normalized"),
                      (dict(projector_pixel_size=100e-4, absorption_coeff=1, container_radius=1, n_rot=3, rot_vel=24),
                       "dose")):
        proj_geo = vam.geometry.ProjectionGeometry(angles, ray_type="parallel", CUDA=True, **kw)
        opts = vam.optimize.Options(method="OSMO", n_iter=10, d_h=0.85, d_l=0.6, filter="hamming", units=units,
                                    verbose=None)
        opt_sino, opt_recon, error = vam.optimize.optimize(target_geo, proj_geo, opts)
        assert opt_sino.array.shape == (n, 120, nz)
        assert np.all(np.isfinite(error)) and error[-1] <= error[0]
        sinos.append(np.asarray(opt_sino.array))
    assert proj_geo.absorption_mask is not None
    # the mask reweights the volume: the optimised sinograms differ
    assert np.abs(sinos[0] - sinos[1]).max() > 1e-3


def test_bclp_basic_and_ray_trace(gpu, tmp_path):
    """examples/BCLPoptimizationbasic.py and BCLPoptimizationraytrace.py
    (output "full", logs.show()): parallel beam over 180 degrees and the ray
    tracer with a homogeneous interpolated index, cylinder attenuation,
    symplectic Euler and the physical-path-length parametrisation."""
    import matplotlib
    matplotlib.use("Agg")
    import vamtoolbox_amd as vam
    target_geo = vam.geometry.TargetGeometry(imagefilename=_png(tmp_path, 64, 1), pixels=48, binarize_image=False)
    angles = np.linspace(0, 180 - 180 / 60, 60)
    response_model = vam.response.ResponseModel(B=25)
    proj_geo = vam.geometry.ProjectionGeometry(angles, ray_type="parallel", CUDA=True)
    opts = vam.optimize.Options(method="BCLP", n_iter=8, filter="ram-lak", verbose=None, p=2, learning_rate=0.00001,
                                response_model=response_model, eps=0.1)
    opt_sino, opt_recon, opt_response, logs = vam.optimize.optimize(target_geo, proj_geo, opts, output="full")
    assert np.all(np.isfinite(logs.loss[:9])) and logs.loss[8] < logs.loss[0]
    logs.show()
    coord_vec = target_geo.constructCoordVec()
    index_model = vam.medium.IndexModel(coord_vec, type="interpolation", form="homogeneous")
    attenuation_model = vam.medium.AttenuationModel(coord_vec, type="analytical", form="homogeneous_cylinder", R=1,
                                                    alpha_internal=1e-3)
    proj_geo = vam.geometry.ProjectionGeometry(
        angles, ray_type="ray_trace", CUDA=True, index_model=index_model, attenuation_model=attenuation_model,
        absorption_model=attenuation_model, ray_trace_method="eikonal",
        eikonal_parametrization="physical_path_length", ray_trace_ode_solver="forward_symplectic_euler",
        ray_trace_ray_config="parallel", inclination_angle=0.0, ray_density=1)
    opts = vam.optimize.Options(method="BCLP", n_iter=6, filter="ram-lak", verbose=None, p=2, learning_rate=10,
                                response_model=response_model, eps=0.1)
    opt_sino, opt_recon, opt_response, logs = vam.optimize.optimize(target_geo, proj_geo, opts, output="full")
    assert np.all(np.isfinite(logs.loss[:7])) and np.all(np.isfinite(np.asarray(opt_recon.array)))


def test_insert_and_zero_dose(gpu):
    """examples/3Doverprint.py (an insert in the target, CUDA=False is the
    same projector here) and 3Dzerodoseconstraint.py (a zero-dose body)."""
    import vamtoolbox_amd as vam
    n, nz = 40, 4
    yy, xx = np.mgrid[:n, :n] - (n - 1) / 2
    ring = (np.hypot(xx, yy) < 12) & (np.hypot(xx, yy) > 6)
    core = np.hypot(xx, yy) < 4
    rep = lambda a: np.repeat(a[:, :, None], nz, axis=2).astype(float)   # noqa: E731
    angles = np.linspace(0, 360 - 360 / 90, 90)
    for kw, cuda in ((dict(insert=rep(core)), False), (dict(zero_dose=rep(core)), True)):
        target_geo = vam.geometry.TargetGeometry(target=rep(ring), **kw)
        proj_geo = vam.geometry.ProjectionGeometry(angles, ray_type="parallel", CUDA=cuda)
        opts = vam.optimize.Options(method="OSMO", n_iter=8, d_h=0.85, d_l=0.6, filter="hamming", verbose=None)
        opt_sino, opt_recon, error = vam.optimize.optimize(target_geo, proj_geo, opts)
        assert np.all(np.isfinite(error)) and opt_sino.array.shape == (n, 90, nz)
        if "zero_dose" in kw:
            assert np.all(np.asarray(opt_sino.array)[np.asarray(proj_geo.zero_dose_sino)] == 0)


def test_image_sequence_example(gpu, tmp_path):
    """examples/imagesequence.py: two arrayed copies, then a 45-degree
    rotation with 2x scaling, on an optimised sinogram; saveAsImages."""
    import vamtoolbox_amd as vam
    n, nz = 64, 16
    yy, xx = np.mgrid[:n, :n] - (n - 1) / 2
    target_geo = vam.geometry.TargetGeometry(target=np.repeat(((xx / 20) ** 2 + (yy / 12) ** 2 <= 1)[:, :, None], nz,
                                                              axis=2).astype(float))
    angles = np.linspace(0, 360 - 360 / 36, 36)
    proj_geo = vam.geometry.ProjectionGeometry(angles, ray_type="parallel", CUDA=True)
    sino, _, _ = vam.optimize.optimize(target_geo, proj_geo, vam.optimize.Options(method="OSMO", n_iter=4))
    iconfig0 = vam.imagesequence.ImageConfig(image_dims=(320, 180), array_num=2, array_offset=40,
                                             normalization_percentile=99)
    iconfig1 = vam.imagesequence.ImageConfig(image_dims=(320, 180), rotate_angle=45, size_scale=2,
                                             normalization_percentile=99)
    for cfg in (iconfig0, iconfig1):
        seq = vam.imagesequence.ImageSeq(image_config=cfg, sinogram=sino)
        assert len(seq.images) == 36 and seq.images[0].shape == (180, 320) and seq.images[0].dtype == np.uint8
        assert seq.images[0].max() > 0
    seq.saveAsImages(str(tmp_path), image_prefix="frame")
    assert len(list(tmp_path.glob("frame*.png"))) == 36
    seq.save(str(tmp_path / "seq"))
    back = vam.imagesequence.loadImageSeq(str(tmp_path / "seq.imgseq"))
    assert all(np.array_equal(a, b) for a, b in zip(back.images, seq.images))
