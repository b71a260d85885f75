"""Algebraic (sparse-matrix) propagation: the HIP CSR kernel against SciPy's
float64 sparse products, the propagator's z-tiling, the ray tracer's recorded
matrix against its own integrate / deposit, and the impulse-response matrix
builder against the exact transpose of the back projection."""

import numpy as np
import pytest
import scipy.sparse

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("z", [1, 3, 64, 100, 256])
def test_csr_spmm_matches_scipy(gpu, z):
    import torch
    from vamtoolbox_amd import _lib
    rng = np.random.default_rng(30 + z)
    m = scipy.sparse.random(700, 500, density=0.02, random_state=rng, format="csr", dtype=np.float64)
    m.data = rng.normal(size=m.data.size)
    x = rng.normal(size=(500, z))
    lib = _lib.load()
    rp = torch.as_tensor(m.indptr.astype(np.int64), device=gpu)
    ci = torch.as_tensor(m.indices.astype(np.int32), device=gpu)
    va = torch.as_tensor(m.data.astype(np.float32), device=gpu)
    xt = torch.as_tensor(x.astype(np.float32).reshape(-1), device=gpu)
    y = torch.empty(700 * z, device=gpu)
    _lib.check(lib.vamx_csr_spmm(rp.data_ptr(), ci.data_ptr(), va.data_ptr(), 700, 500, z, xt.data_ptr(),
                                 y.data_ptr(), _lib.stream_ptr(gpu)), "vamx_csr_spmm")
    e = rel_l2(y.cpu().numpy().reshape(700, z), m @ x)
    print("z %d rel L2 %.2e" % (z, e))
    assert e <= 1e-6


def test_algebraic_propagator_z_tiling(gpu, tmp_path):
    import vamtoolbox_amd as vam
    rng = np.random.default_rng(3)
    n, nz, na = 10, 6, 7
    m = scipy.sparse.random(n * na, n * n, density=0.1, random_state=rng, format="csr", dtype=np.float64)
    path = tmp_path / "P.npz"
    scipy.sparse.save_npz(path, m)
    tg = vam.geometry.TargetGeometry(target=np.ones((n, n, nz)), clip_to_circle=False)
    pg = vam.geometry.ProjectionGeometry(np.arange(na), "algebraic", CUDA=True, loading_path_for_matrix=str(path))
    A = vam.projectorconstructor.projectorconstructor(tg, pg)
    assert A.z_tiling == nz and A.shape == (n * na * nz, n * n * nz)
    x = rng.normal(size=n * n * nz)
    b = rng.normal(size=n * na * nz)
    X, B = x.reshape(n * n, nz), b.reshape(n * na, nz)
    e1 = rel_l2(A.forward(x), (m @ X).reshape(-1))
    e2 = rel_l2(A.backward(b), (m.T @ B).reshape(-1))
    print("algebraic forward %.2e backward %.2e" % (e1, e2))
    assert e1 <= 1e-6 and e2 <= 1e-6


def test_raytrace_recorded_matrix(gpu):
    """P f = integrate(f) and P^T g = deposit(g) for the recorded matrix."""
    import torch
    from tests.test_gpu_raytrace import _setup
    tg, pg, cv, nf, im, am, ab = _setup(14, 6, np.arange(0.0, 360.0, 30.0), "rk4", "canonical")
    import vamtoolbox_amd as vam
    P_op = vam.projectorconstructor.projectorconstructor(tg, pg)
    P_op.output_torch_tensor = True
    P = P_op.buildPropagationMatrix()
    rng = np.random.default_rng(4)
    f = rng.uniform(0, 1, P_op.vol_shape)
    g = rng.uniform(0, 1, P_op.num_rays)
    Pd = torch.sparse_coo_tensor(P.indices().cpu(), P.values().double().cpu(), size=P.shape)
    pf = torch.sparse.mm(Pd, torch.as_tensor(f.reshape(-1, 1))).numpy().ravel()
    ptg = torch.sparse.mm(Pd.t(), torch.as_tensor(g.reshape(-1, 1))).numpy().ravel()
    e1 = rel_l2(P_op.forward(torch.tensor(f, dtype=torch.float32, device=gpu)).cpu().numpy(), pf)
    e2 = rel_l2(P_op.backward(torch.tensor(g, dtype=torch.float32, device=gpu)).cpu().numpy().ravel(), ptg)
    print("recorded matrix nnz %d: P f vs integrate %.2e, P^T g vs deposit %.2e" % (P._nnz(), e1, e2))
    assert e1 <= 1e-5 and e2 <= 1e-5


def test_impulse_matrix_and_algebraic_osmo(gpu, tmp_path):
    """buildMatrix rows = back projections of unit sinogram elements, so
    P = B^T; OSMO through the algebraic propagator with that matrix equals the
    oracle OSMO loop on the same explicit matrix."""
    import torch
    import oracle.optim as oo
    import vamtoolbox_amd as vam
    from vamtoolbox_amd.parallel import ParallelPlan
    from vamtoolbox_amd.util.build_sparse_matrix import buildMatrix
    n, na = 12, 10
    ang = np.linspace(0, 180, na, endpoint=False)
    rng = np.random.default_rng(6)
    t = np.zeros((n, n, 1))
    t[4:8, 3:9, 0] = 1.0
    tg = vam.geometry.TargetGeometry(target=t)
    pg = vam.geometry.ProjectionGeometry(ang, "parallel", CUDA=True)
    path = tmp_path / "P.npz"
    P = buildMatrix(tg, pg, save_path=str(path))
    plan = ParallelPlan(n, 1, ang, device=gpu)
    x = rng.normal(size=(n, n, 1))
    bt = plan.backward_adjoint(torch.tensor(x, dtype=torch.float32, device=gpu)).cpu().numpy().ravel()
    e = rel_l2(P @ x.ravel(), bt)
    print("impulse matrix vs B^T: %.2e (nnz %d)" % (e, P.nnz))
    assert e <= 1e-6
    pga = vam.geometry.ProjectionGeometry(ang, "algebraic", CUDA=True, loading_path_for_matrix=str(path))
    opts = vam.optimize.Options(method="OSMO", n_iter=4, d_h=0.85, d_l=0.6, filter=None)
    sino, recon, err = vam.optimize.optimize(tg, pga, opts)
    Pd = P.toarray()
    b_ref, x_ref, err_ref = oo.osmo(tg.array, lambda v: Pd @ v.ravel(), lambda s: (Pd.T @ s.ravel()).reshape(n, n, 1),
                                    4, d_h=0.85, d_l=0.6, filt=None)
    e_b, e_x = rel_l2(np.ravel(sino.array), np.ravel(b_ref)), rel_l2(np.ravel(recon.array), np.ravel(x_ref))
    print("algebraic OSMO rel L2 sinogram %.2e recon %.2e" % (e_b, e_x))
    assert e_b <= 1e-4 and e_x <= 1e-4
