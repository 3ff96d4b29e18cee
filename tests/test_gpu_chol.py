"""GPU tiled Cholesky solve (the reduced-camera-system solver of ba) against numpy.linalg.solve."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def db():
    from hislam2_amd import backend
    return backend()


@pytest.mark.parametrize('n', [1, 6, 63, 64, 65, 200, 1800])
def test_spd_solve_f64(db, n):
    rng = np.random.default_rng(n)
    M = rng.standard_normal((n, n))
    A = M @ M.T + n * np.eye(n)
    b = rng.standard_normal(n)
    x, ok = db.spd_solve(torch.from_numpy(A).cuda(), torch.from_numpy(b).cuda())
    ref = np.linalg.solve(A, b)
    err = float(np.abs(x.cpu().numpy() - ref).max() / np.abs(ref).max())
    print(n, err)
    assert int(ok) == 1 and err <= 1e-10


def test_spd_solve_ill_conditioned_f64(db):
    """Condition ~1e8, like a damped reduced system with gauge freedom."""
    rng = np.random.default_rng(7)
    n = 300
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (Q * np.logspace(-3, 5, n)) @ Q.T
    b = rng.standard_normal(n)
    x, ok = db.spd_solve(torch.from_numpy(A).cuda(), torch.from_numpy(b).cuda())
    ref = np.linalg.solve(A, b)
    assert int(ok) == 1
    assert np.abs(x.cpu().numpy() - ref).max() / np.abs(ref).max() <= 1e-6


def test_spd_solve_f32(db):
    rng = np.random.default_rng(3)
    n = 500
    M = rng.standard_normal((n, n))
    A = M @ M.T + n * np.eye(n)
    b = rng.standard_normal(n)
    x, ok = db.spd_solve(torch.from_numpy(A).float().cuda(), torch.from_numpy(b).float().cuda())
    ref = np.linalg.solve(A, b)
    assert int(ok) == 1
    assert np.abs(x.cpu().numpy() - ref).max() / np.abs(ref).max() <= 1e-4


@pytest.mark.parametrize('n', [5, 130])
def test_spd_solve_reports_indefinite(db, n):
    """A matrix that is not positive definite gives x = 0 and ok = 0 (the failure contract of
    the reduced solve, src/droid_kernels.cu:1669-1677)."""
    rng = np.random.default_rng(11)
    M = rng.standard_normal((n, n))
    A = M @ M.T + np.eye(n)
    A[n // 2, n // 2] = -5.0
    x, ok = db.spd_solve(torch.from_numpy(A).cuda(), torch.from_numpy(rng.standard_normal(n)).cuda())
    assert int(ok) == 0
    assert torch.count_nonzero(x).item() == 0
