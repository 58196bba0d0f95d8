"""Host-side helpers (CPU only)."""
import numpy as np

from ants_amd.analysis import jacobian_determinant


def test_jacobian_of_affine_field():
    # u = A x: det(I + A) everywhere
    A = np.array([[0.1, 0.02, 0.0], [-0.03, 0.05, 0.01], [0.0, 0.02, -0.04]])
    z, y, x = np.meshgrid(np.arange(6.0), np.arange(7.0), np.arange(8.0), indexing="ij")
    pts = np.stack([x, y, z], -1)
    u = pts @ A.T
    J = jacobian_determinant(u, (1.0, 1.0, 1.0))
    np.testing.assert_allclose(J, np.linalg.det(np.eye(3) + A), rtol=1e-12)
