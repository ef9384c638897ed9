"""Shape functions and masses written by the model generator, against what
the mathematics fixes: partition of unity, linear reproduction, consistency
of derivatives, total mass = density x area."""
import numpy as np
import pytest

from melody2d_amd.preprocessor import shape
from melody2d_amd.preprocessor.bodies import deformable_body, disc_nodes


@pytest.fixture(scope="module")
def body():
    rng = np.random.Generator(np.random.PCG64(7))
    bxy, ixy = disc_nodes(rng, 0.1, -0.2, 1e-3, 32, 0.2e-3)
    return deformable_body(bxy, ixy, "M", 0.2e-3, 1000.0)


def _rows(ptr, vals):
    return np.add.reduceat(vals, ptr[:-1])


def test_mls_partition_of_unity_and_linear_reproduction(body):
    b = body
    assert np.allclose(_rows(b.gp_ptr, b.gp_phi), 1.0, atol=1e-12)
    assert np.allclose(_rows(b.gp_ptr, b.gp_phix), 0.0, atol=1e-6)
    for c in range(2):
        xc = b.X[b.gp_idx, c]
        assert np.allclose(_rows(b.gp_ptr, b.gp_phi * xc), b.gp_xy[:, c], atol=1e-15)
        dx = _rows(b.gp_ptr, b.gp_phix * xc)
        dy = _rows(b.gp_ptr, b.gp_phiy * xc)
        assert np.allclose(dx, 1.0 if c == 0 else 0.0, atol=1e-9)
        assert np.allclose(dy, 1.0 if c == 1 else 0.0, atol=1e-9)


def test_mls_derivative_matches_finite_difference(body):
    b = body
    pts = b.gp_xy[:5]
    h = 1e-9
    ptr, idx = shape.supports(pts, b.X, b.domain)
    phi, phix, phiy = shape.mls(pts, b.X, b.domain, ptr, idx)
    for d in range(2):
        e = np.zeros(2); e[d] = h
        p1 = shape.mls(pts + e, b.X, b.domain, ptr, idx)[0]
        p0 = shape.mls(pts - e, b.X, b.domain, ptr, idx)[0]
        fd = (p1 - p0) / (2 * h)
        an = phix if d == 0 else phiy
        assert np.allclose(fd, an, rtol=1e-5, atol=1e-3)


def test_rpim_reproduces_linear_fields():
    rng = np.random.Generator(np.random.PCG64(3))
    nodes = rng.random((60, 2))
    pts = 0.3 + 0.4 * rng.random((20, 2))
    from scipy.spatial import cKDTree
    _, nn = cKDTree(nodes).query(pts, k=12)
    nn = np.sort(nn, axis=1)
    ptr = np.arange(21) * 12
    idx = nn.reshape(-1)
    phi, phix, phiy = shape.rpim(pts, nodes, ptr, idx, shape_c=1.0)
    assert np.allclose(_rows(ptr, phi), 1.0, atol=1e-9)
    assert np.allclose(_rows(ptr, phi * nodes[idx, 0]), pts[:, 0], atol=1e-9)
    assert np.allclose(_rows(ptr, phix * nodes[idx, 0]), 1.0, atol=1e-6)
    assert np.allclose(_rows(ptr, phiy * nodes[idx, 1]), 1.0, atol=1e-6)


def test_lumped_mass_sums_to_density_times_area(body):
    b = body
    area = b.gp_jac.sum() / 3.0
    assert abs(b.mass[:, 0].sum() - 1000.0 * area) < 1e-12 * 1000.0 * area * 10
    assert np.all(b.mass > 0)


def test_border_is_counter_clockwise_and_closed(body):
    br = body.borders[0]
    assert br.nodes[0] == br.nodes[-1]
    P = body.X[br.nodes[:-1]]
    area2 = np.sum(P[:, 0] * np.roll(P[:, 1], -1) - np.roll(P[:, 0], -1) * P[:, 1])
    assert area2 > 0
