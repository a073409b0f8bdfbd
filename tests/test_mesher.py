"""Pins of the package's own SES mesher (the stand-in for the closed MSMS /
NanoShaper executables the reference shells out to, src/readin.f90:90-246):
area and enclosed volume against closed forms for one atom (the SES of a
single ball is its sphere) and for two overlapping atoms (two spherical caps
joined by the probe's toroidal reentrant patch), converging as the grid
refines; and the mesh is a closed, consistently oriented 2-manifold whose
Euler characteristic matches its genus."""
import numpy as np
import pytest
from scipy import integrate

from tabipb import surface


def area_volume(v, f):
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1).sum()
    vol = np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0   # divergence theorem
    return area, vol


def two_atom_ses(R, d, p):
    """Area and volume of the SES of two balls of radius R, centres d apart on
    the z axis, probe radius p (requires d < 2(R+p) and a probe-centre circle
    of radius rc > p): spherical caps beyond the contact circles at |z| = zc,
    and between them the inner half of the torus swept by the probe, whose
    meridian is rho(z) = rc - sqrt(p^2 - z^2)."""
    rc = np.sqrt((R + p) ** 2 - (d / 2) ** 2)
    zc = (d / 2) * p / (R + p)
    phi0 = np.arcsin((d / 2) / (R + p))
    area_torus = 2 * np.pi * p * (2 * rc * phi0 - 2 * p * np.sin(phi0))
    area_caps = 2 * (2 * np.pi * R * (d / 2 + R - zc))
    v_torus = integrate.quad(lambda z: np.pi * (rc - np.sqrt(p * p - z * z)) ** 2, -zc, zc)[0]
    v_caps = 2 * integrate.quad(lambda z: np.pi * (R * R - (z - d / 2) ** 2), zc, d / 2 + R)[0]
    return area_caps + area_torus, v_caps + v_torus


def manifold_checks(v, f):
    """Every directed edge appears once and its reverse once (closed, oriented,
    2-manifold); returns the Euler characteristic V - E + F."""
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    key = e[:, 0] * (len(v) + 1) + e[:, 1]
    assert len(np.unique(key)) == len(key), "a directed edge repeats"
    rev = e[:, 1] * (len(v) + 1) + e[:, 0]
    assert np.all(np.isin(rev, key)), "an edge has no opposite half-edge (hole or flip)"
    used = np.unique(f)
    assert len(used) == len(v), "unreferenced vertices"
    return len(v) - len(key) // 2 + len(f)


@pytest.mark.parametrize("case", ["one", "two"])
def test_area_and_volume_converge_to_closed_forms(case):
    if case == "one":
        xyz, r = np.zeros((1, 3)), np.array([2.0])
        A0, V0 = 4 * np.pi * 2.0 ** 2, 4 / 3 * np.pi * 2.0 ** 3
    else:
        xyz, r = np.array([[0.0, 0.0, -1.4], [0.0, 0.0, 1.4]]), np.array([1.8, 1.8])
        A0, V0 = two_atom_ses(1.8, 2.8, 1.4)
    errs = []
    for h in (0.3, 0.2):
        v, n, f = surface.ses_mesh(xyz, r, probe=1.4, h=h)
        A, V = area_volume(v, f)
        errs.append((abs(A / A0 - 1), abs(V / V0 - 1)))
        assert manifold_checks(v, f) == 2
    # the grid distance transform places the surface up to ~0.1 h outside
    # the exact SES: a few percent at these spacings, shrinking with h
    (a3, v3), (a2, v2) = errs
    assert a3 < 0.06 and v3 < 0.07 and a2 < 0.04 and v2 < 0.05
    assert a2 < a3 / 1.3 and v2 < v3 / 1.3


def test_single_atom_normals_are_outward_radial():
    v, n, f = surface.ses_mesh(np.zeros((1, 3)), np.array([2.0]), probe=1.4, h=0.25)
    np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-12)
    cosang = np.einsum("ij,ij->i", n, v / np.linalg.norm(v, axis=1)[:, None])
    # every normal outward (area-weighted face normals of a grid mesh carry
    # the staircase of the grid: a few vertices tilt by up to ~25 degrees)
    assert cosang.min() > 0.85 and np.median(cosang) > 0.995 and cosang.mean() > 0.99
    rad = np.linalg.norm(v, axis=1)
    assert abs(rad.mean() - 2.0) < 0.05 and rad.std() < 0.03


def test_molecular_surface_is_a_closed_oriented_manifold():
    xyz, q, r = surface.synth_molecule(150, seed=11)
    v, n, f = surface.ses_mesh(xyz, r, probe=1.4, h=0.5)
    chi = manifold_checks(v, f)
    assert chi <= 2 and chi % 2 == 0          # genus g = (2 - chi) / 2 >= 0
    A, V = area_volume(v, f)
    assert V > 0                              # outward orientation
    # every atom centre lies inside the surface: the solid angle of the mesh
    # seen from a point inside a closed surface is 4 pi
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    for p in xyz[::15]:
        ra, rb, rc = a - p, b - p, c - p
        la, lb, lc = (np.linalg.norm(x, axis=1) for x in (ra, rb, rc))
        num = np.einsum("ij,ij->i", ra, np.cross(rb, rc))
        den = la * lb * lc + np.einsum("ij,ij->i", ra, rb) * lc + \
            np.einsum("ij,ij->i", ra, rc) * lb + np.einsum("ij,ij->i", rb, rc) * la
        omega = 2 * np.arctan2(num, den).sum()
        assert abs(omega - 4 * np.pi) < 1e-6
