"""Pins of the oracle's kinematics, geometry, ray-march and source sampling
against what the mathematics fixes (no GPU).

Each check is chosen so that a plausible slip in the restatement (a wrong
sign or weight in the centre-of-mass transform, a swapped axis, a missing
coincidence rule, an off-by-one bin) fails it:
  * elastic scattering off a target at rest: E' in [alpha E, E], E' uniform
    there (isotropic in CM), the CM speed conserved, and the lab cosine tied
    to E' by the two-body relations (src/physics.cpp:754-818);
  * free-gas target velocities: the sample moments of the constant-cross-section
    free-gas density, computed by quadrature (src/physics.cpp:990-1041);
  * plane / cylinder distances on hand cases (src/surface.cpp:208-218, 400-442);
  * RectLattice distance and get_indices at tile edges (src/lattice.cpp:252-342);
  * per-bin track fractions of the mesh ray-march against brute-force plane
    crossings (src/mesh.cpp:1000-1146);
  * the box-plus-domain source: acceptance fraction and where the sites land
    (src/source.cpp:226-271, 355-417).
"""
import math
import os

import numpy as np
import pytest
from scipy import integrate, stats

from helpers import prepare
from oracle.oracle import Oracle, elastic, surface_distance, target_velocity

INFTY = 1.7976931348623157e308


# ------------------------------------------------------------------ elastic

@pytest.mark.parametrize("awr", [1.5, 11.9, 235.0])
def test_elastic_off_target_at_rest_two_body_relations(awr):
    """E >= 400 kT and A > 1: the target is at rest.  With mu_cm the CM
    cosine, E'/E = (A^2 + 2 A mu_cm + 1) / (A + 1)^2 and the lab cosine is
    (1 + A mu_cm) / sqrt(A^2 + 2 A mu_cm + 1); the neutron's CM speed is
    A/(A+1) sqrt(E) before and after."""
    A = awr
    alpha = ((A - 1) / (A + 1)) ** 2
    rng = np.random.default_rng(int(awr * 10))
    E = 1.0e6
    kT = 0.0253
    seed = 12345
    Ep = []
    for k in range(4000):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        e2, u2, seed = elastic(A, kT, E, u, seed)
        Ep.append(e2)
        assert alpha * E * (1 - 1e-12) <= e2 <= E * (1 + 1e-12)
        # CM speed: v_cm = sqrt(E) u / (A + 1); |v' - v_cm| = A sqrt(E) / (A + 1)
        v_cm = math.sqrt(E) * u / (A + 1)
        v2 = math.sqrt(e2) * np.asarray(u2)
        assert np.linalg.norm(v2 - v_cm) == pytest.approx(A * math.sqrt(E) / (A + 1), rel=1e-10)
        # lab cosine from mu_cm
        mu_cm = ((e2 / E) * (A + 1) ** 2 - A * A - 1) / (2 * A)
        mu_lab = (1 + A * mu_cm) / math.sqrt(A * A + 2 * A * mu_cm + 1)
        assert float(np.dot(u, u2)) == pytest.approx(mu_lab, abs=1e-8)
        assert np.linalg.norm(u2) == pytest.approx(1.0, abs=1e-14)
    # isotropic in CM <=> E' uniform on [alpha E, E]
    p = stats.kstest((np.array(Ep) - alpha * E) / ((1 - alpha) * E), "uniform").pvalue
    assert p > 1e-4


def _free_gas_moments(beta_vn):
    """E[x] and E[mu] of x = (beta v_t)^2 and the target cosine mu under the
    density ~ sqrt(b^2 + x - 2 b sqrt(x) mu) x^(1/2) e^(-x) dx dmu (the
    constant-cross-section free-gas kernel in the variable x), b = beta_vn."""
    def dens(mu, x):
        return math.sqrt(max(beta_vn ** 2 + x - 2 * beta_vn * math.sqrt(x) * mu, 0.0)) * math.sqrt(x) * math.exp(-x)

    z = integrate.dblquad(dens, 0, 40, -1, 1, epsabs=1e-11)[0]
    mx = integrate.dblquad(lambda mu, x: x * dens(mu, x), 0, 40, -1, 1, epsabs=1e-11)[0] / z
    mm = integrate.dblquad(lambda mu, x: mu * dens(mu, x), 0, 40, -1, 1, epsabs=1e-11)[0] / z
    return mx, mm


@pytest.mark.parametrize("beta_vn", [0.3, 1.0, 3.0])
def test_free_gas_target_velocity_moments(beta_vn):
    """Sample means of (beta v_t)^2 and of the target cosine against the
    quadrature of the free-gas density (4.5 sigma)."""
    awr, kT = 2.0, 1.0
    E = beta_vn ** 2 * kT / awr
    u = (0.0, 0.6, 0.8)
    n = 40000
    x = np.empty(n)
    mu = np.empty(n)
    seed = 99
    for k in range(n):
        v, seed = target_velocity(awr, E, u, kT, seed)
        v = np.asarray(v)
        sp = np.linalg.norm(v)
        x[k] = awr * sp * sp / kT
        mu[k] = float(np.dot(v, u)) / sp
    mx, mm = _free_gas_moments(beta_vn)
    assert abs(x.mean() - mx) < 4.5 * x.std() / math.sqrt(n)
    assert abs(mu.mean() - mm) < 4.5 * mu.std() / math.sqrt(n)
    # the plain Maxwellian (no relative-speed weight) would give E[x] = 1.5, E[mu] = 0
    assert abs(mx - 1.5) > 0.02 or abs(mm) > 0.02


# ------------------------------------------------------------------ surfaces

def test_plane_distances_hand_cases():
    # x-plane at 3: from x=1 along +x -> 2; along -x -> never; oblique
    assert surface_distance(1, [3.0], (1, 0, 0), (1, 0, 0)) == 2.0
    assert surface_distance(1, [3.0], (1, 0, 0), (-1, 0, 0)) == INFTY
    assert surface_distance(1, [3.0], (1, 5, 0), (0.6, 0.8, 0)) == pytest.approx(2 / 0.6, rel=1e-15)
    assert surface_distance(1, [3.0], (1, 0, 0), (0, 1, 0)) == INFTY       # parallel
    assert surface_distance(1, [3.0], (3.0, 0, 0), (1, 0, 0)) == INFTY     # on it (|f| < 1e-12)
    assert surface_distance(1, [3.0], (1, 0, 0), (1, 0, 0), True) == INFTY  # coincident flag
    # y- and z-planes use their own axis
    assert surface_distance(2, [-4.0], (9, 0, 9), (0, -1, 0)) == 4.0
    assert surface_distance(3, [10.0], (0, 0, -10), (0, 0.6, 0.8)) == pytest.approx(20 / 0.8, rel=1e-15)
    # general plane x + y + z = 3 from the origin along (1,1,1)/sqrt3 -> sqrt3
    s = 1 / math.sqrt(3)
    assert surface_distance(4, [1, 1, 1, 3], (0, 0, 0), (s, s, s)) == pytest.approx(math.sqrt(3), rel=1e-15)


def test_cylinder_distances_hand_cases():
    c = [0.0, 0.0, 0.4]  # z-cylinder r = 0.4 at the origin
    # inside at the axis, along +x: 0.4
    assert surface_distance(5, c, (0, 0, 0), (1, 0, 0)) == pytest.approx(0.4, rel=1e-15)
    # inside off-axis: from (0.1, 0, 0) along +y -> sqrt(0.16 - 0.01)
    assert surface_distance(5, c, (0.1, 0, 7), (0, 1, 0)) == pytest.approx(math.sqrt(0.15), rel=1e-14)
    # outside, heading in: from (-1, 0, 0) along +x -> 0.6 (near side)
    assert surface_distance(5, c, (-1, 0, 0), (1, 0, 0)) == pytest.approx(0.6, rel=1e-14)
    # outside, heading away -> never
    assert surface_distance(5, c, (-1, 0, 0), (-1, 0, 0)) == INFTY
    # outside, missing: from (-1, 0.5, 0) along +x -> never
    assert surface_distance(5, c, (-1, 0.5, 0), (1, 0, 0)) == INFTY
    # along the axis direction -> never (a = 0)
    assert surface_distance(5, c, (0.1, 0, 0), (0, 0, 1)) == INFTY
    # oblique: from the axis along (0.6, 0, 0.8): the xy-speed is 0.6 -> 0.4 / 0.6
    assert surface_distance(5, c, (0, 0, 0), (0.6, 0, 0.8)) == pytest.approx(0.4 / 0.6, rel=1e-14)
    # on the surface heading out -> never; heading in -> the chord 0.8
    assert surface_distance(5, c, (0.4, 0, 0), (1, 0, 0), True) == INFTY
    assert surface_distance(5, c, (0.4, 0, 0), (-1, 0, 0), True) == pytest.approx(0.8, rel=1e-14)
    # an x-cylinder uses (y, z)
    assert surface_distance(6, [0.0, 0.0, 2.0], (5, 0, 0), (0, 0, 1)) == pytest.approx(2.0, rel=1e-15)


# ------------------------------------------------------------------ lattice

@pytest.fixture(scope="module")
def lattice_oracle():
    d, model, nucs = prepare(1, n_particles=2000, n_batches=1, grid_lo=2000, grid_hi=3000, n_nuclides=6,
                             lattice_n=5, mesh=(5, 5, 4), half_height=20.0)
    return Oracle(os.path.join(d, "oracle_problem.txt"))


def test_lattice_indices_and_distance_at_tile_edges(lattice_oracle):
    o = lattice_oracle
    pitch, n = 1.26, 5
    ll = -0.5 * pitch * n
    # interior points: floor((x - ll) / pitch)
    assert o.lattice_indices(0, (0.0, 0.0, 0.0), (1, 0, 0)) == (2, 2, 0)
    assert o.lattice_indices(0, (ll + 0.1, ll + 4.9 * pitch, 0), (1, 0, 0)) == (0, 4, 0)
    # exactly on the edge x = ll + 3 pitch: the direction decides the tile
    x = ll + 3 * pitch
    assert o.lattice_indices(0, (x, 0.0, 0.0), (1, 0, 0)) == (3, 2, 0)
    assert o.lattice_indices(0, (x, 0.0, 0.0), (-1, 0, 0)) == (2, 2, 0)
    assert o.lattice_indices(0, (x + 1e-13, 0.0, 0.0), (-1, 0, 0)) == (2, 2, 0)  # within FP_COINCIDENT
    assert o.lattice_indices(0, (0.0, x, 0.0), (0, 1, 0)) == (2, 3, 0)
    # distance in the tile's local frame: from the centre along +x -> half a pitch, translation +x
    d, t = o.lattice_distance(0, (0, 0, 0), (1, 0, 0))
    assert d == pytest.approx(0.63, rel=1e-15) and t == (1, 0, 0)
    # along (0.6, -0.8): y-faces first (0.63 / 0.8 < 0.63 / 0.6), translation -y
    d, t = o.lattice_distance(0, (0, 0, 0), (0.6, -0.8, 0))
    assert d == pytest.approx(0.63 / 0.8, rel=1e-14) and t == (0, -1, 0)
    # on the +x face heading +x: that face is skipped (within FP_PRECISION) -> the y face
    d, t = o.lattice_distance(0, (0.63, 0.0, 0), (0.6, 0.8, 0))
    assert d == pytest.approx(0.63 / 0.8, rel=1e-14) and t == (0, 1, 0)


def test_locate_and_distance_to_boundary_in_the_lattice(lattice_oracle):
    """Cell search and distance_to_boundary across levels: from a fuel centre
    along +x the fuel cylinder (r 0.4) comes first; from the moderator near a
    tile edge the lattice boundary does."""
    o = lattice_oracle
    pitch, n = 1.26, 5
    ll = -0.5 * pitch * n
    cx = ll + 2.5 * pitch
    d, info = o.locate((cx, cx, 0.0), (1, 0, 0))
    assert info["found"] and info["n_coord"] == 2 and info["lattice_index"] == (2, 2, 0)
    assert d == pytest.approx(0.4, rel=1e-14) and info["coord_level"] == 2 and info["surface"] > 0
    # moderator at local (0.55, 0): the +x tile face is 0.08 away
    d, info = o.locate((cx + 0.55, cx, 0.0), (1, 0, 0))
    assert d == pytest.approx(0.08, rel=1e-9) and info["surface"] == 0 and info["translation"] == (1, 0, 0)
    # going down from z = 0: the vacuum plane at z = -20
    d, info = o.locate((cx + 0.55, cx, 0.0), (0, 0, -1))
    assert d == pytest.approx(20.0, rel=1e-14) and info["coord_level"] == 1


# ------------------------------------------------------------------ ray-march

def _brute_force_tracks(shape, ll, ur, r0, r1):
    """Bins and track fractions of segment r0 -> r1 from all plane crossings."""
    r0 = np.asarray(r0, float)
    r1 = np.asarray(r1, float)
    seg = r1 - r0
    width = (np.asarray(ur) - np.asarray(ll)) / np.asarray(shape)
    ts = {0.0, 1.0}
    for k in range(3):
        if seg[k] == 0:
            continue
        for i in range(shape[k] + 1):
            t = (ll[k] + i * width[k] - r0[k]) / seg[k]
            if 0 < t < 1:
                ts.add(t)
    ts = sorted(ts)
    out = {}
    for a, b in zip(ts[:-1], ts[1:]):
        if b - a < 1e-12:
            continue
        m = r0 + 0.5 * (a + b) * seg
        # a point on a bin face belongs to the lower bin: the mesh's 1-based
        # index is ceil((x - ll) / width) (src/mesh.cpp:1328-1331)
        ijk = np.ceil((m - ll) / width).astype(int) - 1
        if np.all(ijk >= 0) and np.all(ijk < shape):
            b_ = int(ijk[0] + shape[0] * (ijk[1] + shape[1] * ijk[2]))
            out[b_] = out.get(b_, 0.0) + (b - a)
    return out


@pytest.fixture(scope="module")
def cube_oracle():
    d, model, nucs = prepare(0, n_particles=100, n_batches=1, n_grid=2000)
    return Oracle(os.path.join(d, "oracle_problem.txt"))


@pytest.mark.parametrize("r0,r1", [
    ((-9.5, 0.3, 0.3), (9.5, 0.3, 0.3)),       # along x, through 10 bins
    ((-9.0, -9.0, -9.0), (9.0, 9.0, 9.0)),     # the diagonal: through vertices (corner crossings)
    ((-9.0, -3.0, 1.0), (7.0, 5.0, -8.5)),     # oblique
    ((-15.0, 0.1, 0.2), (15.0, 0.1, 0.2)),     # starts and ends outside the mesh
    ((0.5, 0.5, 0.5), (0.7, 0.6, 0.5)),        # inside one bin
    ((-4.0, 2.0, -12.0), (-4.0, 2.0, 12.0)),   # along z on bin faces x = -4, y = 2
], ids=["x", "diagonal", "oblique", "outside-in", "one-bin", "on-faces"])
def test_raytrace_track_fractions_brute_force(cube_oracle, r0, r1):
    """Per-bin track fractions of the 10^3 mesh over the 20 cm cube."""
    seg = np.asarray(r1) - np.asarray(r0)
    u = seg / np.linalg.norm(seg)
    bins, fr = cube_oracle.raytrace(r0, r1, u)
    got = {}
    for b, f in zip(bins, fr):
        got[b] = got.get(b, 0.0) + f
    want = _brute_force_tracks((10, 10, 10), (-10, -10, -10), (10, 10, 10), r0, r1)
    got = {b: f for b, f in got.items() if f > 1e-9}
    assert set(got) == set(want)
    for b in want:
        assert got[b] == pytest.approx(want[b], abs=2e-9)
    inside = all(abs(c) <= 10 for c in list(r0) + list(r1))
    if inside:
        assert sum(fr) == pytest.approx(1.0, abs=1e-12)


# ------------------------------------------------------------------ source

def test_box_source_with_fuel_domain(lattice_oracle):
    """Box over the lattice with the fuel cells as the domain: the accepted
    fraction of uniform points is the fuel's share of the box,
    25 pi 0.4^2 / (6.3^2 * ...) in x-y; every sampled site lies inside a fuel
    pin, spread evenly over the 25 pins and uniformly in area within them."""
    o = lattice_oracle
    pitch, n, r = 1.26, 5, 0.4
    half = 0.5 * pitch * n
    rng = np.random.default_rng(8)
    m = 20000
    pts = rng.uniform([-half, -half, -20], [half, half, 20], size=(m, 3))
    acc = np.array([o.in_source_domain(p) for p in pts])
    frac = n * n * math.pi * r * r / (2 * half) ** 2
    assert abs(acc.mean() - frac) < 4.5 * math.sqrt(frac * (1 - frac) / m)
    seed = 77
    pins = np.zeros(n * n, int)
    rad2 = []
    for k in range(5000):
        site, seed = o.sample_source(seed)
        x, y = site[0], site[1]
        ix = int(math.floor((x + half) / pitch))
        iy = int(math.floor((y + half) / pitch))
        dx = x - (-half + (ix + 0.5) * pitch)
        dy = y - (-half + (iy + 0.5) * pitch)
        assert dx * dx + dy * dy < r * r
        pins[ix + n * iy] += 1
        rad2.append((dx * dx + dy * dy) / (r * r))
        assert abs(site[2]) <= 20.0 and site[6] > 0
    assert stats.chisquare(pins).pvalue > 1e-4
    assert stats.kstest(rad2, "uniform").pvalue > 1e-4
