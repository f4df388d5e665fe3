"""Pins the periodic Voronoi geometry producer against scipy.spatial.Voronoi (Qhull, an independent implementation) on
periodic images, and against closed forms: the divergence theorem (each closed cell's area vectors sum to zero), the box
volume, the pyramid formula V = sum |A| |d| / (2 D), and the centroid from the pyramids."""
import numpy as np
import pytest
from scipy.spatial import ConvexHull, Voronoi

from luna_hydro import mesh, problems


def build(n, D, jitter=0.25, seed=3):
    pts = problems.jittered_lattice(n, D, jitter, seed)
    return mesh.build(pts, 1.0)


def neighbours(m, i):
    out = {}
    for k in range(m["start"][i], m["start"][i + 1]):
        e = m["list"][k]
        f = e if e >= 0 else ~e
        j = m["neighbour"][f] if e >= 0 else m["owner"][f]
        out[int(j)] = f
    return out


@pytest.mark.parametrize("D,n", [(3, 8), (3, 12), (2, 24)])
def test_closed_cells_and_box_volume(D, n):
    m = build(n, D)
    N = m["cells"]
    clo = np.zeros((N, D))
    np.add.at(clo, m["owner"], m["area"])
    np.add.at(clo, m["neighbour"], -m["area"])
    scale = np.abs(m["area"]).max()
    assert np.abs(clo).max() < 1e-13 * scale
    assert abs(m["volume"].sum() - 1.0) < 1e-13
    # owner bit and list order
    for i in range(N):
        lst = m["list"][m["start"][i]:m["start"][i + 1]]
        faces = [e if e >= 0 else ~e for e in lst]
        assert faces == sorted(faces)
        for e in lst:
            f = e if e >= 0 else ~e
            assert (m["owner"][f] == i) == (e >= 0)
    assert np.all(m["owner"] < m["neighbour"])


@pytest.mark.parametrize("D,n", [(3, 8), (2, 16)])
def test_against_qhull(D, n):
    m = build(n, D)
    x = m["positions"]
    N = m["cells"]
    shifts = np.array(np.meshgrid(*[[-1, 0, 1]] * D, indexing="ij")).reshape(D, -1).T
    images = (x[None, :, :] + shifts[:, None, :]).reshape(-1, D)
    centre = [k for k, s in enumerate(shifts) if not s.any()][0]
    vor = Voronoi(images)
    ridges = {}
    for (a, b), rv in zip(vor.ridge_points, vor.ridge_vertices):
        for p, q in ((a, b), (b, a)):
            if p // N == centre:
                ridges.setdefault(p % N, []).append((q % N, rv))
    for i in range(N):
        mine = neighbours(m, i)
        region = vor.regions[vor.point_region[centre * N + i]]
        verts = vor.vertices[region] - x[i]
        hull = ConvexHull(verts)
        assert abs(hull.volume - m["volume"][i]) < 1e-12 * hull.volume
        # centroid of the region (tetrahedra / triangles from its interior point)
        cent = verts.mean(axis=0)
        tot, mom = 0.0, np.zeros(D)
        for simplex in hull.simplices:
            pts = np.vstack([cent[None, :], verts[simplex]])
            vol = abs(np.linalg.det(pts[1:] - pts[0])) / (6.0 if D == 3 else 2.0)
            tot += vol
            mom += vol * pts.mean(axis=0)
        assert np.abs(mom / tot - m["centroid"][i]).max() < 1e-11 * n ** -1
        # faces: qhull's ridges of positive area are exactly this cell's faces
        qfaces = {}
        for j, rv in ridges[i]:
            pv = vor.vertices[rv] - x[i]
            if D == 2:
                area = np.linalg.norm(pv[1] - pv[0])
            else:
                c = pv.mean(axis=0)
                nrm = np.cross(pv[1] - pv[0], pv[2] - pv[0])
                nrm /= np.linalg.norm(nrm)
                # order the ridge's vertices around their mean
                u = pv[0] - c
                u /= np.linalg.norm(u)
                v = np.cross(nrm, u)
                ang = np.arctan2((pv - c) @ v, (pv - c) @ u)
                pv = pv[np.argsort(ang)]
                area = 0.5 * abs(sum(np.cross(pv[t] - pv[0], pv[t + 1] - pv[0]) @ nrm for t in range(1, len(pv) - 1)))
            if area > 1e-12:
                qfaces[j] = area
        assert set(qfaces) == set(mine), (i, sorted(qfaces), sorted(mine))
        for j, f in mine.items():
            assert abs(np.linalg.norm(m["area"][f]) - qfaces[j]) < 1e-11 * n ** -(D - 1)


@pytest.mark.parametrize("D,n", [(3, 10), (2, 20)])
def test_pyramid_formulas(D, n):
    m = build(n, D)
    x, N = m["positions"], m["cells"]
    d = x[m["neighbour"]] - x[m["owner"]]
    d -= np.round(d)
    dist = np.linalg.norm(d, axis=1)
    a = np.linalg.norm(m["area"], axis=1)
    V = np.zeros(N)
    np.add.at(V, m["owner"], a * dist / (2 * D))
    np.add.at(V, m["neighbour"], a * dist / (2 * D))
    assert np.abs(V - m["volume"]).max() < 1e-13 * m["volume"].max()
    # area vectors are parallel to the generator separation (bisector faces)
    cosang = np.einsum("ij,ij->i", m["area"], d) / (a * dist)
    assert np.abs(cosang - 1).max() < 1e-13
    # face centroids lie on the bisector plane: (fc - d/2) . d = 0 (relative to the owner's generator)
    off = np.einsum("ij,ij->i", m["face_centroid"] - 0.5 * d, d) / dist
    assert np.abs(off).max() < 1e-13


@pytest.mark.parametrize("D,n", [(2, 12), (3, 6)])
def test_geometry_derivative_against_the_producer(D, n):
    """The geometry's derivative at fixed topology (luna_mesh_derivative_*): its tangent equals central differences of the
    producer itself (positions moved by whole grid steps, so the snapping adds nothing; the topology unchanged), and its
    adjoint is the tangent's transpose (dot test)."""
    x0 = problems.jittered_lattice(n, D, 0.25, 3)
    m0 = mesh.build(x0, 1.0)
    x, step = m0["positions"], m0["step"]
    g = mesh.GeometryDerivative(m0)
    rng = np.random.default_rng(5)
    dx = rng.integers(-300, 301, size=x.shape).astype(np.float64) * step
    mp, mm = mesh.build(x + dx, 1.0), mesh.build(x - dx, 1.0)
    for key in ("owner", "neighbour", "start", "list"):
        assert np.array_equal(mp[key], m0[key]) and np.array_equal(mm[key], m0[key])
    dV, dc, dA, dxf = g.tangent(dx)
    for got, key in ((dV, "volume"), (dc, "centroid"), (dA, "area"), (dxf, "face_centroid")):
        fd = (mp[key] - mm[key]) / 2
        assert np.max(np.abs(got - fd)) < 1e-6 * np.max(np.abs(fd)), key
    bars = [rng.normal(size=a.shape) for a in (dV, dc, dA, dxf)]
    xbar = g.adjoint(*bars)
    lhs = sum(np.sum(b * t) for b, t in zip(bars, (dV, dc, dA, dxf)))
    assert abs(lhs - np.sum(xbar * dx)) < 1e-12 * abs(lhs)
    g.close()
