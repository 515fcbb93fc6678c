"""Particle-set pins for the host mesh builder (input stage of the sum).

Reference behaviour: src/LpmMeshSeed.cpp:225-284 (counts),
tests/LpmMeshSeedUnitTests.cpp (counts per level), tests/LpmCoordsUnitTests.cpp:106-139
(sphere area 4 pi), meshSeeds/*.dat (seed geometry).  The geometry is checked
against an independent numpy construction (recursive subdivision with
normalized midpoints, Van Oosterom-Strackee triangle areas), which shares no
code with the C++ builder.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

import lpmcxx_amd as L

A = 0.577350269189626
CUBE_CORNERS = np.array([[1, -1, 1], [1, -1, -1], [1, 1, -1], [1, 1, 1], [-1, 1, -1], [-1, 1, 1],
                         [-1, -1, -1], [-1, -1, 1]], float) * A
CUBE_FACES = [[0, 1, 2, 3], [3, 2, 4, 5], [5, 4, 6, 7], [7, 6, 1, 0], [7, 0, 3, 5], [1, 6, 4, 2]]


def _icosahedron():
    t = (1 + 5 ** 0.5) / 2
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], float)
    v /= np.linalg.norm(v, axis=1)[:, None]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6],
         [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
         [8, 6, 7], [9, 8, 1]]
    return v, f


def _nrm(v):
    return v / np.linalg.norm(v)


def _tri_area(a, b, c):
    # Van Oosterom & Strackee: tan(E/2) = |a.(b x c)| / (1 + a.b + b.c + c.a)
    num = abs(np.dot(a, np.cross(b, c)))
    den = 1 + np.dot(a, b) + np.dot(b, c) + np.dot(c, a)
    return 2 * np.arctan2(num, den)


def _refine(polys, level):
    for _ in range(level):
        out = []
        for p in polys:
            if len(p) == 4:
                v0, v1, v2, v3 = p
                m01, m12, m23, m30 = (_nrm(v0 + v1), _nrm(v1 + v2), _nrm(v2 + v3), _nrm(v3 + v0))
                c = _nrm(v0 + v1 + v2 + v3)
                out += [(v0, m01, c, m30), (m01, v1, m12, c), (c, m12, v2, m23), (m30, c, m23, v3)]
            else:
                v0, v1, v2 = p
                m01, m12, m20 = _nrm(v0 + v1), _nrm(v1 + v2), _nrm(v2 + v0)
                out += [(v0, m01, m20), (m01, v1, m12), (m20, m12, v2), (m01, m12, m20)]
        polys = out
    return polys


def _independent(seed, level):
    if seed == L.LPM_SEED_CUBED:
        polys = [tuple(CUBE_CORNERS[i] for i in f) for f in CUBE_FACES]
    else:
        v, f = _icosahedron()
        polys = [tuple(v[i] for i in t) for t in f]
    polys = _refine(polys, level)
    cents = np.array([_nrm(sum(p)) for p in polys])
    areas = np.array([sum(_tri_area(c, p[k], p[(k + 1) % len(p)]) for k in range(len(p)))
                      for c, p in zip(cents, polys)])
    verts = np.unique(np.round(np.concatenate([np.array(p) for p in polys]), 12), axis=0)
    return cents, areas, verts


@pytest.mark.parametrize("seed,level", [(L.LPM_SEED_ICOS_TRI, k) for k in range(6)] +
                         [(L.LPM_SEED_CUBED, k) for k in range(7)])
def test_counts_and_area(mesh, seed, level):
    m = mesh(seed, level)
    if seed == L.LPM_SEED_CUBED:
        nv, nl, nroot = 2 + 6 * 4 ** level, 6 * 4 ** level, 6
    else:
        nv, nl, nroot = 2 + 10 * 4 ** level, 20 * 4 ** level, 20
    assert m.n_vertices == nv
    assert m.n_leaves == nl
    assert m.n_faces == sum(nroot * 4 ** k for k in range(level + 1))
    # leaves are exactly the faces created by the last division sweep
    assert np.array_equal(m.face_index, np.arange(m.n_faces - nl, m.n_faces))
    assert abs(m.area.sum() - 4 * np.pi) < 1e-12
    assert np.all(m.area > 0)
    r = np.sqrt(m.cx ** 2 + m.cy ** 2 + m.cz ** 2)
    assert np.abs(r - 1).max() < 4e-16
    rv = np.sqrt(m.vx ** 2 + m.vy ** 2 + m.vz ** 2)
    assert np.abs(rv - 1).max() < 4e-16
    assert abs(m.max_radius - rv.max()) < 1e-16


def test_seed_vertices_are_seed_file_values(mesh):
    m = mesh(L.LPM_SEED_CUBED, 0)
    assert np.array_equal(np.stack([m.vx, m.vy, m.vz], 1), CUBE_CORNERS)
    mi = mesh(L.LPM_SEED_ICOS_TRI, 0)
    assert mi.vz[0] == 1.0 and mi.vz[11] == -1.0
    assert mi.vx[1] == float("0.723606797749978969640917366873")
    # cube-face centroids of the level-0 cubed sphere are the axis points
    assert np.allclose(np.abs(np.stack([m.cx, m.cy, m.cz], 1)).max(1), 1.0, atol=1e-15)


@pytest.mark.parametrize("seed,level", [(L.LPM_SEED_CUBED, 3), (L.LPM_SEED_ICOS_TRI, 3)])
def test_geometry_matches_independent_construction(mesh, seed, level):
    m = mesh(seed, level)
    cents, areas, verts = _independent(seed, level)
    mc = np.stack([m.cx, m.cy, m.cz], 1)
    if seed == L.LPM_SEED_ICOS_TRI:
        # the seed file's icosahedron is a rotated copy of the textbook one:
        # compare rotation-invariant quantities (sorted areas, neighbour spacing)
        assert np.allclose(np.sort(m.area), np.sort(areas), rtol=1e-10, atol=0)
        d1 = cKDTree(mc).query(mc, k=2)[0][:, 1]
        d2 = cKDTree(cents).query(cents, k=2)[0][:, 1]
        assert np.allclose(np.sort(d1), np.sort(d2), rtol=1e-10)
        return
    d, idx = cKDTree(cents).query(mc)
    assert d.max() < 1e-12
    assert len(np.unique(idx)) == len(idx)
    assert np.allclose(m.area, areas[idx], rtol=1e-11, atol=0)
    mv = np.stack([m.vx, m.vy, m.vz], 1)
    dv, iv = cKDTree(verts).query(mv)
    assert dv.max() < 1e-11 and len(np.unique(iv)) == len(verts) == len(mv)


def test_children_are_consecutive_and_nested(mesh):
    # every level-(k+1) face centroid lies closer to its parent's centroid
    # (face index p -> children first_child(p) .. +3) than to any other parent
    m1, m2 = mesh(L.LPM_SEED_CUBED, 2), mesh(L.LPM_SEED_CUBED, 3)
    p = np.stack([m1.cx, m1.cy, m1.cz], 1)
    c = np.stack([m2.cx, m2.cy, m2.cz], 1)
    owner = cKDTree(p).query(c)[1]
    assert np.array_equal(owner, np.repeat(np.arange(len(p)), 4))


def test_gpu_generator_counts_match_host_builder(mesh):
    """lpm_mesh_counts (the GPU generator's closed forms, no device needed)
    against the host builder's sizes; out-of-range arguments rejected."""
    for seed in (L.LPM_SEED_CUBED, L.LPM_SEED_ICOS_TRI):
        for level in range(0, 6):
            m = mesh(seed, level)
            assert L.mesh_counts(seed, level) == (m.n_vertices, m.n_faces, m.n_leaves)
        assert L.mesh_counts(seed, 12)[2] == (6 if seed == L.LPM_SEED_CUBED else 20) * 4 ** 12
    for seed, level in ((2, 3), (1, 13), (0, -1)):
        with pytest.raises(L.LpmError):
            L.mesh_counts(seed, level)


def _seed_fixture(name):
    import json
    from pathlib import Path
    d = json.loads((Path(__file__).resolve().parent / "golden" / "mesh_seeds.json").read_text())[name]
    xyz = [np.array([float(v) for v in r]) for r in d["xyz"]]
    # the root faces' vertex lists by the reference's rule (src/LpmFaces.cpp:98-121): the
    # start vertex of each edge in the face's edge list, orig when the face is the edge's left face
    verts = []
    for f, el in enumerate(d["face_edges"]):
        vl = []
        for e in el:
            o, de, left, right = d["edges"][e]
            vl.append(o if left == f else de)
        verts.append(vl)
    assert verts == d["face_verts"]
    return xyz, verts


def _divide_level(xyz, faces):
    """One uniform refinement in the reference's order (src/LpmPolyMesh2d.cpp:58-65:
    leaves in index order; src/LpmQuadFaces.cpp:14-103, src/LpmTriFaces.cpp:13-95:
    each face's edges in order, an undivided edge gets the next vertex (its
    sphere midpoint, src/LpmEdges.cpp:99-120), then a quad's centroid; the
    children placed as the reference's newFaceVertInds)."""
    mid = {}
    out = []
    for F in faces:
        n = len(F)
        m = []
        for j in range(n):
            a, b = F[j], F[(j + 1) % n]
            key = (min(a, b), max(a, b))
            if key not in mid:
                p = 0.5 * (xyz[a] + xyz[b])
                mid[key] = len(xyz)
                xyz.append(p / np.linalg.norm(p))
            m.append(mid[key])
        if n == 4:
            c = sum(xyz[v] for v in F) / 4.0
            ci = len(xyz)
            xyz.append(c / np.linalg.norm(c))
            for j in range(4):
                ch = [0] * 4
                ch[j], ch[(j + 1) % 4], ch[(j + 2) % 4], ch[(j + 3) % 4] = F[j], m[j], ci, m[(j - 1) % 4]
                out.append(ch)
        else:
            out += [[F[0], m[0], m[2]], [m[0], F[1], m[1]], [m[2], m[1], F[2]], [m[1], m[2], m[0]]]
    return out


@pytest.mark.parametrize("seed,name", [(L.LPM_SEED_CUBED, "cubed"), (L.LPM_SEED_ICOS_TRI, "icos")])
def test_index_order_levels_1_2(seed, name):
    """Vertex and leaf-face index order of the mesh builder against the order
    derived from the seed files (tests/golden/mesh_seeds.json) and the
    reference's division rules, levels 0-2: every vertex and every leaf
    centroid at its index."""
    xyz, faces = _seed_fixture(name)
    for level in range(3):
        if level:
            faces = _divide_level(xyz, faces)
        m = L.Mesh(seed, level)
        V = np.stack(xyz, 1)
        assert m.n_vertices == V.shape[1] and m.n_leaves == len(faces)
        assert np.abs(np.stack([m.vx, m.vy, m.vz]) - V).max() < 1e-15
        C = np.stack([sum(V[:, v] for v in F) / len(F) for F in faces], 1)
        C /= np.linalg.norm(C, axis=0)
        assert np.abs(np.stack([m.cx, m.cy, m.cz]) - C).max() < 1e-15
        # the leaves are the last faces, children of one parent consecutive
        assert np.array_equal(m.face_index, np.arange(m.n_faces - m.n_leaves, m.n_faces))
