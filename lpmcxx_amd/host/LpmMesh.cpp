// Mesh construction for the particle sets of the spherical sum.
// See LpmMesh.hpp for the reference behaviour this reproduces.
#include "LpmMesh.hpp"

#include <algorithm>
#include <cstdlib>

#include "../csrc/lpm_seeds.hpp"

namespace Lpm {

scalar_type Field::maxMagnitude() const {
    scalar_type r = 0.0;
    for (index_type i = 0; i < _n; ++i) {
        scalar_type m = 0.0;
        for (int c = 0; c < _nDim; ++c) m += _comp[c][i] * _comp[c][i];
        m = std::sqrt(m);
        if (m > r) r = m;
    }
    return r;
}

// ---------------------------------------------------------------- seed tables
// The tables live in lpm_seeds.hpp, shared with the GPU mesh generator.
using namespace lpm_seed;

void IcosTriSphereSeed::tables(std::vector<XyzVector>& v, std::vector<index_type>& e,
                               std::vector<index_type>& f) const {
    v.clear(); e.clear(); f.clear();
    for (auto& c : kIcosXyz) v.emplace_back(std::strtod(c[0], nullptr), std::strtod(c[1], nullptr), std::strtod(c[2], nullptr));
    for (auto& r : kIcosEdges) for (int k = 0; k < 4; ++k) e.push_back(r[k]);
    for (auto& r : kIcosFaces) for (int k = 0; k < 3; ++k) f.push_back(r[k]);
}

void CubedSphereSeed::tables(std::vector<XyzVector>& v, std::vector<index_type>& e,
                             std::vector<index_type>& f) const {
    v.clear(); e.clear(); f.clear();
    const scalar_type a = std::strtod(kCubeA, nullptr);
    for (auto& s : kCubeSigns) v.emplace_back(s[0] * a, s[1] * a, s[2] * a);
    for (auto& r : kCubeEdges) for (int k = 0; k < 4; ++k) e.push_back(r[k]);
    for (auto& r : kCubeFaces) for (int k = 0; k < 4; ++k) f.push_back(r[k]);
}

// ---------------------------------------------------------------- PolyMesh2d
PolyMesh2d::PolyMesh2d(const MeshSeed& seed, int maxLevel, scalar_type radius)
    : _nepf(seed.nEdgesPerFace()), _radius(radius), _nLeaves(0) {
    if (maxLevel < 0 || maxLevel > 12) throw std::invalid_argument("PolyMesh2d: recursion level must be in [0, 12]");
    index_type nf = 0, ne = 0;
    for (int k = 0; k <= maxLevel; ++k) {
        nf += seed.nFaces(k);
        ne += seed.nEdges(seed.nVertices(k), seed.nFaces(k));
    }
    const index_type nv = seed.nVertices(maxLevel);
    vx.reserve(nv); vy.reserve(nv); vz.reserve(nv);
    eorig.reserve(ne); edest.reserve(ne); eleft.reserve(ne); eright.reserve(ne); echild.reserve(2 * ne);
    fedges.reserve(nf * _nepf); fchild.reserve(4 * nf); fdiv.reserve(nf); farea.reserve(nf);

    std::vector<XyzVector> sv;
    std::vector<index_type> se, sf;
    seed.tables(sv, se, sf);
    for (auto v : sv) { v.scale(radius); insertVertex(v); }
    for (size_t i = 0; i < se.size(); i += 4) insertEdge(se[i], se[i + 1], se[i + 2], se[i + 3]);
    for (size_t i = 0; i < sf.size(); i += _nepf) {
        for (int k = 0; k < _nepf; ++k) fedges.push_back(sf[i + k]);
        for (int k = 0; k < 4; ++k) fchild.push_back(-1);
        fdiv.push_back(0);
        farea.push_back(0.0);
    }
    for (index_type i = 0; i < nFaces(); ++i) computeArea(i);   // resetAreas: all seed faces are leaves

    for (int k = 0; k < maxLevel; ++k) {
        const index_type n = nFaces();
        for (index_type i = 0; i < n; ++i)
            if (!fdiv[i]) divideFace(i);
    }
    _nLeaves = 0;
    for (char d : fdiv) _nLeaves += (d ? 0 : 1);
}

void PolyMesh2d::insertVertex(const XyzVector& v) {
    vx.push_back(v.x); vy.push_back(v.y); vz.push_back(v.z);
}

index_type PolyMesh2d::insertEdge(index_type o, index_type d, index_type l, index_type r) {
    eorig.push_back(o); edest.push_back(d); eleft.push_back(l); eright.push_back(r);
    echild.push_back(-1); echild.push_back(-1);
    return index_type(eorig.size()) - 1;
}

// Split edge e at the normalized chord midpoint; child 0 runs orig->mid,
// child 1 mid->dest, both inheriting e's left/right faces.
void PolyMesh2d::divideEdge(index_type e) {
    const XyzVector mid = sphereMidpoint(getVertex(eorig[e]), getVertex(edest[e]), _radius);
    const index_type m = nVertices();
    insertVertex(mid);
    const index_type c0 = insertEdge(eorig[e], m, eleft[e], eright[e]);
    const index_type c1 = insertEdge(m, edest[e], eleft[e], eright[e]);
    echild[2 * e] = c0;
    echild[2 * e + 1] = c1;
}

void PolyMesh2d::faceVertices(index_type f, index_type* out) const {
    for (int k = 0; k < _nepf; ++k) {
        const index_type e = fedges[f * _nepf + k];
        out[k] = edgePositive(f, e) ? eorig[e] : edest[e];
    }
}

XyzVector PolyMesh2d::faceCentroid(index_type f) const {
    index_type vi[4];
    XyzVector v[4];
    faceVertices(f, vi);
    for (int k = 0; k < _nepf; ++k) v[k] = getVertex(vi[k]);
    return sphereCentroid(v, _nepf);
}

// Sum of L'Huilier areas of the fan (centroid, v_k, v_k+1).
void PolyMesh2d::computeArea(index_type f) {
    const XyzVector c = faceCentroid(f);
    index_type vi[4];
    faceVertices(f, vi);
    scalar_type a = 0.0;
    for (int k = 0; k < _nepf; ++k)
        a += sphereTriArea(c, getVertex(vi[k]), getVertex(vi[(k + 1) % _nepf]), _radius);
    farea[f] = a;
}

// Divide face f into four children appended at the end of the face list.
// Child j keeps parent corner j; for quads a new center vertex (the parent's
// centroid) is the opposite corner of every child.
void PolyMesh2d::divideFace(index_type f) {
    const int n = _nepf;
    const index_type c0 = nFaces();
    index_type ne[4][4], nvx[4][4];
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) ne[a][b] = nvx[a][b] = -1;
    index_type pv[4];
    faceVertices(f, pv);
    for (int j = 0; j < n; ++j) nvx[j][j] = pv[j];

    for (int j = 0; j < n; ++j) {
        const index_type pe = fedges[f * n + j];
        if (echild[2 * pe] < 0) divideEdge(pe);
        const index_type h0 = echild[2 * pe], h1 = echild[2 * pe + 1];
        const int jn = (j + 1) % n;
        if (edgePositive(f, pe)) {
            ne[j][j] = h0;   eleft[h0] = c0 + j;
            ne[jn][j] = h1;  eleft[h1] = c0 + jn;
        } else {
            ne[j][j] = h1;   eright[h1] = c0 + j;
            ne[jn][j] = h0;  eright[h0] = c0 + jn;
        }
        const index_type m = edest[h0];
        if (n == 4) {
            nvx[j][jn] = m;
            nvx[jn][j] = m;
        } else if (j == 0) {
            nvx[0][1] = m; nvx[1][0] = m; nvx[3][2] = m;
        } else if (j == 1) {
            nvx[1][2] = m; nvx[2][1] = m; nvx[3][0] = m;
        } else {
            nvx[2][0] = m; nvx[0][2] = m; nvx[3][1] = m;
        }
    }

    const index_type eb = nEdges();
    if (n == 4) {
        const index_type cv = nVertices();
        insertVertex(faceCentroid(f));
        for (int j = 0; j < 4; ++j) nvx[j][(j + 2) % 4] = cv;
        ne[0][1] = eb;     ne[0][2] = eb + 3;
        ne[1][3] = eb;     ne[1][2] = eb + 2;
        ne[2][3] = eb + 1; ne[2][0] = eb + 2;
        ne[3][1] = eb + 1; ne[3][0] = eb + 3;
        insertEdge(nvx[0][1], nvx[0][2], c0 + 0, c0 + 1);
        insertEdge(nvx[3][1], nvx[3][2], c0 + 3, c0 + 2);
        insertEdge(nvx[1][2], nvx[1][3], c0 + 1, c0 + 2);
        insertEdge(nvx[0][2], nvx[0][3], c0 + 0, c0 + 3);
    } else {
        ne[3][0] = eb; ne[3][1] = eb + 1; ne[3][2] = eb + 2;
        ne[0][1] = eb + 1;
        ne[1][2] = eb + 2;
        ne[2][0] = eb;
        insertEdge(nvx[2][1], nvx[2][0], c0 + 3, c0 + 2);
        insertEdge(nvx[0][2], nvx[0][1], c0 + 3, c0 + 0);
        insertEdge(nvx[1][0], nvx[1][2], c0 + 3, c0 + 1);
    }

    for (int j = 0; j < 4; ++j) {
        for (int k = 0; k < n; ++k) fedges.push_back(ne[j][k]);
        for (int k = 0; k < 4; ++k) fchild.push_back(-1);
        fdiv.push_back(0);
        farea.push_back(0.0);
    }
    fdiv[f] = 1;
    for (int j = 0; j < 4; ++j) fchild[4 * f + j] = c0 + j;
    farea[f] = 0.0;
    for (int j = 0; j < 4; ++j) computeArea(c0 + j);
}

scalar_type PolyMesh2d::surfaceArea() const {
    scalar_type s = 0.0;
    for (scalar_type a : farea) s += a;
    return s;
}

scalar_type PolyMesh2d::maxR() const {
    scalar_type r = 0.0;
    for (index_type i = 0; i < nVertices(); ++i) {
        const scalar_type t = square(vx[i]) + square(vy[i]) + square(vz[i]);
        if (t > r) r = t;
    }
    return std::sqrt(r);
}

// ---------------------------------------------------------------- MeshedParticles
MeshedParticles::MeshedParticles(const MeshSeed& seed, int maxLevel, bool, scalar_type radius, int)
    : mesh2d(std::make_shared<PolyMesh2d>(seed, maxLevel, radius)) {
    auto area = std::make_shared<Field>(mesh2d->nFaces(), 1, "faceArea", "length^2");
    for (index_type i = 0; i < mesh2d->nFaces(); ++i) area->replace(i, mesh2d->faceArea(i));
    _faceFields["faceArea"] = area;
}

void MeshedParticles::createVertexField(const std::string& name, const std::string& units, int dim) {
    _vertexFields[name] = std::make_shared<Field>(nVertices(), dim, name, units);
}
void MeshedParticles::createFaceField(const std::string& name, const std::string& units, int dim) {
    _faceFields[name] = std::make_shared<Field>(nFaces(), dim, name, units);
}
std::shared_ptr<Field> MeshedParticles::getVertexFieldPtr(const std::string& name) const {
    auto it = _vertexFields.find(name);
    if (it == _vertexFields.end()) throw std::out_of_range("no vertex field named " + name);
    return it->second;
}
std::shared_ptr<Field> MeshedParticles::getFaceFieldPtr(const std::string& name) const {
    auto it = _faceFields.find(name);
    if (it == _faceFields.end()) throw std::out_of_range("no face field named " + name);
    return it->second;
}
void MeshedParticles::initializeVertexFieldWithFunction(const std::string& name, const AnalyticFunction* fn) {
    auto f = getVertexFieldPtr(name);
    for (index_type i = 0; i < nVertices(); ++i) {
        if (f->nDim() == 1) f->replace(i, fn->evaluateScalar(getVertex(i)));
        else f->replace(i, fn->evaluateVector(getVertex(i)));
    }
}
void MeshedParticles::initializeFaceFieldWithFunction(const std::string& name, const AnalyticFunction* fn) {
    auto f = getFaceFieldPtr(name);
    for (index_type i = 0; i < nFaces(); ++i) {
        if (faceIsDivided(i)) {
            if (f->nDim() == 1) f->replace(i, 0.0);
            else f->replace(i, XyzVector());
        } else {
            const XyzVector c = getFaceCentroid(i);
            if (f->nDim() == 1) f->replace(i, fn->evaluateScalar(c));
            else f->replace(i, fn->evaluateVector(c));
        }
    }
}

void MeshedParticles::leafFaceArrays(std::vector<scalar_type>& cx, std::vector<scalar_type>& cy,
                                     std::vector<scalar_type>& cz, std::vector<scalar_type>& area,
                                     std::vector<index_type>& faceIndex) const {
    const index_type nl = nLeafFaces();
    cx.resize(nl); cy.resize(nl); cz.resize(nl); area.resize(nl); faceIndex.resize(nl);
    index_type k = 0;
    for (index_type i = 0; i < nFaces(); ++i) {
        if (faceIsDivided(i)) continue;
        const XyzVector c = getFaceCentroid(i);
        cx[k] = c.x; cy[k] = c.y; cz[k] = c.z;
        area[k] = getFaceArea(i);
        faceIndex[k] = i;
        ++k;
    }
}

}  // namespace Lpm
