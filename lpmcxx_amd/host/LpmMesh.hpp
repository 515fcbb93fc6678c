// Spherical polygonal meshes (cubed sphere, icosahedral triangles) that
// materialize the particles of the Biot-Savart / Poisson sum: vertices
// (targets) and leaf-face centroids with areas (targets and sources).
//
// Reference behaviour kept:
//   seeds and per-level counts     src/LpmMeshSeed.cpp:17-67,225-284, meshSeeds/*.dat
//   uniform refinement loop         src/LpmPolyMesh2d.cpp:62-69
//   edge division                   src/LpmEdges.cpp:99-120
//   triangle / quad division        src/LpmTriFaces.cpp:13-142, src/LpmQuadFaces.cpp:14-146
//   face vertices, centroid, area   src/LpmFaces.cpp:63-80,147-167
//   fields on vertices / faces      src/LpmField.h:16-94, src/LpmMeshedParticles.cpp:12-77
// Index order of vertices and faces matches the reference for the same seed
// and level, so fields exchange index-for-index with the reference's.
#pragma once

#include "LpmCore.hpp"

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace Lpm {

// ---------------------------------------------------------------- Field
// SoA field of nDim components (src/LpmField.h:16-94).
class Field {
  public:
    Field(index_type n, int nDim = 1, const std::string& name = "null", const std::string& units = "n/a")
        : _n(n), _nDim(nDim), _name(name), _units(units), _comp(nDim, std::vector<scalar_type>(n, 0.0)) {}

    index_type n() const { return _n; }
    int nDim() const { return _nDim; }
    const std::string& name() const { return _name; }
    const std::string& units() const { return _units; }
    void rename(const std::string& s) { _name = s; }

    scalar_type getScalar(index_type i) const { return _comp[0][i]; }
    XyzVector getVec(index_type i) const {
        return XyzVector(_comp[0][i], _nDim > 1 ? _comp[1][i] : 0.0, _nDim > 2 ? _comp[2][i] : 0.0);
    }
    void replace(index_type i, scalar_type v) { _comp[0][i] = v; }
    void replace(index_type i, const XyzVector& v) {
        _comp[0][i] = v.x;
        if (_nDim > 1) _comp[1][i] = v.y;
        if (_nDim > 2) _comp[2][i] = v.z;
    }
    scalar_type* getPtrToData(int comp = 0) { return _comp.at(comp).data(); }
    const scalar_type* getPtrToData(int comp = 0) const { return _comp.at(comp).data(); }
    void initializeToConstant(scalar_type v) {
        for (auto& c : _comp) std::fill(c.begin(), c.end(), v);
    }
    scalar_type maxMagnitude() const;

  private:
    index_type _n;
    int _nDim;
    std::string _name, _units;
    std::vector<std::vector<scalar_type>> _comp;
};

// ---------------------------------------------------------------- analytic functions
class AnalyticFunction {
  public:
    virtual ~AnalyticFunction() {}
    virtual scalar_type evaluateScalar(const XyzVector&) const { return 0.0; }
    virtual XyzVector evaluateVector(const XyzVector&) const { return XyzVector(); }
};

// zeta = 30 cos(4 lambda) z (z^2-1)^2 (src/LpmAnalyticFunctions.h:87-103).
class SphereHarmonic54 : public AnalyticFunction {
  public:
    scalar_type evaluateScalar(const XyzVector& v) const override {
        const scalar_type lon = atan4(v.y, v.x);
        return 30.0 * std::cos(4.0 * lon) * legendre54(v.z);
    }
    static scalar_type legendre54(scalar_type z) { return z * (z * z - 1.0) * (z * z - 1.0); }
};

// Exact velocity for SphereHarmonic54 (src/LpmAnalyticFunctions.h:105-120).
class RossbyHaurwitz54Velocity : public AnalyticFunction {
  public:
    XyzVector evaluateVector(const XyzVector& v) const override {
        const scalar_type lat = latitude(v);
        const scalar_type lon = longitude(v);
        const scalar_type u = 0.5 * std::cos(4.0 * lon) * cube(std::cos(lat)) * (5.0 * std::cos(2.0 * lat) - 3.0);
        const scalar_type w = 4.0 * cube(std::cos(lat)) * std::sin(lat) * std::sin(4.0 * lon);
        return XyzVector(-u * std::sin(lon) - w * std::sin(lat) * std::cos(lon),
                         u * std::cos(lon) - w * std::sin(lat) * std::sin(lon), w * std::cos(lat));
    }
};

// ---------------------------------------------------------------- seeds
class MeshSeed {
  public:
    virtual ~MeshSeed() {}
    virtual index_type nFaces(int level) const = 0;
    virtual index_type nVertices(int level) const = 0;
    virtual index_type nEdges(index_type nverts, index_type nfaces) const = 0;
    virtual int nRootFaces() const = 0;
    virtual int nEdgesPerFace() const = 0;
    virtual std::string idString() const = 0;
    // Seed tables: vertex coordinates, edges (orig, dest, left, right), face edge lists.
    virtual void tables(std::vector<XyzVector>& verts, std::vector<index_type>& edges,
                        std::vector<index_type>& faceEdges) const = 0;
};

class IcosTriSphereSeed : public MeshSeed {
  public:
    index_type nFaces(int k) const override { return 20 * (index_type(1) << (2 * k)); }
    index_type nVertices(int k) const override { return 2 + 10 * (index_type(1) << (2 * k)); }
    index_type nEdges(index_type nv, index_type nf) const override { return nf + nv - 2; }
    int nRootFaces() const override { return 20; }
    int nEdgesPerFace() const override { return 3; }
    std::string idString() const override { return "icosTriSphere"; }
    void tables(std::vector<XyzVector>&, std::vector<index_type>&, std::vector<index_type>&) const override;
};

class CubedSphereSeed : public MeshSeed {
  public:
    index_type nFaces(int k) const override { return 6 * (index_type(1) << (2 * k)); }
    index_type nVertices(int k) const override { return 2 + 6 * (index_type(1) << (2 * k)); }
    index_type nEdges(index_type nv, index_type nf) const override { return nf + nv - 2; }
    int nRootFaces() const override { return 6; }
    int nEdgesPerFace() const override { return 4; }
    std::string idString() const override { return "cubedSphere"; }
    void tables(std::vector<XyzVector>&, std::vector<index_type>&, std::vector<index_type>&) const override;
};

// ---------------------------------------------------------------- mesh
// Half-edge style mesh with uniform recursive refinement.  Every face,
// divided or not, keeps its index; leaves carry the particles.
class PolyMesh2d {
  public:
    PolyMesh2d(const MeshSeed& seed, int maxRecursionLevel, scalar_type radius = 1.0);

    index_type nVertices() const { return index_type(vx.size()); }
    index_type nFaces() const { return index_type(fdiv.size()); }
    index_type nLeafFaces() const { return _nLeaves; }
    index_type nEdges() const { return index_type(eorig.size()); }
    int nEdgesPerFace() const { return _nepf; }
    scalar_type radius() const { return _radius; }

    XyzVector getVertex(index_type i) const { return XyzVector(vx[i], vy[i], vz[i]); }
    bool faceIsDivided(index_type i) const { return fdiv[i] != 0; }
    scalar_type faceArea(index_type i) const { return farea[i]; }
    XyzVector faceCentroid(index_type i) const;
    void faceVertices(index_type i, index_type* out) const;
    scalar_type surfaceArea() const;
    // sqrt(max |v|^2) over vertices (src/LpmCoords.cpp:136-143).
    scalar_type maxR() const;

    std::vector<scalar_type> vx, vy, vz;                 // vertex coordinates
    std::vector<index_type> eorig, edest, eleft, eright;  // edges
    std::vector<index_type> echild;                       // 2 per edge, -1 if undivided
    std::vector<index_type> fedges;                       // nepf per face
    std::vector<index_type> fchild;                       // 4 per face
    std::vector<char> fdiv;
    std::vector<scalar_type> farea;

  private:
    void insertVertex(const XyzVector& v);
    index_type insertEdge(index_type o, index_type d, index_type l, index_type r);
    void divideEdge(index_type e);
    void divideFace(index_type f);
    void computeArea(index_type f);
    bool edgePositive(index_type f, index_type e) const { return eleft[e] == f; }

    int _nepf;
    scalar_type _radius;
    index_type _nLeaves;
};

// ---------------------------------------------------------------- meshed particles
// Mesh plus named vertex and face fields (src/LpmMeshedParticles.h:13-60).
class MeshedParticles {
  public:
    MeshedParticles(const MeshSeed& seed, int maxRecursionLevel, bool isLagrangian = false,
                    scalar_type domainRadius = 1.0, int prank = 0);

    index_type nVertices() const { return mesh2d->nVertices(); }
    index_type nFaces() const { return mesh2d->nFaces(); }
    index_type nLeafFaces() const { return mesh2d->nLeafFaces(); }
    scalar_type maxR() const { return mesh2d->maxR(); }
    XyzVector getVertex(index_type i) const { return mesh2d->getVertex(i); }
    XyzVector getFaceCentroid(index_type i) const { return mesh2d->faceCentroid(i); }
    scalar_type getFaceArea(index_type i) const { return mesh2d->faceArea(i); }
    bool faceIsDivided(index_type i) const { return mesh2d->faceIsDivided(i); }
    std::shared_ptr<PolyMesh2d> meshPtr() const { return mesh2d; }

    void createVertexField(const std::string& name, const std::string& units, int dim);
    void createFaceField(const std::string& name, const std::string& units, int dim);
    // Unknown names throw std::out_of_range (the reference returns a null
    // pointer and then dereferences it, src/LpmMeshedParticles.cpp:80-129).
    std::shared_ptr<Field> getVertexFieldPtr(const std::string& name) const;
    std::shared_ptr<Field> getFaceFieldPtr(const std::string& name) const;
    void initializeVertexFieldWithFunction(const std::string& name, const AnalyticFunction* fn);
    void initializeFaceFieldWithFunction(const std::string& name, const AnalyticFunction* fn);

    // Leaf-face centroids / areas / indices in ascending face order, and all
    // vertices, as contiguous SoA arrays (the source/target materialization
    // done once per solve).
    void leafFaceArrays(std::vector<scalar_type>& cx, std::vector<scalar_type>& cy, std::vector<scalar_type>& cz,
                        std::vector<scalar_type>& area, std::vector<index_type>& faceIndex) const;

    // Legacy VTK polydata of the mesh with every vertex field as point data
    // and every face field as cell data (src/LpmMeshedParticles.cpp:131-153;
    // writer in LpmVtk.cpp).
    void writeToVtkFile(const std::string& fname, const std::string& desc = "") const;

  private:
    std::shared_ptr<PolyMesh2d> mesh2d;
    std::map<std::string, std::shared_ptr<Field>> _vertexFields, _faceFields;
};

}  // namespace Lpm
