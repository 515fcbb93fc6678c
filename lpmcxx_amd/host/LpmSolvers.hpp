// Host C++ API of the spherical Biot-Savart / Poisson sum, namespace Lpm.
//
// Keeps the reference's class names, constructor arguments and method
// signatures on this path so calling code moves over unchanged:
//   VectorKernel / BiotSavartSphere        src/LpmVectorKernel.h:11-39
//   ScalarKernel / SphereGreensFn          src/LpmScalarKernel.h:11-45
//   MultiIndex                             src/LpmMultiIndex.h:12-72
//   Box3d                                  src/LpmBox3d.h:12-91
//   MPIReplicatedData                      src/LpmMPIReplicatedData.h:11-37
//   Coords (SoA positions)                 src/LpmCoords.h:16-78
//   Tree / Node                            src/LpmOctree.h:21-98
//   DirectSum                              src/LpmDirectSum.h:18-56
//   TreeSum                                src/LpmTreeSum.h:47-97
//   TaylorSeries3d / SphereGreensSeries    src/LpmTaylorSeries3d.h:20-63
// The sums run on the GPU through the C ABI of include/lpm_gpu.h; failures
// come back as the reference's exception types (std::logic_error for an
// unusable kernel, std::bad_alloc for allocation, std::invalid_argument /
// std::runtime_error otherwise).  One process is one rank: ranks of a
// replicated-data partition are blocks of targets solved locally, and the
// broadcast step has nothing to exchange.
#pragma once

#include "LpmCore.hpp"
#include "LpmMesh.hpp"
#include "lpm_gpu.h"

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace Lpm {

// ---------------------------------------------------------------- kernels
class VectorKernel {
  public:
    virtual ~VectorKernel() {}
    virtual XyzVector evaluate(const XyzVector& tgtLoc, const XyzVector& srcLoc) const = 0;
    virtual bool isSingular() const = 0;
    virtual scalar_type sumMultiplier() const { return 1.0; }
};

class BiotSavartSphere : public VectorKernel {
  public:
    BiotSavartSphere(const scalar_type delta = 0.0, const scalar_type sphRadius = 1.0)
        : _delta(delta), _sphRadius(sphRadius) {}
    // (x * y) / (R^2 - x.y + delta^2), zero when |den| <= 1e-13
    XyzVector evaluate(const XyzVector& t, const XyzVector& s) const override {
        const XyzVector c = t.crossProduct(s);
        const scalar_type den = _sphRadius * _sphRadius - t.dotProduct(s) + _delta * _delta;
        return std::abs(den) > ZERO_TOL ? c.scalarMultiply(1.0 / den) : XyzVector(0.0, 0.0, 0.0);
    }
    scalar_type sumMultiplier() const override { return -1.0 / (4.0 * PI * _sphRadius); }
    bool isSingular() const override { return _delta == 0.0; }
    scalar_type delta() const { return _delta; }
    scalar_type radius() const { return _sphRadius; }

  protected:
    scalar_type _delta, _sphRadius;
};

class ScalarKernel {
  public:
    virtual ~ScalarKernel() {}
    virtual scalar_type evaluate(const XyzVector& tgtLoc, const XyzVector& srcLoc) const = 0;
    virtual bool isSingular() const = 0;
};

class SphereGreensFn : public ScalarKernel {
  public:
    SphereGreensFn(const scalar_type r) : _radius(r) {}
    // -log(R^2 - y.x) / (4 pi R)
    scalar_type evaluate(const XyzVector& t, const XyzVector& s) const override {
        return -std::log(_radius * _radius - s.dotProduct(t)) / (4.0 * PI * _radius);
    }
    bool isSingular() const override { return true; }
    scalar_type radius() const { return _radius; }

  protected:
    scalar_type _radius;
};

// ---------------------------------------------------------------- multi-index
inline index_type factorial(const index_type n) { return (n == 1 || n == 0) ? 1 : n * factorial(n - 1); }

struct MultiIndex {
    MultiIndex(const unsigned i1 = 0, const unsigned i2 = 0, const unsigned i3 = 0) : k1(i1), k2(i2), k3(i3) {}
    index_type k1, k2, k3;
    index_type magnitude() const { return k1 + k2 + k3; }
    scalar_type vectorPower(const XyzVector& v) const {
        return std::pow(v.x, double(k1)) * std::pow(v.y, double(k2)) * std::pow(v.z, double(k3));
    }
    index_type factorial() const { return Lpm::factorial(k1) * Lpm::factorial(k2) * Lpm::factorial(k3); }
};
inline bool operator<(const MultiIndex& a, const MultiIndex& b) {
    if (a.k1 != b.k1) return a.k1 < b.k1;
    if (a.k2 != b.k2) return a.k2 < b.k2;
    return a.k3 < b.k3;
}
inline bool operator==(const MultiIndex& a, const MultiIndex& b) { return a.k1 == b.k1 && a.k2 == b.k2 && a.k3 == b.k3; }
inline bool operator!=(const MultiIndex& a, const MultiIndex& b) { return !(a == b); }
inline bool operator>(const MultiIndex& a, const MultiIndex& b) { return b < a; }
inline bool operator<=(const MultiIndex& a, const MultiIndex& b) { return !(b < a); }
inline bool operator>=(const MultiIndex& a, const MultiIndex& b) { return !(a < b); }
inline std::ostream& operator<<(std::ostream& os, const MultiIndex& k) {
    return os << "(" << k.k1 << ", " << k.k2 << ", " << k.k3 << ")";
}

// ---------------------------------------------------------------- Box3d
#define BOX_PADDING_FACTOR 0.00001

class Box3d {
  public:
    Box3d() : xmin(0), xmax(0), ymin(0), ymax(0), zmin(0), zmax(0), minRadius(0), maxRadius(0) {}
    Box3d(scalar_type x0, scalar_type x1, scalar_type y0, scalar_type y1, scalar_type z0, scalar_type z1)
        : xmin(x0), xmax(x1), ymin(y0), ymax(y1), zmin(z0), zmax(z1) { finish(); }
    Box3d(scalar_type maxr, bool pad = true) : xmin(-maxr), xmax(maxr), ymin(-maxr), ymax(maxr), zmin(-maxr), zmax(maxr) {
        if (pad) {
            xmin -= BOX_PADDING_FACTOR; xmax += BOX_PADDING_FACTOR;
            ymin -= BOX_PADDING_FACTOR; ymax += BOX_PADDING_FACTOR;
            zmin -= BOX_PADDING_FACTOR; zmax += BOX_PADDING_FACTOR;
        }
        finish();
    }
    scalar_type volume() const { return (xmax - xmin) * (ymax - ymin) * (zmax - zmin); }
    bool containsPoint(const XyzVector& v) const {
        return (xmin <= v.x && v.x < xmax) && (ymin <= v.y && v.y < ymax) && (zmin <= v.z && v.z < zmax);
    }
    std::vector<XyzVector> corners() const;
    std::vector<Box3d> bisectAll() const;
    scalar_type edgeLength(int dim) const { return dim == 0 ? xmax - xmin : dim == 1 ? ymax - ymin : zmax - zmin; }
    scalar_type longestEdge() const;
    std::string infoString() const;

    scalar_type xmin, xmax, ymin, ymax, zmin, zmax;
    scalar_type minRadius, maxRadius;
    XyzVector centroid;

  private:
    void finish();
};

// ---------------------------------------------------------------- replicated-data partition
class MPIReplicatedData {
  public:
    MPIReplicatedData(const index_type nItems, const int rank = 0, const int nProcs = 1);
    index_type startIndex(const int rank) const { return _start.at(rank); }
    index_type endIndex(const int rank) const { return _end.at(rank); }
    index_type msgSize(const int rank) const { return _size.at(rank); }
    int getRank() const { return _rank; }
    int getSize() const { return _nProcs; }
    index_type nItems() const { return _nItems; }

  private:
    index_type _nItems;
    int _nProcs, _rank;
    std::vector<index_type> _start, _end, _size;
};

// ---------------------------------------------------------------- Coords
class Coords {
  public:
    explicit Coords(index_type nMax = 0) { x.reserve(nMax); y.reserve(nMax); z.reserve(nMax); }
    index_type n() const { return index_type(x.size()); }
    void insert(const XyzVector& v) { x.push_back(v.x); y.push_back(v.y); z.push_back(v.z); }
    void insert(scalar_type a, scalar_type b, scalar_type c) { x.push_back(a); y.push_back(b); z.push_back(c); }
    XyzVector getVec(index_type i) const { return XyzVector(x[i], y[i], z[i]); }
    scalar_type minX() const; scalar_type maxX() const;
    scalar_type minY() const; scalar_type maxY() const;
    scalar_type minZ() const; scalar_type maxZ() const;
    scalar_type maxR() const;
    std::vector<scalar_type> x, y, z;
};

// ---------------------------------------------------------------- particles
// Point particles with named fields for the meshfree direct sum (Particles,
// src/LpmParticles.h:17-79: construction from a coordinate set, fields by
// name).
class Particles {
  public:
    explicit Particles(std::shared_ptr<Coords> crds) : _coords(std::move(crds)) {}
    std::shared_ptr<Coords> getCoordPtr(const bool = false) const { return _coords; }
    index_type n() const { return _coords->n(); }
    void createField(const std::string& name, const std::string& units, const int dim) {
        _fields[name] = std::make_shared<Field>(n(), dim, name, units);
    }
    // null if there is no such field (the reference logs a warning and returns null)
    std::shared_ptr<Field> getFieldPtr(const std::string& name) const {
        auto it = _fields.find(name);
        return it == _fields.end() ? nullptr : it->second;
    }
    void initializeFieldWithFunction(const std::string& name, const AnalyticFunction* fn);

  private:
    std::shared_ptr<Coords> _coords;
    std::map<std::string, std::shared_ptr<Field>> _fields;
};

// ---------------------------------------------------------------- GPU context
// One lpm_ctx per host thread, created on first use.
class GpuContext {
  public:
    explicit GpuContext(int device = 0);
    ~GpuContext();
    GpuContext(const GpuContext&) = delete;
    GpuContext& operator=(const GpuContext&) = delete;
    lpm_ctx* get() const { return _ctx; }
    // Throws the exception type matching an lpm_status.
    void check(int status, const char* where) const;
    static GpuContext& threadDefault();

  private:
    lpm_ctx* _ctx = nullptr;
};

// ---------------------------------------------------------------- octree
struct Node {
    Box3d box;
    int level = 0;
    index_type parent = -1;                  // node id, -1 at the root
    std::vector<index_type> kids;            // node ids in octant order
    std::vector<index_type> coordsContained; // source indices (ascending in leaves)
    bool isLeaf() const { return kids.empty(); }
    bool hasKids() const { return !kids.empty(); }
    index_type nCoords() const { return index_type(coordsContained.size()); }
};

class Tree {
  public:
    enum TREE_DEPTH_CONTROL { MAX_COORDS_PER_NODE, MAX_DEPTH };

    // Point-cloud tree: root box [min - 1e-5 min, max + 1e-5 max) per axis
    // (src/LpmOctree.cpp:16-36).
    Tree(const std::shared_ptr<Coords> crds, const scalar_type maxAspectRatio = 1.0, const int prank = 0);
    virtual ~Tree() {}

    // Builds on the GPU.  do_shrink = true is not supported (invalid_argument).
    virtual void buildTree(const TREE_DEPTH_CONTROL depth_type, const index_type intParam, const bool do_shrink = false);
    int depth() const { return _depth; }
    index_type nNodes() const { return _nnodes; }
    index_type recursiveNNodes() const { return _nnodes; }
    const Node& node(index_type i) const { return _nodes.at(i); }
    const Node& root() const { return _nodes.at(0); }
    std::string infoString() const;
    // Legacy VTK unstructured grid of every node box, depth-first
    // (src/LpmOctree.cpp:245-331).
    virtual void writeToVtk(const std::string& filename, const std::string& desc = "") const;

  protected:
    Tree() {}
    // Reads the GPU node table of the context into _nodes (depth-first ids).
    void importFromGpu(index_type nSources);
    // A GPU context of this tree's own: the device tree, its moments and the
    // order model stay with the object however many trees a thread holds.
    void ownContext();

    int _depth = 0;
    index_type _nnodes = 0;
    std::weak_ptr<Coords> _crds;
    std::vector<Node> _nodes;
    std::vector<index_type> _gpuNode;     // depth-first id -> node id of the device tree
    std::vector<index_type> _gpuSources;  // device source j -> coordinate index (points inside the root box)
    std::shared_ptr<GpuContext> _ctx;
    GpuContext* _gpu = nullptr;
};

// ---------------------------------------------------------------- direct sum
class DirectSum {
  public:
    DirectSum(std::shared_ptr<MeshedParticles> pmesh, const std::shared_ptr<ScalarKernel> kernel,
              const std::string tgtFieldName, const std::string srcFieldName);
    DirectSum(std::shared_ptr<MeshedParticles> pmesh, const std::shared_ptr<VectorKernel> kernel,
              const std::string tgtFieldName, const std::string srcFieldName);

    // Meshfree direct sum over point particles (src/LpmDirectSum.h:22-23):
    // target field = sum_j kernel(x_i, x_j) src_j [weight_j], j != i for a
    // singular kernel.  The GPU evaluates SphereGreensFn.
    DirectSum(std::shared_ptr<Particles> particles, const std::shared_ptr<ScalarKernel> kernel,
              const std::string tgtFieldName, const std::string srcFieldName,
              const std::string srcWeightName = std::string());

    // The solves fill every rank's block [startIndex(r), endIndex(r)], r = 0 ..
    // nProcs - 1, in this process (one GPU holds all of them), so the Fields
    // hold after the solve what the reference's solve + meshBroadcast leave.
    void meshSolve(const MPIReplicatedData& mpiVerts, const MPIReplicatedData& mpiFaces) const;
    void meshfreeSolve(const MPIReplicatedData& mpiParticles) const;
    // Every rank's block is computed in this process: nothing to exchange.
    void meshBroadcast(const MPIReplicatedData&, const MPIReplicatedData&) const {}
    void meshfreeBroadcast(const MPIReplicatedData&) const {}

    // LPM_F64 (default, the reference's pair formula) or LPM_F32C.
    void setPrecision(int prec) { _prec = prec; }

  protected:
    std::weak_ptr<ScalarKernel> _kernel;
    std::weak_ptr<VectorKernel> _vecKernel;
    std::weak_ptr<MeshedParticles> _mesh;
    std::weak_ptr<Field> _vertTgt, _faceSrc, _faceTgt;
    std::weak_ptr<Coords> _srcLocs;
    std::weak_ptr<Field> _srcVals, _tgtVals, _srcWeights;
    bool _useWeights = false;
    int _prec = LPM_F64;
};

// ---------------------------------------------------------------- treecode
// Settings of the learned Taylor order and acceptance of TreeSum (no
// counterpart in the reference).
struct OrderModelConfig {
    scalar_type thetaMax = 0.85;  // pairs passing r_n + r_b < thetaMax |c_b - c_n| go to the model
    scalar_type tau = 3e-7;       // per-pair error bound, in units of sum |gamma| / (4 pi)
    int maxOrder = 8;
    int margin = 0;               // orders added to the model's class
    int epochs = 8;
    int labelStride = 16;         // every labelStride-th target batch is labelled
    scalar_type keepRate = 0.3;   // fraction of its candidate pairs kept
    uint64_t seed = 1;
    // stereographic far field (LPM_OPT_STEREO) in meshSolve: -1 = off, else the
    // extra order of the plane series (the bench uses 1 with stereoOnly,
    // tau 2.5e-6, thetaMax 0.92, stereoRatio 2.0 and leaves of 192); leaves are
    // then priced as stereographic entries
    int stereoExtra = -1;
    scalar_type stereoRatio = 2.0;
    // LPM_OPT_STEREO 2: the stereographic far field alone (Cartesian moments of
    // order 3 for the model's features, pairs no projection takes descend)
    bool stereoOnly = false;
};

class TreeSum : public Tree {
  public:
    // Point-cloud tree (src/LpmTreeSum.h:49-50, src/LpmTreeSum.cpp:212-219,
    // generateTree :240-263): the padded root box of Tree(crds), a node splits
    // iff it holds more than max_particles_per_box points; smoother is the
    // far field's delta.  do_shrink = true is rejected with
    // std::invalid_argument, as by Tree::buildTree.
    TreeSum(const std::shared_ptr<Coords> crds, const int max_series_order, const int max_particles_per_box,
            const scalar_type smoother, const bool do_shrink = false);
    TreeSum(const int max_series_order, const int max_particles_per_box, std::shared_ptr<MeshedParticles> pmesh,
            const std::shared_ptr<VectorKernel> kernel, const std::string tgtFieldName,
            const std::string srcFieldName);
    // Scalar kernel (SphereGreensFn): the streamfunction tree sum the
    // reference declares (initScalarKernel, src/LpmTreeSum.h:55-56) but never
    // implements (computeScalarSum is commented out, :61-65).  meshSolve then
    // fills the scalar target field by lpm_tree_eval_psi: batch acceptance with
    // tolerance mp_tol, log series of order max_series_order (moment sign +1);
    // macCounter counts accepted (target, node) pairs.
    TreeSum(const int max_series_order, const int max_particles_per_box, std::shared_ptr<MeshedParticles> pmesh,
            const std::shared_ptr<ScalarKernel> kernel, const std::string tgtFieldName,
            const std::string srcFieldName);
    void initVectorKernel(VectorKernel* kernel) { vKern = kernel; sKern = nullptr; }
    void initScalarKernel(ScalarKernel* kernel) { sKern = kernel; vKern = nullptr; }

    void computeMeshMoments();
    // Point-cloud moments (src/LpmTreeSum.cpp:173-180, SumNode::calc_moments
    // :139-147): for every node, the sum over the points in its box of
    // s_j y_j h_j^k, |k| < max_series_order, h = c - y (the reference's sign,
    // the default) or y - c after setMomentSign(+1); strength holds one value
    // per coordinate.  Points outside the root box reach no node.
    void computeMoments(const std::shared_ptr<Field> strength);
    // SumNode::moments of node i (depth-first id) after computeMoments or
    // computeMeshMoments: multi-index -> (x, y, z) components, lex order.
    std::map<MultiIndex, std::vector<scalar_type>> nodeMoments(index_type i) const;
    // Fills every rank's block in this process (as DirectSum::meshSolve);
    // macCounter is then the sum over all blocks, the reference's MPI_Reduce
    // total.
    void meshSolve(const MPIReplicatedData& mpiVerts, const MPIReplicatedData& mpiFaces, const scalar_type mp_tol,
                   int& macCounter);
    // Every rank's block is computed in this process: nothing to exchange.
    void meshBroadcast(const MPIReplicatedData&, const MPIReplicatedData&) const {}

    // Defaults reproduce the reference: moment sign -1 (h = centroid - y),
    // the reference series, FP64.  setMomentSign(+1) with
    // setSeries(LPM_SERIES_CARTESIAN) selects the convergent expansion.
    void setMomentSign(int s) { _momentSign = s; _haveMoments = false; }
    void setSeries(int s) { _series = s; }
    void setPrecision(int p) { _prec = p; }
    void setSmoothing(scalar_type d) { _smooth = d; }
    int maxSeriesOrder() const { return _maxSeriesOrder; }

    // Acceptance per batch of <= 64 Morton-sorted targets (lpm_tree_eval_batch):
    // node n is accepted for the whole batch when r_n + r_b < mp_tol |c_b - c_n|;
    // Cartesian series of order max_series_order (moment sign +1 is set).
    // macCounter then counts accepted (target, node) pairs.
    void setBatchAcceptance(bool on) {
        _batch = on;
        if (on && _momentSign != 1) setMomentSign(1);
    }
    // Labels from the FP64 treecode at the leaf faces, then GPU training of
    // the order model; meshSolve uses it afterwards (lpm_tree_eval_learned),
    // mp_tol being ignored.  Returns the final training loss.
    float trainOrderModel(const OrderModelConfig& cfg = OrderModelConfig());
    void setLearned(bool on) { _learned = on; }

  protected:
    int _maxSeriesOrder, _maxParticlesPerNode;
    scalar_type _smooth = 0.0;
    int _momentSign = -1, _series = LPM_SERIES_REFERENCE, _prec = LPM_F64;
    bool _haveMoments = false;
    int _momentOrder = 0;
    bool _batch = false, _learned = false, _trained = false;
    OrderModelConfig _model;
    scalar_type _uref = 0.0;
    void ensureMoments(int order);
    std::weak_ptr<MeshedParticles> _mesh;
    std::weak_ptr<Field> _vertTgt, _faceSrc, _faceTgt;
    VectorKernel* vKern = nullptr;
    ScalarKernel* sKern = nullptr;
    std::vector<index_type> _leafFaces;  // source j -> face index
    std::vector<scalar_type> _moments;   // device-tree moments [node][term][3] (lex), made by nodeMoments
    bool _momentsImported = false;
    void buildFromMesh(const MeshedParticles& pm, int leafMax);
    void scalarSolve(const MPIReplicatedData& mv, const MPIReplicatedData& mf, scalar_type mp_tol, int& macCounter);
};

// ---------------------------------------------------------------- Taylor series (coefficients and moments on the GPU)
class TaylorSeries3d {
  public:
    virtual ~TaylorSeries3d() {}
    scalar_type getCoeff(const MultiIndex& k) const { return coeffs.at(k); }
    scalar_type getMoment(const MultiIndex& k) const { return moments.at(k); }
    scalar_type sum() const;
    std::string infoString() const;
    virtual void computeCoeffs(const XyzVector& tgtVec, const XyzVector& srcVec, const scalar_type param = 0.0) = 0;
    virtual std::unique_ptr<TaylorSeries3d> createEmpty() const = 0;
    void computeMoments(const std::shared_ptr<Coords> crds, const std::vector<index_type>& crdInds,
                        const XyzVector& cntd, const std::shared_ptr<Field> srcStr);
    void computeMoments(const std::shared_ptr<Coords> crds, const std::vector<index_type>& crdInds,
                        const XyzVector& cntd, const std::shared_ptr<Field> srcVals,
                        const std::shared_ptr<Field> srcWeights);
    void computeMoments(const std::shared_ptr<PolyMesh2d> faces, const std::vector<index_type>& crdInds,
                        const XyzVector& cntd, const std::shared_ptr<Field> srcVals);
    int maxOrder() const { return maxP; }

  protected:
    explicit TaylorSeries3d(const int maxSeriesOrder);
    // lpm_series_moments on the gathered points; weightOrder as in lpm_gpu.h
    void momentsOnGpu(const std::vector<scalar_type>& x, const std::vector<scalar_type>& y,
                      const std::vector<scalar_type>& z, const std::vector<scalar_type>& s,
                      const std::vector<scalar_type>* w, int weightOrder, const XyzVector& cntd);
    std::map<MultiIndex, scalar_type> coeffs, moments;
    int maxP;
};

// Coefficients of 1/(1 - x.c - x.h) for |k| <= maxP (src/LpmTaylorSeries3d.cpp:63-117).
class SphereGreensSeries : public TaylorSeries3d {
  public:
    explicit SphereGreensSeries(const int maxSeriesOrder) : TaylorSeries3d(maxSeriesOrder) {}
    // on the GPU (lpm_series_coeffs), the reference's values
    void computeCoeffs(const XyzVector& tgtVec, const XyzVector& srcVec, const scalar_type param = 0.0) override;
    std::unique_ptr<TaylorSeries3d> createEmpty() const override {
        return std::unique_ptr<TaylorSeries3d>(new SphereGreensSeries(maxP));
    }
};

}  // namespace Lpm
