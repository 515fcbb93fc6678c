// Implementation of the host API in LpmSolvers.hpp.
#include "LpmSolvers.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <limits>
#include <new>
#include <sstream>

namespace Lpm {

// ---------------------------------------------------------------- Box3d
void Box3d::finish() {
    centroid = XyzVector(0.5 * (xmax + xmin), 0.5 * (ymax + ymin), 0.5 * (zmax + zmin));
    scalar_type r = 0.0;
    for (const XyzVector& c : corners()) {
        const scalar_type t = distance(centroid, c);
        if (t > r) r = t;
    }
    maxRadius = r;
    const XyzVector fc[6] = {XyzVector(0.5 * (xmin + xmax), ymin, 0.5 * (zmin + zmax)),
                             XyzVector(xmax, 0.5 * (ymin + ymax), 0.5 * (zmin + zmax)),
                             XyzVector(0.5 * (xmin + xmax), ymax, 0.5 * (zmin + zmax)),
                             XyzVector(xmin, 0.5 * (ymin + ymax), 0.5 * (zmin + zmax)),
                             XyzVector(0.5 * (xmin + xmax), 0.5 * (ymin + ymax), zmin),
                             XyzVector(0.5 * (xmin + xmax), 0.5 * (ymin + ymax), zmax)};
    scalar_type m = std::numeric_limits<scalar_type>::max();
    for (const XyzVector& c : fc) {
        const scalar_type t = distance(centroid, c);
        if (t < m) m = t;
    }
    minRadius = m;
}

// corner order (xmin,ymin,zmin), (xmin,ymax,zmin), (xmin,ymin,zmax), ... (src/LpmBox3d.cpp:55-66)
std::vector<XyzVector> Box3d::corners() const {
    return {XyzVector(xmin, ymin, zmin), XyzVector(xmin, ymax, zmin), XyzVector(xmin, ymin, zmax),
            XyzVector(xmin, ymax, zmax), XyzVector(xmax, ymin, zmin), XyzVector(xmax, ymax, zmin),
            XyzVector(xmax, ymin, zmax), XyzVector(xmax, ymax, zmax)};
}

// children c = ix + 2 iy + 4 iz, planes 0.5*(min+max) (src/LpmBox3d.cpp:140-151)
std::vector<Box3d> Box3d::bisectAll() const {
    const scalar_type mx = 0.5 * (xmin + xmax), my = 0.5 * (ymin + ymax), mz = 0.5 * (zmin + zmax);
    std::vector<Box3d> r;
    r.reserve(8);
    for (int c = 0; c < 8; ++c)
        r.emplace_back((c & 1) ? mx : xmin, (c & 1) ? xmax : mx, (c & 2) ? my : ymin, (c & 2) ? ymax : my,
                       (c & 4) ? mz : zmin, (c & 4) ? zmax : mz);
    return r;
}

scalar_type Box3d::longestEdge() const { return std::max(xmax - xmin, std::max(ymax - ymin, zmax - zmin)); }

std::string Box3d::infoString() const {
    std::stringstream ss;
    ss << "x in [" << xmin << ", " << xmax << "], y in [" << ymin << ", " << ymax << "], z in [" << zmin << ", "
       << zmax << "]" << std::endl;
    return ss.str();
}

// ---------------------------------------------------------------- MPIReplicatedData
// start = r floor(n/P), end = (r+1) floor(n/P) - 1, the last rank ends at n-1
// (src/LpmMPIReplicatedData.cpp:12-23).
MPIReplicatedData::MPIReplicatedData(const index_type nItems, const int rank, const int nProcs)
    : _nItems(nItems), _nProcs(nProcs), _rank(rank), _start(nProcs), _end(nProcs), _size(nProcs) {
    if (nProcs < 1 || rank < 0 || rank >= nProcs || nItems < 0)
        throw std::invalid_argument("MPIReplicatedData: bad rank/size");
    const index_type chunk = nItems / nProcs;
    for (int i = 0; i < nProcs; ++i) {
        _start[i] = i * chunk;
        _end[i] = (i + 1) * chunk - 1;
    }
    _end[nProcs - 1] = nItems - 1;
    for (int i = 0; i < nProcs; ++i) _size[i] = _end[i] - _start[i] + 1;
}

// ---------------------------------------------------------------- Coords
scalar_type Coords::minX() const { return *std::min_element(x.begin(), x.end()); }
scalar_type Coords::maxX() const { return *std::max_element(x.begin(), x.end()); }
scalar_type Coords::minY() const { return *std::min_element(y.begin(), y.end()); }
scalar_type Coords::maxY() const { return *std::max_element(y.begin(), y.end()); }
scalar_type Coords::minZ() const { return *std::min_element(z.begin(), z.end()); }
scalar_type Coords::maxZ() const { return *std::max_element(z.begin(), z.end()); }
scalar_type Coords::maxR() const {
    scalar_type r = 0.0;
    for (index_type i = 0; i < n(); ++i) {
        const scalar_type t = square(x[i]) + square(y[i]) + square(z[i]);
        if (t > r) r = t;
    }
    return std::sqrt(r);
}

// ---------------------------------------------------------------- GpuContext
GpuContext::GpuContext(int device) {
    const int st = lpm_create(&_ctx, device);
    if (st != LPM_OK) throw std::runtime_error(std::string("lpm_create: ") + lpm_status_string(st));
}

GpuContext::~GpuContext() { lpm_destroy(_ctx); }

void GpuContext::check(int st, const char* where) const {
    if (st == LPM_OK) return;
    const std::string msg = std::string(where) + ": " + lpm_status_string(st) + ": " + lpm_last_error(_ctx);
    switch (st) {
        case LPM_ERR_ALLOC: throw std::bad_alloc();
        case LPM_ERR_KERNEL:
        case LPM_ERR_STATE: throw std::logic_error(msg);
        case LPM_ERR_INVALID_ARG:
        case LPM_ERR_NULL_PTR: throw std::invalid_argument(msg);
        default: throw std::runtime_error(msg);
    }
}

GpuContext& GpuContext::threadDefault() {
    thread_local std::unique_ptr<GpuContext> ctx;
    if (!ctx) ctx.reset(new GpuContext(0));
    return *ctx;
}

// ---------------------------------------------------------------- Tree
Tree::Tree(const std::shared_ptr<Coords> crds, const scalar_type, const int) : _crds(crds) {
    ownContext();
    _nnodes = 1;
}

void Tree::ownContext() {
    if (!_ctx) _ctx = std::make_shared<GpuContext>(0);
    _gpu = _ctx.get();
}

void Tree::buildTree(const TREE_DEPTH_CONTROL type, const index_type param, const bool do_shrink) {
    if (do_shrink) throw std::invalid_argument("Tree::buildTree: do_shrink is not supported");
    auto c = _crds.lock();
    if (!c) throw std::logic_error("Tree::buildTree: coordinates expired");
    // root box padded relatively by BOX_PADDING_FACTOR (src/LpmOctree.cpp:20-26);
    // for a negative minimum this moves the bound inward, and such extreme
    // points stay in the root's list but in no child, as in the reference.
    double lo[3] = {c->minX() - BOX_PADDING_FACTOR * c->minX(), c->minY() - BOX_PADDING_FACTOR * c->minY(),
                    c->minZ() - BOX_PADDING_FACTOR * c->minZ()};
    double hi[3] = {c->maxX() + BOX_PADDING_FACTOR * c->maxX(), c->maxY() + BOX_PADDING_FACTOR * c->maxY(),
                    c->maxZ() + BOX_PADDING_FACTOR * c->maxZ()};
    const Box3d rootBox(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]);
    std::vector<index_type> inside;
    std::vector<scalar_type> x, y, z;
    for (index_type i = 0; i < c->n(); ++i) {
        if (!rootBox.containsPoint(c->getVec(i))) continue;
        inside.push_back(i);
        x.push_back(c->x[i]); y.push_back(c->y[i]); z.push_back(c->z[i]);
    }
    int control = type == MAX_COORDS_PER_NODE ? LPM_TREE_MAX_PER_NODE : LPM_TREE_MAX_DEPTH_CONTROL;
    int cparam = int(param);
    if (type == MAX_COORDS_PER_NODE && c->n() > param && index_type(inside.size()) <= param) {
        // the root's list counts the points outside it too, so the root splits
        // once while its children (all of them inside) stay leaves
        control = LPM_TREE_MAX_DEPTH_CONTROL;
        cparam = 1;
    }
    int64_t nn = 0;
    int depth = 0;
    _gpu->check(lpm_tree_build_box(_gpu->get(), x.data(), y.data(), z.data(), int64_t(x.size()), lo, hi, control,
                                   cparam, &nn, &depth),
                "Tree::buildTree");
    _gpuSources = inside;
    _nnodes = nn;
    _depth = depth;
    importFromGpu(index_type(x.size()));
    for (Node& n : _nodes)
        for (index_type& j : n.coordsContained) j = inside[j];
    if (index_type(inside.size()) != c->n()) {
        std::vector<index_type>& all = _nodes[0].coordsContained;
        all.resize(c->n());
        for (index_type i = 0; i < c->n(); ++i) all[i] = i;
    }
}

void Tree::importFromGpu(index_type ns) {
    int64_t nn = 0;
    int depth = 0;
    int64_t nl = 0;
    _gpu->check(lpm_tree_info(_gpu->get(), &nn, &depth, &nl), "Tree::importFromGpu");
    std::vector<double> box(6 * nn), cent(4 * nn);
    std::vector<int32_t> level(nn), parent(nn), first(nn), nk(nn), range(2 * nn), perm(std::max<index_type>(ns, 1));
    _gpu->check(lpm_tree_export(_gpu->get(), box.data(), cent.data(), level.data(), parent.data(), first.data(),
                                nk.data(), range.data(), perm.data()),
                "Tree::importFromGpu");
    // depth-first ids, children in octant order
    std::vector<index_type>& order = _gpuNode;
    order.clear();
    std::vector<index_type> dfs_id(nn);
    std::vector<index_type> stack = {0};
    while (!stack.empty()) {
        const index_type i = stack.back();
        stack.pop_back();
        dfs_id[i] = index_type(order.size());
        order.push_back(i);
        for (int k = nk[i] - 1; k >= 0; --k) stack.push_back(first[i] + k);
    }
    _nodes.assign(nn, Node());
    for (index_type d = 0; d < nn; ++d) {
        const index_type i = order[d];
        Node& n = _nodes[d];
        n.box = Box3d(box[6 * i + 0], box[6 * i + 3], box[6 * i + 1], box[6 * i + 4], box[6 * i + 2], box[6 * i + 5]);
        n.level = level[i];
        n.parent = parent[i] < 0 ? -1 : dfs_id[parent[i]];
        for (int k = 0; k < nk[i]; ++k) n.kids.push_back(dfs_id[first[i] + k]);
        n.coordsContained.assign(perm.begin() + range[2 * i], perm.begin() + range[2 * i + 1]);
        std::sort(n.coordsContained.begin(), n.coordsContained.end());
    }
    _nnodes = nn;
    _depth = depth;
}

std::string Tree::infoString() const {
    std::stringstream ss;
    ss << "Tree info:" << std::endl << "\tnNodes = " << _nnodes << std::endl << "\tdepth = " << _depth << std::endl;
    return ss.str();
}

void Tree::writeToVtk(const std::string& filename, const std::string& desc) const {
    std::ofstream fs(filename);
    if (!fs.is_open()) throw std::ios_base::failure("file write error");
    fs << "# vtk DataFile Version 2.0" << std::endl << desc << std::endl << "ASCII" << std::endl;
    fs << "DATASET UNSTRUCTURED_GRID" << std::endl;
    fs << "POINTS " << 8 * _nnodes << " double" << std::endl;
    for (const Node& n : _nodes) {
        const Box3d& b = n.box;
        fs << b.xmin << " " << b.ymin << " " << b.zmin << "\n" << b.xmax << " " << b.ymin << " " << b.zmin << "\n"
           << b.xmax << " " << b.ymax << " " << b.zmin << "\n" << b.xmin << " " << b.ymax << " " << b.zmin << "\n"
           << b.xmin << " " << b.ymin << " " << b.zmax << "\n" << b.xmax << " " << b.ymin << " " << b.zmax << "\n"
           << b.xmax << " " << b.ymax << " " << b.zmax << "\n" << b.xmin << " " << b.ymax << " " << b.zmax << "\n";
    }
    fs << "CELLS " << _nnodes << " " << 9 * _nnodes << std::endl;
    for (index_type i = 0; i < _nnodes; ++i) {
        fs << 8 << " ";
        for (int k = 0; k < 8; ++k) fs << 8 * i + k << " ";
        fs << "\n";
    }
    fs << "CELL_TYPES " << _nnodes << std::endl;
    for (index_type i = 0; i < _nnodes; ++i) fs << "12\n";
    fs << "CELL_DATA " << _nnodes << std::endl << "SCALARS tree_level int 1" << std::endl;
    fs << "LOOKUP_TABLE default" << std::endl;
    for (const Node& n : _nodes) fs << n.level << "\n";
}

// ---------------------------------------------------------------- DirectSum
DirectSum::DirectSum(std::shared_ptr<MeshedParticles> pmesh, const std::shared_ptr<ScalarKernel> kernel,
                     const std::string tgt, const std::string src)
    : _kernel(kernel), _mesh(pmesh), _vertTgt(pmesh->getVertexFieldPtr(tgt)), _faceSrc(pmesh->getFaceFieldPtr(src)),
      _faceTgt(pmesh->getFaceFieldPtr(tgt)) {}

DirectSum::DirectSum(std::shared_ptr<MeshedParticles> pmesh, const std::shared_ptr<VectorKernel> kernel,
                     const std::string tgt, const std::string src)
    : _vecKernel(kernel), _mesh(pmesh), _vertTgt(pmesh->getVertexFieldPtr(tgt)), _faceSrc(pmesh->getFaceFieldPtr(src)),
      _faceTgt(pmesh->getFaceFieldPtr(tgt)) {}

DirectSum::DirectSum(std::shared_ptr<Particles> particles, const std::shared_ptr<ScalarKernel> kernel,
                     const std::string tgt, const std::string src, const std::string weight)
    : _kernel(kernel), _srcLocs(particles->getCoordPtr()), _srcVals(particles->getFieldPtr(src)),
      _tgtVals(particles->getFieldPtr(tgt)), _useWeights(!weight.empty()) {
    if (_useWeights) _srcWeights = particles->getFieldPtr(weight);
    if (_srcVals.expired() || _tgtVals.expired() || (_useWeights && _srcWeights.expired()))
        throw std::invalid_argument("DirectSum: unknown particle field");
}

void Particles::initializeFieldWithFunction(const std::string& name, const AnalyticFunction* fn) {
    auto f = getFieldPtr(name);
    if (!f) throw std::invalid_argument("Particles: unknown field " + name);
    for (index_type i = 0; i < n(); ++i) {
        const XyzVector v = _coords->getVec(i);
        if (f->nDim() == 1) f->replace(i, fn->evaluateScalar(v));
        else f->replace(i, fn->evaluateVector(v));
    }
}

namespace {
// The items of every rank's block, [startIndex(0), endIndex(P - 1)] = all of
// them.  One process runs the solve for all blocks (one GPU holds every
// block), so after it each rank's Field already holds what the reference's
// meshBroadcast would have gathered, and the broadcast has nothing to do.
struct AllBlocks {
    index_type first, last;
    explicit AllBlocks(const MPIReplicatedData& m) : first(m.startIndex(0)), last(m.endIndex(m.getSize() - 1)) {}
};
}  // namespace

// src/LpmDirectSum.cpp:175-254 over every rank's block.  The singular branch
// there stops at endIndex(i) instead of endIndex(myRank); not reproduced.
void DirectSum::meshfreeSolve(const MPIReplicatedData& mp) const {
    auto crd = _srcLocs.lock();
    auto src = _srcVals.lock();
    auto tgt = _tgtVals.lock();
    auto k = _kernel.lock();
    if (!crd || !src || !tgt || !k) throw std::logic_error("DirectSum::meshfreeSolve: particles or fields expired");
    auto gk = std::dynamic_pointer_cast<SphereGreensFn>(k);
    if (!gk) throw std::logic_error("invalid kernel");
    auto w = _useWeights ? _srcWeights.lock() : nullptr;
    if (_useWeights && !w) throw std::logic_error("DirectSum::meshfreeSolve: weight field expired");
    const int64_t ns = int64_t(crd->n());
    std::vector<scalar_type> g(ns);
    for (int64_t j = 0; j < ns; ++j) g[j] = src->getScalar(j) * (w ? w->getScalar(j) : 1.0);
    const AllBlocks blk(mp);
    const index_type b = blk.first, e = blk.last;
    const int64_t nt = std::max<int64_t>(0, int64_t(e - b + 1));
    std::vector<int64_t> skip(nt);
    for (int64_t i = 0; i < nt; ++i) skip[i] = gk->isSingular() ? int64_t(b + i) : -1;
    std::vector<scalar_type> psi(nt);
    GpuContext& gpu = GpuContext::threadDefault();
    gpu.check(lpm_direct_sum(gpu.get(), LPM_K_GREENS_SPHERE, _prec, gk->radius(), 0.0, crd->x.data() + b,
                             crd->y.data() + b, crd->z.data() + b, nt, crd->x.data(), crd->y.data(), crd->z.data(),
                             g.data(), ns, skip.data(), nullptr, nullptr, nullptr, psi.data()),
              "DirectSum::meshfreeSolve");
    for (int64_t i = 0; i < nt; ++i) tgt->replace(b + i, psi[i]);
}

namespace {
// Leaf-face sources (ascending face index) with strength zeta * area and the
// targets of every rank's blocks: all vertices, then all leaf faces.
struct MeshProblem {
    std::vector<scalar_type> sx, sy, sz, g, tx, ty, tz;
    std::vector<int64_t> skip;
    std::vector<index_type> vtx, face;  // target -> vertex / face index
};

MeshProblem materialize(const MeshedParticles& pm, const Field& src, const MPIReplicatedData& mv,
                        const MPIReplicatedData& mf) {
    MeshProblem p;
    std::vector<index_type> leafOf(pm.nFaces(), -1);
    for (index_type j = 0; j < pm.nFaces(); ++j) {
        if (pm.faceIsDivided(j)) continue;
        leafOf[j] = index_type(p.sx.size());
        const XyzVector c = pm.getFaceCentroid(j);
        p.sx.push_back(c.x); p.sy.push_back(c.y); p.sz.push_back(c.z);
        p.g.push_back(src.getScalar(j) * pm.getFaceArea(j));
    }
    const AllBlocks bv(mv), bf(mf);
    for (index_type i = bv.first; i <= bv.last; ++i) {
        const XyzVector v = pm.getVertex(i);
        p.tx.push_back(v.x); p.ty.push_back(v.y); p.tz.push_back(v.z);
        p.skip.push_back(-1);
        p.vtx.push_back(i);
    }
    for (index_type i = bf.first; i <= bf.last; ++i) {
        if (pm.faceIsDivided(i)) continue;
        p.tx.push_back(p.sx[leafOf[i]]); p.ty.push_back(p.sy[leafOf[i]]); p.tz.push_back(p.sz[leafOf[i]]);
        p.skip.push_back(leafOf[i]);
        p.face.push_back(i);
    }
    return p;
}
}  // namespace

void DirectSum::meshSolve(const MPIReplicatedData& mv, const MPIReplicatedData& mf) const {
    auto pm = _mesh.lock();
    auto src = _faceSrc.lock();
    auto vt = _vertTgt.lock();
    auto ft = _faceTgt.lock();
    if (!pm || !src || !vt || !ft) throw std::logic_error("DirectSum::meshSolve: mesh or fields expired");
    GpuContext& gpu = GpuContext::threadDefault();
    MeshProblem p = materialize(*pm, *src, mv, mf);
    const int64_t nt = int64_t(p.tx.size()), ns = int64_t(p.sx.size());
    const index_type nv = index_type(p.vtx.size());
    if (auto k = _kernel.lock()) {
        auto gk = std::dynamic_pointer_cast<SphereGreensFn>(k);
        if (!gk) throw std::logic_error("invalid kernel");
        std::vector<scalar_type> psi(nt);
        gpu.check(lpm_direct_sum(gpu.get(), LPM_K_GREENS_SPHERE, _prec, gk->radius(), 0.0, p.tx.data(), p.ty.data(),
                                 p.tz.data(), nt, p.sx.data(), p.sy.data(), p.sz.data(), p.g.data(), ns,
                                 p.skip.data(), nullptr, nullptr, nullptr, psi.data()),
                  "DirectSum::meshSolve");
        for (index_type i = 0; i < nv; ++i) vt->replace(p.vtx[i], psi[i]);
        for (size_t i = 0; i < p.face.size(); ++i) ft->replace(p.face[i], psi[nv + i]);
    } else if (auto k = _vecKernel.lock()) {
        auto bk = std::dynamic_pointer_cast<BiotSavartSphere>(k);
        if (!bk) throw std::logic_error("invalid kernel");
        std::vector<scalar_type> ux(nt), uy(nt), uz(nt);
        gpu.check(lpm_direct_sum(gpu.get(), LPM_K_BIOT_SAVART_SPHERE, _prec, bk->radius(), bk->delta(), p.tx.data(),
                                 p.ty.data(), p.tz.data(), nt, p.sx.data(), p.sy.data(), p.sz.data(), p.g.data(), ns,
                                 nullptr, ux.data(), uy.data(), uz.data(), nullptr),
                  "DirectSum::meshSolve");
        for (index_type i = 0; i < nv; ++i) vt->replace(p.vtx[i], XyzVector(ux[i], uy[i], uz[i]));
        for (size_t i = 0; i < p.face.size(); ++i)
            ft->replace(p.face[i], XyzVector(ux[nv + i], uy[nv + i], uz[nv + i]));
    } else {
        throw std::logic_error("invalid kernel");
    }
}

// ---------------------------------------------------------------- TreeSum
TreeSum::TreeSum(const int max_series_order, const int max_particles_per_box, std::shared_ptr<MeshedParticles> pmesh,
                 const std::shared_ptr<VectorKernel> kernel, const std::string tgt, const std::string src)
    : _maxSeriesOrder(max_series_order), _maxParticlesPerNode(max_particles_per_box), _mesh(pmesh),
      _vertTgt(pmesh->getVertexFieldPtr(tgt)), _faceSrc(pmesh->getFaceFieldPtr(src)),
      _faceTgt(pmesh->getFaceFieldPtr(tgt)), vKern(kernel.get()) {
    auto bk = dynamic_cast<BiotSavartSphere*>(vKern);
    if (!bk) throw std::logic_error("invalid kernel");
    if (bk->radius() != 1.0) throw std::invalid_argument("TreeSum: the series assumes the unit sphere");
    _smooth = bk->delta();
    buildFromMesh(*pmesh, max_particles_per_box);
}

TreeSum::TreeSum(const int max_series_order, const int max_particles_per_box, std::shared_ptr<MeshedParticles> pmesh,
                 const std::shared_ptr<ScalarKernel> kernel, const std::string tgt, const std::string src)
    : _maxSeriesOrder(max_series_order), _maxParticlesPerNode(max_particles_per_box), _mesh(pmesh),
      _vertTgt(pmesh->getVertexFieldPtr(tgt)), _faceSrc(pmesh->getFaceFieldPtr(src)),
      _faceTgt(pmesh->getFaceFieldPtr(tgt)), sKern(kernel.get()) {
    auto gk = dynamic_cast<SphereGreensFn*>(sKern);
    if (!gk) throw std::logic_error("invalid kernel");
    if (gk->radius() != 1.0) throw std::invalid_argument("TreeSum: the series assumes the unit sphere");
    _momentSign = 1;
    buildFromMesh(*pmesh, max_particles_per_box);
}

TreeSum::TreeSum(const std::shared_ptr<Coords> crds, const int max_series_order, const int max_particles_per_box,
                 const scalar_type smoother, const bool do_shrink)
    : Tree(crds), _maxSeriesOrder(max_series_order), _maxParticlesPerNode(max_particles_per_box), _smooth(smoother) {
    if (max_series_order < 1 || max_series_order > 12)
        throw std::invalid_argument("TreeSum: max_series_order must be in [1, 12]");
    buildTree(MAX_COORDS_PER_NODE, max_particles_per_box, do_shrink);
}

void TreeSum::computeMoments(const std::shared_ptr<Field> strength) {
    auto c = _crds.lock();
    if (!c) throw std::logic_error("TreeSum::computeMoments: coordinates expired (point-cloud tree only)");
    if (!strength || strength->n() < c->n()) throw std::invalid_argument("TreeSum::computeMoments: strength field");
    std::vector<scalar_type> g;
    g.reserve(_gpuSources.size());
    for (index_type i : _gpuSources) g.push_back(strength->getScalar(i));
    _gpu->check(lpm_tree_moments(_gpu->get(), g.data(), _maxSeriesOrder, _momentSign), "TreeSum::computeMoments");
    _haveMoments = true;
    _momentOrder = _maxSeriesOrder;
    _momentsImported = false;
}

std::map<MultiIndex, std::vector<scalar_type>> TreeSum::nodeMoments(index_type i) const {
    if (!_haveMoments) throw std::logic_error("TreeSum::nodeMoments: no moments (computeMoments first)");
    if (i < 0 || i >= _nnodes) throw std::out_of_range("TreeSum::nodeMoments: node id");
    const int p = _momentOrder, nt = p * (p + 1) * (p + 2) / 6;
    auto* self = const_cast<TreeSum*>(this);
    if (!_momentsImported) {
        self->_moments.assign(size_t(_nnodes) * nt * 3, 0.0);
        _gpu->check(lpm_tree_export_moments(_gpu->get(), self->_moments.data()), "TreeSum::nodeMoments");
        self->_momentsImported = true;
    }
    std::map<MultiIndex, std::vector<scalar_type>> out;
    const scalar_type* m = _moments.data() + size_t(_gpuNode[i]) * nt * 3;
    int t = 0;
    for (int a = 0; a < p; ++a)
        for (int b = 0; a + b < p; ++b)
            for (int c = 0; a + b + c < p; ++c, ++t) out[MultiIndex(a, b, c)] = {m[3 * t], m[3 * t + 1], m[3 * t + 2]};
    return out;
}

void TreeSum::buildFromMesh(const MeshedParticles& pm, int leafMax) {
    ownContext();
    std::vector<scalar_type> x, y, z;
    for (index_type i = 0; i < pm.nFaces(); ++i) {
        if (pm.faceIsDivided(i)) continue;
        const XyzVector c = pm.getFaceCentroid(i);
        x.push_back(c.x); y.push_back(c.y); z.push_back(c.z);
        _leafFaces.push_back(i);
    }
    int64_t nn = 0;
    int depth = 0;
    _gpu->check(lpm_tree_build(_gpu->get(), x.data(), y.data(), z.data(), int64_t(x.size()), pm.maxR(), leafMax, &nn,
                               &depth),
                "TreeSum::TreeSum");
    _nnodes = nn;
    _depth = depth;
    importFromGpu(index_type(x.size()));
    // node lists hold source positions among the leaf faces; report face indices
    for (Node& n : _nodes)
        for (index_type& j : n.coordsContained) j = _leafFaces[j];
}

void TreeSum::scalarSolve(const MPIReplicatedData& mv, const MPIReplicatedData& mf, const scalar_type mp_tol,
                          int& macCounter) {
    auto pm = _mesh.lock();
    auto vt = _vertTgt.lock();
    auto ft = _faceTgt.lock();
    auto src = _faceSrc.lock();
    if (!pm || !vt || !ft || !src) throw std::logic_error("TreeSum::meshSolve: mesh or fields expired");
    std::vector<scalar_type> g;
    g.reserve(_leafFaces.size());
    for (index_type f : _leafFaces) g.push_back(src->getScalar(f) * pm->getFaceArea(f));
    _gpu->check(lpm_tree_moments_psi(_gpu->get(), g.data(), _maxSeriesOrder), "TreeSum::meshSolve");
    _haveMoments = false;  // the near-field strengths were reset
    std::vector<scalar_type> tx, ty, tz;
    std::vector<index_type> vtx, face;
    const AllBlocks bv(mv), bf(mf);
    for (index_type i = bv.first; i <= bv.last; ++i) {
        const XyzVector v = pm->getVertex(i);
        tx.push_back(v.x); ty.push_back(v.y); tz.push_back(v.z);
        vtx.push_back(i);
    }
    for (index_type i = bf.first; i <= bf.last; ++i) {
        if (pm->faceIsDivided(i)) continue;
        const XyzVector c = pm->getFaceCentroid(i);
        tx.push_back(c.x); ty.push_back(c.y); tz.push_back(c.z);
        face.push_back(i);
    }
    const int64_t nt = int64_t(tx.size());
    std::vector<scalar_type> psi(nt);
    _gpu->check(lpm_tree_eval_psi(_gpu->get(), _prec, _maxSeriesOrder, mp_tol, tx.data(), ty.data(), tz.data(), nt,
                                  psi.data()),
                "TreeSum::meshSolve");
    int64_t st[LPM_TREE_NSTATS] = {};
    _gpu->check(lpm_tree_stats(_gpu->get(), st), "TreeSum::meshSolve");
    const size_t nv = vtx.size();
    for (size_t i = 0; i < nv; ++i) vt->replace(vtx[i], psi[i]);
    for (size_t i = 0; i < face.size(); ++i) ft->replace(face[i], psi[nv + i]);
    macCounter = int(st[LPM_STAT_FAR]);
}

void TreeSum::computeMeshMoments() { ensureMoments(_maxSeriesOrder); }

void TreeSum::ensureMoments(int order) {
    if (_haveMoments && _momentOrder == order) return;
    auto pm = _mesh.lock();
    auto src = _faceSrc.lock();
    if (!pm || !src) throw std::logic_error("TreeSum::computeMeshMoments: mesh or field expired");
    std::vector<scalar_type> g;
    g.reserve(_leafFaces.size());
    _uref = 0.0;
    for (index_type f : _leafFaces) {
        g.push_back(src->getScalar(f) * pm->getFaceArea(f));
        _uref += std::abs(g.back());
    }
    _uref /= 4.0 * PI;
    const bool stereo = _learned && _trained && _model.stereoExtra >= 0;
    _gpu->check(lpm_set_option(_gpu->get(), LPM_OPT_STEREO, stereo ? (_model.stereoOnly ? 2.0 : 1.0) : 0.0),
                "TreeSum::computeMeshMoments");
    if (stereo) {
        _gpu->check(lpm_set_option(_gpu->get(), LPM_OPT_STEREO_EXTRA_ORDER, double(_model.stereoExtra)),
                    "TreeSum::computeMeshMoments");
        _gpu->check(lpm_set_option(_gpu->get(), LPM_OPT_STEREO_RATIO, _model.stereoRatio),
                    "TreeSum::computeMeshMoments");
    }
    _gpu->check(lpm_tree_moments(_gpu->get(), g.data(), order, _momentSign), "TreeSum::computeMeshMoments");
    _haveMoments = true;
    _momentOrder = order;
    _momentsImported = false;
}

float TreeSum::trainOrderModel(const OrderModelConfig& cfg) {
    auto pm = _mesh.lock();
    if (!pm) throw std::logic_error("TreeSum::trainOrderModel: mesh expired");
    if (cfg.maxOrder < 1 || cfg.maxOrder > 12 || !(cfg.tau > 0.0))
        throw std::invalid_argument("TreeSum::trainOrderModel: maxOrder in 1..12 and tau > 0");
    if (cfg.stereoExtra < -1 || cfg.stereoExtra > 14 || !(cfg.stereoRatio > 0.0 && cfg.stereoRatio <= 4.0))
        throw std::invalid_argument("TreeSum::trainOrderModel: stereoExtra in -1..14, stereoRatio in (0, 4]");
    if (_momentSign != 1) setMomentSign(1);
    ensureMoments(12);
    std::vector<scalar_type> tx, ty, tz;
    for (index_type f : _leafFaces) {
        const XyzVector c = pm->getFaceCentroid(f);
        tx.push_back(c.x); ty.push_back(c.y); tz.push_back(c.z);
    }
    const int64_t cap = int64_t(1) << 21;
    std::vector<float> feat(size_t(cap) * LPM_MODEL_INPUTS), err(size_t(cap) * LPM_LABEL_ORDERS);
    int64_t ns = 0;
    _gpu->check(lpm_tree_labels(_gpu->get(), cfg.thetaMax, _smooth, _uref, cfg.labelStride, cfg.keepRate, 2, cfg.seed,
                                cap, tx.data(), ty.data(), tz.data(), int64_t(tx.size()), feat.data(), err.data(),
                                nullptr, &ns),
                "TreeSum::trainOrderModel");
    if (ns == 0) throw std::runtime_error("TreeSum::trainOrderModel: no candidate pairs (mesh too small?)");
    // class = lowest order meeting tau; 0 = descend, also at a leaf whose
    // direct sum is cheaper than the series (Cartesian: 9 instructions per
    // term, 11 per pair; stereographic: 25 per entry + 4 per complex term, 13
    // per pair)
    std::vector<int32_t> y(ns);
    for (int64_t i = 0; i < ns; ++i) {
        int p = 0;
        for (int q = 1; q <= cfg.maxOrder; ++q)
            if (err[size_t(i) * LPM_LABEL_ORDERS + q - 1] <= cfg.tau) { p = q; break; }
        const float* f = &feat[size_t(i) * LPM_MODEL_INPUTS];
        const double npart = std::exp2(double(f[6]));
        const double far = cfg.stereoExtra >= 0 ? 25.0 + 4.0 * (p + cfg.stereoExtra + 1)
                                                : 9.0 * (p * (p + 1) * (p + 2) / 6);
        const double near = (cfg.stereoExtra >= 0 ? 13.0 : 11.0) * npart;
        if (p > 0 && f[9] > 0.5f && far >= near) p = 0;
        y[i] = p;
    }
    float mean[LPM_MODEL_INPUTS] = {}, inv[LPM_MODEL_INPUTS];
    for (int k = 0; k < LPM_MODEL_INPUTS; ++k) inv[k] = 1.0f;
    for (int k = 0; k < 10; ++k) {
        double s = 0.0, s2 = 0.0;
        for (int64_t i = 0; i < ns; ++i) {
            const double v = feat[size_t(i) * LPM_MODEL_INPUTS + k];
            s += v;
            s2 += v * v;
        }
        const double m = s / double(ns), var = std::max(0.0, s2 / double(ns) - m * m);
        mean[k] = float(m);
        inv[k] = float(1.0 / std::max(std::sqrt(var), 1e-3));
    }
    _gpu->check(lpm_model_init(_gpu->get(), cfg.seed), "TreeSum::trainOrderModel");
    std::vector<float> params(LPM_MODEL_NPARAMS);
    _gpu->check(lpm_model_get_params(_gpu->get(), params.data(), nullptr, nullptr), "TreeSum::trainOrderModel");
    _gpu->check(lpm_model_set_params(_gpu->get(), params.data(), mean, inv), "TreeSum::trainOrderModel");
    float loss = 0.f;
    _gpu->check(lpm_model_train(_gpu->get(), feat.data(), y.data(), ns, cfg.epochs, 4096, 2e-3f, cfg.seed, &loss),
                "TreeSum::trainOrderModel");
    _model = cfg;
    _trained = true;
    _learned = true;
    _haveMoments = false;  // meshSolve makes the moments of its order (and the stereographic ones)
    return loss;
}

void TreeSum::meshSolve(const MPIReplicatedData& mv, const MPIReplicatedData& mf, const scalar_type mp_tol,
                        int& macCounter) {
    if (sKern) return scalarSolve(mv, mf, mp_tol, macCounter);
    if (!vKern) throw std::logic_error("TreeSum::meshSolve: no kernel");
    auto pm = _mesh.lock();
    auto vt = _vertTgt.lock();
    auto ft = _faceTgt.lock();
    if (!pm || !vt || !ft) throw std::logic_error("TreeSum::meshSolve: mesh or fields expired");
    if (_learned && !_trained) throw std::logic_error("TreeSum::meshSolve: trainOrderModel first");
    ensureMoments(_learned ? std::max(_model.maxOrder, 1) : _maxSeriesOrder);
    std::vector<scalar_type> tx, ty, tz;
    std::vector<index_type> vtx, face;
    const AllBlocks bv(mv), bf(mf);
    for (index_type i = bv.first; i <= bv.last; ++i) {
        const XyzVector v = pm->getVertex(i);
        tx.push_back(v.x); ty.push_back(v.y); tz.push_back(v.z);
        vtx.push_back(i);
    }
    for (index_type i = bf.first; i <= bf.last; ++i) {
        if (pm->faceIsDivided(i)) continue;
        const XyzVector c = pm->getFaceCentroid(i);
        tx.push_back(c.x); ty.push_back(c.y); tz.push_back(c.z);
        face.push_back(i);
    }
    const int64_t nt = int64_t(tx.size());
    std::vector<scalar_type> ux(nt), uy(nt), uz(nt);
    int64_t mac = 0;
    if (_learned) {
        _gpu->check(lpm_tree_eval_learned(_gpu->get(), _prec, _model.thetaMax, _model.margin, _model.maxOrder, _uref,
                                          _smooth, tx.data(), ty.data(), tz.data(), nt, ux.data(), uy.data(),
                                          uz.data()),
                    "TreeSum::meshSolve");
    } else if (_batch) {
        _gpu->check(lpm_tree_eval_batch(_gpu->get(), _prec, _maxSeriesOrder, mp_tol, _smooth, tx.data(), ty.data(),
                                        tz.data(), nt, ux.data(), uy.data(), uz.data()),
                    "TreeSum::meshSolve");
    } else {
        _gpu->check(lpm_tree_eval(_gpu->get(), _prec, _series, mp_tol, _smooth, tx.data(), ty.data(), tz.data(), nt,
                                  ux.data(), uy.data(), uz.data(), &mac),
                    "TreeSum::meshSolve");
    }
    if (_learned || _batch) {
        int64_t st[LPM_TREE_NSTATS] = {};
        _gpu->check(lpm_tree_stats(_gpu->get(), st), "TreeSum::meshSolve");
        mac = st[LPM_STAT_FAR];
    }
    const size_t nv = vtx.size();
    for (size_t i = 0; i < nv; ++i) vt->replace(vtx[i], XyzVector(ux[i], uy[i], uz[i]));
    for (size_t i = 0; i < face.size(); ++i) ft->replace(face[i], XyzVector(ux[nv + i], uy[nv + i], uz[nv + i]));
    macCounter = int(mac);
}

// ---------------------------------------------------------------- TaylorSeries3d
TaylorSeries3d::TaylorSeries3d(const int maxSeriesOrder) : maxP(maxSeriesOrder) {
    const scalar_type fill = std::numeric_limits<scalar_type>::max();
    for (int a = 0; a <= maxP; ++a)
        for (int b = 0; a + b <= maxP; ++b)
            for (int c = 0; a + b + c <= maxP; ++c) {
                coeffs.emplace(MultiIndex(a, b, c), fill);
                moments.emplace(MultiIndex(a, b, c), fill);
            }
}

scalar_type TaylorSeries3d::sum() const {
    scalar_type r = 0.0;
    for (const auto& e : coeffs) r += e.second * moments.at(e.first);
    return r;
}

std::string TaylorSeries3d::infoString() const {
    std::stringstream ss;
    ss << "TaylorSeries info:" << std::endl << "\t(multi-index) -- coeff -- moment : " << std::endl;
    for (const auto& e : coeffs) ss << "\tindex " << e.first << " -- " << e.second << " -- " << moments.at(e.first) << "\n";
    return ss.str();
}

// Coefficients on the GPU (lpm_series_coeffs): the loop order and FP64
// operations of src/LpmTaylorSeries3d.cpp:63-117, so the map holds the
// reference's values.  The map iterates in lex order, the ABI's order.
void SphereGreensSeries::computeCoeffs(const XyzVector& t, const XyzVector& s, const scalar_type) {
    std::vector<scalar_type> out(coeffs.size());
    GpuContext& gpu = GpuContext::threadDefault();
    gpu.check(lpm_series_coeffs(gpu.get(), maxP, 1, &t.x, &t.y, &t.z, &s.x, &s.y, &s.z, out.data()),
              "SphereGreensSeries::computeCoeffs");
    size_t i = 0;
    for (auto& e : coeffs) e.second = out[i++];
}

// m_k = sum_i (y_i - c)^k * strength_i on the GPU (lpm_series_moments,
// src/LpmTaylorSeries3d.cpp:119-154), points in the order of inds.
void TaylorSeries3d::momentsOnGpu(const std::vector<scalar_type>& x, const std::vector<scalar_type>& y,
                                  const std::vector<scalar_type>& z, const std::vector<scalar_type>& s,
                                  const std::vector<scalar_type>* w, int weightOrder, const XyzVector& cntd) {
    std::vector<scalar_type> out(moments.size());
    GpuContext& gpu = GpuContext::threadDefault();
    gpu.check(lpm_series_moments(gpu.get(), maxP, int64_t(x.size()), x.data(), y.data(), z.data(), s.data(),
                                 w ? w->data() : nullptr, weightOrder, cntd.x, cntd.y, cntd.z, out.data()),
              "TaylorSeries3d::computeMoments");
    size_t i = 0;
    for (auto& e : moments) e.second = out[i++];
}

void TaylorSeries3d::computeMoments(const std::shared_ptr<Coords> crds, const std::vector<index_type>& inds,
                                    const XyzVector& cntd, const std::shared_ptr<Field> str) {
    std::vector<scalar_type> x, y, z, s;
    for (index_type i : inds) {
        x.push_back(crds->x[i]); y.push_back(crds->y[i]); z.push_back(crds->z[i]);
        s.push_back(str->getScalar(i));
    }
    momentsOnGpu(x, y, z, s, nullptr, 0, cntd);
}

void TaylorSeries3d::computeMoments(const std::shared_ptr<Coords> crds, const std::vector<index_type>& inds,
                                    const XyzVector& cntd, const std::shared_ptr<Field> vals,
                                    const std::shared_ptr<Field> wts) {
    std::vector<scalar_type> x, y, z, s, w;
    for (index_type i : inds) {
        x.push_back(crds->x[i]); y.push_back(crds->y[i]); z.push_back(crds->z[i]);
        s.push_back(vals->getScalar(i));
        w.push_back(wts->getScalar(i));
    }
    momentsOnGpu(x, y, z, s, &w, 0, cntd);
}

void TaylorSeries3d::computeMoments(const std::shared_ptr<PolyMesh2d> faces, const std::vector<index_type>& inds,
                                    const XyzVector& cntd, const std::shared_ptr<Field> vals) {
    std::vector<scalar_type> x, y, z, s, a;
    for (index_type i : inds) {
        const XyzVector c = faces->faceCentroid(i);
        x.push_back(c.x); y.push_back(c.y); z.push_back(c.z);
        s.push_back(vals->getScalar(i));
        a.push_back(faces->faceArea(i));
    }
    momentsOnGpu(x, y, z, s, &a, 1, cntd);
}

}  // namespace Lpm
