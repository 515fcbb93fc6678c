// Implementation of the Lagrangian BVE classes of LpmBVE.hpp.
#include "LpmBVE.hpp"

#include "LpmVtk.hpp"

#include <fstream>
#include <iomanip>
#include <ios>
#include <stdexcept>

namespace Lpm {

GaussianVortexSphere::GaussianVortexSphere(scalar_type strength, scalar_type shape, scalar_type lat0,
                                           scalar_type lon0)
    : _strength(strength), _shape(shape),
      _x0(std::cos(lat0) * std::cos(lon0), std::cos(lat0) * std::sin(lon0), std::sin(lat0)) {}

scalar_type GaussianVortexSphere::evaluateScalar(const XyzVector& x) const {
    return _strength * std::exp(-_shape * (x - _x0).magnitudeSquared());
}

// ---------------------------------------------------------------- SphereBVE
SphereBVE::SphereBVE(const MeshSeed& seed, int maxRecursionLevel, scalar_type omega)
    : _mesh(std::make_shared<PolyMesh2d>(seed, maxRecursionLevel, 1.0)), _omega(omega) {
    const PolyMesh2d& m = *_mesh;
    _nv = m.nVertices();
    for (index_type f = 0; f < m.nFaces(); ++f)
        if (!m.faceIsDivided(f)) _leaf.push_back(f);
    _nf = index_type(_leaf.size());
    const index_type n = _nf + _nv;
    _x.resize(n);
    _y.resize(n);
    _z.resize(n);
    _area.resize(_nf);
    for (index_type i = 0; i < _nf; ++i) {
        const XyzVector c = m.faceCentroid(_leaf[i]);
        _x[i] = c.x;
        _y[i] = c.y;
        _z[i] = c.z;
        _area[i] = m.faceArea(_leaf[i]);
    }
    for (index_type i = 0; i < _nv; ++i) {
        _x[_nf + i] = m.vx[i];
        _y[_nf + i] = m.vy[i];
        _z[_nf + i] = m.vz[i];
    }
    _x0 = _x;
    _y0 = _y;
    _z0 = _z;
    _zabs.assign(n, 0.0);
    for (index_type i = 0; i < n; ++i) _zabs[i] = 2.0 * _omega * _z[i];
}

void SphereBVE::initializeRelVorticity(const AnalyticFunction& fn, bool removeMean) {
    const index_type n = _nf + _nv;
    std::vector<scalar_type> rel(n);
    for (index_type i = 0; i < n; ++i) rel[i] = fn.evaluateScalar(XyzVector(_x[i], _y[i], _z[i]));
    if (removeMean) {
        scalar_type s = 0.0, a = 0.0;
        for (index_type i = 0; i < _nf; ++i) {
            s += rel[i] * _area[i];
            a += _area[i];
        }
        const scalar_type mean = s / a;
        for (auto& r : rel) r -= mean;
    }
    for (index_type i = 0; i < n; ++i) _zabs[i] = rel[i] + 2.0 * _omega * _z[i];
}

scalar_type SphereBVE::totalRelCirculation() const {
    scalar_type s = 0.0;
    for (index_type i = 0; i < _nf; ++i) s += faceRelVorticity(i) * _area[i];
    return s;
}

scalar_type SphereBVE::absoluteEnstrophy() const {
    scalar_type s = 0.0;
    for (index_type i = 0; i < _nf; ++i) s += 0.5 * _zabs[i] * _zabs[i] * _area[i];
    return s;
}

void SphereBVE::writeToVtkFile(const std::string& fname, const std::string& desc) const {
    std::ofstream fs(fname);
    if (!fs.is_open()) throw std::ios_base::failure("SphereBVE::writeToVtkFile: cannot open " + fname);
    const PolyMesh2d& m = *_mesh;
    VtkWriter w;
    w.writeVTKHeader(fs, desc.empty() ? "SphereBVE t = " + std::to_string(_time) : desc);
    w.writePointsToVTK(fs, _x.data() + _nf, _y.data() + _nf, _z.data() + _nf, _nv);
    w.writeFacesToVTKPolygons(fs, m);
    w.writeVTKPointDataHeader(fs, _nv);
    w.writeCoordsToVTKPointData(fs, _x0.data() + _nf, _y0.data() + _nf, _z0.data() + _nf, _nv, true);
    Field vrel(_nv, 1, "relVort", "1/s"), vabs(_nv, 1, "absVort", "1/s");
    for (index_type i = 0; i < _nv; ++i) {
        vrel.replace(i, vertexRelVorticity(i));
        vabs.replace(i, vertexAbsVorticity(i));
    }
    w.writeFieldToVTKData(fs, vabs);
    w.writeFieldToVTKData(fs, vrel);
    // face fields indexed by mesh face (parents stay 0 and are not written)
    Field frel(m.nFaces(), 1, "relVort", "1/s"), fabs_(m.nFaces(), 1, "absVort", "1/s"),
        fpos(m.nFaces(), 3, "facePosition", "m");
    for (index_type i = 0; i < _nf; ++i) {
        frel.replace(_leaf[i], faceRelVorticity(i));
        fabs_.replace(_leaf[i], faceAbsVorticity(i));
        fpos.replace(_leaf[i], facePosition(i));
    }
    w.writeVTKCellDataHeader(fs, _nf);
    w.writeFaceFieldToVTKCellData(fs, m, fabs_);
    w.writeFaceFieldToVTKCellData(fs, m, fpos);
    w.writeFaceFieldToVTKCellData(fs, m, frel);
    w.writeFaceAreaToVTKCellData(fs, m);
}

void SphereBVE::writeToCsv(const std::string& fname, bool faces) const {
    std::ofstream fs(fname);
    if (!fs.is_open()) throw std::ios_base::failure("SphereBVE::writeToCsv: cannot open " + fname);
    const index_type n = faces ? _nf : _nv, o = faces ? 0 : _nf;
    auto rel = std::make_shared<Field>(n, 1, "relVort", "1/s");
    auto abs_ = std::make_shared<Field>(n, 1, "absVort", "1/s");
    auto area = std::make_shared<Field>(n, 1, "area", "m^2");
    for (index_type i = 0; i < n; ++i) {
        rel->replace(i, _zabs[o + i] - 2.0 * _omega * _z[o + i]);
        abs_->replace(i, _zabs[o + i]);
        if (faces) area->replace(i, _area[i]);
    }
    std::vector<std::shared_ptr<Field>> fields = {rel, abs_};
    if (faces) fields.push_back(area);
    VtkWriter().writePointsAndFieldsToCSV(fs, _x.data() + o, _y.data() + o, _z.data() + o, n, fields);
}

// ---------------------------------------------------------------- SphereBVESolver
SphereBVESolver::SphereBVESolver(std::shared_ptr<SphereBVE> bve, const lpm_solver_config& cfg, scalar_type dt)
    : _bve(bve), _cfg(cfg), _dt(dt), _gpu(&GpuContext::threadDefault()) {
    if (!_bve) throw std::invalid_argument("SphereBVESolver: null SphereBVE");
}

void SphereBVESolver::timestep(int nsteps) {
    SphereBVE& b = *_bve;
    _gpu->check(lpm_bve_steps(_gpu->get(), &_cfg, _dt, b._omega, nsteps, b._nf, b._nv, b._x.data(), b._y.data(),
                              b._z.data(), b._zabs.data(), b._area.data()),
                "SphereBVESolver::timestep");
    b._time += nsteps * _dt;
}

scalar_type SphereBVESolver::kineticEnergy() const {
    const SphereBVE& b = *_bve;
    const index_type nf = b._nf;
    std::vector<scalar_type> g(nf), ux(nf), uy(nf), uz(nf);
    std::vector<int64_t> self(nf);
    for (index_type i = 0; i < nf; ++i) {
        g[i] = b.faceRelVorticity(i) * b._area[i];
        self[i] = i;
    }
    _gpu->check(lpm_direct_sum(_gpu->get(), LPM_K_BIOT_SAVART_SPHERE, LPM_F64, 1.0, 0.0, b._x.data(), b._y.data(),
                               b._z.data(), nf, b._x.data(), b._y.data(), b._z.data(), g.data(), nf, self.data(),
                               ux.data(), uy.data(), uz.data(), nullptr),
                "SphereBVESolver::kineticEnergy");
    scalar_type ke = 0.0;
    for (index_type i = 0; i < nf; ++i) ke += 0.5 * (ux[i] * ux[i] + uy[i] * uy[i] + uz[i] * uz[i]) * b._area[i];
    return ke;
}

}  // namespace Lpm
