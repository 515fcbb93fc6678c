// Barotropic vorticity equation (BVE) on the rotating unit sphere, solved by
// Lagrangian particles: the leaf faces of a spherical mesh carry absolute
// vorticity and area (the vortex particles), the vertices are passive
// particles that follow the same flow.  Absolute vorticity is conserved along
// trajectories, areas are conserved (incompressible flow), and
// zeta_rel = zeta_abs - 2 Omega z gives the circulation whose Biot-Savart
// velocity moves everything.  No remeshing.
//
// The reference C++ has no time stepping; its Fortran code organizes the same
// problem as a SphereBVE mesh type advanced by a SphereBVESolver with RK4
// (README.md:48-81, the BVESingleGaussianVortex driver).  These classes keep
// those names; the velocity and the RK4 step run on the GPU through
// lpm_bve_steps (include/lpm_gpu.h).
#pragma once

#include "LpmMesh.hpp"
#include "LpmSolvers.hpp"

#include <memory>
#include <string>
#include <vector>

namespace Lpm {

// Relative vorticity of one Gaussian vortex, strength * exp(-shape |x - x0|^2),
// centred at (lat0, lon0) in radians.
class GaussianVortexSphere : public AnalyticFunction {
  public:
    GaussianVortexSphere(scalar_type strength, scalar_type shape, scalar_type lat0, scalar_type lon0);
    scalar_type evaluateScalar(const XyzVector& x) const override;

  private:
    scalar_type _strength, _shape;
    XyzVector _x0;
};

class SphereBVE {
  public:
    SphereBVE(const MeshSeed& seed, int maxRecursionLevel, scalar_type omega);

    // zeta_rel = fn at the face centroids and vertices; with removeMean the
    // area-weighted face mean is subtracted from both (zero net circulation,
    // as a velocity on the closed sphere requires).  zeta_abs = zeta_rel + 2 Omega z.
    void initializeRelVorticity(const AnalyticFunction& fn, bool removeMean = true);

    index_type nFaces() const { return _nf; }       // leaf faces (vortex particles)
    index_type nVertices() const { return _nv; }
    scalar_type omega() const { return _omega; }
    scalar_type time() const { return _time; }
    const PolyMesh2d& mesh() const { return *_mesh; }

    XyzVector facePosition(index_type i) const { return XyzVector(_x[i], _y[i], _z[i]); }
    XyzVector vertexPosition(index_type i) const { return XyzVector(_x[_nf + i], _y[_nf + i], _z[_nf + i]); }
    scalar_type faceArea(index_type i) const { return _area[i]; }
    scalar_type faceAbsVorticity(index_type i) const { return _zabs[i]; }
    scalar_type vertexAbsVorticity(index_type i) const { return _zabs[_nf + i]; }
    scalar_type faceRelVorticity(index_type i) const { return _zabs[i] - 2.0 * _omega * _z[i]; }
    scalar_type vertexRelVorticity(index_type i) const { return _zabs[_nf + i] - 2.0 * _omega * _z[_nf + i]; }

    // sum over faces of zeta_rel A (zero for a consistent state) and of
    // zeta_abs^2 A / 2 (absolute enstrophy, invariant)
    scalar_type totalRelCirculation() const;
    scalar_type absoluteEnstrophy() const;

    // Legacy VTK polydata: current vertex positions, leaf faces; point data
    // lagrangian_coords (initial vertex positions) and the vertex vorticities;
    // cell data the face vorticities, face particle positions and areas.
    void writeToVtkFile(const std::string& fname, const std::string& desc = "") const;
    // CSV of the face particles (x, y, z, relVort, absVort, area) or of the
    // vertices (x, y, z, relVort, absVort).
    void writeToCsv(const std::string& fname, bool faces = true) const;

  private:
    friend class SphereBVESolver;
    std::shared_ptr<PolyMesh2d> _mesh;
    index_type _nf = 0, _nv = 0;
    std::vector<index_type> _leaf;                  // mesh face index of every leaf face
    std::vector<scalar_type> _x, _y, _z;            // faces [0, nf), then vertices [nf, nf + nv)
    std::vector<scalar_type> _x0, _y0, _z0;         // initial (Lagrangian) positions, same layout
    std::vector<scalar_type> _zabs;                 // absolute vorticity, same layout
    std::vector<scalar_type> _area;                 // face areas [nf]
    scalar_type _omega = 0.0, _time = 0.0;
};

class SphereBVESolver {
  public:
    // cfg: velocity solver of every RK4 stage (lpm_solver_config of
    // include/lpm_gpu.h: direct sum, treecode, batch or learned treecode).
    SphereBVESolver(std::shared_ptr<SphereBVE> bve, const lpm_solver_config& cfg, scalar_type dt);

    // nsteps RK4 steps of faces and vertices
    void timestep(int nsteps = 1);
    // kinetic energy (1/2) sum_f |u_f|^2 A_f of the face velocities (FP64
    // direct sum), conserved by the BVE
    scalar_type kineticEnergy() const;
    scalar_type dt() const { return _dt; }

  private:
    std::shared_ptr<SphereBVE> _bve;
    lpm_solver_config _cfg;
    scalar_type _dt;
    GpuContext* _gpu;
};

}  // namespace Lpm
