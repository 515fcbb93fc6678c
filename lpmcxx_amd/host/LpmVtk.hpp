// Legacy-VTK and CSV output of meshed particles: the file layout of
// src/LpmVtkFileIO.cpp:12-239 and MeshedParticles::writeToVtkFile
// (src/LpmMeshedParticles.cpp:131-153), so files open in the same
// post-processing:
//   POLYDATA; POINTS = the vertices; POLYGONS = the leaf faces (vertex
//   indices in edge order); POINT_DATA = vertex fields; CELL_DATA = leaf-face
//   fields; every field is a SCALARS block named "<name>_<units>" with nDim
//   components; 19 significant digits.
// CSV: header "x,y,z,<field columns>" (vector fields as <name>0, <name>1, ...)
// and one row per point.
#pragma once

#include "LpmMesh.hpp"

#include <memory>
#include <ostream>
#include <string>
#include <vector>

namespace Lpm {

class VtkWriter {
  public:
    void writeVTKHeader(std::ostream& os, const std::string& title) const;
    // POINTS block of n points (SoA coordinates)
    void writePointsToVTK(std::ostream& os, const scalar_type* x, const scalar_type* y, const scalar_type* z,
                          index_type n) const;
    // POLYGONS block: the leaf faces of the mesh
    void writeFacesToVTKPolygons(std::ostream& os, const PolyMesh2d& mesh) const;
    void writeVTKPointDataHeader(std::ostream& os, index_type nPoints) const;
    // SCALARS block "lagrangian_coords" / "physical_coords" of 3 components
    void writeCoordsToVTKPointData(std::ostream& os, const scalar_type* x, const scalar_type* y,
                                   const scalar_type* z, index_type n, bool lagrangian) const;
    // SCALARS block of a field defined at every point
    void writeFieldToVTKData(std::ostream& os, const Field& field) const;
    void writeVTKCellDataHeader(std::ostream& os, index_type nCells) const;
    // SCALARS block of a face field, leaf faces only (in face order)
    void writeFaceFieldToVTKCellData(std::ostream& os, const PolyMesh2d& mesh, const Field& field) const;
    void writeFaceAreaToVTKCellData(std::ostream& os, const PolyMesh2d& mesh) const;
    void writePointsAndFieldsToCSV(std::ostream& os, const scalar_type* x, const scalar_type* y,
                                   const scalar_type* z, index_type n,
                                   const std::vector<std::shared_ptr<Field>>& fields) const;
};

}  // namespace Lpm
