#include "LpmVtk.hpp"

#include <fstream>
#include <iomanip>
#include <ios>

namespace Lpm {

namespace {

constexpr int kDigits = 19;  // significant digits of every real written

void write_components(std::ostream& os, const Field& f, index_type i, const char* sep) {
    const XyzVector v = f.getVec(i);
    os << v.x;
    if (f.nDim() > 1) os << sep << v.y;
    if (f.nDim() > 2) os << sep << v.z;
}

void scalars_header(std::ostream& os, const Field& f) {
    os << "SCALARS " << f.name() << "_" << f.units() << " double " << f.nDim() << "\n";
    os << "LOOKUP_TABLE default\n";
}

}  // namespace

void VtkWriter::writeVTKHeader(std::ostream& os, const std::string& title) const {
    os << "# vtk DataFile Version 2.0\n" << title << "\nASCII\nDATASET POLYDATA\n";
}

void VtkWriter::writePointsToVTK(std::ostream& os, const scalar_type* x, const scalar_type* y, const scalar_type* z,
                                 index_type n) const {
    os << "POINTS " << n << " double\n" << std::setprecision(kDigits);
    for (index_type i = 0; i < n; ++i) os << x[i] << " " << y[i] << " " << z[i] << "\n";
}

void VtkWriter::writeFacesToVTKPolygons(std::ostream& os, const PolyMesh2d& mesh) const {
    const int nv = mesh.nEdgesPerFace();
    os << "POLYGONS " << mesh.nLeafFaces() << " " << (nv + 1) * mesh.nLeafFaces() << "\n";
    std::vector<index_type> vi(nv);
    for (index_type f = 0; f < mesh.nFaces(); ++f) {
        if (mesh.faceIsDivided(f)) continue;
        mesh.faceVertices(f, vi.data());
        os << nv;
        for (int k = 0; k < nv; ++k) os << " " << vi[k];
        os << "\n";
    }
}

void VtkWriter::writeVTKPointDataHeader(std::ostream& os, index_type nPoints) const {
    os << "POINT_DATA " << nPoints << "\n";
}

void VtkWriter::writeCoordsToVTKPointData(std::ostream& os, const scalar_type* x, const scalar_type* y,
                                          const scalar_type* z, index_type n, bool lagrangian) const {
    os << "SCALARS " << (lagrangian ? "lagrangian_coords" : "physical_coords") << " double 3\n";
    os << "LOOKUP_TABLE default\n" << std::setprecision(kDigits);
    for (index_type i = 0; i < n; ++i) os << x[i] << " " << y[i] << " " << z[i] << "\n";
}

void VtkWriter::writeFieldToVTKData(std::ostream& os, const Field& field) const {
    scalars_header(os, field);
    os << std::setprecision(kDigits);
    for (index_type i = 0; i < field.n(); ++i) {
        write_components(os, field, i, " ");
        os << "\n";
    }
}

void VtkWriter::writeVTKCellDataHeader(std::ostream& os, index_type nCells) const {
    os << "CELL_DATA " << nCells << "\n";
}

void VtkWriter::writeFaceFieldToVTKCellData(std::ostream& os, const PolyMesh2d& mesh, const Field& field) const {
    scalars_header(os, field);
    os << std::setprecision(kDigits);
    for (index_type f = 0; f < mesh.nFaces(); ++f) {
        if (mesh.faceIsDivided(f)) continue;
        write_components(os, field, f, " ");
        os << "\n";
    }
}

void VtkWriter::writeFaceAreaToVTKCellData(std::ostream& os, const PolyMesh2d& mesh) const {
    os << "SCALARS faceArea double 1\nLOOKUP_TABLE default\n" << std::setprecision(kDigits);
    for (index_type f = 0; f < mesh.nFaces(); ++f)
        if (!mesh.faceIsDivided(f)) os << mesh.faceArea(f) << "\n";
}

void VtkWriter::writePointsAndFieldsToCSV(std::ostream& os, const scalar_type* x, const scalar_type* y,
                                          const scalar_type* z, index_type n,
                                          const std::vector<std::shared_ptr<Field>>& fields) const {
    os << "x,y,z";
    for (const auto& f : fields) {
        if (f->nDim() == 1) os << "," << f->name();
        else
            for (int c = 0; c < f->nDim(); ++c) os << "," << f->name() << c;
    }
    os << "\n" << std::setprecision(kDigits);
    for (index_type i = 0; i < n; ++i) {
        os << x[i] << "," << y[i] << "," << z[i];
        for (const auto& f : fields) {
            os << ",";
            write_components(os, *f, i, ",");
        }
        os << "\n";
    }
}

void MeshedParticles::writeToVtkFile(const std::string& fname, const std::string& desc) const {
    std::ofstream fs(fname);
    if (!fs.is_open()) throw std::ios_base::failure("MeshedParticles::writeToVtkFile: cannot open " + fname);
    const PolyMesh2d& m = *mesh2d;
    VtkWriter w;
    w.writeVTKHeader(fs, desc);
    w.writePointsToVTK(fs, m.vx.data(), m.vy.data(), m.vz.data(), m.nVertices());
    w.writeFacesToVTKPolygons(fs, m);
    w.writeVTKPointDataHeader(fs, m.nVertices());
    for (const auto& kv : _vertexFields) w.writeFieldToVTKData(fs, *kv.second);
    w.writeVTKCellDataHeader(fs, m.nLeafFaces());
    for (const auto& kv : _faceFields) w.writeFaceFieldToVTKCellData(fs, m, *kv.second);
}

}  // namespace Lpm
