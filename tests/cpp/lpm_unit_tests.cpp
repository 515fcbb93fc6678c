// Unit checks of the host Lpm API that need no GPU: multi-indices, boxes,
// replicated-data partition, kernels, mesh fields, VTK / CSV output.
// Mirrors tests/LpmMultiIndexUnitTests.cpp, tests/LpmBox3dUnitTests.cpp and
// tests/LpmMPIReplicatedDataUnitTests.cpp of the reference, with assertions
// (the Taylor series, whose coefficients and moments run on the GPU, are in
// lpm_series_tree.cpp).  Exit status = number of failures.
#include "LpmSolvers.hpp"
#include "LpmVtk.hpp"

#include <fstream>

#include <cstdio>
#include <iostream>

using namespace Lpm;

static int failures = 0;
#define CHECK(cond)                                                         \
    do {                                                                    \
        if (!(cond)) {                                                      \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                                     \
        }                                                                   \
    } while (0)

static double relerr(double a, double b) { return std::abs(a - b) / std::abs(b); }

int main(int argc, char** argv) {
    // MultiIndex known answers (tests/LpmMultiIndexUnitTests.cpp:28-35)
    const MultiIndex k(4, 3, 1);
    CHECK(k.magnitude() == 8);
    CHECK(k.vectorPower(XyzVector(2.0, 3.0, 0.5)) == 216.0);
    CHECK(k.factorial() == 144);
    CHECK(MultiIndex(0, 5, 0) < MultiIndex(1, 0, 0) && MultiIndex(1, 0, 2) < MultiIndex(1, 1, 0));

    // Box3d: padded root, bisection order, half-open containment, radii
    const Box3d root(1.0);
    CHECK(root.xmin == -1.0 - 0.00001 && root.xmax == 1.0 + 0.00001);
    CHECK(root.centroid.x == 0.0 && root.centroid.y == 0.0);
    const std::vector<Box3d> kids = root.bisectAll();
    CHECK(kids.size() == 8);
    CHECK(kids[1].xmin == 0.0 && kids[1].ymax == 0.0 && kids[2].ymin == 0.0 && kids[4].zmin == 0.0);
    CHECK(kids[0].containsPoint(XyzVector(-0.5, -0.5, -0.5)) && !kids[0].containsPoint(XyzVector(0.0, -0.5, -0.5)));
    CHECK(kids[7].containsPoint(XyzVector(0.0, 0.0, 0.0)));
    CHECK(relerr(kids[3].maxRadius, std::sqrt(3.0) * 0.5 * (1.0 + 0.00001)) < 1e-14);
    CHECK(relerr(kids[3].minRadius, 0.5 * (1.0 + 0.00001)) < 1e-14);

    // replicated-data partition (src/LpmMPIReplicatedData.cpp:12-23)
    MPIReplicatedData part(10, 1, 3);
    CHECK(part.startIndex(0) == 0 && part.endIndex(0) == 2);
    CHECK(part.startIndex(2) == 6 && part.endIndex(2) == 9 && part.msgSize(2) == 4);
    MPIReplicatedData one(7, 0, 1);
    CHECK(one.startIndex(0) == 0 && one.endIndex(0) == 6);

    // kernels: closed forms
    const BiotSavartSphere bs;
    const double th = 0.9;
    const XyzVector np(0, 0, 1), src(std::sin(th), 0, std::cos(th));
    const XyzVector kv = bs.evaluate(np, src);
    CHECK(kv.x == 0.0 && relerr(kv.y, std::sin(th) / (1 - std::cos(th))) < 1e-14);
    CHECK(bs.evaluate(np, np).y == 0.0 && bs.isSingular());
    CHECK(relerr(bs.sumMultiplier(), -1.0 / (4.0 * PI)) < 1e-15);
    const SphereGreensFn gf(1.0);
    CHECK(relerr(gf.evaluate(np, src), -std::log(1 - std::cos(th)) / (4.0 * PI)) < 1e-14);

    // mesh fields
    CubedSphereSeed seed;
    auto pm = std::make_shared<MeshedParticles>(seed, 2);
    pm->createFaceField("vorticity", "1/time", 1);
    SphereHarmonic54 y54;
    pm->initializeFaceFieldWithFunction("vorticity", &y54);
    CHECK(pm->getFaceFieldPtr("vorticity")->getScalar(0) == 0.0);  // parent face
    bool threw = false;
    try { pm->getVertexFieldPtr("missing"); } catch (const std::out_of_range&) { threw = true; }
    CHECK(threw);
    double area = 0.0;
    for (index_type i = 0; i < pm->nFaces(); ++i) area += pm->getFaceArea(i);
    CHECK(std::abs(area - 4 * PI) < 1e-12);

    // VTK polydata and CSV output of the mesh (argv[1], argv[2])
    pm->createVertexField("streamFn", "m^2/s", 1);
    pm->initializeVertexFieldWithFunction("streamFn", &y54);
    pm->createFaceField("velocity", "m/s", 3);
    RossbyHaurwitz54Velocity rh;
    pm->initializeFaceFieldWithFunction("velocity", &rh);
    if (argc > 1) pm->writeToVtkFile(argv[1], "cubed sphere level 2");
    if (argc > 2) {
        std::ofstream csv(argv[2]);
        const PolyMesh2d& m = *pm->meshPtr();
        VtkWriter().writePointsAndFieldsToCSV(csv, m.vx.data(), m.vy.data(), m.vz.data(), m.nVertices(),
                                              {pm->getVertexFieldPtr("streamFn")});
    }

    std::printf("lpm_unit_tests: %d failure(s)\n", failures);
    return failures;
}
