// Streamfunction by direct sum through the Lpm API: psi = zeta/30 for
// zeta = Y_5^4 on cubed-sphere and icosahedral meshes; the asserting
// counterpart of tests/LpmPoissonSolverTests.cpp:254-401.  Also runs the
// point-cloud Tree of tests/LpmOctreeUnitTests.cpp:35-127 and writes its VTK.
#include "LpmSolvers.hpp"

#include <cstdio>
#include <random>

using namespace Lpm;

static int failures = 0;
#define CHECK(cond)                                                                \
    do {                                                                           \
        if (!(cond)) {                                                             \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                                            \
        }                                                                          \
    } while (0)

int main(int argc, char** argv) {
    SphereHarmonic54 y54;
    auto kernel = std::make_shared<SphereGreensFn>(1.0);
    CubedSphereSeed cubed;
    IcosTriSphereSeed icos;
    for (const MeshSeed* seed : {static_cast<const MeshSeed*>(&cubed), static_cast<const MeshSeed*>(&icos)}) {
        double prev = 1e9;
        for (int k = 2; k <= 5; ++k) {
            auto pm = std::make_shared<MeshedParticles>(*seed, k);
            pm->createVertexField("psi", "", 1);
            pm->createFaceField("psi", "", 1);
            pm->createFaceField("vorticity", "", 1);
            pm->initializeFaceFieldWithFunction("vorticity", &y54);
            DirectSum solver(pm, kernel, "psi", "vorticity");
            solver.meshSolve(MPIReplicatedData(pm->nVertices()), MPIReplicatedData(pm->nFaces()));
            double num = 0.0, den = 0.0;
            auto psi = pm->getFaceFieldPtr("psi");
            for (index_type i = 0; i < pm->nFaces(); ++i) {
                if (pm->faceIsDivided(i)) continue;
                const double ex = y54.evaluateScalar(pm->getFaceCentroid(i)) / 30.0;
                num = std::max(num, std::abs(psi->getScalar(i) - ex));
                den = std::max(den, std::abs(ex));
            }
            std::printf("%s level %d: psi max rel err %.3e\n", seed->idString().c_str(), k, num / den);
            CHECK(num / den < prev);
            prev = num / den;
            if (k == 4) {
                // meshfree sum over the leaf-face particles weighted by area equals the mesh sum at the faces
                auto crd = std::make_shared<Coords>();
                std::vector<index_type> ids;
                for (index_type i = 0; i < pm->nFaces(); ++i) {
                    if (pm->faceIsDivided(i)) continue;
                    crd->insert(pm->getFaceCentroid(i));
                    ids.push_back(i);
                }
                auto parts = std::make_shared<Particles>(crd);
                parts->createField("zeta", "", 1);
                parts->createField("area", "", 1);
                parts->createField("psi", "", 1);
                parts->initializeFieldWithFunction("zeta", &y54);
                for (size_t j = 0; j < ids.size(); ++j)
                    parts->getFieldPtr("area")->replace(index_type(j), pm->getFaceArea(ids[j]));
                DirectSum mf(parts, kernel, "psi", "zeta", "area");
                const MPIReplicatedData blocks(parts->n(), 0, 1);
                mf.meshfreeSolve(blocks);
                mf.meshfreeBroadcast(blocks);
                double d = 0.0, m = 0.0;
                for (size_t j = 0; j < ids.size(); ++j) {
                    d = std::max(d, std::abs(parts->getFieldPtr("psi")->getScalar(index_type(j)) -
                                             psi->getScalar(ids[j])));
                    m = std::max(m, std::abs(psi->getScalar(ids[j])));
                }
                std::printf("  meshfree vs mesh psi: max rel diff %.3e\n", d / m);
                CHECK(d / m < 1e-12);
                bool threw = false;
                try {
                    DirectSum bad(parts, kernel, "psi", "nosuchfield");
                } catch (const std::invalid_argument&) { threw = true; }
                CHECK(threw);
                // streamfunction treecode (TreeSum with the scalar kernel) against the direct sum
                pm->createVertexField("psiTree", "", 1);
                pm->createFaceField("psiTree", "", 1);
                TreeSum tree(10, 64, pm, kernel, "psiTree", "vorticity");
                int mac = 0;
                tree.meshSolve(MPIReplicatedData(pm->nVertices()), MPIReplicatedData(pm->nFaces()), 0.5, mac);
                double td = 0.0, tm = 0.0;
                auto pt = pm->getFaceFieldPtr("psiTree");
                auto vt = pm->getVertexFieldPtr("psiTree"), vd = pm->getVertexFieldPtr("psi");
                for (index_type i = 0; i < pm->nFaces(); ++i) {
                    if (pm->faceIsDivided(i)) continue;
                    td = std::max(td, std::abs(pt->getScalar(i) - psi->getScalar(i)));
                    tm = std::max(tm, std::abs(psi->getScalar(i)));
                }
                for (index_type i = 0; i < pm->nVertices(); ++i) {
                    td = std::max(td, std::abs(vt->getScalar(i) - vd->getScalar(i)));
                    tm = std::max(tm, std::abs(vd->getScalar(i)));
                }
                std::printf("  psi treecode (p 10, theta 0.5) vs direct: max rel diff %.3e, %d far pairs\n", td / tm,
                            mac);
                CHECK(td / tm < 1e-5 && mac > 0);
            }
        }
        CHECK(prev < 0.06);
    }
    // standalone octree on random points (leaf <= 10), as the reference's octree test
    auto crds = std::make_shared<Coords>(8000);
    std::mt19937_64 gen(1234);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    for (int i = 0; i < 8000; ++i) crds->insert(U(gen), U(gen), U(gen));
    Tree tree(crds);
    tree.buildTree(Tree::MAX_COORDS_PER_NODE, 10);
    index_type covered = 0;
    for (index_type i = 0; i < tree.nNodes(); ++i) {
        const Node& n = tree.node(i);
        if (!n.isLeaf()) continue;
        CHECK(n.nCoords() <= 10);
        for (index_type j : n.coordsContained) CHECK(n.box.containsPoint(crds->getVec(j)));
        covered += n.nCoords();
    }
    // points at the extremes can fall outside the relatively padded root box;
    // they stay in the root's list only
    CHECK(covered <= 8000 && covered > 7990 && tree.root().nCoords() == 8000);
    std::printf("octree: %ld nodes, depth %d, %ld of 8000 points in leaves\n", (long)tree.nNodes(), tree.depth(),
                (long)covered);
    if (argc > 1) tree.writeToVtk(argv[1], "octree of 8000 random points");
    Tree byDepth(crds);
    byDepth.buildTree(Tree::MAX_DEPTH, 4);
    CHECK(byDepth.depth() == 4);
    std::printf("lpm_poisson: %d failure(s)\n", failures);
    return failures;
}
