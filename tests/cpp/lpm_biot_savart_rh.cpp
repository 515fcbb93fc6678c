// Rossby-Haurwitz wave velocity by direct sum and treecode through the Lpm
// API on the GPU; the asserting counterpart of
// tests/LpmBiotSavartRHWaveTests.cpp:46-148 (icosahedral levels, zeta = Y_5^4,
// exact velocity rossbyHaurwitz54Velocity).  Exit status = failures.
#include "LpmSolvers.hpp"

#include <algorithm>
#include <cstdio>

using namespace Lpm;

static int failures = 0;
#define CHECK(cond)                                                                \
    do {                                                                           \
        if (!(cond)) {                                                             \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                                            \
        }                                                                          \
    } while (0)

static double maxRelErr(const Field& a, const Field& b, const MeshedParticles& pm, bool faces) {
    double num = 0.0, den = 0.0;
    for (index_type i = 0; i < a.n(); ++i) {
        if (faces && pm.faceIsDivided(i)) continue;
        const XyzVector d = a.getVec(i) - b.getVec(i);
        num = std::max(num, d.magnitude());
        den = std::max(den, b.getVec(i).magnitude());
    }
    return num / den;
}

int main() {
    IcosTriSphereSeed seed;
    SphereHarmonic54 y54;
    RossbyHaurwitz54Velocity rh;
    auto kernel = std::make_shared<BiotSavartSphere>();
    double prev = 1e9;
    for (int k = 2; k <= 5; ++k) {
        auto pm = std::make_shared<MeshedParticles>(seed, k, true, 1.0, 0);
        for (const char* n : {"velocity", "treeVelocity", "exact", "treeVelocityRef"}) {
            pm->createVertexField(n, "length/time", 3);
            pm->createFaceField(n, "length/time", 3);
        }
        pm->createVertexField("vorticity", "1/time", 1);
        pm->createFaceField("vorticity", "1/time", 1);
        pm->initializeVertexFieldWithFunction("vorticity", &y54);
        pm->initializeFaceFieldWithFunction("vorticity", &y54);
        pm->initializeVertexFieldWithFunction("exact", &rh);
        pm->initializeFaceFieldWithFunction("exact", &rh);
        MPIReplicatedData mpiVerts(pm->nVertices(), 0, 1);
        MPIReplicatedData mpiFaces(pm->nFaces(), 0, 1);

        DirectSum solver(pm, kernel, "velocity", "vorticity");
        solver.meshSolve(mpiVerts, mpiFaces);
        solver.meshBroadcast(mpiVerts, mpiFaces);
        const double err = maxRelErr(*pm->getFaceFieldPtr("velocity"), *pm->getFaceFieldPtr("exact"), *pm, true);

        // reference defaults: moment sign -1, reference series (parity mode)
        TreeSum refTree(6, 500, pm, kernel, "treeVelocityRef", "vorticity");
        refTree.computeMeshMoments();
        int macRef = 0;
        refTree.meshSolve(mpiVerts, mpiFaces, 0.33, macRef);

        // convergent expansion
        TreeSum tree(8, 64, pm, kernel, "treeVelocity", "vorticity");
        tree.setMomentSign(1);
        tree.setSeries(LPM_SERIES_CARTESIAN);
        tree.computeMeshMoments();
        int mac = 0;
        tree.meshSolve(mpiVerts, mpiFaces, 0.5, mac);
        const double terr = maxRelErr(*pm->getFaceFieldPtr("treeVelocity"), *pm->getFaceFieldPtr("velocity"), *pm, true);
        std::printf("icos level %d: faces %ld, max rel err direct vs exact %.3e, tree(p=8,theta=0.5) vs direct %.3e, "
                    "mac %d, reference-mode mac %d, nodes %ld depth %d\n",
                    k, (long)pm->nLeafFaces(), err, terr, mac, macRef, (long)tree.nNodes(), tree.depth());
        CHECK(err < prev);
        prev = err;
        CHECK(terr < 1e-4);
        if (k >= 4) CHECK(mac > 0);
        // the tree's node lists partition the leaf faces
        index_type covered = 0;
        for (index_type i = 0; i < tree.nNodes(); ++i)
            if (tree.node(i).isLeaf()) covered += tree.node(i).nCoords();
        CHECK(covered == pm->nLeafFaces());
    }
    CHECK(prev < 0.02);
    {
        // batch acceptance and the learned order model through the same API
        CubedSphereSeed cseed;
        auto pm = std::make_shared<MeshedParticles>(cseed, 6, true, 1.0, 0);
        for (const char* n : {"velocity", "batchVelocity", "learnedVelocity", "stereoVelocity", "stereoOnlyVelocity"}) {
            pm->createVertexField(n, "length/time", 3);
            pm->createFaceField(n, "length/time", 3);
        }
        pm->createVertexField("vorticity", "1/time", 1);
        pm->createFaceField("vorticity", "1/time", 1);
        pm->initializeVertexFieldWithFunction("vorticity", &y54);
        pm->initializeFaceFieldWithFunction("vorticity", &y54);
        MPIReplicatedData mpiVerts(pm->nVertices(), 0, 1), mpiFaces(pm->nFaces(), 0, 1);
        DirectSum ds(pm, kernel, "velocity", "vorticity");
        ds.meshSolve(mpiVerts, mpiFaces);
        TreeSum bt(8, 64, pm, kernel, "batchVelocity", "vorticity");
        bt.setBatchAcceptance(true);
        int macB = 0;
        bt.meshSolve(mpiVerts, mpiFaces, 0.5, macB);
        const double berr = maxRelErr(*pm->getFaceFieldPtr("batchVelocity"), *pm->getFaceFieldPtr("velocity"), *pm, true);
        TreeSum lt(8, 64, pm, kernel, "learnedVelocity", "vorticity");
        OrderModelConfig cfg;
        cfg.labelStride = 2;
        cfg.keepRate = 1.0;
        const float loss = lt.trainOrderModel(cfg);
        int macL = 0;
        lt.meshSolve(mpiVerts, mpiFaces, 0.0, macL);
        const double lerr = maxRelErr(*pm->getFaceFieldPtr("learnedVelocity"), *pm->getFaceFieldPtr("velocity"), *pm, true);
        // the same with the stereographic far field
        TreeSum st(8, 64, pm, kernel, "stereoVelocity", "vorticity");
        cfg.stereoExtra = 1;
        st.trainOrderModel(cfg);
        st.setPrecision(LPM_F32C);
        int macS = 0;
        st.meshSolve(mpiVerts, mpiFaces, 0.0, macS);
        const double serr = maxRelErr(*pm->getFaceFieldPtr("stereoVelocity"), *pm->getFaceFieldPtr("velocity"), *pm, true);
        // and with the stereographic far field alone (LPM_OPT_STEREO 2)
        TreeSum so(8, 64, pm, kernel, "stereoOnlyVelocity", "vorticity");
        cfg.stereoOnly = true;
        so.trainOrderModel(cfg);
        so.setPrecision(LPM_F32C);
        int macO = 0;
        so.meshSolve(mpiVerts, mpiFaces, 0.0, macO);
        const double oerr = maxRelErr(*pm->getFaceFieldPtr("stereoOnlyVelocity"), *pm->getFaceFieldPtr("velocity"), *pm, true);
        std::printf("cubed level 6: batch tree vs direct %.3e (far pairs %d), learned vs direct %.3e (far pairs %d, "
                    "training loss %.3f), learned FP32C with the stereographic far field %.3e (far pairs %d), "
                    "stereographic only %.3e (far pairs %d)\n",
                    berr, macB, lerr, macL, loss, serr, macS, oerr, macO);
        CHECK(berr < 1e-4 && macB > 0);
        CHECK(lerr < 1e-3 && macL > 0);
        CHECK(serr < 1e-3 && macS > 0);
        CHECK(oerr < 1e-3 && macO > 0);
        CHECK(loss < 1.5f);
    }
    {
        // several ranks in one process: every rank's solve fills every block,
        // so each rank's Fields equal the one-rank result (the reference's
        // solve + meshBroadcast), and macCounter is the total over all blocks
        auto pm = std::make_shared<MeshedParticles>(seed, 4, true, 1.0, 0);
        for (const char* n : {"u1", "t1", "uP", "tP"}) {
            pm->createVertexField(n, "length/time", 3);
            pm->createFaceField(n, "length/time", 3);
        }
        pm->createVertexField("vorticity", "1/time", 1);
        pm->createFaceField("vorticity", "1/time", 1);
        pm->initializeVertexFieldWithFunction("vorticity", &y54);
        pm->initializeFaceFieldWithFunction("vorticity", &y54);
        MPIReplicatedData v1(pm->nVertices(), 0, 1), f1(pm->nFaces(), 0, 1);
        DirectSum(pm, kernel, "u1", "vorticity").meshSolve(v1, f1);
        TreeSum t1(6, 64, pm, kernel, "t1", "vorticity");
        t1.computeMeshMoments();
        int mac1 = 0;
        t1.meshSolve(v1, f1, 0.5, mac1);
        int same = 0, ranks = 0;
        for (int P : {2, 3}) {
            for (int r = 0; r < P; ++r) {
                MPIReplicatedData vP(pm->nVertices(), r, P), fP(pm->nFaces(), r, P);
                for (const char* f : {"uP", "tP"}) {
                    pm->getVertexFieldPtr(f)->initializeToConstant(0.0);
                    pm->getFaceFieldPtr(f)->initializeToConstant(0.0);
                }
                DirectSum ds(pm, kernel, "uP", "vorticity");
                ds.meshSolve(vP, fP);
                ds.meshBroadcast(vP, fP);
                TreeSum tP(6, 64, pm, kernel, "tP", "vorticity");
                tP.computeMeshMoments();
                int macP = 0;
                tP.meshSolve(vP, fP, 0.5, macP);
                tP.meshBroadcast(vP, fP);
                bool eq = macP == mac1;
                for (const char* a : {"u1", "t1"}) {
                    const char* b = a[0] == 'u' ? "uP" : "tP";
                    for (int c = 0; c < 3; ++c) {
                        const Field &va = *pm->getVertexFieldPtr(a), &vb = *pm->getVertexFieldPtr(b);
                        const Field &fa = *pm->getFaceFieldPtr(a), &fb = *pm->getFaceFieldPtr(b);
                        eq = eq && std::equal(va.getPtrToData(c), va.getPtrToData(c) + va.n(), vb.getPtrToData(c));
                        eq = eq && std::equal(fa.getPtrToData(c), fa.getPtrToData(c) + fa.n(), fb.getPtrToData(c));
                    }
                }
                same += eq;
                ++ranks;
            }
        }
        std::printf("icos level 4, 2 and 3 ranks in one process: %d of %d rank solves equal the one-rank solve "
                    "(mac %d)\n", same, ranks, mac1);
        CHECK(same == ranks && mac1 > 0);
    }
    // a kernel the GPU cannot evaluate is the reference's "invalid kernel"
    struct Other : VectorKernel {
        XyzVector evaluate(const XyzVector&, const XyzVector&) const override { return XyzVector(); }
        bool isSingular() const override { return false; }
    };
    auto pm = std::make_shared<MeshedParticles>(seed, 1);
    pm->createVertexField("u", "", 3);
    pm->createFaceField("u", "", 3);
    pm->createFaceField("z", "", 1);
    bool threw = false;
    try {
        DirectSum bad(pm, std::shared_ptr<VectorKernel>(new Other()), "u", "z");
        bad.meshSolve(MPIReplicatedData(pm->nVertices()), MPIReplicatedData(pm->nFaces()));
    } catch (const std::logic_error&) { threw = true; }
    CHECK(threw);
    std::printf("lpm_biot_savart_rh: %d failure(s)\n", failures);
    return failures;
}
