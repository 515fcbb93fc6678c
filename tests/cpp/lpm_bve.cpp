// Driver of the Lagrangian BVE classes (SphereBVE, SphereBVESolver): one
// Gaussian vortex on the rotating sphere, the problem of the reference's
// Fortran BVESingleGaussianVortex driver (README.md:68-81).  Checks the
// invariants of the BVE and the physics of the vortex drift, compares the
// direct-sum and treecode solvers, and writes VTK / CSV output to the paths
// given on the command line.  Exit status = number of failures.
#include "LpmBVE.hpp"

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

// vorticity-weighted centre of the faces where zeta_rel > half its maximum
static XyzVector vortex_centre(const SphereBVE& b) {
    scalar_type zmax = 0.0;
    for (index_type i = 0; i < b.nFaces(); ++i) zmax = std::max(zmax, b.faceRelVorticity(i));
    XyzVector c;
    for (index_type i = 0; i < b.nFaces(); ++i) {
        const scalar_type w = b.faceRelVorticity(i);
        if (w > 0.5 * zmax) c += b.facePosition(i).scalarMultiply(w * b.faceArea(i));
    }
    c.normalize();
    return c;
}

int main(int argc, char** argv) {
    const scalar_type omega = 2.0 * PI, dt = 0.01;
    const int nsteps = 20;
    const GaussianVortexSphere vortex(4.0 * PI, 16.0, PI / 6.0, 0.0);  // cyclone at 30 N

    auto bve = std::make_shared<SphereBVE>(CubedSphereSeed(), 5, omega);
    bve->initializeRelVorticity(vortex);
    std::cout << "cubed L5: " << bve->nFaces() << " faces, " << bve->nVertices() << " vertices" << std::endl;
    CHECK(bve->nFaces() == 6144 && bve->nVertices() == 6146);
    CHECK(std::abs(bve->totalRelCirculation()) < 1e-12);
    const scalar_type ens0 = bve->absoluteEnstrophy();
    const XyzVector c0 = vortex_centre(*bve);

    lpm_solver_config direct;
    lpm_solver_config_default(&direct);
    SphereBVESolver solver(bve, direct, dt);
    const scalar_type ke0 = solver.kineticEnergy();
    solver.timestep(nsteps);
    const scalar_type ke1 = solver.kineticEnergy();
    const XyzVector c1 = vortex_centre(*bve);
    std::cout << "t = " << bve->time() << ": kinetic energy " << ke0 << " -> " << ke1 << " (rel change "
              << (ke1 - ke0) / ke0 << "), circulation " << bve->totalRelCirculation() << std::endl;
    std::cout << "vortex centre lat/lon " << std::asin(c0.z) * 180 / PI << "/" << atan4(c0.y, c0.x) * 180 / PI
              << " -> " << std::asin(c1.z) * 180 / PI << "/" << atan4(c1.y, c1.x) * 180 / PI << std::endl;
    CHECK(std::abs(bve->time() - nsteps * dt) < 1e-12);
    // energy is conserved by the BVE; its particle quadrature drifts slowly (2e-4 here)
    CHECK(std::abs(ke1 - ke0) < 1e-3 * ke0);
    CHECK(std::abs(bve->absoluteEnstrophy() - ens0) < 1e-13 * ens0);  // carried by the particles
    // sum of zeta_rel A = sum zeta_abs A - 2 Omega sum z A: the last sum is a
    // moving quadrature of the integral of z over the sphere (0)
    CHECK(std::abs(bve->totalRelCirculation()) < 1e-4);
    scalar_type rmax = 0.0;
    for (index_type i = 0; i < bve->nFaces(); ++i) rmax = std::max(rmax, std::abs(bve->facePosition(i).magnitude() - 1.0));
    for (index_type i = 0; i < bve->nVertices(); ++i)
        rmax = std::max(rmax, std::abs(bve->vertexPosition(i).magnitude() - 1.0));
    CHECK(rmax < 1e-14);
    // a northern-hemisphere cyclone drifts poleward and westward (beta drift)
    CHECK(c1.z > c0.z);
    CHECK(atan4(c1.y, c1.x) > PI);  // started at longitude 0: moved west, so now just below 2 pi

    // the FP64 treecode and the FP32C batch treecode follow the direct sum
    lpm_solver_config tree = direct;
    tree.method = LPM_METHOD_TREE;
    tree.order = 10;
    tree.theta = 0.4;
    tree.leaf_max = 64;
    lpm_solver_config batch = tree;
    batch.method = LPM_METHOD_TREE_BATCH;
    batch.prec = LPM_F32C;
    for (const lpm_solver_config* cfg : {&tree, &batch}) {
        auto b2 = std::make_shared<SphereBVE>(CubedSphereSeed(), 5, omega);
        b2->initializeRelVorticity(vortex);
        SphereBVESolver(b2, *cfg, dt).timestep(nsteps);
        scalar_type d = 0.0;
        for (index_type i = 0; i < bve->nFaces(); ++i)
            d = std::max(d, (b2->facePosition(i) - bve->facePosition(i)).magnitude());
        for (index_type i = 0; i < bve->nVertices(); ++i)
            d = std::max(d, (b2->vertexPosition(i) - bve->vertexPosition(i)).magnitude());
        std::cout << "method " << cfg->method << " prec " << cfg->prec << ": max position difference to direct "
                  << d << std::endl;
        CHECK(d < 1e-6);
    }
    if (argc > 1) bve->writeToVtkFile(argv[1]);
    if (argc > 2) bve->writeToCsv(argv[2], true);
    if (argc > 3) bve->writeToCsv(argv[3], false);
    std::cout << failures << " failure(s)" << std::endl;
    return failures;
}
