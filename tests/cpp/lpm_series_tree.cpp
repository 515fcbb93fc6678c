// TaylorSeries3d / SphereGreensSeries and the point-cloud TreeSum through the
// Lpm API, on the GPU.  Asserting counterparts of
// tests/LpmTaylorSeries3dUnitTests.cpp and tests/LpmOctreeUnitTests.cpp:35-127:
//   * coefficients = |k|!/k! t^k / D^(|k|+1) (closed form), series sum ->
//     1/(1 - t.y) for a nearby source;
//   * the three computeMoments overloads against a host sum of
//     s (y - c)^k in index order (src/LpmTaylorSeries3d.cpp:119-154);
//   * the point-cloud TreeSum (src/LpmTreeSum.cpp:212-263): node structure
//     against a host restatement of generateTree (root list of every point,
//     split iff count > max, bisectAll children, half-open containment), and
//     every node's moments (src/LpmTreeSum.cpp:139-147,173-180, both signs)
//     against a host sum over the points in its box.
// Exit status = failures.
#include "LpmSolvers.hpp"

#include <cmath>
#include <cstdio>
#include <functional>
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

static double relerr(double a, double b) { return std::abs(a - b) / std::abs(b); }

struct RefNode {  // host restatement of the reference's recursive point-cloud tree
    Box3d box;
    std::vector<index_type> pts;
    std::vector<RefNode> kids;
};

static void refSplit(RefNode& n, const Coords& c, index_type maxPer) {
    if (index_type(n.pts.size()) <= maxPer) return;
    for (const Box3d& b : n.box.bisectAll()) {
        RefNode k;
        k.box = b;
        for (index_type j : n.pts)
            if (b.containsPoint(c.getVec(j))) k.pts.push_back(j);
        if (!k.pts.empty()) n.kids.push_back(k);
    }
    for (RefNode& k : n.kids) refSplit(k, c, maxPer);
}

static void refFlatten(const RefNode& n, std::vector<const RefNode*>& out) {
    out.push_back(&n);
    for (const RefNode& k : n.kids) refFlatten(k, out);
}

int main() {
    // ---- SphereGreensSeries coefficients (lpm_series_coeffs)
    for (int P : {0, 1, 6, 12}) {
        SphereGreensSeries ser(P);
        const XyzVector t(0.3, -0.5, 0.81), s(0.1, 0.2, -0.3);
        ser.computeCoeffs(t, s);
        const double D = 1.0 - t.dotProduct(s);
        double worst = 0.0;
        for (int a = 0; a <= P; ++a)
            for (int b = 0; a + b <= P; ++b)
                for (int c = 0; a + b + c <= P; ++c) {
                    const MultiIndex m(a, b, c);
                    const double exact = double(factorial(a + b + c)) / m.factorial() * std::pow(t.x, a) *
                                         std::pow(t.y, b) * std::pow(t.z, c) / std::pow(D, a + b + c + 1);
                    worst = std::max(worst, relerr(ser.getCoeff(m), exact));
                }
        std::printf("SphereGreensSeries(%d): coefficients vs closed form, max rel err %.2e\n", P, worst);
        CHECK(worst < 1e-13);
        CHECK(ser.createEmpty()->maxOrder() == P);
    }
    {
        SphereGreensSeries ser(10);
        const XyzVector t(0.3, -0.5, 0.81), s(0.1, 0.2, -0.3);
        ser.computeCoeffs(t, s);
        auto crds = std::make_shared<Coords>();
        crds->insert(s + XyzVector(0.01, -0.02, 0.015));
        auto str = std::make_shared<Field>(1, 1, "str");
        str->replace(0, 1.0);
        ser.computeMoments(crds, {0}, s, str);
        const double direct = 1.0 / (1.0 - t.dotProduct(crds->getVec(0)));
        std::printf("series sum vs 1/(1 - t.y): %.2e\n", relerr(ser.sum(), direct));
        CHECK(relerr(ser.sum(), direct) < 1e-12);
    }

    // ---- computeMoments: three overloads against a host sum in index order
    std::mt19937_64 rng(5);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    {
        const int P = 8, n = 300;
        auto crds = std::make_shared<Coords>();
        auto vals = std::make_shared<Field>(n, 1, "v"), wts = std::make_shared<Field>(n, 1, "w");
        for (int i = 0; i < n; ++i) {
            crds->insert(U(rng), U(rng), U(rng));
            vals->replace(i, U(rng));
            wts->replace(i, 1.0 + 0.5 * U(rng));
        }
        std::vector<index_type> inds;
        for (int i = 0; i < n; i += 2) inds.push_back(i);
        const XyzVector cntd(0.05, -0.1, 0.2);
        SphereGreensSeries a(P), b(P);
        a.computeMoments(crds, inds, cntd, vals);
        b.computeMoments(crds, inds, cntd, vals, wts);
        double wa = 0.0, wb = 0.0;
        for (int k1 = 0; k1 <= P; ++k1)
            for (int k2 = 0; k1 + k2 <= P; ++k2)
                for (int k3 = 0; k1 + k2 + k3 <= P; ++k3) {
                    const MultiIndex k(k1, k2, k3);
                    double ma = 0.0, mb = 0.0, sa = 0.0;
                    for (index_type i : inds) {
                        const double vp = k.vectorPower(crds->getVec(i) - cntd);
                        ma += vp * vals->getScalar(i);
                        mb += vp * (vals->getScalar(i) * wts->getScalar(i));
                        sa += std::abs(vp * vals->getScalar(i));
                    }
                    wa = std::max(wa, std::abs(a.getMoment(k) - ma) / sa);
                    wb = std::max(wb, std::abs(b.getMoment(k) - mb) / (2 * sa));
                }
        std::printf("computeMoments (values; values x weights) vs host sums: %.2e %.2e\n", wa, wb);
        CHECK(wa < 1e-14 && wb < 1e-14);
    }
    {
        IcosTriSphereSeed seed;
        auto pm = std::make_shared<MeshedParticles>(seed, 2);
        pm->createFaceField("vorticity", "1/time", 1);
        SphereHarmonic54 y54;
        pm->initializeFaceFieldWithFunction("vorticity", &y54);
        std::vector<index_type> inds;
        for (index_type i = 0; i < pm->nFaces(); ++i)
            if (!pm->faceIsDivided(i)) inds.push_back(i);
        const XyzVector cntd(0.0, 0.0, 0.3);
        SphereGreensSeries f(6);
        f.computeMoments(pm->meshPtr(), inds, cntd, pm->getFaceFieldPtr("vorticity"));
        double wf = 0.0;
        for (int k1 = 0; k1 <= 6; ++k1)
            for (int k2 = 0; k1 + k2 <= 6; ++k2)
                for (int k3 = 0; k1 + k2 + k3 <= 6; ++k3) {
                    const MultiIndex k(k1, k2, k3);
                    double m = 0.0, sa = 0.0;
                    for (index_type i : inds) {
                        const double t = k.vectorPower(pm->getFaceCentroid(i) - cntd) *
                                         pm->getFaceFieldPtr("vorticity")->getScalar(i) * pm->getFaceArea(i);
                        m += t;
                        sa += std::abs(t);
                    }
                    wf = std::max(wf, std::abs(f.getMoment(k) - m) / sa);
                }
        std::printf("computeMoments (faces, values x areas) vs host sum: %.2e\n", wf);
        CHECK(wf < 1e-14);
    }

    // ---- point-cloud TreeSum
    for (const auto& cs : std::vector<std::pair<int, int>>{{4000, 40}, {9, 8}, {600, 1}}) {
        const int n = cs.first, leaf = cs.second, P = 4;
        auto crds = std::make_shared<Coords>();
        auto str = std::make_shared<Field>(n, 1, "strength");
        for (int i = 0; i < n; ++i) {
            crds->insert(U(rng), 0.7 * U(rng), 0.4 * U(rng));
            str->replace(i, U(rng));
        }
        // the reference's root: [min - 1e-5 min, max + 1e-5 max) per axis, every point in its list
        RefNode root;
        root.box = Box3d(crds->minX() - 0.00001 * crds->minX(), crds->maxX() + 0.00001 * crds->maxX(),
                         crds->minY() - 0.00001 * crds->minY(), crds->maxY() + 0.00001 * crds->maxY(),
                         crds->minZ() - 0.00001 * crds->minZ(), crds->maxZ() + 0.00001 * crds->maxZ());
        for (int i = 0; i < n; ++i) root.pts.push_back(i);
        refSplit(root, *crds, leaf);
        std::vector<const RefNode*> ref;
        refFlatten(root, ref);
        bool shrinkRejected = false;
        try {
            TreeSum bad(crds, P, leaf, 0.0, true);
        } catch (const std::invalid_argument&) {
            shrinkRejected = true;
        }
        CHECK(shrinkRejected);
        for (int sign : {-1, 1}) {
            TreeSum ts(crds, P, leaf, 0.0);
            if (sign == 1) ts.setMomentSign(1);
            ts.computeMoments(str);
            bool same = ts.nNodes() == index_type(ref.size());
            double worst = 0.0;
            for (index_type i = 0; same && i < ts.nNodes(); ++i) {
                const Node& nd = ts.node(i);
                const RefNode& rn = *ref[i];
                same = nd.box.xmin == rn.box.xmin && nd.box.ymax == rn.box.ymax && nd.box.zmin == rn.box.zmin &&
                       nd.coordsContained == rn.pts && int(nd.kids.size()) == int(rn.kids.size());
                // moments: points inside the node's box (src/LpmTreeSum.cpp:199-206); the
                // error scale sum |s_j| maxRadius^|k| bounds the terms of the upward
                // translation the device makes them with
                const auto mom = ts.nodeMoments(i);
                for (const auto& e : mom) {
                    double m[3] = {0, 0, 0}, sc = 1e-300;
                    const double rk = std::pow(rn.box.maxRadius, double(e.first.magnitude()));
                    for (index_type j : rn.pts) {
                        const XyzVector y = crds->getVec(j);
                        if (!rn.box.containsPoint(y)) continue;
                        const XyzVector h = sign < 0 ? rn.box.centroid - y : y - rn.box.centroid;
                        const double vp = e.first.vectorPower(h), sj = str->getScalar(j);
                        m[0] += sj * y.x * vp; m[1] += sj * y.y * vp; m[2] += sj * y.z * vp;
                        sc += std::abs(sj) * rk;
                    }
                    for (int c = 0; c < 3; ++c) worst = std::max(worst, std::abs(e.second[c] - m[c]) / sc);
                }
            }
            std::printf("point-cloud TreeSum n=%d leaf=%d sign %+d: %ld nodes, structure %s the reference "
                        "recursion, moments max err %.2e\n", n, leaf, sign, (long)ts.nNodes(),
                        same ? "equals" : "DIFFERS FROM", worst);
            CHECK(same);
            CHECK(worst < 1e-13);
        }
    }
    std::printf("lpm_series_tree: %d failure(s)\n", failures);
    return failures;
}
