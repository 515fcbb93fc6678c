/*
 * CPU ORACLE -- test infrastructure only.
 *
 * A plain-C FP64 restatement of the reference algorithms on the spherical
 * Biot-Savart / Poisson path.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline; the product never calls it.
 *
 * Each function cites the reference lines it follows.  Built with gcc
 * -ffp-contract=off and without -ffast-math, so every expression rounds in the
 * reference's operation order.  Outer loops over independent targets may run
 * under OpenMP; each target's sum is sequential in ascending source order as
 * in the reference, so results do not depend on the thread count.
 *
 * Pins (tests/test_oracle_*.py): closed-form kernels on hand-picked pairs,
 * long-double brute force on small sets, the analytic Rossby-Haurwitz
 * velocity and psi = zeta/30 on refined meshes, series closed forms, octree
 * brute-force descent, treecode -> direct sum as theta -> 0.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static const double kPi = 3.1415926535897932384626433832795027975;
static const double kZeroTol = 1.0e-13; /* src/LpmTypeDefs.h:35 */

/*
 * BiotSavartSphere::evaluate (src/LpmVectorKernel.h:24-28) summed as in
 * DirectSum::meshSolve's vector branch (src/LpmDirectSum.cpp:104-118,120-147):
 * u_i = mult * sum_j [(x_i x y_j) * (1/den)] * gamma_j, j ascending,
 * den = R*R - x_i.y_j + delta*delta, 0 if |den| <= 1e-13,
 * mult = -1/(4 pi R) applied after the sum (src/LpmVectorKernel.h:30-32).
 * skip: per-target source index left out (face self pair), -1 none, NULL none.
 */
void oracle_bs_direct(const double* tx, const double* ty, const double* tz, int64_t nt, const double* sx,
                      const double* sy, const double* sz, const double* gamma, int64_t ns, const int64_t* skip,
                      double radius, double delta, double* ux, double* uy, double* uz) {
    const double mult = -1.0 / (4.0 * kPi * radius);
    const double R2 = radius * radius, d2 = delta * delta;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < nt; ++i) {
        const double x = tx[i], y = ty[i], z = tz[i];
        const int64_t sk = skip ? skip[i] : -1;
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int64_t j = 0; j < ns; ++j) {
            if (j == sk) continue;
            const double cx = y * sz[j] - z * sy[j];
            const double cy = z * sx[j] - x * sz[j];
            const double cz = x * sy[j] - y * sx[j];
            const double den = R2 - (x * sx[j] + y * sy[j] + z * sz[j]) + d2;
            if (fabs(den) > kZeroTol) {
                const double inv = 1.0 / den;
                ax += (cx * inv) * gamma[j];
                ay += (cy * inv) * gamma[j];
                az += (cz * inv) * gamma[j];
            } else {
                ax += 0.0;
                ay += 0.0;
                az += 0.0;
            }
        }
        ux[i] = ax * mult;
        uy[i] = ay * mult;
        uz[i] = az * mult;
    }
}

/*
 * SphereGreensFn::evaluate (src/LpmScalarKernel.h:35-38) summed as in the
 * scalar branch of DirectSum::meshSolve (src/LpmDirectSum.cpp:37-75):
 * psi_i = sum_{j != skip_i} (-log(R*R - y_j.x_i) / (4 pi R)) * gamma_j.
 */
void oracle_greens_direct(const double* tx, const double* ty, const double* tz, int64_t nt, const double* sx,
                          const double* sy, const double* sz, const double* gamma, int64_t ns,
                          const int64_t* skip, double radius, double* psi) {
    const double R2 = radius * radius;
    const double den = 4.0 * kPi * radius;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < nt; ++i) {
        const double x = tx[i], y = ty[i], z = tz[i];
        const int64_t sk = skip ? skip[i] : -1;
        double a = 0.0;
        for (int64_t j = 0; j < ns; ++j) {
            if (j == sk) continue;
            const double inarg = R2 - (sx[j] * x + sy[j] * y + sz[j] * z);
            const double kval = -log(inarg) / den;
            a += kval * gamma[j];
        }
        psi[i] = a;
    }
}

/* Number of OpenMP threads the oracle runs with (for the CPU baseline record). */
int oracle_num_threads(void) {
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}
