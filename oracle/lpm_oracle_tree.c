/*
 * CPU ORACLE -- test infrastructure only (see lpm_oracle.c for the rules).
 *
 * FP64 restatement of the reference treecode on the mesh path:
 *   Box3d geometry      src/LpmBox3d.h:15-39,88, src/LpmBox3d.cpp:32-66,140-151
 *   tree generation     src/LpmTreeSum.cpp:221-238 (root), :266-290 (recursion)
 *   multi-indices       src/LpmTreeSum.cpp:7-20 (|k| < p), src/LpmMultiIndex.h:25-43
 *   moments             src/LpmTreeSum.cpp:139-147,182-197 (sign knob: the
 *                       reference uses h = c - y; moment_sign=+1 uses y - c)
 *   coefficients        src/LpmTreeSum.cpp:149-171 (lex-order last-writer recurrence)
 *   traversal / MAC     src/LpmTreeSum.cpp:22-24,55-86,106-137
 * One tree lives in module state (single-threaded build; evaluation runs
 * OpenMP over targets, each target's traversal sequential as in the reference).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static const double kPi = 3.1415926535897932384626433832795027975;
static const double kZeroTol = 1.0e-13;
static const double kBoxPad = 0.00001; /* BOX_PADDING_FACTOR, src/LpmBox3d.h:8 */

typedef struct {
    double lo[3], hi[3], c[3], max_radius;
    int level;
    int64_t parent, first_kid, n_kids; /* kids are consecutive node ids */
    int64_t begin, end;                /* range in the permutation */
} onode;

static onode* g_nodes = NULL;
static int64_t g_nn = 0, g_cap = 0;
static int64_t* g_perm = NULL; /* particle indices, each node a contiguous range */
static int64_t g_np = 0;
static int g_depth = 0;

static int64_t new_node(void) {
    if (g_nn == g_cap) {
        g_cap = g_cap ? 2 * g_cap : 1024;
        g_nodes = (onode*)realloc(g_nodes, sizeof(onode) * g_cap);
    }
    memset(&g_nodes[g_nn], 0, sizeof(onode));
    return g_nn++;
}

/* Box3d(x0,x1,y0,y1,z0,z1): centroid 0.5*(max+min); maxRadius = max over the
 * 8 corners (src/LpmBox3d.cpp:55-66 order) of distance(centroid, corner). */
static void box_finish(onode* n) {
    for (int d = 0; d < 3; ++d) n->c[d] = 0.5 * (n->hi[d] + n->lo[d]);
    const double cx[8] = {n->lo[0], n->lo[0], n->lo[0], n->lo[0], n->hi[0], n->hi[0], n->hi[0], n->hi[0]};
    const double cy[8] = {n->lo[1], n->hi[1], n->lo[1], n->hi[1], n->lo[1], n->hi[1], n->lo[1], n->hi[1]};
    const double cz[8] = {n->lo[2], n->lo[2], n->hi[2], n->hi[2], n->lo[2], n->lo[2], n->hi[2], n->hi[2]};
    double r = 0.0;
    for (int i = 0; i < 8; ++i) {
        const double t = sqrt((cx[i] - n->c[0]) * (cx[i] - n->c[0]) + (cy[i] - n->c[1]) * (cy[i] - n->c[1]) +
                              (cz[i] - n->c[2]) * (cz[i] - n->c[2]));
        if (t > r) r = t;
    }
    n->max_radius = r;
}

static const double* g_x;
static const double* g_y;
static const double* g_z;
static int g_leaf_max;
static int g_control; /* 0: split iff count > g_leaf_max; 1: split iff level < g_leaf_max */
static int g_fail;

/* Recursive split of node `id` whose particles are g_perm[begin, end) in
 * ascending index order (src/LpmTreeSum.cpp:266-290). */
static void split(int64_t id, int64_t* scratch) {
    onode n = g_nodes[id];
    if (g_control == 0 && n.end - n.begin <= g_leaf_max) return;
    if (g_control == 1 && (n.level >= g_leaf_max || n.end == n.begin)) return;
    if (n.level >= 60) { g_fail = 1; return; }
    double mid[3];
    for (int d = 0; d < 3; ++d) mid[d] = 0.5 * (n.lo[d] + n.hi[d]);
    int64_t cnt[8] = {0}, off[9];
    /* child c = ix + 2 iy + 4 iz; the half-open containment picks the high half iff v >= mid */
    for (int64_t i = n.begin; i < n.end; ++i) {
        const int64_t j = g_perm[i];
        const int c = (g_x[j] >= mid[0]) | ((g_y[j] >= mid[1]) << 1) | ((g_z[j] >= mid[2]) << 2);
        cnt[c]++;
    }
    off[0] = n.begin;
    for (int c = 0; c < 8; ++c) off[c + 1] = off[c] + cnt[c];
    int64_t pos[8];
    for (int c = 0; c < 8; ++c) pos[c] = off[c];
    for (int64_t i = n.begin; i < n.end; ++i) {
        const int64_t j = g_perm[i];
        const int c = (g_x[j] >= mid[0]) | ((g_y[j] >= mid[1]) << 1) | ((g_z[j] >= mid[2]) << 2);
        scratch[pos[c]++] = j;
    }
    memcpy(g_perm + n.begin, scratch + n.begin, sizeof(int64_t) * (n.end - n.begin));
    int64_t first = -1, nk = 0;
    for (int c = 0; c < 8; ++c) {
        if (cnt[c] == 0) continue;
        const int64_t k = new_node();
        onode* kn = &g_nodes[k];
        for (int d = 0; d < 3; ++d) {
            const int hi = (c >> d) & 1;
            kn->lo[d] = hi ? mid[d] : g_nodes[id].lo[d];
            kn->hi[d] = hi ? g_nodes[id].hi[d] : mid[d];
        }
        box_finish(kn);
        kn->level = g_nodes[id].level + 1;
        kn->parent = id;
        kn->begin = off[c];
        kn->end = off[c + 1];
        kn->first_kid = -1;
        if (first < 0) first = k;
        ++nk;
        if (kn->level > g_depth) g_depth = kn->level;
    }
    g_nodes[id].first_kid = first;
    g_nodes[id].n_kids = nk;
    for (int64_t k = first; k < first + nk; ++k) split(k, scratch);
}

void oracle_tree_free(void) {
    free(g_nodes);
    free(g_perm);
    g_nodes = NULL;
    g_perm = NULL;
    g_nn = g_cap = g_np = 0;
    g_depth = 0;
}

/* General root box [lo, hi) per axis with depth control (Tree::buildTree,
 * src/LpmOctree.cpp:73-180; control 0 = MAX_COORDS_PER_NODE, 1 = MAX_DEPTH).
 * Returns the node count, or -1 if a particle lies outside the root box or the
 * recursion does not terminate (more than `param` coincident particles). */
int64_t oracle_tree_build_box(const double* x, const double* y, const double* z, int64_t n, const double* lo,
                              const double* hi, int control, int param) {
    oracle_tree_free();
    g_x = x; g_y = y; g_z = z;
    g_leaf_max = param;
    g_control = control;
    g_fail = 0;
    g_np = n;
    g_perm = (int64_t*)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    int64_t* scratch = (int64_t*)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    const int64_t r = new_node();
    onode* rn = &g_nodes[r];
    for (int d = 0; d < 3; ++d) { rn->lo[d] = lo[d]; rn->hi[d] = hi[d]; }
    box_finish(rn);
    rn->parent = -1;
    rn->first_kid = -1;
    rn->begin = 0;
    rn->end = n;
    for (int64_t i = 0; i < n; ++i) {
        g_perm[i] = i;
        if (!(lo[0] <= x[i] && x[i] < hi[0] && lo[1] <= y[i] && y[i] < hi[1] && lo[2] <= z[i] && z[i] < hi[2]))
            g_fail = 1;
    }
    if (!g_fail) split(r, scratch);
    free(scratch);
    return g_fail ? -1 : g_nn;
}

/* Mesh tree: root Box3d(maxR, pad=true) = [-maxR - 1e-5, maxR + 1e-5]^3
 * (src/LpmBox3d.h:21-33), split iff count > leaf_max. */
int64_t oracle_tree_build(const double* x, const double* y, const double* z, int64_t n, double max_r, int leaf_max,
                          int unused) {
    (void)unused;
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = -max_r;
        lo[d] -= kBoxPad;
        hi[d] = max_r;
        hi[d] += kBoxPad;
    }
    return oracle_tree_build_box(x, y, z, n, lo, hi, 0, leaf_max);
}

/* Export the node table (any pointer may be NULL). */
void oracle_tree_get(double* box /* [nn][8]: lo3 hi3 ... */, double* cent /* [nn][4]: c3 maxR */, int64_t* level,
                     int64_t* parent, int64_t* first_kid, int64_t* n_kids, int64_t* begin, int64_t* end,
                     int64_t* perm) {
    for (int64_t i = 0; i < g_nn; ++i) {
        const onode* n = &g_nodes[i];
        if (box) for (int d = 0; d < 3; ++d) { box[8 * i + d] = n->lo[d]; box[8 * i + 3 + d] = n->hi[d]; }
        if (cent) { for (int d = 0; d < 3; ++d) cent[4 * i + d] = n->c[d]; cent[4 * i + 3] = n->max_radius; }
        if (level) level[i] = n->level;
        if (parent) parent[i] = n->parent;
        if (first_kid) first_kid[i] = n->first_kid;
        if (n_kids) n_kids[i] = n->n_kids;
        if (begin) begin[i] = n->begin;
        if (end) end[i] = n->end;
    }
    if (perm) memcpy(perm, g_perm, sizeof(int64_t) * g_np);
}

/* ---------------------------------------------------------------- series */
/* Number of multi-indices with |k| < p, lex order (k1, k2, k3) ascending. */
static int nterms(int p) { return p * (p + 1) * (p + 2) / 6; }

static int cmp_i64(const void* a, const void* b) {
    const int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return (x > y) - (x < y);
}

/* Moments of every node (src/LpmTreeSum.cpp:139-147 via :188-197): for each
 * particle j of the node, ascending j, m_k[c] += (gamma_j * y_j[c]) * pow(h)^k
 * with h = sign * (y_j - centroid).  out: [nn][nterms(p)][3], lex term order. */
void oracle_tree_moments(const double* x, const double* y, const double* z, const double* gamma, int p,
                         int moment_sign, double* out) {
    const int nt = nterms(p);
    int (*ks)[3] = malloc(sizeof(int[3]) * nt);
    int t = 0;
    for (int a = 0; a < p; ++a)
        for (int b = 0; b < p; ++b)
            for (int c = 0; c < p; ++c)
                if (a + b + c < p) { ks[t][0] = a; ks[t][1] = b; ks[t][2] = c; ++t; }
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t i = 0; i < g_nn; ++i) {
        const onode* n = &g_nodes[i];
        const int64_t cnt = n->end - n->begin;
        int64_t* idx = malloc(sizeof(int64_t) * (cnt > 0 ? cnt : 1));
        memcpy(idx, g_perm + n->begin, sizeof(int64_t) * cnt);
        qsort(idx, cnt, sizeof(int64_t), cmp_i64);
        double* m = out + (size_t)i * nt * 3;
        memset(m, 0, sizeof(double) * nt * 3);
        for (int64_t q = 0; q < cnt; ++q) {
            const int64_t j = idx[q];
            double h[3];
            if (moment_sign < 0) {
                h[0] = n->c[0] - x[j]; h[1] = n->c[1] - y[j]; h[2] = n->c[2] - z[j];
            } else {
                h[0] = x[j] - n->c[0]; h[1] = y[j] - n->c[1]; h[2] = z[j] - n->c[2];
            }
            for (int k = 0; k < nt; ++k) {
                const double P = pow(h[0], (double)ks[k][0]) * pow(h[1], (double)ks[k][1]) * pow(h[2], (double)ks[k][2]);
                m[3 * k + 0] += gamma[j] * x[j] * P;
                m[3 * k + 1] += gamma[j] * y[j] * P;
                m[3 * k + 2] += gamma[j] * z[j] * P;
            }
        }
        free(idx);
    }
    free(ks);
}

/* SumNode::calc_coeffs (src/LpmTreeSum.cpp:149-171) for target t and centroid c:
 * denom = 1 + delta^2 - t.c; a_000 = 1/denom; visiting k in lex order, every
 * existing successor k+e_i is overwritten with (|k|+1)/denom * t_i * a_k / (k_i+1).
 * out: nterms(p) coefficients in lex order.  Returns nterms(p). */
int oracle_coeffs_lex(double tx, double ty, double tz, double cx, double cy, double cz, double delta, int p,
                      double* out) {
    const int nt = nterms(p);
    /* dense index of (a,b,c) in lex order */
    int* id = malloc(sizeof(int) * p * p * p);
    int t = 0;
    for (int a = 0; a < p; ++a)
        for (int b = 0; b < p; ++b)
            for (int c = 0; c < p; ++c) id[(a * p + b) * p + c] = (a + b + c < p) ? t++ : -1;
    const double denom = 1.0 + delta * delta - (tx * cx + ty * cy + tz * cz);
    for (int k = 0; k < nt; ++k) out[k] = 0.0;
    out[0] = 1.0 / denom;
    const double tv[3] = {tx, ty, tz};
    for (int a = 0; a < p; ++a)
        for (int b = 0; b < p; ++b)
            for (int c = 0; c < p; ++c) {
                if (a + b + c >= p) continue;
                const double ak = out[id[(a * p + b) * p + c]];
                const double numer = a + b + c + 1;
                const int kk[3] = {a, b, c};
                for (int d = 0; d < 3; ++d) {
                    int n1[3] = {a, b, c};
                    n1[d] += 1;
                    if (n1[0] + n1[1] + n1[2] >= p) continue;
                    out[id[(n1[0] * p + n1[1]) * p + n1[2]]] = numer / denom * tv[d] * ak / (double)(kk[d] + 1);
                }
            }
    free(id);
    return nt;
}

/* SphereGreensSeries::computeCoeffs (src/LpmTaylorSeries3d.cpp:63-117): the
 * same expansion for |k| <= maxP with the reference's explicit loop order.
 * out: lex order over |k| <= maxP.  Returns the term count. */
int oracle_coeffs_series(double tx, double ty, double tz, double sx, double sy, double sz, int maxP, double* out) {
    const int q = maxP + 1;
    const int nt = nterms(q);
    int* id = malloc(sizeof(int) * q * q * q);
    int t = 0;
    for (int a = 0; a < q; ++a)
        for (int b = 0; b < q; ++b)
            for (int c = 0; c < q; ++c) id[(a * q + b) * q + c] = (a + b + c <= maxP) ? t++ : -1;
#define A(a, b, c) out[id[((a) * q + (b)) * q + (c)]]
    const double denom = 1.0 - (tx * sx + ty * sy + tz * sz);
    A(0, 0, 0) = 1.0 / denom;
    for (int k1 = 1; k1 <= maxP; ++k1) A(k1, 0, 0) = tx * A(k1 - 1, 0, 0) / denom;
    for (int k2 = 1; k2 <= maxP; ++k2) A(0, k2, 0) = ty * A(0, k2 - 1, 0) / denom;
    for (int k1 = 1; k1 <= maxP; ++k1)
        for (int k2 = 1; k2 <= maxP - k1; ++k2)
            A(k1, k2, 0) = ty * A(k1, k2 - 1, 0) * ((double)(k1 + k2) / k2) / denom;
    for (int k3 = 1; k3 <= maxP; ++k3) {
        A(0, 0, k3) = tz * A(0, 0, k3 - 1) / denom;
        for (int k1 = 1; k1 <= maxP - k3; ++k1) A(k1, 0, k3) = tx * A(k1 - 1, 0, k3) * ((double)(k1 + k3) / k1) / denom;
        for (int k2 = 1; k2 <= maxP - k3; ++k2) A(0, k2, k3) = ty * A(0, k2 - 1, k3) * ((double)(k2 + k3) / k2) / denom;
        for (int k1 = 1; k1 <= maxP - k3; ++k1)
            for (int k2 = 1; k2 <= maxP - k3 - k1; ++k2)
                A(k1, k2, k3) = ty * A(k1, k2 - 1, k3) * ((double)(k1 + k2 + k3) / k2) / denom;
    }
#undef A
    free(id);
    return nt;
}

/* ---------------------------------------------------------------- evaluation */
typedef struct {
    const double *sx, *sy, *sz, *g, *mom;
    int p, nt;
    const int (*ks)[3];
    const int* succ; /* [nt][3] successor term ids or -1 */
    const int* pred; /* [nt][6] ids of k - e_i (0..2) and k - 2 e_i (3..5) or -1 */
    double theta, delta, radius;
    int series;      /* 0: reference series, 1: Cartesian series */
} eval_env;

/* Taylor coefficients in h of 2 / (|d - h|^2 + 2 delta^2), d = t - c:
 * b_k = 2 a_k with a_0 = 1/R2, R2 = |d|^2 + 2 delta^2 and, in lex order,
 * a_k = (2 sum_i d_i a_{k-e_i} - sum_i a_{k-2e_i}) / R2
 * (from Q grad F = -F grad Q for F = 1/Q, Q = |d - h|^2 + 2 delta^2). */
static void coeffs_cartesian(const eval_env* e, const double d[3], double* coef) {
    const double R2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + 2.0 * e->delta * e->delta;
    coef[0] = 1.0 / R2;
    for (int k = 1; k < e->nt; ++k) {
        double s = 0.0;
        for (int i = 0; i < 3; ++i) {
            const int a = e->pred[6 * k + i], b = e->pred[6 * k + 3 + i];
            if (a >= 0) s += 2.0 * d[i] * coef[a];
            if (b >= 0) s -= coef[b];
        }
        coef[k] = s / R2;
    }
    for (int k = 0; k < e->nt; ++k) coef[k] *= 2.0;
}

static void visit(const eval_env* e, int64_t id, const double t[3], double sol[3], int64_t* mac, double* coef) {
    const onode* n = &g_nodes[id];
    const double dist = sqrt((t[0] - n->c[0]) * (t[0] - n->c[0]) + (t[1] - n->c[1]) * (t[1] - n->c[1]) +
                             (t[2] - n->c[2]) * (t[2] - n->c[2]));
    if (n->max_radius / dist < e->theta && e->series == 1) {
        const double d[3] = {t[0] - n->c[0], t[1] - n->c[1], t[2] - n->c[2]};
        coeffs_cartesian(e, d, coef);
        const double* m = e->mom + (size_t)id * e->nt * 3;
        double S[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < e->nt; ++k) {
            S[0] += coef[k] * m[3 * k + 0];
            S[1] += coef[k] * m[3 * k + 1];
            S[2] += coef[k] * m[3 * k + 2];
        }
        sol[0] += t[1] * S[2] - t[2] * S[1];
        sol[1] += t[2] * S[0] - t[0] * S[2];
        sol[2] += t[0] * S[1] - t[1] * S[0];
        *mac += 1;
        return;
    }
    if (n->max_radius / dist < e->theta) {
        const double denom = 1.0 + e->delta * e->delta - (t[0] * n->c[0] + t[1] * n->c[1] + t[2] * n->c[2]);
        for (int k = 0; k < e->nt; ++k) coef[k] = 0.0;
        coef[0] = 1.0 / denom;
        for (int k = 0; k < e->nt; ++k) {
            const double numer = e->ks[k][0] + e->ks[k][1] + e->ks[k][2] + 1;
            for (int d = 0; d < 3; ++d) {
                const int s = e->succ[3 * k + d];
                if (s >= 0) coef[s] = numer / denom * t[d] * coef[k] / (double)(e->ks[k][d] + 1);
            }
        }
        const double* m = e->mom + (size_t)id * e->nt * 3;
        for (int k = 0; k < e->nt; ++k) {
            sol[0] += coef[k] * (t[1] * m[3 * k + 2] - t[2] * m[3 * k + 1]);
            sol[1] += coef[k] * (t[2] * m[3 * k + 0] - t[0] * m[3 * k + 2]);
            sol[2] += coef[k] * (t[0] * m[3 * k + 1] - t[1] * m[3 * k + 0]);
        }
        *mac += 1;
        return;
    }
    if (n->n_kids == 0) {
        /* leaf: ascending particle indices, kernel of src/LpmVectorKernel.h:24-28 */
        const double R2 = e->radius * e->radius, d2 = e->delta * e->delta;
        for (int64_t q = n->begin; q < n->end; ++q) {
            const int64_t j = g_perm[q];
            const double den = R2 - (t[0] * e->sx[j] + t[1] * e->sy[j] + t[2] * e->sz[j]) + d2;
            if (fabs(den) > kZeroTol) {
                const double inv = 1.0 / den;
                sol[0] += ((t[1] * e->sz[j] - t[2] * e->sy[j]) * inv) * e->g[j];
                sol[1] += ((t[2] * e->sx[j] - t[0] * e->sz[j]) * inv) * e->g[j];
                sol[2] += ((t[0] * e->sy[j] - t[1] * e->sx[j]) * inv) * e->g[j];
            }
        }
        return;
    }
    for (int64_t k = n->first_kid; k < n->first_kid + n->n_kids; ++k) visit(e, k, t, sol, mac, coef);
}

/* TreeSum::meshSolve for arbitrary targets (src/LpmTreeSum.cpp:55-86): returns
 * the total multipole-acceptance count; velocities are scaled by -1/(4 pi R). */
int64_t oracle_tree_eval(const double* sx, const double* sy, const double* sz, const double* gamma,
                         const double* moments, int p, double theta, double delta, double radius, const double* tx,
                         const double* ty, const double* tz, int64_t ntg, double* ux, double* uy, double* uz,
                         int64_t* mac_per_target, int64_t series) {
    const int nt = nterms(p);
    int (*ks)[3] = malloc(sizeof(int[3]) * nt);
    int* succ = malloc(sizeof(int) * 3 * nt);
    int* id = malloc(sizeof(int) * p * p * p);
    int t = 0;
    for (int a = 0; a < p; ++a)
        for (int b = 0; b < p; ++b)
            for (int c = 0; c < p; ++c) {
                id[(a * p + b) * p + c] = (a + b + c < p) ? t : -1;
                if (a + b + c < p) { ks[t][0] = a; ks[t][1] = b; ks[t][2] = c; ++t; }
            }
    for (int k = 0; k < nt; ++k)
        for (int d = 0; d < 3; ++d) {
            int n1[3] = {ks[k][0], ks[k][1], ks[k][2]};
            n1[d] += 1;
            succ[3 * k + d] = (n1[0] + n1[1] + n1[2] < p) ? id[(n1[0] * p + n1[1]) * p + n1[2]] : -1;
        }
    int* pred = malloc(sizeof(int) * 6 * nt);
    for (int k = 0; k < nt; ++k)
        for (int d = 0; d < 3; ++d) {
            int n1[3] = {ks[k][0], ks[k][1], ks[k][2]};
            n1[d] -= 1;
            pred[6 * k + d] = (n1[d] >= 0) ? id[(n1[0] * p + n1[1]) * p + n1[2]] : -1;
            n1[d] -= 1;
            pred[6 * k + 3 + d] = (n1[d] >= 0) ? id[(n1[0] * p + n1[1]) * p + n1[2]] : -1;
        }
    eval_env e = {sx, sy, sz, gamma, moments, p, nt, (const int(*)[3])ks, succ, pred, theta, delta, radius,
                  (int)series};
    const double mult = -1.0 / (4.0 * kPi * radius);
    int64_t total = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
    for (int64_t i = 0; i < ntg; ++i) {
        double* coef = malloc(sizeof(double) * nt);
        double tv[3] = {tx[i], ty[i], tz[i]}, sol[3] = {0.0, 0.0, 0.0};
        int64_t mac = 0;
        if (g_nn > 0) visit(&e, 0, tv, sol, &mac, coef);
        ux[i] = sol[0] * mult;
        uy[i] = sol[1] * mult;
        uz[i] = sol[2] * mult;
        if (mac_per_target) mac_per_target[i] = mac;
        total += mac;
        free(coef);
    }
    free(ks);
    free(succ);
    free(pred);
    free(id);
    return total;
}

/* Far field of one box (centroid c, any particles) at target t, unscaled
 * (the sum before the -1/(4 pi R) multiplier), with moments about c of sign
 * moment_sign and the given series (0 reference, 1 Cartesian).  For the
 * single-box series checks. */
void oracle_node_far(double cx, double cy, double cz, double tx, double ty, double tz, const double* sx,
                     const double* sy, const double* sz, const double* gamma, int64_t n, int p, int moment_sign,
                     int series, double delta, double* out3) {
    onode node;
    memset(&node, 0, sizeof(node));
    node.c[0] = cx; node.c[1] = cy; node.c[2] = cz;
    node.max_radius = 0.0;
    const int nt = nterms(p);
    double* mom = calloc((size_t)nt * 3, sizeof(double));
    int (*ks)[3] = malloc(sizeof(int[3]) * nt);
    int* succ = malloc(sizeof(int) * 3 * nt);
    int* pred = malloc(sizeof(int) * 6 * nt);
    int* id = malloc(sizeof(int) * p * p * p);
    int t = 0;
    for (int a = 0; a < p; ++a)
        for (int b = 0; b < p; ++b)
            for (int c = 0; c < p; ++c) {
                id[(a * p + b) * p + c] = (a + b + c < p) ? t : -1;
                if (a + b + c < p) { ks[t][0] = a; ks[t][1] = b; ks[t][2] = c; ++t; }
            }
    for (int k = 0; k < nt; ++k)
        for (int d = 0; d < 3; ++d) {
            int n1[3] = {ks[k][0], ks[k][1], ks[k][2]};
            n1[d] += 1;
            succ[3 * k + d] = (n1[0] + n1[1] + n1[2] < p) ? id[(n1[0] * p + n1[1]) * p + n1[2]] : -1;
            int m1[3] = {ks[k][0], ks[k][1], ks[k][2]};
            m1[d] -= 1;
            pred[6 * k + d] = (m1[d] >= 0) ? id[(m1[0] * p + m1[1]) * p + m1[2]] : -1;
            m1[d] -= 1;
            pred[6 * k + 3 + d] = (m1[d] >= 0) ? id[(m1[0] * p + m1[1]) * p + m1[2]] : -1;
        }
    for (int64_t j = 0; j < n; ++j) {
        double h[3];
        if (moment_sign < 0) { h[0] = cx - sx[j]; h[1] = cy - sy[j]; h[2] = cz - sz[j]; }
        else { h[0] = sx[j] - cx; h[1] = sy[j] - cy; h[2] = sz[j] - cz; }
        for (int k = 0; k < nt; ++k) {
            const double P = pow(h[0], (double)ks[k][0]) * pow(h[1], (double)ks[k][1]) * pow(h[2], (double)ks[k][2]);
            mom[3 * k + 0] += gamma[j] * sx[j] * P;
            mom[3 * k + 1] += gamma[j] * sy[j] * P;
            mom[3 * k + 2] += gamma[j] * sz[j] * P;
        }
    }
    /* reuse the traversal's far-field code with a one-node table */
    onode* saved = g_nodes;
    int64_t saved_nn = g_nn;
    g_nodes = &node;
    g_nn = 1;
    eval_env e = {sx, sy, sz, gamma, mom, p, nt, (const int(*)[3])ks, succ, pred, 1e300, delta, 1.0, series};
    double tv[3] = {tx, ty, tz}, sol[3] = {0.0, 0.0, 0.0};
    double* coef = malloc(sizeof(double) * nt);
    int64_t mac = 0;
    node.max_radius = 0.0; /* MAC ratio 0 < theta: always accepted */
    visit(&e, 0, tv, sol, &mac, coef);
    g_nodes = saved;
    g_nn = saved_nn;
    out3[0] = sol[0]; out3[1] = sol[1]; out3[2] = sol[2];
    free(coef); free(mom); free(ks); free(succ); free(pred); free(id);
}

/* ---------------------------------------------------------------- batch acceptance
 * Batch-MAC treecode (no counterpart in the reference; the specification of
 * lpm_tree_eval_batch in include/lpm_gpu.h).  Targets are ordered by the
 * top levels of 63-bit Morton keys of 21 bisections of the root box (clamped
 * into it), ties by index; batches are runs of at most 64 sorted targets inside one octree
 * cell (oracle_target_batches); node n is accepted when r_n + r_b <
 * theta |c_b - c_n| with r_n the source radius of node_radius.  Batch
 * centroid and radius are
 * formed with the xor-butterfly summation order of a 64-lane wave reduction
 * so the acceptance decisions reproduce the GPU's exactly. */
typedef struct {
    uint64_t key;
    int64_t idx;
} keyed;

/* The targets are ordered by the top g_target_levels key levels (ties in
 * index order: a stable sort), and the batch level is at most that: the
 * GPU sorts general targets on 12 levels and reuses the tree build's key
 * order (10 levels unless the tree is deeper) for targets that are the
 * sources.  Below the sorted levels neighbouring keys keep the input order,
 * so cells there would not be contiguous runs. */
static int g_target_levels = 12;

void oracle_set_target_levels(int levels) { g_target_levels = levels < 1 ? 1 : (levels > 21 ? 21 : levels); }

static int cmp_keyed(const void* a, const void* b) {
    const keyed *x = (const keyed*)a, *y = (const keyed*)b;
    const int sh = 3 * (21 - g_target_levels);
    const uint64_t kx = x->key >> sh, ky = y->key >> sh;
    if (kx != ky) return kx < ky ? -1 : 1;
    return (x->idx > y->idx) - (x->idx < y->idx);
}

/* Morton keys of the targets in the current tree's root box, sorted
 * (order[i] = target at sorted position i, keys[i] its key; keys may be NULL). */
static void sorted_keys(const double* tx, const double* ty, const double* tz, int64_t n, int64_t* order,
                        uint64_t* keys) {
    keyed* k = (keyed*)malloc(sizeof(keyed) * (n > 0 ? n : 1));
    const onode* r = &g_nodes[0];
    for (int64_t i = 0; i < n; ++i) {
        const double p[3] = {fmin(fmax(tx[i], r->lo[0]), r->hi[0]), fmin(fmax(ty[i], r->lo[1]), r->hi[1]),
                             fmin(fmax(tz[i], r->lo[2]), r->hi[2])};
        double lo[3] = {r->lo[0], r->lo[1], r->lo[2]}, hi[3] = {r->hi[0], r->hi[1], r->hi[2]};
        uint64_t key = 0;
        for (int l = 0; l < 21; ++l) {
            int bits = 0;
            for (int d = 0; d < 3; ++d) {
                const double m = 0.5 * (lo[d] + hi[d]);
                const int b = p[d] >= m;
                if (b) lo[d] = m; else hi[d] = m;
                bits |= b << d;
            }
            key = (key << 3) | (uint64_t)bits;
        }
        k[i].key = key;
        k[i].idx = i;
    }
    qsort(k, (size_t)n, sizeof(keyed), cmp_keyed);
    for (int64_t i = 0; i < n; ++i) {
        order[i] = k[i].idx;
        if (keys) keys[i] = k[i].key;
    }
    free(k);
}

void oracle_target_order(const double* tx, const double* ty, const double* tz, int64_t n, int64_t* order) {
    sorted_keys(tx, ty, tz, n, order, NULL);
}

/* leading equal octree levels of two 63-bit keys */
static int key_lcp(uint64_t a, uint64_t b) {
    const uint64_t x = a ^ b;
    if (!x) return 21;
    return (62 - (63 - __builtin_clzll(x))) / 3;
}

/* Target batches: the sorted targets split where the level-L cell changes,
 * L the shallowest level with at least max(1, n / 256) occupied cells and at
 * most g_target_levels, and cut into runs of at most 64.  Fills order and
 * bstart; returns the count. */
int64_t oracle_target_batches(const double* tx, const double* ty, const double* tz, int64_t n, int64_t* order,
                              int64_t* bstart) {
    if (n == 0) return 0;
    uint64_t* keys = malloc(sizeof(uint64_t) * n);
    sorted_keys(tx, ty, tz, n, order, keys);
    int64_t hist[22] = {0};
    for (int64_t i = 1; i < n; ++i) hist[key_lcp(keys[i - 1], keys[i])]++;
    const int64_t want = n / 256 > 1 ? n / 256 : 1;
    int L = g_target_levels;
    int64_t cells = 1;
    for (int l = 0; l < g_target_levels; ++l) {
        if (cells >= want) { L = l; break; }
        cells += hist[l];
    }
    int64_t nb = 0, cs = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (i == 0 || key_lcp(keys[i - 1], keys[i]) < L) cs = i;
        if ((i - cs) % 64 == 0) bstart[nb++] = i;
    }
    free(keys);
    return nb;
}

static double butterfly(double* v, int is_max) {
    double w[64];
    for (int o = 32; o > 0; o >>= 1) {
        for (int l = 0; l < 64; ++l) w[l] = is_max ? fmax(v[l], v[l ^ o]) : v[l] + v[l ^ o];
        memcpy(v, w, sizeof(w));
    }
    return v[0];
}

/* Centroid (c[0..2]) and radius (c[3]) of a batch of nv <= 64 targets t[64][3]. */
static void batch_geometry(const double (*t)[3], int nv, double c[4]) {
    double v[64];
    for (int d = 0; d < 3; ++d) {
        for (int l = 0; l < 64; ++l) v[l] = l < nv ? t[l][d] : 0.0;
        c[d] = butterfly(v, 0) * (1.0 / (double)nv);
    }
    for (int l = 0; l < 64; ++l) {
        const double dx = t[l][0] - c[0], dy = t[l][1] - c[1], dz = t[l][2] - c[2];
        v[l] = l < nv ? dx * dx + dy * dy + dz * dz : 0.0;
    }
    c[3] = sqrt(butterfly(v, 1));
}

/* Source radius about the box centroid: at a leaf the largest distance of
 * its particles, above it max over children of |c_child - c| + r_child. */
static double node_radius(int64_t id, const double* sx, const double* sy, const double* sz, double* nrad) {
    const onode* n = &g_nodes[id];
    double r = 0.0;
    if (n->n_kids == 0) {
        double r2 = 0.0;
        for (int64_t q = n->begin; q < n->end; ++q) {
            const int64_t j = g_perm[q];
            const double dx = sx[j] - n->c[0], dy = sy[j] - n->c[1], dz = sz[j] - n->c[2];
            r2 = fmax(r2, dx * dx + dy * dy + dz * dz);
        }
        r = sqrt(r2);
    } else {
        for (int64_t k = n->first_kid; k < n->first_kid + n->n_kids; ++k) {
            const double rk = node_radius(k, sx, sy, sz, nrad);
            const double dx = g_nodes[k].c[0] - n->c[0], dy = g_nodes[k].c[1] - n->c[1], dz = g_nodes[k].c[2] - n->c[2];
            r = fmax(r, sqrt(dx * dx + dy * dy + dz * dz) + rk);
        }
    }
    nrad[id] = r;
    return r;
}

/* Lex-order Cartesian setup shared by the batch functions. */
static void cartesian_env(eval_env* e, int p) {
    const int nt = nterms(p);
    int (*ks)[3] = malloc(sizeof(int[3]) * nt);
    int* id = malloc(sizeof(int) * p * p * p);
    int* pred = malloc(sizeof(int) * 6 * nt);
    int t = 0;
    for (int a = 0; a < p; ++a)
        for (int b = 0; b < p; ++b)
            for (int c = 0; c < p; ++c) {
                id[(a * p + b) * p + c] = (a + b + c < p) ? t : -1;
                if (a + b + c < p) { ks[t][0] = a; ks[t][1] = b; ks[t][2] = c; ++t; }
            }
    for (int k = 0; k < nt; ++k)
        for (int d = 0; d < 3; ++d) {
            int n1[3] = {ks[k][0], ks[k][1], ks[k][2]};
            n1[d] -= 1;
            pred[6 * k + d] = (n1[d] >= 0) ? id[(n1[0] * p + n1[1]) * p + n1[2]] : -1;
            n1[d] -= 1;
            pred[6 * k + 3 + d] = (n1[d] >= 0) ? id[(n1[0] * p + n1[1]) * p + n1[2]] : -1;
        }
    free(id);
    e->p = p;
    e->nt = nt;
    e->ks = (const int(*)[3])ks;
    e->pred = pred;
    e->succ = NULL;
    e->series = 1;
}

static void cartesian_env_free(eval_env* e) {
    free((void*)e->ks);
    free((void*)e->pred);
}

/* Batch treecode velocity (FP64, Cartesian series of order p with lex
 * moments of order p and sign +1).  stats[0] += accepted pairs x batch
 * size, stats[1] += near-field pairs x batch size. */
void oracle_tree_eval_batch(const double* sx, const double* sy, const double* sz, const double* gamma,
                            const double* moments, int p, double theta, double delta, const double* tx,
                            const double* ty, const double* tz, int64_t ntg, double* ux, double* uy, double* uz,
                            int64_t* stats) {
    eval_env e;
    memset(&e, 0, sizeof(e));
    e.sx = sx; e.sy = sy; e.sz = sz; e.g = gamma; e.mom = moments;
    e.theta = theta; e.delta = delta; e.radius = 1.0;
    cartesian_env(&e, p);
    int64_t* order = malloc(sizeof(int64_t) * (ntg > 0 ? ntg : 1));
    int64_t* bstart = malloc(sizeof(int64_t) * (ntg > 0 ? ntg : 1));
    const int64_t nb = oracle_target_batches(tx, ty, tz, ntg, order, bstart);
    double* nrad = malloc(sizeof(double) * (g_nn > 0 ? g_nn : 1));
    if (g_nn > 0) node_radius(0, sx, sy, sz, nrad);
    const double mult = -1.0 / (4.0 * kPi);
    int64_t far_tot = 0, near_tot = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : far_tot, near_tot)
    for (int64_t b = 0; b < nb; ++b) {
        double t[64][3], sol[64][3], bc[4];
        double* coef = malloc(sizeof(double) * e.nt);
        int64_t* stack = malloc(sizeof(int64_t) * 8 * 72); /* >= 7 x depth + 1 */
        const int64_t b0 = bstart[b], b1 = b + 1 < nb ? bstart[b + 1] : ntg;
        const int nv = (int)(b1 - b0);
        for (int l = 0; l < 64; ++l) {
            const int64_t i = l < nv ? order[b0 + l] : 0;
            t[l][0] = l < nv ? tx[i] : 0.0;
            t[l][1] = l < nv ? ty[i] : 0.0;
            t[l][2] = l < nv ? tz[i] : 0.0;
            sol[l][0] = sol[l][1] = sol[l][2] = 0.0;
        }
        batch_geometry((const double(*)[3])t, nv, bc);
        int64_t sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const onode* n = &g_nodes[stack[--sp]];
            const int64_t id = n - g_nodes;
            const double dx = bc[0] - n->c[0], dy = bc[1] - n->c[1], dz = bc[2] - n->c[2];
            const double s = sqrt(dx * dx + dy * dy + dz * dz);
            if (nrad[id] + bc[3] < theta * s) {
                const double* m = moments + (size_t)id * e.nt * 3;
                for (int l = 0; l < nv; ++l) {
                    const double d[3] = {t[l][0] - n->c[0], t[l][1] - n->c[1], t[l][2] - n->c[2]};
                    coeffs_cartesian(&e, d, coef);
                    double S[3] = {0.0, 0.0, 0.0};
                    for (int k = 0; k < e.nt; ++k)
                        for (int c = 0; c < 3; ++c) S[c] += coef[k] * m[3 * k + c];
                    sol[l][0] += t[l][1] * S[2] - t[l][2] * S[1];
                    sol[l][1] += t[l][2] * S[0] - t[l][0] * S[2];
                    sol[l][2] += t[l][0] * S[1] - t[l][1] * S[0];
                }
                far_tot += nv;
                continue;
            }
            if (n->n_kids == 0) {
                const double d2 = delta * delta;
                for (int l = 0; l < nv; ++l)
                    for (int64_t q = n->begin; q < n->end; ++q) {
                        const int64_t j = g_perm[q];
                        const double den = 1.0 - (t[l][0] * sx[j] + t[l][1] * sy[j] + t[l][2] * sz[j]) + d2;
                        if (fabs(den) > kZeroTol) {
                            const double inv = 1.0 / den;
                            sol[l][0] += ((t[l][1] * sz[j] - t[l][2] * sy[j]) * inv) * gamma[j];
                            sol[l][1] += ((t[l][2] * sx[j] - t[l][0] * sz[j]) * inv) * gamma[j];
                            sol[l][2] += ((t[l][0] * sy[j] - t[l][1] * sx[j]) * inv) * gamma[j];
                        }
                    }
                near_tot += (int64_t)nv * (n->end - n->begin);
                continue;
            }
            for (int64_t k = n->first_kid + n->n_kids - 1; k >= n->first_kid; --k) stack[sp++] = k;
        }
        for (int l = 0; l < nv; ++l) {
            const int64_t i = order[b0 + l];
            ux[i] = sol[l][0] * mult;
            uy[i] = sol[l][1] * mult;
            uz[i] = sol[l][2] * mult;
        }
        free(coef);
        free(stack);
    }
    if (stats) { stats[0] = far_tot; stats[1] = near_tot; }
    free(order);
    free(bstart);
    free(nrad);
    cartesian_env_free(&e);
}

/* Label of one (batch, node) pair: for every order q = 1..p the largest, over
 * the batch's targets, of |t x (S_q - S_exact)| / (4 pi u_ref), where S_q is
 * the Cartesian far field truncated to |k| < q (lex moments of order p, sign
 * +1) and S_exact = sum_j gamma_j y_j / (1 - t.y_j + delta^2) over the node. */
void oracle_pair_errors(const double* sx, const double* sy, const double* sz, const double* gamma,
                        const double* moments, int p, double delta, double u_ref, const double* tx,
                        const double* ty, const double* tz, int64_t ntg, int64_t batch, int64_t node,
                        double* err /* [p] */) {
    eval_env e;
    memset(&e, 0, sizeof(e));
    e.delta = delta;
    cartesian_env(&e, p);
    int64_t* order = malloc(sizeof(int64_t) * (ntg > 0 ? ntg : 1));
    int64_t* bstart = malloc(sizeof(int64_t) * (ntg > 0 ? ntg : 1));
    const int64_t nb = oracle_target_batches(tx, ty, tz, ntg, order, bstart);
    const onode* n = &g_nodes[node];
    const double* m = moments + (size_t)node * e.nt * 3;
    double* coef = malloc(sizeof(double) * e.nt);
    for (int q = 0; q < p; ++q) err[q] = 0.0;
    const int64_t b0 = bstart[batch], b1 = batch + 1 < nb ? bstart[batch + 1] : ntg;
    const int nv = (int)(b1 - b0);
    for (int l = 0; l < nv; ++l) {
        const int64_t i = order[b0 + l];
        const double t[3] = {tx[i], ty[i], tz[i]};
        double E[3] = {0.0, 0.0, 0.0};
        for (int64_t qq = n->begin; qq < n->end; ++qq) {
            const int64_t j = g_perm[qq];
            const double den = 1.0 - (t[0] * sx[j] + t[1] * sy[j] + t[2] * sz[j]) + delta * delta;
            if (fabs(den) > kZeroTol) {
                E[0] += gamma[j] / den * sx[j];
                E[1] += gamma[j] / den * sy[j];
                E[2] += gamma[j] / den * sz[j];
            }
        }
        const double d[3] = {t[0] - n->c[0], t[1] - n->c[1], t[2] - n->c[2]};
        coeffs_cartesian(&e, d, coef);
        for (int q = 1; q <= p; ++q) {
            double S[3] = {0.0, 0.0, 0.0};
            for (int k = 0; k < e.nt; ++k) {
                if (e.ks[k][0] + e.ks[k][1] + e.ks[k][2] >= q) continue;
                for (int c = 0; c < 3; ++c) S[c] += coef[k] * m[3 * k + c];
            }
            const double a[3] = {S[0] - E[0], S[1] - E[1], S[2] - E[2]};
            const double v[3] = {t[1] * a[2] - t[2] * a[1], t[2] * a[0] - t[0] * a[2], t[0] * a[1] - t[1] * a[0]};
            const double r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / (4.0 * kPi * u_ref);
            if (r > err[q - 1]) err[q - 1] = r;
        }
    }
    free(coef);
    free(order);
    free(bstart);
    cartesian_env_free(&e);
}
