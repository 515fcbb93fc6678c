// GPU mesh generation: uniform recursive refinement of the cubed-sphere and
// icosahedral seeds into the particle sets of the spherical sum (vertices,
// leaf-face centroids and areas), in the host builder's index order.
//
// The host builder (PolyMesh2d, after src/LpmPolyMesh2d.cpp:62-69) divides
// the leaf faces of a level one after the other in face order.  A face first
// divides each of its edges that no earlier face has divided (appending the
// midpoint vertex and two child edges, src/LpmEdges.cpp:99-120), then, for a
// quad, appends its centre vertex, then its interior edges and four child
// faces (src/LpmQuadFaces.cpp:14-146, src/LpmTriFaces.cpp:13-142).  Every leaf
// edge lies between two leaf faces of the level, so the face that divides
// it is the one of lower index.  That fixes every index without running the
// loop in order:
//
//   mesh_count      per face: edges it divides first, vertices and edges it
//                   appends (packed into one 64-bit word);
//   exclusive scan  of the packed words: the face's first new vertex and
//                   first new edge;
//   mesh_divide     per face: midpoints of the edges it divides first, the
//                   quad centre, the two child edges of each of those edges;
//   mesh_children   per face: its side (left or right) of the child edges of
//                   all its edges, its interior edges and its four children.
//
// Faces, edges and (implicitly) vertex order therefore equal the host's.
// Indices of one level are level-local (the leaf faces of level k are
// 0..F_k-1); vertex indices are global.  The last level's faces feed
// mesh_leaves: centroid and L'Huilier area in the host's operation order
// (src/LpmFaces.cpp:63-80,147-167, src/LpmXyzVector.cpp:76-107).  Midpoints
// and centroids use only IEEE-exact operations (+, *, /, sqrt), so vertices
// and centroids equal the host's bit for bit; areas go through atan2 and tan
// of the device math library and agree to a few units in the last place.
// As in the reference, face centroids -- the quad centre vertices among them
// -- lie on the unit sphere whatever the radius: SphericalCoords::centroid
// calls sphereCentroid without it (src/LpmSphericalCoords.cpp:29-34), while
// midpoints and seed corners are scaled to the radius.
//
// Nothing is read back: every level's size is known in closed form
// (src/LpmMeshSeed.cpp:233-284), so the whole build is enqueued on the
// context stream without a host wait.
#include "lpm_internal.hpp"
#include "lpm_seeds.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace lpm {
namespace {

constexpr int kMeshBlock = 256;
constexpr int kMaxMeshLevel = 12;  // the host builder's limit

struct V3 {
    double x, y, z;
};

__device__ inline V3 load_v(const double* vx, const double* vy, const double* vz, int i) {
    return V3{vx[i], vy[i], vz[i]};
}

__device__ inline void store_v(double* vx, double* vy, double* vz, int i, V3 v) {
    vx[i] = v.x;
    vy[i] = v.y;
    vz[i] = v.z;
}

// v / |v| * radius with |v| = sqrt((x^2 + y^2) + z^2) (normalize, scale)
__device__ inline V3 to_sphere(V3 v, double radius) {
    const double n = sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    v.x /= n;
    v.y /= n;
    v.z /= n;
    return V3{v.x * radius, v.y * radius, v.z * radius};
}

// normalized chord midpoint (src/LpmXyzVector.cpp:102-107)
__device__ inline V3 sphere_midpoint(V3 a, V3 b, double radius) {
    return to_sphere(V3{0.5 * (a.x + b.x), 0.5 * (a.y + b.y), 0.5 * (a.z + b.z)}, radius);
}

// normalized vertex mean: sum from zero, times 1/n (src/LpmXyzVector.cpp:92-100)
template <int N>
__device__ inline V3 sphere_centroid(const V3* v, double radius) {
    V3 c{0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < N; ++k) {
        c.x += v[k].x;
        c.y += v[k].y;
        c.z += v[k].z;
    }
    const double s = 1.0 / N;
    return to_sphere(V3{c.x * s, c.y * s, c.z * s}, radius);
}

// great-circle angle atan2(|a x b|, a . b) (src/LpmXyzVector.cpp:76-80);
// equal bit for bit under a <-> b (each cross component only changes sign)
__device__ inline double sphere_angle(V3 a, V3 b) {
    const double cx = a.y * b.z - a.z * b.y;
    const double cy = a.z * b.x - a.x * b.z;
    const double cz = a.x * b.y - a.y * b.x;
    const double m = sqrt(cx * cx + cy * cy + cz * cz);
    return atan2(m, a.x * b.x + a.y * b.y + a.z * b.z);
}

// L'Huilier area from the three side angles (src/LpmXyzVector.cpp:82-90)
__device__ inline double lhuilier(double s1, double s2, double s3, double radius) {
    const double hp = 0.5 * (s1 + s2 + s3);
    const double zz = tan(0.5 * hp) * tan(0.5 * (hp - s1)) * tan(0.5 * (hp - s2)) * tan(0.5 * (hp - s3));
    return 4.0 * atan2(sqrt(zz), 1.0) * radius * radius;
}

// One level of the mesh: edges (orig, dest, left face, right face) and the
// edge list of every face.
struct Level {
    int *eo, *ed, *el, *er, *fe;
};

struct SeedArgs {
    double x[12], y[12], z[12];
    int edge[30][4];
    int face[20][4];
    int nv, ne, nf;
};

__global__ void mesh_seed(SeedArgs s, int nepf, Level L, double* vx, double* vy, double* vz) {
    const int t = threadIdx.x;
    if (t < s.nv) {
        vx[t] = s.x[t];
        vy[t] = s.y[t];
        vz[t] = s.z[t];
    }
    if (t < s.ne) {
        L.eo[t] = s.edge[t][0];
        L.ed[t] = s.edge[t][1];
        L.el[t] = s.edge[t][2];
        L.er[t] = s.edge[t][3];
    }
    if (t < s.nf)
        for (int k = 0; k < nepf; ++k) L.fe[t * nepf + k] = s.face[t][k];
}

// vertices appended by face i: the edges it divides first (+ the quad centre);
// edges appended: two per divided edge + the interior ones
template <int N>
__global__ __launch_bounds__(kMeshBlock) void mesh_count(Level L, int nf, unsigned long long* cnt) {
    const int i = blockIdx.x * kMeshBlock + threadIdx.x;
    if (i >= nf) return;
    int first = 0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const int e = L.fe[size_t(i) * N + j];
        first += (min(L.el[e], L.er[e]) == i);
    }
    const unsigned long long nv = first + (N == 4 ? 1 : 0), ne = 2 * first + (N == 4 ? 4 : 3);
    cnt[i] = nv | (ne << 32);
}

template <int N>
__device__ inline void face_vertices(const Level& L, int i, int* pv) {
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const int e = L.fe[size_t(i) * N + j];
        pv[j] = (L.el[e] == i) ? L.eo[e] : L.ed[e];
    }
}

// midpoints and child edges of the edges face i divides first, then the quad centre
template <int N>
__global__ __launch_bounds__(kMeshBlock) void mesh_divide(Level L, Level Nx, int nf, const unsigned long long* off,
                                                          int v0, double* vx, double* vy, double* vz, double radius,
                                                          int* emid, int* ech) {
    const int i = blockIdx.x * kMeshBlock + threadIdx.x;
    if (i >= nf) return;
    const unsigned long long o = off[i];
    const int vb = v0 + int(o & 0xffffffffu), eb = int(o >> 32);
    int r = 0;
    V3 pv[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const int e = L.fe[size_t(i) * N + j];
        const int a = L.eo[e], b = L.ed[e], l = L.el[e], rt = L.er[e];
        const V3 va = load_v(vx, vy, vz, a), vb_ = load_v(vx, vy, vz, b);
        pv[j] = (l == i) ? va : vb_;
        if (min(l, rt) == i) {
            const int m = vb + r, c0 = eb + 2 * r;
            store_v(vx, vy, vz, m, sphere_midpoint(va, vb_, radius));
            emid[e] = m;
            ech[e] = c0;
            Nx.eo[c0] = a;
            Nx.ed[c0] = m;
            Nx.eo[c0 + 1] = m;
            Nx.ed[c0 + 1] = b;
            ++r;
        }
    }
    if (N == 4) store_v(vx, vy, vz, vb + r, sphere_centroid<N>(pv, 1.0));  // unit sphere, as the reference
}

// face i's side of its edges' children, its interior edges, its four children
template <int N>
__global__ __launch_bounds__(kMeshBlock) void mesh_children(Level L, Level Nx, int nf, const unsigned long long* off,
                                                            int v0, const int* emid, const int* ech) {
    const int i = blockIdx.x * kMeshBlock + threadIdx.x;
    if (i >= nf) return;
    const unsigned long long o = off[i];
    const int vb = v0 + int(o & 0xffffffffu), eb = int(o >> 32);
    const int c0 = 4 * i;
    int ne[4][4], nvx[4][4];
    int first = 0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const int pe = L.fe[size_t(i) * N + j];
        const int jn = (j + 1) % N;
        const int l = L.el[pe], rt = L.er[pe];
        const int h0 = ech[pe], h1 = h0 + 1, m = emid[pe];
        nvx[j][j] = (l == i) ? L.eo[pe] : L.ed[pe];
        if (l == i) {
            ne[j][j] = h0;
            Nx.el[h0] = c0 + j;
            ne[jn][j] = h1;
            Nx.el[h1] = c0 + jn;
        } else {
            ne[j][j] = h1;
            Nx.er[h1] = c0 + j;
            ne[jn][j] = h0;
            Nx.er[h0] = c0 + jn;
        }
        if (N == 4) {
            nvx[j][jn] = m;
            nvx[jn][j] = m;
        } else if (j == 0) {
            nvx[0][1] = m;
            nvx[1][0] = m;
            nvx[3][2] = m;
        } else if (j == 1) {
            nvx[1][2] = m;
            nvx[2][1] = m;
            nvx[3][0] = m;
        } else {
            nvx[2][0] = m;
            nvx[0][2] = m;
            nvx[3][1] = m;
        }
        first += (min(l, rt) == i);
    }
    const int ei = eb + 2 * first;  // first interior edge
    auto edge = [&](int k, int a, int b, int fl, int fr) {
        Nx.eo[ei + k] = a;
        Nx.ed[ei + k] = b;
        Nx.el[ei + k] = fl;
        Nx.er[ei + k] = fr;
    };
    if (N == 4) {
        const int cv = vb + first;
#pragma unroll
        for (int j = 0; j < 4; ++j) nvx[j][(j + 2) % 4] = cv;
        ne[0][1] = ei;
        ne[0][2] = ei + 3;
        ne[1][3] = ei;
        ne[1][2] = ei + 2;
        ne[2][3] = ei + 1;
        ne[2][0] = ei + 2;
        ne[3][1] = ei + 1;
        ne[3][0] = ei + 3;
        edge(0, nvx[0][1], nvx[0][2], c0 + 0, c0 + 1);
        edge(1, nvx[3][1], nvx[3][2], c0 + 3, c0 + 2);
        edge(2, nvx[1][2], nvx[1][3], c0 + 1, c0 + 2);
        edge(3, nvx[0][2], nvx[0][3], c0 + 0, c0 + 3);
    } else {
        ne[3][0] = ei;
        ne[3][1] = ei + 1;
        ne[3][2] = ei + 2;
        ne[0][1] = ei + 1;
        ne[1][2] = ei + 2;
        ne[2][0] = ei;
        edge(0, nvx[2][1], nvx[2][0], c0 + 3, c0 + 2);
        edge(1, nvx[0][2], nvx[0][1], c0 + 3, c0 + 0);
        edge(2, nvx[1][0], nvx[1][2], c0 + 3, c0 + 1);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int k = 0; k < N; ++k) Nx.fe[(size_t(c0) + j) * N + k] = ne[j][k];
}

// centroid and area of every face of the last level; the fan triangle k is
// (centroid, v_k, v_k+1), whose side angles are shared with its neighbours
template <int N>
__global__ __launch_bounds__(kMeshBlock) void mesh_leaves(Level L, int nf, const double* vx, const double* vy,
                                                          const double* vz, double radius, double* cx, double* cy,
                                                          double* cz, double* area, int64_t* leafv) {
    const int i = blockIdx.x * kMeshBlock + threadIdx.x;
    if (i >= nf) return;
    int pv[N];
    face_vertices<N>(L, i, pv);
    V3 v[N];
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = load_v(vx, vy, vz, pv[k]);
    const V3 c = sphere_centroid<N>(v, 1.0);  // unit sphere, as the reference
    double sc[N], se[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        sc[k] = sphere_angle(c, v[k]);
        se[k] = sphere_angle(v[k], v[(k + 1) % N]);
    }
    double a = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) a += lhuilier(sc[k], se[k], sc[(k + 1) % N], radius);
    cx[i] = c.x;
    cy[i] = c.y;
    cz[i] = c.z;
    area[i] = a;
    if (leafv)
#pragma unroll
        for (int k = 0; k < N; ++k) leafv[size_t(i) * N + k] = pv[k];
}

inline unsigned grid_of(int64_t n) { return unsigned((n + kMeshBlock - 1) / kMeshBlock); }
inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

struct Counts {
    int nepf;
    int64_t v[kMaxMeshLevel + 1], f[kMaxMeshLevel + 1], e[kMaxMeshLevel + 1], f_all;
};

// closed forms of src/LpmMeshSeed.cpp:233-284 (edges by Euler: E = F + V - 2)
bool mesh_counts(int seed, int level, Counts& n) {
    if ((seed != LPM_SEED_ICOS_TRI && seed != LPM_SEED_CUBED) || level < 0 || level > kMaxMeshLevel) return false;
    const int64_t f0 = seed == LPM_SEED_CUBED ? 6 : 20, vmul = seed == LPM_SEED_CUBED ? 6 : 10;
    n.nepf = seed == LPM_SEED_CUBED ? 4 : 3;
    n.f_all = 0;
    for (int k = 0; k <= level; ++k) {
        n.f[k] = f0 << (2 * k);
        n.v[k] = 2 + (vmul << (2 * k));
        n.e[k] = n.f[k] + n.v[k] - 2;
        n.f_all += n.f[k];
    }
    return true;
}

template <int N>
int generate(lpm_ctx* c, const Counts& n, int level, double radius, double* vx, double* vy, double* vz, double* cx,
             double* cy, double* cz, double* area, int64_t* leafv) {
    hipStream_t st = c->stream;
    // two level regions (even and odd levels), each sized for the largest
    // level of its parity; per-level scratch sized for the last divided level
    size_t reg_bytes[2] = {0, 0};
    for (int k = 0; k <= level; ++k) {
        const size_t b = 4 * align256(size_t(n.e[k]) * sizeof(int)) + align256(size_t(n.f[k]) * N * sizeof(int));
        reg_bytes[k & 1] = std::max(reg_bytes[k & 1], b);
    }
    const int64_t ef = level > 0 ? n.e[level - 1] : 0, ff = level > 0 ? n.f[level - 1] : 0;
    size_t scan_bytes = 0;
    if (level > 0)
        LPM_HIP(c, hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (unsigned long long*)nullptr,
                                                    (unsigned long long*)nullptr, int(ff), st));
    const size_t aux_bytes = 2 * align256(size_t(ef) * sizeof(int)) +
                             2 * align256(size_t(ff) * sizeof(unsigned long long)) + align256(scan_bytes);
    LPM_CHECK(c->mesh_work.reserve(reg_bytes[0] + reg_bytes[1] + aux_bytes));
    char* base = c->mesh_work.as<char>();
    auto level_at = [&](int k) {
        char* p = base + ((k & 1) ? reg_bytes[0] : 0);
        Level L;
        const size_t eb = align256(size_t(n.e[k]) * sizeof(int));
        L.eo = reinterpret_cast<int*>(p);
        L.ed = reinterpret_cast<int*>(p + eb);
        L.el = reinterpret_cast<int*>(p + 2 * eb);
        L.er = reinterpret_cast<int*>(p + 3 * eb);
        L.fe = reinterpret_cast<int*>(p + 4 * eb);
        return L;
    };
    char* aux = base + reg_bytes[0] + reg_bytes[1];
    int* emid = reinterpret_cast<int*>(aux);
    int* ech = reinterpret_cast<int*>(aux + align256(size_t(ef) * sizeof(int)));
    auto* cnt = reinterpret_cast<unsigned long long*>(aux + 2 * align256(size_t(ef) * sizeof(int)));
    auto* off = cnt + align256(size_t(ff) * sizeof(unsigned long long)) / sizeof(unsigned long long);
    void* scan_tmp = aux + 2 * align256(size_t(ef) * sizeof(int)) + 2 * align256(size_t(ff) * sizeof(unsigned long long));

    // seed: vertices scaled by the radius, edges and faces as tabulated
    SeedArgs s{};
    if (N == 4) {
        const double a = std::strtod(lpm_seed::kCubeA, nullptr);
        s.nv = 8;
        s.ne = 12;
        s.nf = 6;
        for (int i = 0; i < 8; ++i) {
            s.x[i] = (lpm_seed::kCubeSigns[i][0] * a) * radius;
            s.y[i] = (lpm_seed::kCubeSigns[i][1] * a) * radius;
            s.z[i] = (lpm_seed::kCubeSigns[i][2] * a) * radius;
        }
        for (int i = 0; i < 12; ++i)
            for (int k = 0; k < 4; ++k) s.edge[i][k] = lpm_seed::kCubeEdges[i][k];
        for (int i = 0; i < 6; ++i)
            for (int k = 0; k < 4; ++k) s.face[i][k] = lpm_seed::kCubeFaces[i][k];
    } else {
        s.nv = 12;
        s.ne = 30;
        s.nf = 20;
        for (int i = 0; i < 12; ++i) {
            s.x[i] = std::strtod(lpm_seed::kIcosXyz[i][0], nullptr) * radius;
            s.y[i] = std::strtod(lpm_seed::kIcosXyz[i][1], nullptr) * radius;
            s.z[i] = std::strtod(lpm_seed::kIcosXyz[i][2], nullptr) * radius;
        }
        for (int i = 0; i < 30; ++i)
            for (int k = 0; k < 4; ++k) s.edge[i][k] = lpm_seed::kIcosEdges[i][k];
        for (int i = 0; i < 20; ++i)
            for (int k = 0; k < 3; ++k) s.face[i][k] = lpm_seed::kIcosFaces[i][k];
    }
    hipLaunchKernelGGL(mesh_seed, dim3(1), dim3(64), 0, st, s, N, level_at(0), vx, vy, vz);
    LPM_HIP(c, hipGetLastError());

    for (int k = 0; k < level; ++k) {
        const Level L = level_at(k), Nx = level_at(k + 1);
        const int nf = int(n.f[k]), v0 = int(n.v[k]);
        const unsigned g = grid_of(nf);
        hipLaunchKernelGGL(mesh_count<N>, dim3(g), dim3(kMeshBlock), 0, st, L, nf, cnt);
        size_t sb = scan_bytes;
        LPM_HIP(c, hipcub::DeviceScan::ExclusiveSum(scan_tmp, sb, cnt, off, nf, st));
        hipLaunchKernelGGL(mesh_divide<N>, dim3(g), dim3(kMeshBlock), 0, st, L, Nx, nf, off, v0, vx, vy, vz, radius,
                           emid, ech);
        hipLaunchKernelGGL(mesh_children<N>, dim3(g), dim3(kMeshBlock), 0, st, L, Nx, nf, off, v0, emid, ech);
        LPM_HIP(c, hipGetLastError());
    }
    const int nl = int(n.f[level]);
    mark_begin(c);
    hipLaunchKernelGGL(mesh_leaves<N>, dim3(grid_of(nl)), dim3(kMeshBlock), 0, st, level_at(level), nl, vx, vy, vz,
                       radius, cx, cy, cz, area, leafv);
    mark_end(c);
    LPM_HIP(c, hipGetLastError());
    return LPM_OK;
}

int check_generate(lpm_ctx* c, int seed, int level, double radius, Counts& n, const void* vx, const void* vy,
                   const void* vz, const void* cx, const void* cy, const void* cz, const void* area) {
    if (!c) return LPM_ERR_NULL_PTR;
    c->err.clear();
    if (!mesh_counts(seed, level, n))
        return fail(c, LPM_ERR_INVALID_ARG, "mesh seed %d / level %d out of range (levels 0..%d)", seed, level,
                    kMaxMeshLevel);
    if (!(radius > 0.0) || !std::isfinite(radius)) return fail(c, LPM_ERR_INVALID_ARG, "mesh radius must be > 0");
    if (!vx || !vy || !vz || !cx || !cy || !cz || !area)
        return fail(c, LPM_ERR_NULL_PTR, "mesh output arrays must not be NULL");
    return LPM_OK;
}

}  // namespace
}  // namespace lpm

using namespace lpm;

extern "C" {

int lpm_mesh_counts(int seed, int level, int64_t* n_vertices, int64_t* n_faces, int64_t* n_leaf_faces) {
    Counts n;
    if (!mesh_counts(seed, level, n)) return LPM_ERR_INVALID_ARG;
    if (n_vertices) *n_vertices = n.v[level];
    if (n_faces) *n_faces = n.f_all;
    if (n_leaf_faces) *n_leaf_faces = n.f[level];
    return LPM_OK;
}

int lpm_mesh_generate_dev(lpm_ctx* c, int seed, int level, double radius, double* vx, double* vy, double* vz,
                          double* cx, double* cy, double* cz, double* area, int64_t* leaf_vertices) {
    Counts n;
    LPM_CHECK(check_generate(c, seed, level, radius, n, vx, vy, vz, cx, cy, cz, area));
    LPM_HIP(c, hipSetDevice(c->device));
    if (n.nepf == 4) return generate<4>(c, n, level, radius, vx, vy, vz, cx, cy, cz, area, leaf_vertices);
    return generate<3>(c, n, level, radius, vx, vy, vz, cx, cy, cz, area, leaf_vertices);
}

int lpm_mesh_generate(lpm_ctx* c, int seed, int level, double radius, double* vx, double* vy, double* vz, double* cx,
                      double* cy, double* cz, double* area, int64_t* leaf_vertices) {
    Counts n;
    LPM_CHECK(check_generate(c, seed, level, radius, n, vx, vy, vz, cx, cy, cz, area));
    LPM_HIP(c, hipSetDevice(c->device));
    const size_t nv = size_t(n.v[level]), nl = size_t(n.f[level]);
    const size_t b = (3 * nv + 4 * nl) * sizeof(double) + (leaf_vertices ? nl * n.nepf * sizeof(int64_t) : 0);
    LPM_CHECK(c->h_out.reserve(b));
    double* d = c->h_out.as<double>();
    double *dcx = d + 3 * nv, *dcy = dcx + nl, *dcz = dcy + nl, *da = dcz + nl;
    int64_t* dlv = leaf_vertices ? reinterpret_cast<int64_t*>(da + nl) : nullptr;
    const int s = n.nepf == 4 ? generate<4>(c, n, level, radius, d, d + nv, d + 2 * nv, dcx, dcy, dcz, da, dlv)
                              : generate<3>(c, n, level, radius, d, d + nv, d + 2 * nv, dcx, dcy, dcz, da, dlv);
    LPM_CHECK(s);
    hipStream_t st = c->stream;
    const size_t bv = nv * sizeof(double), bl = nl * sizeof(double);
    LPM_HIP(c, hipMemcpyAsync(vx, d, bv, hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipMemcpyAsync(vy, d + nv, bv, hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipMemcpyAsync(vz, d + 2 * nv, bv, hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipMemcpyAsync(cx, dcx, bl, hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipMemcpyAsync(cy, dcy, bl, hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipMemcpyAsync(cz, dcz, bl, hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipMemcpyAsync(area, da, bl, hipMemcpyDeviceToHost, st));
    if (leaf_vertices)
        LPM_HIP(c, hipMemcpyAsync(leaf_vertices, dlv, nl * n.nepf * sizeof(int64_t), hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipStreamSynchronize(st));
    return LPM_OK;
}

}  // extern "C"
