// GPU octree over the treecode sources, reproducing the reference's recursive
// pointer tree node for node (src/LpmTreeSum.cpp:221-290, src/LpmBox3d.*):
//
//   root   = [-maxR - 1e-5, maxR + 1e-5]^3             (Box3d(maxR, pad=true))
//   split  iff count > leaf_max; children in bisectAll order c = ix + 2 iy + 4 iz,
//          planes 0.5*(min+max) computed by the same FP64 recursion,
//          half-open containment, empty children dropped
//   leaves list their particles in ascending source index.
//
// Build:
//   K4 tree_keys     per source, 21-level descent of the FP64 box recursion
//                    -> 63-bit key (the bits of the child index at each level)
//   K5 radix sort    stable (key, index) sort (hipCUB onesweep)
//   K6 tree_count /  level-by-level split of every node with count > leaf_max;
//      tree_emit     children found by binary search in the sorted keys,
//                    written compactly (exclusive scan) so the node table is in
//                    BFS order with each node's children consecutive
//      tree_geom     box bounds, centroid and maxRadius in the reference's
//                    operation order (-ffp-contract=off)
//      leaf resort   keys truncated to each particle's leaf depth and sorted
//                    again from source order, so every leaf lists its
//                    particles in ascending index -- the permutation equals
//                    the reference's leaf-by-leaf index lists exactly.
#include "lpm_tree.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

namespace lpm {
namespace {

__global__ __launch_bounds__(256) void tree_keys(const double* __restrict__ x, const double* __restrict__ y,
                                                 const double* __restrict__ z, int64_t n, double x0, double x1,
                                                 double y0, double y1, double z0, double z1,
                                                 uint64_t* __restrict__ keys, int* __restrict__ vals,
                                                 int* __restrict__ outside) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double px = x[i], py = y[i], pz = z[i];
    if (!(x0 <= px && px < x1 && y0 <= py && py < y1 && z0 <= pz && pz < z1)) atomicOr(outside, 1);
    double xl = x0, xh = x1, yl = y0, yh = y1, zl = z0, zh = z1;
    uint64_t key = 0;
    for (int l = 0; l < kMaxDepth; ++l) {
        const double mx = 0.5 * (xl + xh), my = 0.5 * (yl + yh), mz = 0.5 * (zl + zh);
        const int bx = px >= mx, by = py >= my, bz = pz >= mz;
        if (bx) xl = mx; else xh = mx;
        if (by) yl = my; else yh = my;
        if (bz) zl = mz; else zh = mz;
        key = (key << 3) | uint64_t(bx | (by << 1) | (bz << 2));
    }
    keys[i] = key;
    vals[i] = int(i);
}

__device__ inline int child_code(uint64_t key, int level) {  // child index taken at depth level+1
    return int((key >> (3 * (kMaxDepth - level - 1))) & 7);
}

// first position in [b, e) whose child code at `level` is >= c
__device__ inline int lower_code(const uint64_t* keys, int b, int e, int level, int c) {
    while (b < e) {
        const int m = b + ((e - b) >> 1);
        if (child_code(keys[m], level) < c) b = m + 1;
        else e = m;
    }
    return b;
}

// control 0 (MAX_COORDS_PER_NODE): split iff count > param
// control 1 (MAX_DEPTH):           split iff level < param (src/LpmOctree.cpp:145-180)
__global__ __launch_bounds__(256) void tree_count(const uint64_t* __restrict__ keys, int lvl_start, int lvl_count,
                                                  int level, int control, int param, const int* __restrict__ nbeg,
                                                  const int* __restrict__ nend, int* __restrict__ cnt,
                                                  int* __restrict__ nfirst, int* __restrict__ nkids) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= lvl_count) return;
    const int i = lvl_start + t;
    const int b = nbeg[i], e = nend[i];
    const bool want = (control == 0) ? (e - b > param) : (level < param && e > b);
    int c = 0;
    if (want && level < kMaxDepth) {
        int prev = b;
        for (int code = 1; code <= 8; ++code) {
            const int nxt = (code == 8) ? e : lower_code(keys, prev, e, level, code);
            c += (nxt > prev);
            prev = nxt;
        }
    }
    cnt[t] = c;
    nfirst[i] = -1;
    nkids[i] = 0;
    if (want && level >= kMaxDepth) cnt[t] = -1;  // cannot split further
}

__global__ __launch_bounds__(256) void tree_emit(const uint64_t* __restrict__ keys, int lvl_start, int lvl_count,
                                                 int level, const int* __restrict__ cnt, const int* __restrict__ off,
                                                 int base, int* nbeg, int* nend, int* nfirst, int* nkids,
                                                 int* nlevel, int* nparent, uint64_t* nprefix) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= lvl_count || cnt[t] <= 0) return;
    const int i = lvl_start + t;
    const int b = nbeg[i], e = nend[i];
    int k = base + off[t];
    nfirst[i] = k;
    nkids[i] = cnt[t];
    int prev = b;
    for (int code = 0; code < 8; ++code) {
        const int nxt = (code == 7) ? e : lower_code(keys, prev, e, level, code + 1);
        if (nxt > prev) {
            nbeg[k] = prev;
            nend[k] = nxt;
            nlevel[k] = level + 1;
            nparent[k] = i;
            nprefix[k] = (nprefix[i] << 3) | uint64_t(code);
            nfirst[k] = -1;
            nkids[k] = 0;
            ++k;
        }
        prev = nxt;
    }
}

// Box bounds by the FP64 bisection recursion, centroid 0.5*(max+min), and
// maxRadius = max over corners (xmin,ymin,zmin), (xmin,ymax,zmin), ... in the
// order of src/LpmBox3d.cpp:55-66 of distance(centroid, corner).
__global__ __launch_bounds__(256) void tree_geom(int nn, double x0, double y0, double z0, double x1, double y1,
                                                 double z1, const int* __restrict__ nlevel,
                                                 const uint64_t* __restrict__ nprefix, double4* __restrict__ ncent,
                                                 double* __restrict__ nbox) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const int L = nlevel[i];
    const uint64_t pre = nprefix[i];
    double b[6] = {x0, y0, z0, x1, y1, z1};
    for (int l = 0; l < L; ++l) {
        const int c = int((pre >> (3 * (L - 1 - l))) & 7);
        for (int d = 0; d < 3; ++d) {
            const double m = 0.5 * (b[d] + b[d + 3]);
            if ((c >> d) & 1) b[d] = m; else b[d + 3] = m;
        }
    }
    const double cx = 0.5 * (b[3] + b[0]), cy = 0.5 * (b[4] + b[1]), cz = 0.5 * (b[5] + b[2]);
    double r = 0.0;
    for (int k = 0; k < 8; ++k) {
        const double px = (k & 4) ? b[3] : b[0];
        const double py = (k & 1) ? b[4] : b[1];
        const double pz = (k & 2) ? b[5] : b[2];
        const double t = sqrt((px - cx) * (px - cx) + (py - cy) * (py - cy) + (pz - cz) * (pz - cz));
        if (t > r) r = t;
    }
    ncent[i] = make_double4(cx, cy, cz, r);
    for (int d = 0; d < 6; ++d) nbox[6 * int64_t(i) + d] = b[d];
}

// Truncate each particle's key to its leaf depth (keyed by source index).
__global__ __launch_bounds__(256) void tree_leaf_keys(int nn, const int* __restrict__ nkids,
                                                      const int* __restrict__ nbeg, const int* __restrict__ nend,
                                                      const int* __restrict__ nlevel, const uint64_t* __restrict__ keys,
                                                      const int* __restrict__ perm, uint64_t* __restrict__ tkeys) {
    const int i = blockIdx.x;
    if (i >= nn || nkids[i] != 0) return;
    const int L = nlevel[i];
    const uint64_t mask = (L == 0) ? 0ull : (~0ull << (3 * (kMaxDepth - L)));
    for (int p = nbeg[i] + threadIdx.x; p < nend[i]; p += blockDim.x) tkeys[perm[p]] = keys[p] & mask;
}

__global__ __launch_bounds__(256) void iota_i32(int* v, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) v[i] = int(i);
}

// Source radius of every leaf about its box centroid: the largest distance
// of its sources (one block per node; internal nodes are set by
// tree_radius_up).
__global__ __launch_bounds__(256) void tree_radius_leaf(int nn, const int* __restrict__ nkids,
                                                        const int* __restrict__ nbeg, const int* __restrict__ nend,
                                                        const double4* __restrict__ ncent,
                                                        const double4* __restrict__ spos, double* __restrict__ nrad) {
    __shared__ double red[256];
    const int i = blockIdx.x;
    if (i >= nn || nkids[i] != 0) return;
    const double4 c = ncent[i];
    double r2 = 0.0;
    for (int q = nbeg[i] + threadIdx.x; q < nend[i]; q += blockDim.x) {
        const double4 s = spos[q];
        const double dx = s.x - c.x, dy = s.y - c.y, dz = s.z - c.z;
        r2 = fmax(r2, dx * dx + dy * dy + dz * dz);
    }
    red[threadIdx.x] = r2;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) nrad[i] = sqrt(red[0]);
}

// Internal nodes of one level: r = max over children of |c_child - c| + r_child.
__global__ __launch_bounds__(256) void tree_radius_up(int first, int count, const int* __restrict__ nfirst,
                                                      const int* __restrict__ nkids,
                                                      const double4* __restrict__ ncent, double* __restrict__ nrad) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int i = first + t;
    const int nk = nkids[i];
    if (nk == 0) return;
    const double4 c = ncent[i];
    double r = 0.0;
    for (int k = nfirst[i]; k < nfirst[i] + nk; ++k) {
        const double4 ck = ncent[k];
        const double dx = ck.x - c.x, dy = ck.y - c.y, dz = ck.z - c.z;
        r = fmax(r, sqrt(dx * dx + dy * dy + dz * dz) + nrad[k]);
    }
    nrad[i] = r;
}

__global__ __launch_bounds__(256) void gather_pos(const double* __restrict__ x, const double* __restrict__ y,
                                                  const double* __restrict__ z, const int* __restrict__ perm,
                                                  int64_t n, double4* __restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        const int j = perm[i];
        out[i] = make_double4(x[j], y[j], z[j], 0.0);
    }
}

inline unsigned blocks(int64_t n, int b = 256) { return unsigned((n + b - 1) / b); }

int grow_nodes(lpm_ctx* c, Tree* t, int64_t need) {
    if (need <= t->cap) return LPM_OK;
    int64_t cap = std::max<int64_t>(need, 2 * t->cap);
    DevBuf* bufs32[] = {&t->nbeg, &t->nend, &t->nfirst, &t->nkids, &t->nlevel, &t->nparent};
    for (DevBuf* b : bufs32) {
        DevBuf nb;
        LPM_CHECK(nb.reserve(size_t(cap) * sizeof(int)));
        if (b->ptr)
            LPM_HIP(c, hipMemcpyAsync(nb.ptr, b->ptr, size_t(t->cap) * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
        LPM_HIP(c, hipStreamSynchronize(c->stream));
        b->release();
        *b = nb;
    }
    DevBuf np;
    LPM_CHECK(np.reserve(size_t(cap) * sizeof(uint64_t)));
    if (t->nprefix.ptr)
        LPM_HIP(c, hipMemcpyAsync(np.ptr, t->nprefix.ptr, size_t(t->cap) * sizeof(uint64_t), hipMemcpyDeviceToDevice,
                                  c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    t->nprefix.release();
    t->nprefix = np;
    t->cap = cap;
    return LPM_OK;
}

template <class K, class V>
int sort_pairs(lpm_ctx* c, Tree* t, const K* kin, K* kout, const V* vin, V* vout, int64_t n, int end_bit) {
    size_t bytes = 0;
    LPM_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, int(n), 0, end_bit, c->stream));
    LPM_CHECK(t->sort_tmp.reserve(bytes + 256));
    LPM_HIP(c, hipcub::DeviceRadixSort::SortPairs(t->sort_tmp.ptr, bytes, kin, kout, vin, vout, int(n), 0, end_bit,
                                                  c->stream));
    return LPM_OK;
}

}  // namespace

void tree_free(Tree* t) {
    if (!t) return;
    for (DevBuf* b : {&t->perm, &t->spos, &t->keys, &t->nbeg, &t->nend, &t->nfirst, &t->nkids, &t->nlevel,
                      &t->nparent, &t->nprefix, &t->ncent, &t->nbox, &t->nrad, &t->mom, &t->momg, &t->mom32, &t->soff32, &t->nstat, &t->tmp0,
                      &t->tmp1, &t->tmp2, &t->tmp3, &t->sort_tmp})
        b->release();
    delete t;
}

int tree_build_device(lpm_ctx* c, const double* x, const double* y, const double* z, int64_t n,
                      const double* blo, const double* bhi, int control, int param) {
    if (!c->tree) c->tree = new Tree();
    Tree* t = c->tree;
    hipStream_t st = c->stream;
    t->ns = n;
    t->nn = 0;
    t->order = 0;
    t->leaf_max = (control == 0) ? param : 0;
    for (int d = 0; d < 3; ++d) { t->root_box[d] = blo[d]; t->root_box[3 + d] = bhi[d]; }
    const int leaf_max = (control == 0) ? param : 1;
    const int64_t nalloc = std::max<int64_t>(n, 1);

    LPM_CHECK(t->keys.reserve(nalloc * sizeof(uint64_t)));
    LPM_CHECK(t->perm.reserve(nalloc * sizeof(int)));
    LPM_CHECK(t->spos.reserve(nalloc * sizeof(double4)));
    LPM_CHECK(t->tmp0.reserve(nalloc * sizeof(uint64_t)));   // unsorted keys / truncated keys
    LPM_CHECK(t->tmp1.reserve(nalloc * sizeof(int) + 64));   // unsorted values / flags
    LPM_CHECK(t->tmp2.reserve(nalloc * sizeof(uint64_t)));
    int* outside = reinterpret_cast<int*>(t->tmp1.as<char>() + nalloc * sizeof(int));
    LPM_HIP(c, hipMemsetAsync(outside, 0, sizeof(int), st));
    if (n > 0) {
        hipLaunchKernelGGL(tree_keys, dim3(blocks(n)), dim3(256), 0, st, x, y, z, n, blo[0], bhi[0], blo[1], bhi[1],
                           blo[2], bhi[2], t->tmp0.as<uint64_t>(), t->tmp1.as<int>(), outside);
        LPM_CHECK(sort_pairs(c, t, t->tmp0.as<uint64_t>(), t->keys.as<uint64_t>(), t->tmp1.as<int>(),
                             t->perm.as<int>(), n, 3 * kMaxDepth));
    }
    int h_out = 0;
    LPM_HIP(c, hipMemcpyAsync(&h_out, outside, sizeof(int), hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipStreamSynchronize(st));
    if (h_out) return fail(c, LPM_ERR_INVALID_ARG, "a source lies outside the root box");

    // root
    LPM_CHECK(grow_nodes(c, t, std::max<int64_t>(1024, 4 * (n / std::max(1, leaf_max)) + 64)));
    const int root_vals[6] = {0, int(n), -1, 0, 0, -1};
    LPM_HIP(c, hipMemcpyAsync(t->nbeg.as<int>(), &root_vals[0], sizeof(int), hipMemcpyHostToDevice, st));
    LPM_HIP(c, hipMemcpyAsync(t->nend.as<int>(), &root_vals[1], sizeof(int), hipMemcpyHostToDevice, st));
    LPM_HIP(c, hipMemcpyAsync(t->nlevel.as<int>(), &root_vals[3], sizeof(int), hipMemcpyHostToDevice, st));
    LPM_HIP(c, hipMemcpyAsync(t->nparent.as<int>(), &root_vals[5], sizeof(int), hipMemcpyHostToDevice, st));
    LPM_HIP(c, hipMemsetAsync(t->nprefix.as<uint64_t>(), 0, sizeof(uint64_t), st));
    LPM_HIP(c, hipStreamSynchronize(st));
    int64_t nn = 1, lvl_start = 0, lvl_count = 1;
    int depth = 0;
    t->level_start.assign(1, 0);
    std::vector<int> hcnt(2);
    for (int level = 0;; ++level) {
        LPM_CHECK(t->tmp3.reserve(size_t(2 * lvl_count + 2) * sizeof(int)));
        int* cnt = t->tmp3.as<int>();
        int* off = cnt + lvl_count + 1;
        hipLaunchKernelGGL(tree_count, dim3(blocks(lvl_count)), dim3(256), 0, st, t->keys.as<uint64_t>(),
                           int(lvl_start), int(lvl_count), level, control, param, t->nbeg.as<int>(),
                           t->nend.as<int>(), cnt, t->nfirst.as<int>(), t->nkids.as<int>());
        // detect nodes that cannot be split (count > leaf_max at the depth limit)
        std::vector<int> hc(lvl_count);
        LPM_HIP(c, hipMemcpyAsync(hc.data(), cnt, lvl_count * sizeof(int), hipMemcpyDeviceToHost, st));
        LPM_HIP(c, hipStreamSynchronize(st));
        int64_t total = 0;
        std::vector<int> ho(lvl_count);
        for (int64_t i = 0; i < lvl_count; ++i) {
            if (hc[i] < 0) return fail(c, LPM_ERR_TREE_DEPTH, "tree deeper than %d levels (coincident sources)", kMaxDepth);
            ho[i] = int(total);
            total += hc[i];
        }
        if (total == 0) break;
        LPM_HIP(c, hipMemcpyAsync(off, ho.data(), lvl_count * sizeof(int), hipMemcpyHostToDevice, st));
        LPM_CHECK(grow_nodes(c, t, nn + total));
        hipLaunchKernelGGL(tree_emit, dim3(blocks(lvl_count)), dim3(256), 0, st, t->keys.as<uint64_t>(),
                           int(lvl_start), int(lvl_count), level, cnt, off, int(nn), t->nbeg.as<int>(),
                           t->nend.as<int>(), t->nfirst.as<int>(), t->nkids.as<int>(), t->nlevel.as<int>(),
                           t->nparent.as<int>(), t->nprefix.as<uint64_t>());
        LPM_HIP(c, hipGetLastError());
        lvl_start = nn;
        lvl_count = total;
        nn += total;
        depth = level + 1;
        t->level_start.push_back(lvl_start);
    }
    t->level_start.push_back(nn);
    t->nn = nn;
    t->depth = depth;

    LPM_CHECK(t->ncent.reserve(size_t(nn) * sizeof(double4)));
    LPM_CHECK(t->nbox.reserve(size_t(nn) * 6 * sizeof(double)));
    hipLaunchKernelGGL(tree_geom, dim3(blocks(nn)), dim3(256), 0, st, int(nn), blo[0], blo[1], blo[2], bhi[0], bhi[1],
                       bhi[2], t->nlevel.as<int>(), t->nprefix.as<uint64_t>(), t->ncent.as<double4>(),
                       t->nbox.as<double>());
    if (n > 0) {
        // leaf-local ascending source order
        hipLaunchKernelGGL(tree_leaf_keys, dim3(unsigned(nn)), dim3(256), 0, st, int(nn), t->nkids.as<int>(),
                           t->nbeg.as<int>(), t->nend.as<int>(), t->nlevel.as<int>(), t->keys.as<uint64_t>(),
                           t->perm.as<int>(), t->tmp0.as<uint64_t>());
        hipLaunchKernelGGL(iota_i32, dim3(blocks(n)), dim3(256), 0, st, t->tmp1.as<int>(), n);
        LPM_CHECK(sort_pairs(c, t, t->tmp0.as<uint64_t>(), t->tmp2.as<uint64_t>(), t->tmp1.as<int>(),
                             t->perm.as<int>(), n, 3 * kMaxDepth));
        hipLaunchKernelGGL(gather_pos, dim3(blocks(n)), dim3(256), 0, st, x, y, z, t->perm.as<int>(), n,
                           t->spos.as<double4>());
    }
    LPM_CHECK(t->nrad.reserve(size_t(nn) * sizeof(double)));
    LPM_HIP(c, hipMemsetAsync(t->nrad.ptr, 0, size_t(nn) * sizeof(double), st));
    hipLaunchKernelGGL(tree_radius_leaf, dim3(unsigned(nn)), dim3(256), 0, st, int(nn), t->nkids.as<int>(),
                       t->nbeg.as<int>(), t->nend.as<int>(), t->ncent.as<double4>(), t->spos.as<double4>(),
                       t->nrad.as<double>());
    for (int l = depth - 1; l >= 0; --l) {
        const int64_t f = t->level_start[l], cnt = t->level_start[l + 1] - f;
        hipLaunchKernelGGL(tree_radius_up, dim3(blocks(cnt)), dim3(256), 0, st, int(f), int(cnt),
                           t->nfirst.as<int>(), t->nkids.as<int>(), t->ncent.as<double4>(), t->nrad.as<double>());
    }
    LPM_HIP(c, hipGetLastError());
    std::vector<int> hk(nn);
    LPM_HIP(c, hipMemcpyAsync(hk.data(), t->nkids.ptr, nn * sizeof(int), hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipStreamSynchronize(st));
    t->nleaves = std::count(hk.begin(), hk.end(), 0);
    return LPM_OK;
}

}  // namespace lpm

using namespace lpm;

extern "C" {

static int check_tree_args(lpm_ctx* c, const double* x, const double* y, const double* z, int64_t n,
                           const double* lo, const double* hi, int control, int param) {
    if (!c) return LPM_ERR_NULL_PTR;
    c->err.clear();
    if (n < 0 || n > (int64_t(1) << 30)) return fail(c, LPM_ERR_INVALID_ARG, "invalid source count");
    if (n > 0 && (!x || !y || !z)) return fail(c, LPM_ERR_NULL_PTR, "source coordinates are NULL");
    if (control != LPM_TREE_MAX_PER_NODE && control != LPM_TREE_MAX_DEPTH_CONTROL)
        return fail(c, LPM_ERR_INVALID_ARG, "invalid depth control %d", control);
    if (control == LPM_TREE_MAX_PER_NODE && param < 1) return fail(c, LPM_ERR_INVALID_ARG, "leaf_max must be >= 1");
    if (control == LPM_TREE_MAX_DEPTH_CONTROL && (param < 0 || param > kMaxDepth))
        return fail(c, LPM_ERR_INVALID_ARG, "max depth must be in [0, %d]", kMaxDepth);
    for (int d = 0; d < 3; ++d)
        if (!(lo[d] < hi[d]) || !std::isfinite(lo[d]) || !std::isfinite(hi[d]))
            return fail(c, LPM_ERR_INVALID_ARG, "root box must have min < max");
    return LPM_OK;
}

static int build_common(lpm_ctx* c, const double* x, const double* y, const double* z, int64_t n, const double* lo,
                        const double* hi, int control, int param, bool host, int64_t* n_nodes, int* depth) {
    LPM_CHECK(check_tree_args(c, x, y, z, n, lo, hi, control, param));
    LPM_HIP(c, hipSetDevice(c->device));
    if (host) {
        const int64_t na = std::max<int64_t>(n, 1);
        LPM_CHECK(c->h_in.reserve(3 * na * sizeof(double)));
        double* d = c->h_in.as<double>();
        if (n > 0) {
            LPM_HIP(c, hipMemcpyAsync(d, x, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
            LPM_HIP(c, hipMemcpyAsync(d + na, y, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
            LPM_HIP(c, hipMemcpyAsync(d + 2 * na, z, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        }
        x = d; y = d + na; z = d + 2 * na;
    }
    LPM_CHECK(tree_build_device(c, x, y, z, n, lo, hi, control, param));
    if (n_nodes) *n_nodes = c->tree->nn;
    if (depth) *depth = c->tree->depth;
    return LPM_OK;
}

static void mesh_root(double max_r, double* lo, double* hi) {
    // Box3d(maxR, pad=true): xmin = -maxr, xmin -= 1e-5; xmax = maxr, xmax += 1e-5
    for (int d = 0; d < 3; ++d) {
        lo[d] = -max_r;
        lo[d] -= 0.00001;
        hi[d] = max_r;
        hi[d] += 0.00001;
    }
}

int lpm_tree_build_dev(lpm_ctx* c, const double* x, const double* y, const double* z, int64_t n, double max_r,
                       int leaf_max, int64_t* n_nodes, int* depth) {
    if (!(max_r > 0.0) || !std::isfinite(max_r)) return fail(c, LPM_ERR_INVALID_ARG, "max_radius must be > 0");
    double lo[3], hi[3];
    mesh_root(max_r, lo, hi);
    return build_common(c, x, y, z, n, lo, hi, LPM_TREE_MAX_PER_NODE, leaf_max, false, n_nodes, depth);
}

int lpm_tree_build(lpm_ctx* c, const double* x, const double* y, const double* z, int64_t n, double max_r,
                   int leaf_max, int64_t* n_nodes, int* depth) {
    if (!(max_r > 0.0) || !std::isfinite(max_r)) return fail(c, LPM_ERR_INVALID_ARG, "max_radius must be > 0");
    double lo[3], hi[3];
    mesh_root(max_r, lo, hi);
    return build_common(c, x, y, z, n, lo, hi, LPM_TREE_MAX_PER_NODE, leaf_max, true, n_nodes, depth);
}

int lpm_tree_build_box(lpm_ctx* c, const double* x, const double* y, const double* z, int64_t n,
                       const double* box_min, const double* box_max, int control, int param, int64_t* n_nodes,
                       int* depth) {
    if (!box_min || !box_max) return fail(c, LPM_ERR_NULL_PTR, "root box is NULL");
    return build_common(c, x, y, z, n, box_min, box_max, control, param, true, n_nodes, depth);
}

int lpm_tree_info(lpm_ctx* c, int64_t* n_nodes, int* depth, int64_t* n_leaves) {
    if (!c) return LPM_ERR_NULL_PTR;
    if (!c->tree) return fail(c, LPM_ERR_STATE, "no tree built");
    if (n_nodes) *n_nodes = c->tree->nn;
    if (depth) *depth = c->tree->depth;
    if (n_leaves) *n_leaves = c->tree->nleaves;
    return LPM_OK;
}

int lpm_tree_export(lpm_ctx* c, double* box, double* cent, int32_t* level, int32_t* parent, int32_t* first_kid,
                    int32_t* n_kids, int32_t* range, int32_t* perm) {
    if (!c) return LPM_ERR_NULL_PTR;
    if (!c->tree) return fail(c, LPM_ERR_STATE, "no tree built");
    Tree* t = c->tree;
    hipStream_t st = c->stream;
    const size_t nn = size_t(t->nn);
    LPM_HIP(c, hipSetDevice(c->device));
    if (box) LPM_HIP(c, hipMemcpyAsync(box, t->nbox.ptr, nn * 6 * sizeof(double), hipMemcpyDeviceToHost, st));
    if (cent) LPM_HIP(c, hipMemcpyAsync(cent, t->ncent.ptr, nn * 4 * sizeof(double), hipMemcpyDeviceToHost, st));
    if (level) LPM_HIP(c, hipMemcpyAsync(level, t->nlevel.ptr, nn * sizeof(int), hipMemcpyDeviceToHost, st));
    if (parent) LPM_HIP(c, hipMemcpyAsync(parent, t->nparent.ptr, nn * sizeof(int), hipMemcpyDeviceToHost, st));
    if (first_kid) LPM_HIP(c, hipMemcpyAsync(first_kid, t->nfirst.ptr, nn * sizeof(int), hipMemcpyDeviceToHost, st));
    if (n_kids) LPM_HIP(c, hipMemcpyAsync(n_kids, t->nkids.ptr, nn * sizeof(int), hipMemcpyDeviceToHost, st));
    if (range) {
        std::vector<int> b(nn), e(nn);
        LPM_HIP(c, hipMemcpyAsync(b.data(), t->nbeg.ptr, nn * sizeof(int), hipMemcpyDeviceToHost, st));
        LPM_HIP(c, hipMemcpyAsync(e.data(), t->nend.ptr, nn * sizeof(int), hipMemcpyDeviceToHost, st));
        LPM_HIP(c, hipStreamSynchronize(st));
        for (size_t i = 0; i < nn; ++i) { range[2 * i] = b[i]; range[2 * i + 1] = e[i]; }
    }
    if (perm && t->ns > 0)
        LPM_HIP(c, hipMemcpyAsync(perm, t->perm.ptr, size_t(t->ns) * sizeof(int), hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipStreamSynchronize(st));
    return LPM_OK;
}

}  // extern "C"
