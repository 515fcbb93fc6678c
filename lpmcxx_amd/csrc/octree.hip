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
#include <climits>
#include <vector>

namespace lpm {
namespace {

// kKeyBlocks blocks stride over the sources, so that the range of |x|^2
// folds into the build table with few atomics.  The sources are also counted
// per cell of level kTopLevels (kTopCells cells in key order), so that
// tree_top lays out the first levels of the tree without searching the
// sorted keys.
constexpr int kKeyBlocks = 2048;
__global__ __launch_bounds__(256) void tree_keys(const double* __restrict__ x, const double* __restrict__ y,
                                                 const double* __restrict__ z, int64_t n, double x0, double x1,
                                                 double y0, double y1, double z0, double z1,
                                                 uint64_t* __restrict__ keys, int* __restrict__ vals,
                                                 int* __restrict__ lvt, unsigned* __restrict__ top_hist) {
    __shared__ unsigned s_hist[kTopCells];
    for (int b = threadIdx.x; b < kTopCells; b += blockDim.x) s_hist[b] = 0u;
    __syncthreads();
    int* outside = lvt + kLvOutside;
    double lo = 0.0, hi = 0.0;
    bool any = false;
    const int lane = threadIdx.x & 63;
    // the loop bound is wave-uniform (whole waves take the same trips) for the ballots
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t w0 = int64_t(blockIdx.x) * blockDim.x + (threadIdx.x & ~63); w0 < n; w0 += stride) {
        const int64_t i = w0 + lane;
        int cell = -1;
        if (i < n) {
            const double px = x[i], py = y[i], pz = z[i];
            const double r2 = px * px + py * py + pz * pz;
            // NaN must win both ends: compare on the bit patterns the words use
            const unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(r2));
            if (!any || bits < static_cast<unsigned long long>(__double_as_longlong(lo))) lo = r2;
            if (!any || bits > static_cast<unsigned long long>(__double_as_longlong(hi))) hi = r2;
            any = true;
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
            cell = int(key >> (3 * (kMaxDepth - kTopLevels)));
        }
        // one LDS add per distinct cell of the wave (neighbouring sources
        // mostly share a cell)
        uint64_t left = __ballot(cell >= 0);
        while (left) {
            const int u = __builtin_amdgcn_readlane(cell, __builtin_ctzll(left));
            const uint64_t m = __ballot(cell == u);
            if (lane == __builtin_ctzll(m)) atomicAdd(&s_hist[u], unsigned(__popcll(m)));
            left &= ~m;
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kTopCells; b += blockDim.x)
        if (s_hist[b]) atomicAdd(&top_hist[b], s_hist[b]);
    unsigned long long* w = reinterpret_cast<unsigned long long*>(lvt + kLvR2) + 2 * (blockIdx.x % kR2Slots);
    r2_block_range(lo, any, w);
    __syncthreads();  // the second call reuses the block's reduction words
    r2_block_range(hi, any, w);
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

// Levels 0 .. kTopLevels of the tree from the per-cell counts of tree_keys.
// Cell b of level l covers the sorted positions [start, start + count), the
// starts being the prefix sums of the level-kTopLevels counts (the keys are
// sorted on at least kTopLevels levels, so a cell is one run of them).  A
// cell is a node when its parent cell is a node that splits (control 0:
// more than param sources; control 1: level below param), and the nodes of
// each level are numbered in key order after those of the levels above (BFS
// order, children consecutive): the same table the level-by-level build
// makes, without searching the keys.  Also resets the level table, the flags
// and the leaf / chunk counters, and counts the leaves of levels below
// kTopLevels (those of level kTopLevels are counted by tree_count).  One
// block.
constexpr int kTopThreads = 1024;
__host__ __device__ constexpr int top_offset(int l) { return ((1 << (3 * l)) - 1) / 7; }  // cells above level l
constexpr int kTopAll = top_offset(kTopLevels + 1);  // cells of levels 0 .. kTopLevels
static_assert(kTopCells == 1 << (3 * kTopLevels) && kTopCells == kTopThreads * 4, "four level-kTopLevels cells per thread");

__global__ __launch_bounds__(kTopThreads) void tree_top(int n, int control, int param,
                                                        const unsigned* __restrict__ top_hist, int* __restrict__ lvt,
                                                        int* __restrict__ nbeg, int* __restrict__ nend,
                                                        int* __restrict__ nfirst, int* __restrict__ nkids,
                                                        int* __restrict__ nlevel, int* __restrict__ nparent,
                                                        uint64_t* __restrict__ nprefix) {
    using Scan = hipcub::BlockScan<int, kTopThreads>;
    __shared__ typename Scan::TempStorage tmp;
    __shared__ int s_cnt[kTopAll], s_start[kTopAll], s_idx[kTopAll];
    __shared__ int s_lv[kTopLevels + 2];
    __shared__ int s_leaves, s_chunks;
    const int tid = threadIdx.x;
    for (int w = tid; w < kLvWords; w += kTopThreads) lvt[w] = 0;
    for (int w = tid; w < 2 * kLvSpread; w += kTopThreads) lvt[kLvLeaves + w] = 0;
    if (tid == 0) {
        lvt[kLvTooDeep] = 0;
        lvt[kLvKeys] = 0;
        s_leaves = 0;
        s_chunks = 0;
        s_lv[0] = 0;
        s_lv[1] = 1;  // the root
    }
    // counts and starts of the deepest top level: four cells per thread
    constexpr int o4 = top_offset(kTopLevels);
    int c4[4], sum = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        c4[j] = int(top_hist[4 * tid + j]);
        sum += c4[j];
    }
    int ex = 0;
    Scan(tmp).ExclusiveSum(sum, ex);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        s_cnt[o4 + 4 * tid + j] = c4[j];
        s_start[o4 + 4 * tid + j] = ex;
        ex += c4[j];
    }
    __syncthreads();
    // the levels above: sums of eight children, start of the first
    for (int l = kTopLevels - 1; l >= 0; --l) {
        const int o = top_offset(l), oc = top_offset(l + 1);
        for (int b = tid; b < (1 << (3 * l)); b += kTopThreads) {
            int c = 0;
            for (int j = 0; j < 8; ++j) c += s_cnt[oc + 8 * b + j];
            s_cnt[o + b] = c;
            s_start[o + b] = s_start[oc + 8 * b];
        }
        __syncthreads();
    }
    // nodes, level by level from the root
    auto splits = [&](int c, int l) { return control == 0 ? c > param : (l < param && c > 0); };
    if (tid == 0) s_idx[0] = 0;
    for (int l = 1; l <= kTopLevels; ++l) {
        const int o = top_offset(l), op = top_offset(l - 1), ncell = 1 << (3 * l);
        __syncthreads();  // s_idx of level l - 1, s_lv[l]
        // cells 4 tid .. 4 tid + 3 (levels with fewer cells use the first threads)
        int flag[4], f = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int b = 4 * tid + j;
            flag[j] = 0;
            if (b < ncell) {
                const int pb = b >> 3;
                flag[j] = s_cnt[o + b] > 0 && s_idx[op + pb] >= 0 && splits(s_cnt[op + pb], l - 1);
            }
            f += flag[j];
        }
        int rank = 0, total = 0;
        __syncthreads();  // tmp
        Scan(tmp).ExclusiveSum(f, rank, total);
        const int first = s_lv[l];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int b = 4 * tid + j;
            if (b < ncell) s_idx[o + b] = flag[j] ? first + rank : -1;
            rank += flag[j];
        }
        if (tid == 0) s_lv[l + 1] = first + total;
    }
    __syncthreads();
    // node words; the children of a node of level l < kTopLevels are its
    // nonempty child cells, consecutive in the numbering
    for (int l = 0; l <= kTopLevels; ++l) {
        const int o = top_offset(l), ncell = 1 << (3 * l);
        for (int b = tid; b < ncell; b += kTopThreads) {
            const int i = s_idx[o + b];
            if (i < 0) continue;
            const int c = s_cnt[o + b], st = s_start[o + b];
            nbeg[i] = st;
            nend[i] = st + c;
            nlevel[i] = l;
            nparent[i] = l == 0 ? -1 : s_idx[top_offset(l - 1) + (b >> 3)];
            nprefix[i] = uint64_t(b);
            int kf = -1, nk = 0;
            if (l < kTopLevels && splits(c, l)) {
                const int oc = top_offset(l + 1);
                for (int j = 0; j < 8; ++j) {
                    const int ci = s_idx[oc + 8 * b + j];
                    if (ci >= 0) {
                        if (nk == 0) kf = ci;
                        ++nk;
                    }
                }
            }
            nfirst[i] = kf;
            nkids[i] = nk;
            if (l < kTopLevels && nk == 0) {
                atomicAdd(&s_leaves, 1);
                atomicAdd(&s_chunks, (c + kChunk - 1) / kChunk);
            }
        }
    }
    __syncthreads();
    if (tid <= kTopLevels) {
        lvt[2 * tid] = s_lv[tid];
        lvt[2 * tid + 1] = s_lv[tid + 1] - s_lv[tid];
    }
    if (tid == 0) {
        lvt[kLvLeaves] = s_leaves;
        lvt[kLvChunks] = s_chunks;
    }
}

// The child intervals of node [b, e) at `level`, found by the 8 lanes of an
// aligned lane group at once: lane j holds the first position whose child
// code is >= j (j = 0: b) and the start of the next interval (lane 7: e), so
// child j is [lo, hi) when it is not empty.  Every lane of the group calls
// this (the shuffles need all eight).
__device__ inline void child_interval(const uint64_t* keys, int b, int e, int level, bool search, int& lo, int& hi) {
    const int j = threadIdx.x & 7;
    lo = (search && j > 0) ? lower_code(keys, b, e, level, j) : b;
    const int nx = __shfl_down(lo, 1, 8);
    hi = (j == 7) ? e : nx;
}

// Children count of every node of `level` (first node and count from the
// level table), 8 lanes per node, 32 nodes per block: cnt[t] for t < count
// and blk[b] = the children of block b's nodes (so the emit pass finds its
// offsets without a separate scan).  Leaves are counted with their moment
// chunks.
// control 0 (MAX_COORDS_PER_NODE): split iff count > param
// control 1 (MAX_DEPTH):           split iff level < param (src/LpmOctree.cpp:145-180)
__global__ __launch_bounds__(256) void tree_count(const uint64_t* __restrict__ keys, int level, int sorted_levels,
                                                  int control, int param, int* __restrict__ lvt,
                                                  const int* __restrict__ nbeg, const int* __restrict__ nend,
                                                  int* __restrict__ cnt, int* __restrict__ blk,
                                                  int* __restrict__ nfirst, int* __restrict__ nkids) {
    __shared__ int s_w[4];
    const int t = (blockIdx.x * blockDim.x + threadIdx.x) >> 3;
    const int j = threadIdx.x & 7;
    const int start = lvt[2 * level], count = lvt[2 * level + 1];
    int c = 0;
    if (t < count) {  // whole lane groups: the group's shuffles see all eight lanes
        const int i = start + t;
        const int b = nbeg[i], e = nend[i];
        const bool want = (control == 0) ? (e - b > param) : (level < param && e > b);
        bool search = false;
        if (want) {
            if (level >= kMaxDepth) {
                if (j == 0) atomicOr(lvt + kLvTooDeep, 1);         // cannot split further
            } else if (level >= sorted_levels) {
                if (j == 0) atomicOr(lvt + kLvKeys, 1);            // keys sorted too coarsely
            } else {
                search = true;
            }
        }
        int lo, hi;
        child_interval(keys, b, e, level, search, lo, hi);
        const unsigned grp = unsigned(__ballot(search && hi > lo) >> (threadIdx.x & 56)) & 0xffu;
        c = __popc(grp);
        if (j == 0) {
            cnt[t] = c;
            nfirst[i] = -1;
            nkids[i] = 0;
        }
        // leaves and their moment chunks, counted over the wave's active lanes by
        // ballots before one atomic each (the deep levels have thousands of
        // leaves on the same two words); a leaf of several chunks (leaf_max >
        // kChunk) adds the chunks beyond its first itself
        const bool is_leaf = j == 0 && c == 0;
        const uint64_t lm = __ballot(is_leaf);
        if (lm != 0) {
            const int nch = is_leaf ? (e - b + kChunk - 1) / kChunk : 0;
            const uint64_t cm = __ballot(nch > 0);
            const int slot = blockIdx.x % kLvSpread;
            if (nch > 1) atomicAdd(lvt + kLvChunks + slot, nch - 1);
            if ((threadIdx.x & 63) == __builtin_ctzll(__ballot(true))) {
                atomicAdd(lvt + kLvLeaves + slot, __popcll(lm));
                atomicAdd(lvt + kLvChunks + slot, __popcll(cm));
            }
        }
    }
    // children of the block's 32 nodes (lane 8k holds node k's count)
    int v = (j == 0) ? c : 0;
    v += __shfl_xor(v, 8, 64);
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) blk[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// Children of the nodes of `level`, written compactly at first + count +
// (children of the earlier nodes), and the next level's entry of the level
// table; 8 lanes per node, lane j writes child j.  The offset of a block's
// first node is the sum of the earlier blocks' blk words.
__global__ __launch_bounds__(256) void tree_emit(const uint64_t* __restrict__ keys, int level, int* lvt,
                                                 const int* __restrict__ cnt, const int* __restrict__ blk,
                                                 int* nbeg, int* nend, int* nfirst, int* nkids, int* nlevel,
                                                 int* nparent, uint64_t* nprefix) {
    __shared__ int s_w[4];
    __shared__ int s_c[32];
    const int start = lvt[2 * level], count = lvt[2 * level + 1];
    if (count == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            lvt[2 * level + 2] = start;
            lvt[2 * level + 3] = 0;
        }
        return;
    }
    if (int(blockIdx.x) * 32 >= count) return;  // the whole block: no node of this level
    // children of the earlier blocks
    int a = 0;
    for (int k = threadIdx.x; k < int(blockIdx.x); k += blockDim.x) a += blk[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    const int t = (blockIdx.x * blockDim.x + threadIdx.x) >> 3;
    const int j = threadIdx.x & 7, slot = threadIdx.x >> 3;
    const int c = t < count ? cnt[t] : 0;
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = a;
    if (j == 0) s_c[slot] = c;
    __syncthreads();
    int off = s_w[0] + s_w[1] + s_w[2] + s_w[3];
    for (int k = 0; k < slot; ++k) off += s_c[k];
    if (t == count - 1 && j == 0) {
        lvt[2 * level + 2] = start + count;
        lvt[2 * level + 3] = off + c;
    }
    const bool split = t < count && c > 0;
    const int i = start + t;
    const int b = split ? nbeg[i] : 0, e = split ? nend[i] : 0;
    int lo, hi;
    child_interval(keys, b, e, level, split, lo, hi);
    const bool mine = split && hi > lo;
    const unsigned grp = unsigned(__ballot(mine) >> (threadIdx.x & 56)) & 0xffu;
    if (!split) return;
    const int first = start + count + off;
    if (j == 0) {
        nfirst[i] = first;
        nkids[i] = c;
    }
    if (!mine) return;
    const int k = first + __popc(grp & ((1u << j) - 1u));
    nbeg[k] = lo;
    nend[k] = hi;
    nlevel[k] = level + 1;
    nparent[k] = i;
    nprefix[k] = (nprefix[i] << 3) | uint64_t(j);
    nfirst[k] = -1;
    nkids[k] = 0;
}

// Leaf chunk list of the moment kernel: node i owns chunks n0[i] .. n0[i] +
// nc[i] - 1 (n0 = exclusive scan of nc), each of at most kChunk particles.
__global__ __launch_bounds__(256) void chunk_count(int nn, const int* __restrict__ nkids,
                                                   const int* __restrict__ nbeg, const int* __restrict__ nend,
                                                   int* __restrict__ nc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nn) nc[i] = nkids[i] == 0 ? (nend[i] - nbeg[i] + kChunk - 1) / kChunk : 0;
}

__global__ __launch_bounds__(256) void chunk_fill(int nn, const int* __restrict__ n0, const int* __restrict__ nc,
                                                  const int* __restrict__ nbeg, const int* __restrict__ nend,
                                                  int* __restrict__ cn, int* __restrict__ cb, int* __restrict__ ce) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const int b = nbeg[i], e = nend[i];
    for (int k = 0; k < nc[i]; ++k) {
        const int q = n0[i] + k;
        cn[q] = i;
        cb[q] = b + k * kChunk;
        ce[q] = min(e, b + (k + 1) * kChunk);
    }
}

// Node count of a finished build table: the start of the first empty level
// (the emit of the level above writes it).  While levels are still to come,
// the entry after the last written one is still 0, and so is the result.
__device__ inline int table_nodes(const int* lvt) {
    for (int l = 1; l <= kMaxDepth + 1; ++l)
        if (lvt[2 * l + 1] == 0) return lvt[2 * l];
    return 0;
}

// Box bounds by the FP64 bisection recursion, centroid 0.5*(max+min), and
// maxRadius = max over corners (xmin,ymin,zmin), (xmin,ymax,zmin), ... in the
// order of src/LpmBox3d.cpp:55-66 of distance(centroid, corner).  The node
// count comes from the build table, so the kernel can be enqueued before the
// host has read it.
// Also packs each node's begin, end, first child and child count into one
// int4 (nodew).
__global__ __launch_bounds__(256) void tree_geom(const int* __restrict__ lvt, double x0, double y0, double z0,
                                                 double x1, double y1, double z1, const int* __restrict__ nlevel,
                                                 const uint64_t* __restrict__ nprefix, const int* __restrict__ nbeg,
                                                 const int* __restrict__ nend, const int* __restrict__ nfirst,
                                                 const int* __restrict__ nkids, double4* __restrict__ ncent,
                                                 double* __restrict__ nbox, int4* __restrict__ nodew) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= table_nodes(lvt)) return;
    nodew[i] = make_int4(nbeg[i], nend[i], nfirst[i], nkids[i]);
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

// Leaf-local ascending source order without a sort: block per leaf (at most
// kRankMax sources), each source's position is the number of the leaf's
// sources with a smaller index; the source's position goes to spos there.
constexpr int kRankMax = 1024;
__global__ __launch_bounds__(256) void tree_leaf_rank(const int* __restrict__ lvt, const int* __restrict__ nkids,
                                                      const int* __restrict__ nbeg, const int* __restrict__ nend,
                                                      const int* __restrict__ perm_in, int* __restrict__ perm_out,
                                                      const double* __restrict__ x, const double* __restrict__ y,
                                                      const double* __restrict__ z, double4* __restrict__ spos) {
    __shared__ __align__(16) int v[kRankMax];
    const int nn = table_nodes(lvt);
    // blocks stride over the nodes (the grid is sized before the node count is known)
    for (int i = blockIdx.x; i < nn; i += gridDim.x) {
        if (nkids[i] != 0) continue;
        const int b = nbeg[i], m = nend[i] - b;
        const int m4 = (m + 3) & ~3;
        __syncthreads();  // the previous leaf's ranks are read
        // padded to a multiple of 4 with INT_MAX, which ranks below no index
        for (int t = threadIdx.x; t < m4; t += blockDim.x) v[t] = t < m ? perm_in[b + t] : INT_MAX;
        __syncthreads();
        const int4* v4 = reinterpret_cast<const int4*>(v);
        for (int t = threadIdx.x; t < m; t += blockDim.x) {
            const int s = v[t];
            int r = 0;
            // four indices per LDS read, several reads in flight
#pragma unroll 4
            for (int k = 0; k < m4 / 4; ++k) {
                const int4 q = v4[k];
                r += (q.x < s ? 1 : 0) + (q.y < s ? 1 : 0) + (q.z < s ? 1 : 0) + (q.w < s ? 1 : 0);
            }
            perm_out[b + r] = s;
            spos[b + r] = make_double4(x[s], y[s], z[s], 0.0);
        }
    }
}

__global__ __launch_bounds__(256) void iota_i32(int* v, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) v[i] = int(i);
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

// stable sort of (key, value) pairs by key bits [begin_bit, 3 kMaxDepth)
template <class K, class V>
int sort_pairs(lpm_ctx* c, Tree* t, const K* kin, K* kout, const V* vin, V* vout, int64_t n, int begin_bit) {
    size_t bytes = 0;
    LPM_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, int(n), begin_bit,
                                                  3 * kMaxDepth, c->stream));
    LPM_CHECK(t->sort_tmp.reserve(bytes + 256));
    LPM_HIP(c, hipcub::DeviceRadixSort::SortPairs(t->sort_tmp.ptr, bytes, kin, kout, vin, vout, int(n), begin_bit,
                                                  3 * kMaxDepth, c->stream));
    return LPM_OK;
}

static_assert(kLvTotal <= kPinWords, "the build table is read through the pinned words");
constexpr int kFastLevels = 10;  // key levels sorted before the first level loop (30 bits: 4 radix passes)
constexpr int kShortLevels = 8;  // 24 bits: 3 radix passes
constexpr int64_t kBuildNodeBudget = int64_t(16) << 20;  // node bound of the levels built between two table reads

}  // namespace

void tree_free(Tree* t) {
    if (!t) return;
    for (DevBuf* b : {&t->perm, &t->kperm, &t->spos, &t->keys, &t->nbeg, &t->nend, &t->nfirst, &t->nkids, &t->nlevel,
                      &t->nparent, &t->nprefix, &t->nodew, &t->ncent, &t->nbox, &t->nrad, &t->mom, &t->momg, &t->mom32,
                      &t->soff32, &t->nstat, &t->mpart, &t->pmom, &t->pmomg, &t->pmom32, &t->tmp0, &t->tmp1,
                      &t->tmp2, &t->tmp3, &t->sort_tmp, &t->scan_tmp, &t->lvt, &t->chunks})
        b->release();
    delete t;
}

int tree_build_device(lpm_ctx* c, const double* x, const double* y, const double* z, int64_t n,
                      const double* blo, const double* bhi, int control, int param) {
    if (!c->tree) c->tree = new Tree();
    Tree* t = c->tree;
    hipStream_t st = c->stream;
    LPM_HIP(c, join_aux(c));  // the second stream may still read the old tree
    const int64_t prev_nn = t->nn;
    t->ns = n;
    t->nn = 0;
    t->sx = t->sy = t->sz = nullptr;  // set once the build has succeeded
    t->order = 0;
    t->psi_order = 0;
    t->stereo_terms = 0;
    t->leaf_max = (control == 0) ? param : 0;
    for (int d = 0; d < 3; ++d) { t->root_box[d] = blo[d]; t->root_box[3 + d] = bhi[d]; }
    const int leaf_max = (control == 0) ? param : 1;
    const int64_t nalloc = std::max<int64_t>(n, 1);

    LPM_CHECK(t->keys.reserve(nalloc * sizeof(uint64_t)));
    LPM_CHECK(t->perm.reserve(nalloc * sizeof(int)));
    LPM_CHECK(t->kperm.reserve(nalloc * sizeof(int)));
    LPM_CHECK(t->spos.reserve(nalloc * sizeof(double4)));
    LPM_CHECK(t->tmp0.reserve(nalloc * sizeof(uint64_t)));   // unsorted keys / truncated keys
    LPM_CHECK(t->tmp1.reserve(nalloc * sizeof(int) + 64));   // unsorted values
    LPM_CHECK(t->tmp2.reserve(nalloc * sizeof(uint64_t)));
    LPM_CHECK(t->tmp3.reserve(size_t(2 * nalloc + 2) * sizeof(int)));  // children counts | their scan
    // the build table, then the per-cell counts of the top levels
    LPM_CHECK(t->lvt.reserve((kLvTotal + kTopCells) * sizeof(int)));
    int* cnt = t->tmp3.as<int>();
    int* off = cnt + nalloc + 1;
    int* lvt = t->lvt.as<int>();
    unsigned* top_hist = reinterpret_cast<unsigned*>(lvt + kLvTotal);
    LPM_CHECK(grow_nodes(c, t, std::max<int64_t>(std::max(1024, kTopAll), 4 * (n / std::max(1, leaf_max)) + 64)));
    LPM_HIP(c, hipMemsetAsync(lvt, 0, (kLvTotal + kTopCells) * sizeof(int), st));
    if (n > 0)
        hipLaunchKernelGGL(tree_keys, dim3(unsigned(std::min<int64_t>(kKeyBlocks, blocks(n)))), dim3(256), 0, st, x, y,
                           z, n, blo[0], bhi[0], blo[1], bhi[1], blo[2], bhi[2], t->tmp0.as<uint64_t>(),
                           t->tmp1.as<int>(), lvt, top_hist);

    // Levels 0 .. kTopLevels from the top-level cell counts (tree_top); then
    // level by level: count children, then emit them (offsets from the
    // per-block child counts) -- every level up to the sorted key levels in
    // one run of launches, grids sized by the node bound of each level, the
    // build table read once at the end.  The keys are first sorted on their
    // top kFastLevels levels only; a node below them that must split sends
    // the build round again with fully sorted keys.  A run whose node bound
    // would pass kBuildNodeBudget stops at that level and reads the table
    // before going on.
    const int64_t split_bound = (control == 0) ? std::min<int64_t>(n, 8 * (n / (int64_t(leaf_max) + 1))) : n;
    // leaves of at most kRankMax sources (max-per-node control) get their
    // ascending source order by ranks within the leaf; others by keys
    // truncated to the leaf depth, sorted from source order
    const bool ranked = control == 0 && leaf_max <= kRankMax;
    const int64_t rank_grid = std::max<int64_t>(4096, prev_nn + prev_nn / 4);
    if (!c->ev_build) LPM_HIP(c, hipEventCreateWithFlags(&c->ev_build, hipEventDisableTiming));
    int* h = c->pin;
    // key levels sorted at first: kFastLevels, or 8 (24 bits, three radix
    // passes) when the previous build's tree was at most 7 levels deep (a
    // rebuild of moved particles mostly keeps its depth); a node below them
    // that must split sends the build round again with more levels sorted
    int sorted_levels = (t->depth > 0 && t->depth < kShortLevels) ? kShortLevels : kFastLevels;
    int64_t nn = 1;
    int* blk = off;  // per-block child counts (at most one word per 32 nodes)
    // The first run of levels stops at the depth of the previous build (a
    // rebuild of moved particles mostly keeps it): levels below it would be
    // launched empty.  A deeper tree continues in the next run.
    int run_cap = t->depth > 0 ? std::max(t->depth, kTopLevels) : kMaxDepth;
    for (;;) {
        if (n > 0)
            LPM_CHECK(sort_pairs(c, t, t->tmp0.as<uint64_t>(), t->keys.as<uint64_t>(), t->tmp1.as<int>(),
                                 t->perm.as<int>(), n, 3 * (kMaxDepth - sorted_levels)));
        hipLaunchKernelGGL(tree_top, dim3(1), dim3(kTopThreads), 0, st, int(n), control, param,
                           static_cast<const unsigned*>(top_hist), lvt, t->nbeg.as<int>(), t->nend.as<int>(),
                           t->nfirst.as<int>(), t->nkids.as<int>(), t->nlevel.as<int>(), t->nparent.as<int>(),
                           t->nprefix.as<uint64_t>());
        // deepest level that may hold nodes to split
        const int last = std::min(sorted_levels, control == 1 ? std::max(0, std::min(param, kMaxDepth)) : kMaxDepth);
        // the count kernels start at the deepest top level (its nodes may
        // split); its node count is known on the host only by its bound
        int lk = kTopLevels;
        int64_t ck = std::min<int64_t>(std::min<int64_t>(kTopCells, std::max<int64_t>(n, 1)), std::max<int64_t>(split_bound, 1));
        nn = top_offset(kTopLevels) + ck;
        bool resort = false;
        for (;;) {
            // levels lk .. lk + G - 1 in this run; need = the nodes so far plus
            // the bound of each level's children
            int G = 0;
            int64_t need = nn, b = ck;
            const int run_last = std::min(last, run_cap);
            run_cap = kMaxDepth;
            while (lk + G <= run_last) {
                const int64_t nb = std::min<int64_t>(split_bound, 8 * b);
                if (G > 0 && need + nb > kBuildNodeBudget) break;
                need += nb;
                b = nb;
                ++G;
            }
            LPM_CHECK(grow_nodes(c, t, need));
            int64_t ub = ck;
            for (int g2 = 0; g2 < G; ++g2) {
                const int level = lk + g2;
                const unsigned nbk = blocks(std::max<int64_t>(1, ub) * 8);
                hipLaunchKernelGGL(tree_count, dim3(nbk), dim3(256), 0, st, t->keys.as<uint64_t>(), level,
                                   sorted_levels, control, param, lvt, t->nbeg.as<int>(), t->nend.as<int>(), cnt, blk,
                                   t->nfirst.as<int>(), t->nkids.as<int>());
                hipLaunchKernelGGL(tree_emit, dim3(nbk), dim3(256), 0, st, t->keys.as<uint64_t>(), level, lvt, cnt,
                                   blk, t->nbeg.as<int>(), t->nend.as<int>(), t->nfirst.as<int>(),
                                   t->nkids.as<int>(), t->nlevel.as<int>(), t->nparent.as<int>(),
                                   t->nprefix.as<uint64_t>());
                ub = std::min<int64_t>(split_bound, 8 * ub);
            }
            LPM_HIP(c, hipGetLastError());
            LPM_HIP(c, hipMemcpyAsync(h, lvt, kLvTotal * sizeof(int), hipMemcpyDeviceToHost, st));
            LPM_HIP(c, hipEventRecord(c->ev_build, st));
            // the node geometry and the leaves' source order of the tree as the
            // table stands: the whole tree when this run finished it (the
            // kernels find nothing to do otherwise), computed while the host
            // waits for the table
            LPM_CHECK(t->ncent.reserve(size_t(t->cap) * sizeof(double4)));
            LPM_CHECK(t->nbox.reserve(size_t(t->cap) * 6 * sizeof(double)));
            LPM_CHECK(t->nodew.reserve(size_t(t->cap) * sizeof(int4)));
            hipLaunchKernelGGL(tree_geom, dim3(blocks(t->cap)), dim3(256), 0, st, static_cast<const int*>(lvt), blo[0],
                               blo[1], blo[2], bhi[0], bhi[1], bhi[2], t->nlevel.as<int>(), t->nprefix.as<uint64_t>(),
                               t->nbeg.as<int>(), t->nend.as<int>(), t->nfirst.as<int>(), t->nkids.as<int>(),
                               t->ncent.as<double4>(), t->nbox.as<double>(), t->nodew.as<int4>());
            // one block per node of the previous tree (and a few more), striding
            if (ranked && n > 0)
                hipLaunchKernelGGL(tree_leaf_rank, dim3(unsigned(std::min<int64_t>(t->cap, rank_grid))), dim3(256), 0,
                                   st, static_cast<const int*>(lvt),
                                   t->nkids.as<int>(), t->nbeg.as<int>(), t->nend.as<int>(), t->perm.as<int>(),
                                   t->kperm.as<int>(), x, y, z, t->spos.as<double4>());
            LPM_HIP(c, hipGetLastError());
            LPM_HIP(c, hipEventSynchronize(c->ev_build));
            if (h[kLvOutside]) return fail(c, LPM_ERR_INVALID_ARG, "a source lies outside the root box");
            if (h[kLvTooDeep]) return fail(c, LPM_ERR_TREE_DEPTH, "tree deeper than %d levels (coincident sources)", kMaxDepth);
            if (h[kLvKeys]) { resort = true; break; }
            lk += G;
            ck = h[2 * lk + 1];
            nn = int64_t(h[2 * lk]) + ck;
            if (ck == 0 || lk > last) break;
        }
        if (!resort) break;
        sorted_levels = sorted_levels < kFastLevels ? kFastLevels : kMaxDepth;
    }
    int depth = 0;
    t->level_start.assign(1, 0);
    for (int l = 1; l <= kMaxDepth && h[2 * l + 1] > 0; ++l) {
        t->level_start.push_back(h[2 * l]);
        depth = l;
    }
    t->level_start.push_back(nn);
    t->nn = nn;
    t->depth = depth;
    t->sorted_levels = sorted_levels;
    t->r2min = r2_from_words(reinterpret_cast<const unsigned long long*>(h + kLvR2), 0, kR2Slots);
    t->r2max = r2_from_words(reinterpret_cast<const unsigned long long*>(h + kLvR2), 1, kR2Slots);
    t->nleaves = 0;
    t->nch = 0;
    for (int w = 0; w < kLvSpread; ++w) {
        t->nleaves += h[kLvLeaves + w];
        t->nch += h[kLvChunks + w];
    }

    // tree_geom (and, for ranked leaves, tree_leaf_rank) ran on the finished
    // table after the last run
    if (n > 0) {
        // leaf-local ascending source order goes to kperm, then the two swap:
        // perm = leaf order, kperm = the order of the sorted keys
        if (!ranked) {
            hipLaunchKernelGGL(tree_leaf_keys, dim3(unsigned(nn)), dim3(256), 0, st, int(nn), t->nkids.as<int>(),
                               t->nbeg.as<int>(), t->nend.as<int>(), t->nlevel.as<int>(), t->keys.as<uint64_t>(),
                               t->perm.as<int>(), t->tmp0.as<uint64_t>());
            hipLaunchKernelGGL(iota_i32, dim3(blocks(n)), dim3(256), 0, st, t->tmp1.as<int>(), n);
            LPM_CHECK(sort_pairs(c, t, t->tmp0.as<uint64_t>(), t->tmp2.as<uint64_t>(), t->tmp1.as<int>(),
                                 t->kperm.as<int>(), n, std::min(3 * (kMaxDepth - depth), 3 * kMaxDepth - 1)));
        }
        std::swap(t->perm, t->kperm);
        if (!ranked)
            hipLaunchKernelGGL(gather_pos, dim3(blocks(n)), dim3(256), 0, st, x, y, z, t->perm.as<int>(), n,
                               t->spos.as<double4>());
    }
    // node source radii: leaves with the strengths, internal nodes in the
    // upward pass of the moments
    LPM_CHECK(t->nrad.reserve(size_t(nn) * sizeof(double)));
    // leaf chunk list of the moment kernel, needed only when a leaf holds
    // more than kChunk particles (else the kernel takes one block per node)
    const int64_t nch = t->nch;
    if (nch != t->nleaves && !(control == 0 && leaf_max <= kDirectLeafMax)) {
        LPM_CHECK(t->chunks.reserve(size_t(3 * std::max<int64_t>(nch, 1) + 2 * nn) * sizeof(int)));
        int* ch = t->chunks.as<int>();
        int *cn = ch, *cb = ch + nch, *ce = ch + 2 * nch, *n0 = ch + 3 * nch, *nc = n0 + nn;
        hipLaunchKernelGGL(chunk_count, dim3(blocks(nn)), dim3(256), 0, st, int(nn), t->nkids.as<int>(),
                           t->nbeg.as<int>(), t->nend.as<int>(), nc);
        size_t sb = 0;
        LPM_HIP(c, hipcub::DeviceScan::ExclusiveSum(nullptr, sb, nc, n0, int(nn), st));
        LPM_CHECK(t->scan_tmp.reserve(sb + 256));
        LPM_HIP(c, hipcub::DeviceScan::ExclusiveSum(t->scan_tmp.ptr, sb, nc, n0, int(nn), st));
        hipLaunchKernelGGL(chunk_fill, dim3(blocks(nn)), dim3(256), 0, st, int(nn), n0, nc, t->nbeg.as<int>(),
                           t->nend.as<int>(), cn, cb, ce);
    }
    LPM_HIP(c, hipGetLastError());
    if (n > 0) {
        t->sx = x;
        t->sy = y;
        t->sz = z;
    }
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
    // staged host sources: later staged targets may reuse the same buffer
    if (host) c->tree->sx = c->tree->sy = c->tree->sz = nullptr;
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
