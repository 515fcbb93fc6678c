// Device octree over the sources of the treecode (shared by octree.hip and
// treecode.hip).
#pragma once

#include "lpm_internal.hpp"

#include <vector>

namespace lpm {

constexpr int kMaxDepth = LPM_TREE_MAX_DEPTH;  // 21 levels x 3 bits = 63-bit keys
constexpr int kMaxOrder = 16;                  // series order limit (|k| < p)
constexpr int kChunk = 128;                    // leaf particles staged at once by the leaf moment kernel
constexpr int kDirectLeafMax = 4 * kChunk;     // leaves up to this size take one moment block each
constexpr int kTopLevels = 4;                  // levels laid out from per-cell counts (tree_top)
constexpr int kTopCells = 1 << (3 * kTopLevels);

// Build table in device memory (int32): first node and node count of levels
// 0 .. kMaxDepth + 1, then flags and counters written by the build kernels.
constexpr int kLvWords = 2 * (kMaxDepth + 2);
constexpr int kLvOutside = kLvWords;      // a source lies outside the root box
constexpr int kLvTooDeep = kLvWords + 1;  // a node at depth kMaxDepth must split
constexpr int kLvKeys = kLvWords + 2;     // a node below the sorted key levels must split
// leaves and leaf chunks of kChunk particles, each counted in kLvSpread
// words (block b adds to word b % kLvSpread; the host sums them) so the
// thousands of leaves of a deep level do not queue on one address
constexpr int kLvSpread = 64;
constexpr int kLvLeaves = kLvWords + 3;
constexpr int kLvChunks = kLvLeaves + kLvSpread;
// kR2Slots pairs of 64-bit words (8-byte aligned): the range of |x|^2 over
// the sources (r2_block_range, block b in pair b % kR2Slots), for the sphere
// precondition of the series
constexpr int kR2Slots = 64;
constexpr int kLvR2 = (kLvChunks + kLvSpread + 1) & ~1;
constexpr int kLvTotal = kLvR2 + 4 * kR2Slots;

struct Tree {
    int64_t ns = 0;       // sources
    int64_t nn = 0;       // nodes (BFS order; children of a node are consecutive)
    int64_t nleaves = 0;
    int depth = 0;
    int leaf_max = 0;
    double r2min = 0.0, r2max = 0.0;  // range of |x|^2 over the sources
    double root_box[6] = {0, 0, 0, 0, 0, 0};  // xmin ymin zmin xmax ymax zmax
    // source data in tree order (device)
    DevBuf perm;      // int32 [ns]  tree position -> source index
    DevBuf spos;      // double4 [ns] (x, y, z, gamma) in tree order (gamma set by moments)
    DevBuf keys;      // uint64 [ns] Morton keys, sorted on their top kFastLevels levels (or fully)
    DevBuf kperm;     // int32 [ns]  source index in the order of `keys` (before the leaf resort)
    int sorted_levels = 0;  // top key levels by which `keys` is sorted (kFastLevels or kMaxDepth)
    // source arrays of the last build (device), so that targets which are the
    // sources can reuse the key order instead of sorting again
    const double *sx = nullptr, *sy = nullptr, *sz = nullptr;
    // node table (device)
    DevBuf nbeg, nend, nfirst, nkids, nlevel, nparent;  // int32 [cap]
    DevBuf nprefix;   // uint64 [cap] key prefix (top 3*level bits)
    DevBuf nodew;     // int4 [nn] begin, end, first child, child count (one load in the evaluation kernels)
    DevBuf ncent;     // double4 [nn] centroid + maxRadius
    DevBuf nbox;      // double [nn][6] lo3 hi3
    DevBuf nrad;      // double [nn] source radius about the box centroid: exact at leaves,
                      // max over children of |c_child - c| + r_child above (0 if empty)
    std::vector<int64_t> level_start;  // host: first node of every level, then nn
    int64_t cap = 0;
    DevBuf lvt;           // int32 [kLvTotal] build table
    DevBuf chunks;        // int32: chunk node [nch], begin [nch], end [nch], first chunk [nn], count [nn]
    int64_t nch = 0;      // leaf chunks (moment kernel blocks)
    // moments
    int order = 0;        // p (|k| < p), 0 = none
    int moment_sign = 1;
    DevBuf mom;           // double [nn][nterms][3], lex order of k (the reference's layout)
    DevBuf momg;          // double [nn][nterms][3], graded order of k (Cartesian series)
    DevBuf mom32;         // float  [nn][nterms][3], graded order scaled by rho^-|k|, rho = box radius (FP32 far field)
    DevBuf soff32;        // float4 [ns] FP32 source offsets from their leaf centroid + gamma
    DevBuf nstat;         // float4 [nn] |m| of the degree-0, 1, 2 moment shells, particle count
    // streamfunction moments mu_k = sum_j gamma_j h_j^k (one component): lex, graded, FP32 graded scaled
    int psi_order = 0;    // p of the streamfunction moments, 0 = none
    DevBuf pmom, pmomg, pmom32;
    // stereographic node data (lpm_stereo.hpp), [2 projections][nn]; made
    // with the moments when LPM_OPT_STEREO is set
    int stereo_terms = 0;  // terms per node and projection, 0 = none
    int stereo_order = 0;  // the Cartesian order p they serve (entries of order p + extra)
    DevBuf sgeo;          // double4: centre (re, im), source radius about it (+inf: unusable), unused
    DevBuf smom;          // double2 [.][kStereoTerms]: a_k
    DevBuf smom32;        // float2  [.][kStereoTerms]: a_k / rho^k
    // scratch
    DevBuf mpart;         // double [nch][nterms][3] leaf-chunk moment partials
    DevBuf tmp0, tmp1, tmp2, tmp3, sort_tmp, scan_tmp;
};

inline int nterms(int p) { return p * (p + 1) * (p + 2) / 6; }

}  // namespace lpm
