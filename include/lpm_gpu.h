/*
 * lpm_gpu.h -- C ABI of the MI355X (gfx950) spherical Biot-Savart / Poisson
 * sum.  Plain pointers and sizes only; no exceptions cross this boundary;
 * every entry point returns an lpm_status (0 = OK, < 0 = error) and leaves a
 * message retrievable with lpm_last_error().
 *
 * Particle data is SoA FP64 (x[], y[], z[]) exactly as the reference's Field
 * and Coords containers expose it (src/LpmField.h:34-51 getPtrToData).
 * Entry points without the _dev suffix take HOST pointers, copy to device,
 * run, copy back and synchronize.  The _dev variants take DEVICE pointers,
 * enqueue on the context's stream and return without synchronizing, with
 * these exceptions: lpm_tree_build_dev waits for the device to size the node
 * table (once per 4 tree levels; it returns the node count and depth),
 * lpm_tree_eval_learned_dev waits once at its end to check that every
 * batch's interaction lists fit (they are grown and planned again if not),
 * and the sphere check below waits once for the points it reads.
 *
 * Sphere precondition.  LPM_F32C forms R^2 - x.y as |x - y|^2 / 2, and the
 * Cartesian, stereographic and logarithmic series expand kernels of
 * |x - y|^2; both are equal to the reference's R^2 - x.y only when |x| = R
 * for every target and source.  The entry points using these forms (the
 * direct sum with LPM_F32C; every treecode form but LPM_F64 with
 * LPM_SERIES_REFERENCE, where R = 1; the RK4 / BVE steps with such a
 * solver) check that | |x|^2 - R^2 | <= 1e-12 R^2 for every point and
 * return LPM_ERR_INVALID_ARG otherwise.  Targets that are the tree's own
 * source arrays are not read again (the build records its sources' range).
 *
 * A context owns one HIP stream, one growable device workspace and at most
 * one source octree.  Use one context per host thread.
 */
#ifndef LPM_GPU_H
#define LPM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lpm_ctx lpm_ctx;

enum lpm_status {
    LPM_OK = 0,
    LPM_ERR_INVALID_ARG = -1,  /* sizes, enums or parameters out of range */
    LPM_ERR_NULL_PTR = -2,     /* a required pointer is NULL */
    LPM_ERR_HIP = -3,          /* a HIP runtime call failed */
    LPM_ERR_ALLOC = -4,        /* device allocation failed (reference: std::bad_alloc) */
    LPM_ERR_STATE = -5,        /* call order violated (e.g. eval before build) */
    LPM_ERR_TREE_DEPTH = -6,   /* octree would exceed LPM_TREE_MAX_DEPTH levels */
    LPM_ERR_KERNEL = -7        /* unknown kernel id (reference: std::logic_error("invalid kernel")) */
};

/* kernel ids (bit mask for the direct sum) */
#define LPM_K_BIOT_SAVART_SPHERE 1 /* velocity, src/LpmVectorKernel.h:19-39 */
#define LPM_K_GREENS_SPHERE 2      /* streamfunction, src/LpmScalarKernel.h:31-45 */

/* arithmetic forms */
#define LPM_F64 0  /* FP64, the reference's per-pair formula */
#define LPM_F32C 1 /* FP32 pair math on tile-local offsets, |x-y|^2/2 denominator,
                      compensated FP32 tile sums, FP64 final reduction */
#define LPM_F64F 2 /* FP64, the reference's pair regrouped for the FP64 units:
                      x x sum_j w_j (y_j - x) with the reference's denominator and
                      zero rule, refined hardware reciprocal (about an ulp), fused
                      multiply-adds; agrees with LPM_F64 to rounding.  Accepted by
                      the direct sum (velocity; the potential is as LPM_F64), by
                      lpm_tree_eval (near field; the far field is as LPM_F64) and
                      by the RK4 / BVE steps with those methods */

/* far-field series of the treecode */
#define LPM_SERIES_REFERENCE 0 /* SumNode::calc_coeffs: Taylor series of 1/(1 + delta^2 - t.y) about the box
                                  centroid (src/LpmTreeSum.cpp:149-171); converges only while
                                  |t.(y - c)| < |1 + delta^2 - t.c| */
#define LPM_SERIES_CARTESIAN 1 /* Taylor series of the same kernel on the sphere, 2/(|t - y|^2 + 2 delta^2),
                                  about the box centroid; converges for maxRadius/|t - c| < 1 */

#define LPM_TREE_MAX_DEPTH 21

/* tree depth control (Tree::TREE_DEPTH_CONTROL, src/LpmOctree.h:51) */
#define LPM_TREE_MAX_PER_NODE 0      /* split iff a node holds more than param sources */
#define LPM_TREE_MAX_DEPTH_CONTROL 1 /* split every non-empty node above level param */

/* ------------------------------------------------------------------ context */
int lpm_create(lpm_ctx** ctx, int device);
void lpm_destroy(lpm_ctx* ctx);
const char* lpm_status_string(int status);
const char* lpm_last_error(const lpm_ctx* ctx);
int lpm_synchronize(lpm_ctx* ctx);
/* hipStream_t of the context (as void*), for callers that enqueue around it */
void* lpm_stream(lpm_ctx* ctx);

/* ------------------------------------------------------------------ direct sum
 * Replaces the pair loops of DirectSum::meshSolve (src/LpmDirectSum.cpp:25-173):
 *   velocity  u_i   = -1/(4 pi R) sum_j gamma_j (x_i x y_j) / (R^2 - x_i.y_j + delta^2)
 *             with a pair contributing 0 when |denominator| <= 1e-13
 *             (src/LpmVectorKernel.h:24-32);
 *   potential psi_i = -1/(4 pi R) sum_{j != skip_i} gamma_j log(R^2 - x_i.y_j)
 *             (src/LpmScalarKernel.h:35-38).
 * kernels  : LPM_K_BIOT_SAVART_SPHERE and/or LPM_K_GREENS_SPHERE (bit mask).
 * prec     : LPM_F64, LPM_F64F or LPM_F32C (LPM_F32C: every point on |x| = radius,
 *             see the sphere precondition above).
 * gamma_j  : source strength (vorticity x area in the mesh solve).
 * tgt_self_src : per-target source index excluded from the potential sum
 *             (the face-target self pair of src/LpmDirectSum.cpp:52-75), -1 for
 *             none (any other value outside [0, ns) is invalid); NULL means no
 *             exclusions.  The velocity needs none: its self pair is exactly
 *             zero.
 * ux,uy,uz : required when the velocity kernel is requested, psi when the
 *             potential kernel is requested; others may be NULL.
 * nt or ns may be 0 (outputs are then zero / untouched respectively).
 * With LPM_F32C and 65,536 or more targets or sources, the sum runs on
 * copies of the targets and sources sorted along the sphere (cube face, then
 * 2D Hilbert index; workspace about 40 bytes per target and 36 per source
 * beyond the unsorted sum's), so its speed does not depend on the caller's
 * order; outputs are in the caller's order.
 */
int lpm_direct_sum(lpm_ctx* ctx, int kernels, int prec, double radius, double delta,
                   const double* tx, const double* ty, const double* tz, int64_t nt,
                   const double* sx, const double* sy, const double* sz, const double* gamma, int64_t ns,
                   const int64_t* tgt_self_src,
                   double* ux, double* uy, double* uz, double* psi);
int lpm_direct_sum_dev(lpm_ctx* ctx, int kernels, int prec, double radius, double delta,
                       const double* tx, const double* ty, const double* tz, int64_t nt,
                       const double* sx, const double* sy, const double* sz, const double* gamma, int64_t ns,
                       const int64_t* tgt_self_src,
                       double* ux, double* uy, double* uz, double* psi);
/* lpm_direct_sum_dev without allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as lpm_direct_sum_dev; its workspace (the
 * prepared sources and the per-split partials) lives in one static device
 * array of 32 MiB per GPU (made when the library is loaded on that GPU, held
 * until the process ends), so it calls no hipMalloc / hipFree.  Fits when
 * 256 + ns x 48 (LPM_F32C; 32 for LPM_F64[F], rounded up to 256 sources and 256
 * bytes) + nt x 32 bytes <= 32 MiB, e.g. 98,304 sources with up to 900,000
 * targets; larger problems return LPM_ERR_INVALID_ARG.  Unlike
 * lpm_direct_sum_dev it does not sort large LPM_F32C problems: its speed
 * needs targets and sources in a spatially coherent order (mesh order is).
 * When the partials of
 * every split do not fit, fewer splits are used (the sum over splits then
 * runs in another order: results agree with lpm_direct_sum_dev to rounding).
 * With LPM_F32C it waits once for the stream to check the sphere
 * precondition; with LPM_F64 it does not wait.  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (several contexts or
 * threads, several streams) race and must be serialized by the caller. */
int lpm_direct_sum_dev_noalloc(lpm_ctx* ctx, int kernels, int prec, double radius, double delta,
                               const double* tx, const double* ty, const double* tz, int64_t nt,
                               const double* sx, const double* sy, const double* sz, const double* gamma,
                               int64_t ns, const int64_t* tgt_self_src,
                               double* ux, double* uy, double* uz, double* psi);

/* ------------------------------------------------------------------ Taylor series
 * TaylorSeries3d / SphereGreensSeries (src/LpmTaylorSeries3d.h:20-63).  The
 * series keeps the multi-indices |k| <= max_p (0 <= max_p <= 20), in the lex
 * order of MultiIndex (src/LpmMultiIndex.h:34-43, the order of the
 * reference's std::map), nterms = (max_p+1)(max_p+2)(max_p+3)/6.  Host
 * pointers; both wait for the device.
 *
 * lpm_series_coeffs: SphereGreensSeries::computeCoeffs
 * (src/LpmTaylorSeries3d.cpp:63-117) for n (target t_i, expansion centre c_i)
 * pairs: coeffs[i][k] of 1/(1 - t.(c + h)) in h, with the reference's loop
 * order and FP64 operations (its bit pattern).
 * Replaces: SphereGreensSeries::computeCoeffs(tgtVec, srcVec, param). */
int lpm_series_coeffs(lpm_ctx* ctx, int max_p, int64_t n, const double* tx, const double* ty, const double* tz,
                      const double* cx, const double* cy, const double* cz, double* coeffs);
/* lpm_series_moments: TaylorSeries3d::computeMoments
 * (src/LpmTaylorSeries3d.cpp:119-154): moments[k] = sum_i (y_i - c)^k s_i over
 * the n points in index order, the power pow(h_x, k1) pow(h_y, k2) pow(h_z, k3)
 * (MultiIndex::vectorPower, src/LpmMultiIndex.h:25-27).  w == NULL: strength
 * s_i; weight_order 0: power x (s_i w_i) (values x weights); weight_order 1:
 * (power x s_i) x w_i (face values x face areas).
 * Replaces: TaylorSeries3d::computeMoments (its three overloads). */
int lpm_series_moments(lpm_ctx* ctx, int max_p, int64_t n, const double* x, const double* y, const double* z,
                       const double* s, const double* w, int weight_order, double cx, double cy, double cz,
                       double* moments);

/* ------------------------------------------------------------------ treecode
 * Source octree (TreeSum mesh constructor, src/LpmTreeSum.cpp:221-238 and
 * generateTreeFromMesh :266-290): root [-max_radius-1e-5, max_radius+1e-5)^3,
 * a node splits into its non-empty octants iff it holds more than leaf_max
 * sources.  Node boxes, leaf lists and the octant order equal the reference's.
 * Returns LPM_ERR_INVALID_ARG if a source lies outside the root box and
 * LPM_ERR_TREE_DEPTH if more than leaf_max sources coincide.
 * Replaces: TreeSum::TreeSum(int, int, shared_ptr<MeshedParticles>, ...) src/LpmTreeSum.h:52-53. */
int lpm_tree_build(lpm_ctx* ctx, const double* sx, const double* sy, const double* sz, int64_t ns,
                   double max_radius, int leaf_max, int64_t* n_nodes, int* depth);
int lpm_tree_build_dev(lpm_ctx* ctx, const double* sx, const double* sy, const double* sz, int64_t ns,
                       double max_radius, int leaf_max, int64_t* n_nodes, int* depth);
/* General root box (Tree(crds) + buildTree, src/LpmOctree.cpp:16-36,73-180):
 * box_min/box_max[3] are the root bounds, control/param as above.  Children
 * are the non-empty octants, as for the mesh tree.  A source outside the root
 * box is an error (the reference keeps such a source in the root's list and
 * drops it from every child).  Host pointers.
 * Replaces: Tree::Tree(shared_ptr<Coords>, ...) and Tree::buildTree(TREE_DEPTH_CONTROL,
 * index_type, bool) src/LpmOctree.h:53,68 (do_shrink = false only). */
int lpm_tree_build_box(lpm_ctx* ctx, const double* sx, const double* sy, const double* sz, int64_t ns,
                       const double* box_min, const double* box_max, int control, int param,
                       int64_t* n_nodes, int* depth);
/* Moments m_k = sum_j gamma_j y_j h_j^k for every node and every |k| < order,
 * h_j = moment_sign * (y_j - centroid); moment_sign = -1 reproduces the
 * reference (h = centroid - y, src/LpmTreeSum.cpp:140), +1 is the expansion
 * the coefficients assume.  order in [1, 12].
 * Replaces: TreeSum::computeMeshMoments() src/LpmTreeSum.h:59. */
int lpm_tree_moments(lpm_ctx* ctx, const double* gamma, int order, int moment_sign);
int lpm_tree_moments_dev(lpm_ctx* ctx, const double* gamma, int order, int moment_sign);
/* Velocity at nt targets: multipole acceptance maxRadius/|t - c| < theta
 * (src/LpmTreeSum.cpp:22-24), far field by `series`, near field by the
 * kernel at unaccepted leaves; result scaled by -1/(4 pi).  mac_count (may be
 * NULL) receives the number of accepted (target, node) pairs, the reference's
 * macCounter.  LPM_SERIES_CARTESIAN requires moments with moment_sign = +1.
 * prec: LPM_F64 (reference pair formula at the leaves), LPM_F64F (its fused
 * form) or LPM_F32C (FP32 leaf sums); the far field is FP64 in all three.
 * Replaces: TreeSum::meshSolve(verts, faces, mp_tol, macCounter) src/LpmTreeSum.h:67. */
int lpm_tree_eval(lpm_ctx* ctx, int prec, int series, double theta, double delta,
                  const double* tx, const double* ty, const double* tz, int64_t nt,
                  double* ux, double* uy, double* uz, int64_t* mac_count);
int lpm_tree_eval_dev(lpm_ctx* ctx, int prec, int series, double theta, double delta,
                      const double* tx, const double* ty, const double* tz, int64_t nt,
                      double* ux, double* uy, double* uz, int64_t* mac_count);
/* Work of the last lpm_tree_eval: accepted (target, node) pairs (far field)
 * and (target, source) pairs summed at leaves (near field).  Synchronizes. */
int lpm_tree_last_stats(lpm_ctx* ctx, int64_t* far_pairs, int64_t* near_pairs);
/* node count, depth (root = 0) and leaf count of the current tree */
int lpm_tree_info(lpm_ctx* ctx, int64_t* n_nodes, int* depth, int64_t* n_leaves);
/* Node table in breadth-first order (children of a node are consecutive):
 * box [nn][6] = (xmin, ymin, zmin, xmax, ymax, zmax), cent [nn][4] =
 * (centroid, maxRadius), level, parent (-1 at the root), first_kid (-1 at a
 * leaf), n_kids, range [nn][2] = [begin, end) into perm, perm [ns] = source
 * index at each tree position (leaves list ascending indices).  Host pointers;
 * any may be NULL.  Replaces: Tree::writeToVtk data, src/LpmOctree.cpp:245-268. */
int lpm_tree_export(lpm_ctx* ctx, double* box, double* cent, int32_t* level, int32_t* parent,
                    int32_t* first_kid, int32_t* n_kids, int32_t* range, int32_t* perm);
/* Moments of the last lpm_tree_moments call, [nn][nterms(order)][3] in the
 * node order of lpm_tree_export and lex order of the multi-indices
 * (nterms(p) = p(p+1)(p+2)/6).  Host pointer.
 * Replaces: SumNode::moments / momentString(), src/LpmTreeSum.h:36-44. */
int lpm_tree_export_moments(lpm_ctx* ctx, double* moments);

/* Tuning options of a context (no counterpart in the reference).
 * LPM_OPT_BATCH_CELL_TARGETS: target batches of the batch and learned
 * treecodes are cut at the octree cells of the shallowest level with at least
 * n_targets / value occupied cells (default 256). */
#define LPM_OPT_BATCH_CELL_TARGETS 1
/* LPM_OPT_STEREO (0 or 1, default 0): lpm_tree_moments* of order p also makes
 * the stereographic moments of every node (orders up to p +
 * LPM_OPT_STEREO_EXTRA_ORDER, at most 15), and the FP32C learned evaluation with
 * delta = 0 evaluates a far pair that the model accepts at order p as the
 * plane Cauchy series of the sphere kernel in stereographic coordinates, of
 * order p + LPM_OPT_STEREO_EXTRA_ORDER (default extra 1, capped by the
 * moments made), when
 * the pair's plane ratio (rho_n + r_b) / |z_b - z_c| is at most
 * LPM_OPT_STEREO_RATIO (default 1.0) times its ratio (r_n + r_b) / s on the
 * sphere.  Every target batch uses the projection from the pole opposite its
 * hemisphere; a pair that fails the ratio test there is tested in the other
 * projection.  P + 1 complex terms replace the (p + 1)(p + 2)(p + 3)/6
 * Cartesian terms of the pair.  LPM_OPT_STEREO 2 (stereographic only): the
 * moments call makes the Cartesian moments only up to order 3 (the degrees 0
 * to 2 that the order model reads) beside the stereographic ones of order p,
 * and a learned evaluation (FP32C, delta = 0, max_order <= p) lets an
 * accepted pair that neither projection takes descend instead of using the
 * Cartesian series; the other evaluations then need moments made without it.
 * Options take effect at the next lpm_tree_moments* call (the stereographic
 * moments) and evaluation. */
#define LPM_OPT_STEREO 2
#define LPM_OPT_STEREO_EXTRA_ORDER 3
#define LPM_OPT_STEREO_RATIO 4
int lpm_set_option(lpm_ctx* ctx, int option, double value);

/* Duration of the last pair or evaluation kernel, measured with HIP events
 * recorded on the context stream around it alone: ds_f64 / ds_f32c of
 * lpm_direct_sum*, tree_eval of lpm_tree_eval*, tree_eval_batch of
 * lpm_tree_eval_batch*, and the low-order interaction-list kernel
 * (tree_eval_lists of the near leaves and the far entries of order <= 7; not
 * the planning pass, not the order 8..12 list kernel) of
 * lpm_tree_eval_learned*, and the leaf kernel (centroids and areas of the
 * last level) of lpm_mesh_generate*.  Waits for the kernel.
 * No counterpart in the reference. */
int lpm_last_kernel_ms(lpm_ctx* ctx, float* ms);

/* Counters of the last tree evaluation (any of the lpm_tree_eval* calls):
 * out[0] far-field (target, node) pairs, out[1] near-field (target, source)
 * pairs, out[2] far-field series terms summed over the far pairs
 * (p(p+1)(p+2)/6 each), out[3] order-model evaluations, and for
 * lpm_tree_eval_learned* the part of out[0] (out[4]) and of out[2] (out[5])
 * evaluated by the order 8..12 list kernel (zero otherwise), the part of
 * out[0] evaluated as stereographic series (out[6]) and their complex terms
 * (out[7], P + 1 each, not in out[2]).  out holds LPM_TREE_NSTATS values.
 * Synchronizes.
 * No counterpart in the reference (its macCounter is out[0] of lpm_tree_eval). */
#define LPM_STAT_FAR 0
#define LPM_STAT_NEAR 1
#define LPM_STAT_TERMS 2
#define LPM_STAT_MODEL 3
#define LPM_STAT_FAR_HIGH 4
#define LPM_STAT_TERMS_HIGH 5
#define LPM_STAT_FAR_STEREO 6
#define LPM_STAT_TERMS_STEREO 7
#define LPM_TREE_NSTATS 8
int lpm_tree_stats(lpm_ctx* ctx, int64_t* out);

/* ------------------------------------------------------------------ streamfunction treecode
 * psi(t) = -(1/4 pi) sum_j gamma_j log(1 - t.y_j) (SphereGreensFn,
 * src/LpmScalarKernel.h:31-45) by treecode.  The reference declares the
 * scalar tree sum (TreeSum::initScalarKernel, src/LpmTreeSum.h:55-56) but
 * never implements it (computeScalarSum is commented out, :61-65); this is
 * that sum.  Moments: mu_k = sum_j gamma_j (y_j - c)^k about every node's box
 * centroid, |k| < p, on the current tree (moment sign +1; gamma also becomes
 * the strength of the tree's near fields).  Evaluation: the batch acceptance
 * of lpm_tree_eval_batch (maxRadius_n + r_b < theta |c_b - c_n| for batches
 * of <= 64 Morton-sorted targets); far nodes by the Taylor series of
 * log(|t - y|^2) of order `order` (<= the moments' order), near leaves
 * summed directly; pairs at zero distance contribute nothing.  prec LPM_F64:
 * FP64 with the reference pair formula in the near field; LPM_F32C: FP32 far
 * field from node-scaled FP32 moments, FP32 log2 near field, FP64 per-target
 * sums. */
int lpm_tree_moments_psi(lpm_ctx* ctx, const double* gamma, int p);
int lpm_tree_moments_psi_dev(lpm_ctx* ctx, const double* gamma, int p);
int lpm_tree_eval_psi(lpm_ctx* ctx, int prec, int order, double theta, const double* tx, const double* ty,
                      const double* tz, int64_t nt, double* psi);
int lpm_tree_eval_psi_dev(lpm_ctx* ctx, int prec, int order, double theta, const double* tx, const double* ty,
                          const double* tz, int64_t nt, double* psi);

/* ------------------------------------------------------------------ batch treecode
 * The treecode velocity of lpm_tree_eval with the acceptance decided per
 * batch of 64 Morton-sorted targets: node n is accepted for the batch when
 * maxRadius_n + r_b < theta |c_b - c_n| (c_b, r_b: centroid and radius of the
 * batch), with the Cartesian series of order `order` (<= the moments' order,
 * moment_sign +1).  prec LPM_F64: FP64 throughout, reference pair formula in
 * the near field; LPM_F32C: FP32 far field from FP32 moments and FP32
 * near field with |d|^2/2 denominators, FP64 per-target sums.  The result
 * differs from lpm_tree_eval only by the series truncation of the pairs each
 * accepts.  Replaces: TreeSum::meshSolve src/LpmTreeSum.h:67 (same role,
 * batch acceptance instead of the per-target criterion of
 * src/LpmTreeSum.cpp:22-24). */
int lpm_tree_eval_batch(lpm_ctx* ctx, int prec, int order, double theta, double delta,
                        const double* tx, const double* ty, const double* tz, int64_t nt,
                        double* ux, double* uy, double* uz);
int lpm_tree_eval_batch_dev(lpm_ctx* ctx, int prec, int order, double theta, double delta,
                            const double* tx, const double* ty, const double* tz, int64_t nt,
                            double* ux, double* uy, double* uz);

/* ------------------------------------------------------------------ learned order and acceptance
 * No counterpart in the reference.  The order model is a 16-64-64-16 ReLU
 * perceptron (10 used inputs, 13 classes) with BF16 weights and FP32
 * accumulation on the MFMA units; class 0 = descend, class q = accept with
 * order q.  Parameters are a flat FP32 array of LPM_MODEL_NPARAMS values:
 * W1[16][64] b1[64] W2[64][64] b2[64] W3[64][16] b3[16] (row = input).
 * Inputs are clamped to the model's input range [lo, hi] and normalized by
 * (x - mean) * inv_std (16 values each), in lpm_model_forward, the gradient
 * and training calls and the treecode's planning pass alike. */
#define LPM_MODEL_NPARAMS 6288
#define LPM_MODEL_INPUTS 16
#define LPM_MODEL_CLASSES 13
#define LPM_LABEL_ORDERS 12
/* He-initialized weights from `seed`, zero biases, identity normalization,
 * unbounded input range. */
int lpm_model_init(lpm_ctx* ctx, uint64_t seed);
/* Set the FP32 master parameters and the normalization (mean / inv_std may be
 * NULL: 0 / 1); resets the Adam state and the input range (unbounded). */
int lpm_model_set_params(lpm_ctx* ctx, const float* params, const float* mean, const float* inv_std);
int lpm_model_get_params(lpm_ctx* ctx, float* params, float* mean, float* inv_std);
/* Input range of the model (16 values each; lo and hi both NULL: unbounded).
 * lpm_model_train sets it to the per-input minimum and maximum of its rows,
 * so that inputs the model never saw in training (a one-target batch of
 * radius 0, a node of zero circulation) are answered as the nearest inputs it
 * did see.  Errors: lo > hi or NaN (LPM_ERR_INVALID_ARG). */
int lpm_model_set_range(lpm_ctx* ctx, const float* lo, const float* hi);
int lpm_model_get_range(lpm_ctx* ctx, float* lo, float* hi);
/* Gradient of the mean softmax cross entropy over n rows (x[n][16] raw
 * inputs, y[n] classes) with respect to the parameters, computed by the
 * forward and backward MFMA kernels; loss may be NULL. */
int lpm_model_gradient(lpm_ctx* ctx, const float* x, const int32_t* y, int64_t n, float* grad, float* loss);
/* Adam training (beta 0.9 / 0.999, eps 1e-8) for `epochs` shuffled passes in
 * mini-batches of `batch` rows; final_loss (may be NULL) is the mean loss of
 * the last epoch.  Sets the input range to that of the rows (which must be
 * finite). */
int lpm_model_train(lpm_ctx* ctx, const float* x, const int32_t* y, int64_t n, int epochs, int batch,
                    float lr, uint64_t seed, float* final_loss);
/* Logits [n][16] (the training forward) and classes [n] (the inference of
 * the planning pass, which sums the two K = 64 layers in 32-wide blocks:
 * it may differ from the argmax of the logits where two classes are within
 * FP32 rounding); either may be NULL. */
int lpm_model_forward(lpm_ctx* ctx, const float* x, int64_t n, float* logits, int32_t* cls);
/* Training data from the FP64 treecode: for every batch_stride-th target
 * batch the wave walks the tree; a node passing maxRadius_n + r_b <
 * theta_max |c_b - c_n| is a candidate pair when at most frontier_levels of
 * its ancestors passed the same test (frontier_levels < 0: any), and a
 * candidate is kept with probability keep_rate (times 65536 / particles for
 * larger nodes).  For each kept pair, features[16] and errors[12] are
 * written: the model inputs and, for orders 1..12, the largest error over the
 * batch's targets of the far-field velocity against the exact sum over the
 * node's particles, divided by u_ref (|du| / u_ref); ids (may be NULL)
 * receives [n][2] = (batch index, node index of lpm_tree_export).  Needs
 * moments of order >= 12 with moment_sign +1.  Host pointers; n_samples
 * receives the number written (<= max_samples). */
int lpm_tree_labels(lpm_ctx* ctx, double theta_max, double delta, double u_ref, int64_t batch_stride,
                    double keep_rate, int frontier_levels, uint64_t seed, int64_t max_samples,
                    const double* tx, const double* ty, const double* tz, int64_t nt,
                    float* features, float* errors, int32_t* ids, int64_t* n_samples);
/* Batch treecode with the order model deciding every pair that passes
 * maxRadius_n + r_b < theta_max |c_b - c_n|: class 0 (or a class above
 * max_order) descends, class q accepts with order min(q + margin, max_order).
 * u_ref scales the moment feature (the value used for the labels).  Needs
 * moments of order >= max_order (<= 12) with moment_sign +1 and a model. */
int lpm_tree_eval_learned(lpm_ctx* ctx, int prec, double theta_max, int margin, int max_order, double u_ref,
                          double delta, const double* tx, const double* ty, const double* tz, int64_t nt,
                          double* ux, double* uy, double* uz);
int lpm_tree_eval_learned_dev(lpm_ctx* ctx, int prec, double theta_max, int margin, int max_order,
                              double u_ref, double delta, const double* tx, const double* ty,
                              const double* tz, int64_t nt, double* ux, double* uy, double* uz);

/* ------------------------------------------------------------------ time stepping
 * Velocity solver used by the RK4 step. */
#define LPM_METHOD_DIRECT 0
#define LPM_METHOD_TREE 1          /* lpm_tree_eval: per-target acceptance */
#define LPM_METHOD_TREE_BATCH 2    /* lpm_tree_eval_batch (Cartesian series) */
#define LPM_METHOD_TREE_LEARNED 3  /* lpm_tree_eval_learned: theta = theta_max, order = max_order */
typedef struct lpm_solver_config {
    int method;              /* LPM_METHOD_DIRECT or LPM_METHOD_TREE */
    int prec;                /* LPM_F64 or LPM_F32C; LPM_F64F with LPM_METHOD_DIRECT / LPM_METHOD_TREE */
    int series;              /* tree: LPM_SERIES_CARTESIAN (moments use sign +1) or LPM_SERIES_REFERENCE */
    int order;               /* tree: series order p, |k| < p, in [1, 12] */
    int leaf_max;            /* tree: sources per leaf */
    double theta;            /* tree: multipole acceptance tolerance */
    double delta;            /* smoothing length (0 = singular kernel) */
    double root_box_min[3];  /* tree: root box; must contain every particle at every stage */
    double root_box_max[3];
    int margin;              /* learned: orders added to the model's class */
    double u_ref;            /* learned: velocity scale of the model's features */
} lpm_solver_config;
/* direct FP64, tree defaults p = 8, leaf 128, theta 0.5, Cartesian series, root box [-1.00101, 1.00101)^3 */
void lpm_solver_config_default(lpm_solver_config* cfg);
/* One classical RK4 step of the barotropic vorticity equation on the rotating
 * unit sphere for n Lagrangian particles (no counterpart in the reference C++,
 * whose tests only evaluate velocities).  Per stage the relative circulation is
 * gamma = (zeta_abs - 2 omega z) area, the velocity comes from the configured
 * solver, and stage/final positions are renormalized to |x| = 1 in FP64.
 * (x, y, z) are updated in place; zeta_abs and area are conserved.  Device
 * pointers; enqueued on the context stream. */
int lpm_rk4_step_dev(lpm_ctx* ctx, const lpm_solver_config* cfg, double dt, double omega, int64_t n,
                     double* x, double* y, double* z, const double* zeta_abs, const double* area);
/* nsteps RK4 steps on host arrays (copied in once, out once). */
int lpm_rk4_steps(lpm_ctx* ctx, const lpm_solver_config* cfg, double dt, double omega, int nsteps, int64_t n,
                  double* x, double* y, double* z, const double* zeta_abs, const double* area);

/* Barotropic vorticity step of a meshed particle set: the RK4 step above for
 * n_src vortex particles (mesh faces: zeta_abs[n_src], area[n_src]) followed
 * in (x, y, z) by n_tracer passive particles (mesh vertices) that move with
 * the velocity the n_src particles induce and carry no circulation.  Their
 * absolute vorticity is conserved along trajectories like the faces'.
 * No counterpart in the reference C++; its Fortran drivers advance faces and
 * vertices of a SphereBVE mesh with a SphereBVESolver (README.md:68-81).
 * Device pointers; enqueued on the context stream. */
int lpm_bve_step_dev(lpm_ctx* ctx, const lpm_solver_config* cfg, double dt, double omega, int64_t n_src,
                     int64_t n_tracer, double* x, double* y, double* z, const double* zeta_abs, const double* area);
/* nsteps BVE steps on host arrays (copied in once, out once). */
int lpm_bve_steps(lpm_ctx* ctx, const lpm_solver_config* cfg, double dt, double omega, int nsteps, int64_t n_src,
                  int64_t n_tracer, double* x, double* y, double* z, const double* zeta_abs, const double* area);

/* ------------------------------------------------------------------ mesh generation
 * Uniform refinement of a seed mesh on the GPU: the particle sets of
 * MeshedParticles(seed, level, ..., radius) (src/LpmMeshedParticles.cpp:12-21,
 * PolyMesh2d src/LpmPolyMesh2d.cpp:20-93 dividing faces by
 * src/LpmQuadFaces.cpp:14-146 / src/LpmTriFaces.cpp:13-142 and edges by
 * src/LpmEdges.cpp:99-120), in the same index order as the host builder of
 * lpm_mesh.h: all vertices, and the leaf faces in ascending face index
 * (centroid, area).  Vertices and centroids equal the host builder's bit for
 * bit; areas agree to a few units in the last place (device atan2 / tan).
 * seed: LPM_SEED_ICOS_TRI (0) or LPM_SEED_CUBED (1); level 0..12. */
#ifndef LPM_SEED_ICOS_TRI
#define LPM_SEED_ICOS_TRI 0  /* as in lpm_mesh.h */
#define LPM_SEED_CUBED 1
#endif
/* vertex, face (parents included) and leaf-face counts of seed / level
 * (src/LpmMeshSeed.cpp:233-284); LPM_ERR_INVALID_ARG out of range */
int lpm_mesh_counts(int seed, int level, int64_t* n_vertices, int64_t* n_faces, int64_t* n_leaf_faces);
/* Device pointers: vx, vy, vz[n_vertices]; cx, cy, cz, area[n_leaf_faces];
 * leaf_vertices[n_leaf_faces * (3 or 4)] the corners of each leaf face
 * counter-clockwise (may be NULL).  Enqueued on the context stream. */
int lpm_mesh_generate_dev(lpm_ctx* ctx, int seed, int level, double radius, double* vx, double* vy, double* vz,
                          double* cx, double* cy, double* cz, double* area, int64_t* leaf_vertices);
/* The same into host arrays (returns when they are filled). */
int lpm_mesh_generate(lpm_ctx* ctx, int seed, int level, double radius, double* vx, double* vy, double* vz,
                      double* cx, double* cy, double* cz, double* area, int64_t* leaf_vertices);

#ifdef __cplusplus
}
#endif

#endif /* LPM_GPU_H */
