/*
 * td3.h - the C ABI of libtd3_hip.so, the gfx950 (MI355X) backend of the
 * field-evaluation path of tree-d-3d (direct sums, flat quadtree, Barnes-Hut walk, FMM).
 *
 * Conventions (shared by every entry point):
 *  - Array arguments are DEVICE pointers (HIP global memory), e.g. torch tensors'
 *    data_ptr(), in the reference's layouts: points (2, N) component-major float64,
 *    per-node arrays node-major, indices int64 unless stated.
 *  - `stream` is a hipStream_t passed as void*; every kernel is enqueued on it.  Entry
 *    points that must size outputs read a few counters back and synchronise the stream
 *    (stated per entry).
 *  - The caller allocates every output and workspace; the library never allocates or
 *    frees device memory.  `*_workspace_bytes` queries give workspace sizes.
 *  - Return codes: TD3_OK (0); TD3_ERR_CAPACITY (-2, a caller buffer is too small: the
 *    needed sizes come back through the entry's out-parameter); TD3_ERR_ARG (-3);
 *    TD3_ERR_HIP (-4).  td3_last_error() describes the last failure of this thread.
 *  - Kernel failures are reported in a device `status` int64[1] (zero on entry), as the
 *    reference's kernels do: > 0 is 1 + the smallest index of a target that met an
 *    unsoftened coincident charge (the reference raises ValueError), < 0 a traversal
 *    stack or depth overflow (RuntimeError).
 *  - Reference interfaces cited as src/tree_d_3d/<file>:<line>.
 */
#ifndef TD3_H
#define TD3_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TD3_ABI_VERSION 1

#define TD3_OK 0
#define TD3_ERR_CAPACITY (-2)
#define TD3_ERR_ARG (-3)
#define TD3_ERR_HIP (-4)

/* kernel */
#define TD3_COULOMB 0
#define TD3_LOG 1

/* precision: all-FP64, or FP32 arithmetic with FP64 accumulation and FP64 tree data */
#define TD3_FP64 0
#define TD3_MIXED 1

int td3_abi_version(void);
const char *td3_last_error(void);

/*
 * Direct (all-pairs) potential and field at M targets from N sources, without k_e.
 * Replaces the kernels direct_sum / direct_log_sum
 * (kernels/direct_sum/direct_sum_cuda.cu:7-44, kernels/direct_log_sum/direct_log_sum_cuda.cu:7-42,
 * called by flat_tree.py:153-158): coulomb phi = sum q/sqrt(r^2+eps^2), E = sum q d/(r^2+eps^2)^1.5;
 * log phi = -sum q ln(r^2+eps^2), E = 2 sum q d/(r^2+eps^2).  Source exclude[i] and zero
 * charges are skipped.  targets (2, M), exclude (M,), pos (2, N), charge (N,),
 * phi (M,), field (2, M) written (not accumulated).  workspace: td3_direct_workspace_bytes.
 */
size_t td3_direct_workspace_bytes(int64_t m, int64_t n);
int td3_direct(int kernel, int precision, const double *targets, int64_t m,
               const int64_t *exclude, const double *pos, int64_t n, const double *charge,
               double softening, double *phi, double *field, int64_t *status, void *workspace,
               size_t workspace_bytes, void *stream);

/*
 * td3_direct for callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc); no counterpart in the reference.  Same
 * arguments and results as td3_direct (bit for bit), without the workspace: the chunk
 * partials live in a scratch array of the library, one per GPU, of
 * td3_direct_noalloc_scratch_bytes() bytes (16 MiB), held until the process ends.  A call
 * whose td3_direct_workspace_bytes(m, n) exceeds it returns TD3_ERR_CAPACITY before any
 * launch.  Calls no allocator.  Every call on a GPU uses that GPU's one scratch array:
 * calls that overlap on one GPU (other streams or threads) race.
 */
size_t td3_direct_noalloc_scratch_bytes(void);
int td3_direct_noalloc(int kernel, int precision, const double *targets, int64_t m,
                       const int64_t *exclude, const double *pos, int64_t n,
                       const double *charge, double softening, double *phi, double *field,
                       int64_t *status, void *stream);

/*
 * Flat quadtree build, bit-identical to the reference's (the arrays of
 * kernels/quadtree_build/quadtree_build_pyccel.py:12-137 and of the per-level GPU build
 * flat_tree.py:532-625 via kernels/quadtree_quadrants, kernels/quadtree_scatter).
 * pos (2, N) in input order; root (xmid, ymid, half).  Node arrays hold `capacity`
 * entries: node_center (C, 2), node_half, node_start, node_end, node_child, node_depth;
 * levels int64[67] (levels[d]:levels[d+1] are the nodes of depth d); perm (N,) int64;
 * sorted_pos (2, N) = pos[:, perm] (may be NULL); node_info (C,) int4 {start, end, child,
 * depth} for the other kernels (may be NULL).  out_sizes (HOST int64[3]) gets
 * {node_count, n_levels, needed capacity}; levels_host (HOST int64[67], may be NULL) the
 * same level starts as levels, so that the caller needs no read-back of its own.  Returns
 * TD3_ERR_CAPACITY when capacity <
 * node_count (call again with out_sizes[2]), and TD3_ERR_ARG with the reference's
 * messages ("Particle position must be finite", "Particle lies outside tree bounds",
 * flat_tree.py:785-794) for a position that is not finite or lies outside the closed root
 * square.  Synchronises the stream once (one read-back of the split counts and of the
 * positions' check).  Supports N < 2^31 and trees of depth <= 64.
 */
size_t td3_tree_workspace_bytes(int64_t n, int64_t capacity);

/*
 * The trees of S independent systems at once (an ensemble; one tree per
 * ElectrostaticSimulation in the reference, simulation.py:641-667).  pos (2, N) holds the
 * systems one after another; particle_system (N,) int32 (non-decreasing) and system_start
 * (S + 1,) int64 (device) say which; roots (S, 3) f64 (device) are the systems' root
 * squares (xmid, ymid, half).  Each system's tree is the one td3_tree_build makes from
 * its own particles and root (same cells, same input order in the leaves); node s is the
 * root of system s, and below the roots the numbering is breadth first over all systems
 * (levels[d]:levels[d+1] are the depth-d nodes of every system).  perm and node ranges
 * index the concatenated particles.  Otherwise as td3_tree_build (same workspace size).
 * No counterpart in the reference (its systems are built one by one).
 */
int td3_forest_build(const double *pos, int64_t n, int64_t systems,
                     const int32_t *particle_system, const int64_t *system_start,
                     const double *roots, int64_t leaf_capacity, int64_t max_depth,
                     int64_t capacity, double *node_center, double *node_half,
                     int64_t *node_start, int64_t *node_end, int64_t *node_child,
                     int64_t *node_depth, int64_t *levels, int64_t *perm, double *sorted_pos,
                     void *node_info, void *workspace, size_t workspace_bytes,
                     int64_t *out_sizes, int64_t *levels_host, void *stream);
int td3_tree_build(const double *pos, int64_t n, double xmid, double ymid, double half,
                   int64_t leaf_capacity, int64_t max_depth, int64_t capacity, double *node_center,
                   double *node_half, int64_t *node_start, int64_t *node_end, int64_t *node_child,
                   int64_t *node_depth, int64_t *levels, int64_t *perm, double *sorted_pos,
                   void *node_info, void *workspace, size_t workspace_bytes, int64_t *out_sizes,
                   int64_t *levels_host, void *stream);

/*
 * td3_forest_build without a host read-back or a memset, so that an ensemble step can be
 * recorded into a HIP graph and replayed (simulation.py:800-835 steps one system with a
 * host-synchronous build; no counterpart in the reference).  Node arrays, node_info and
 * levels (int64[67], device) are at `capacity`.  The keys hold key_levels levels (at most
 * the key's level count): a caller that saw the tree's depth passes a few more levels than
 * it had, so no stage is chosen on the host.  sizes (int64[6], device): node count, level
 * count, the needed capacity when the nodes overflow `capacity` (-1 when the key is too
 * short, else 0), the largest such need of any call so far, 1 when a cell would split
 * below the key's levels (else 0), and 1 when that happened in any call so far (the "so
 * far" entries are never cleared here: the caller zeroes them and reads them after its
 * replays).  levels[d] past the deepest level equal the node count (empty levels).  On an
 * overflow or a short key the forest is the roots alone, each a leaf, so the walk then
 * sums every system directly.  Same tree as td3_forest_build otherwise (no cell is deeper
 * than key_levels); same workspace size.
 */
int td3_forest_build_nosync(const double *pos, int64_t n, int64_t systems,
                            const int32_t *particle_system, const int64_t *system_start,
                            const double *roots, int64_t leaf_capacity, int64_t max_depth,
                            int64_t capacity, int64_t key_levels, double *node_center,
                            double *node_half, int64_t *node_start, int64_t *node_end,
                            int64_t *node_child, int64_t *node_depth, int64_t *levels,
                            int64_t *perm, double *sorted_pos, void *node_info, void *workspace,
                            size_t workspace_bytes, int64_t *sizes, void *stream);

/* dst[r, k] = src[r * ld_src + perm[k]] for r < rows (<= 8), k < n: tree-order copies
 * (flat_tree.py:879-889 _sorted_sources). */
int td3_permute(const double *src, int64_t rows, int64_t ld_src, const int64_t *perm, int64_t n,
                double *dst, void *stream);

/*
 * Cell moments from the sources in tree order (sorted_pos (2, N), sorted_charge (N,)).
 * Replaces kernels/quadtree_moments, kernels/log_quadtree_moments (flat_tree.py:663-744)
 * with the upward pass of kernels/quadtree_moments_upward/..._pyccel.py:13-83 and
 * kernels/log_moments_upward/..._pyccel.py:14-93.  coulomb: node_q, node_qabs (C,),
 * node_p (C, 2), node_s (C, 3) raw (xx, xy, yy) about the cell centre.  log: node_qabs,
 * moment_re/moment_im (C, order+1) with a_n = sum q (z - c)^n (node_q/node_p/node_s
 * untouched).  levels_host: HOST copy of levels (n_levels + 1 entries); binomial (bsz,
 * bsz) device table C(n, k) with bsz >= order + 1.
 */
int td3_moments(int kernel, int64_t order, const double *sorted_pos, const double *sorted_charge,
                int64_t n, const double *node_center, const void *node_info, int64_t node_count,
                const int64_t *levels_host, int64_t n_levels, const double *binomial, int64_t bsz,
                double *node_q, double *node_qabs, double *node_p, double *node_s,
                double *moment_re, double *moment_im, void *stream);

/*
 * Coulomb moments of td3_moments (same values, bit for bit) with the level starts read on
 * the device (levels int64[67] from td3_forest_build_nosync): a fixed grid per level for
 * the depths max_depth - 1 .. 0 (max_depth: a bound on the tree's depth, such as the
 * build's key_levels), no host read-back, so it can be recorded into a HIP graph.  Arrays at `capacity` nodes.  Replaces kernels/quadtree_moments
 * (flat_tree.py:663-744) for a captured ensemble step.
 */
int td3_coulomb_moments_nosync(const double *sorted_pos, const double *sorted_charge, int64_t n,
                               const double *node_center, const void *node_info,
                               int64_t capacity, const int64_t *levels, int64_t max_depth,
                               double *node_q, double *node_qabs, double *node_p,
                               double *node_s, void *stream);

/*
 * Locality order of M targets (2, M): a stable sort by the reference's 16-bit-per-axis
 * Morton key ((t - lo) * scale, x in the high bit of each pair; flat_tree.py:891-912).
 * order (M,) int64 receives the target indices.
 */
size_t td3_target_order_workspace_bytes(int64_t m);
int td3_target_order(const double *targets, int64_t m, double lox, double loy, double scale,
                     int64_t *order, void *workspace, size_t workspace_bytes, void *stream);

/*
 * Barnes-Hut walk at M targets, without k_e.  Replaces kernels/evaluate_quadtree
 * (evaluate_quadtree_cuda.cu:9-107) and kernels/evaluate_log_quadtree
 * (evaluate_log_quadtree_cuda.cu:9-104) behind flat_tree.py:988-1061, with the per-target
 * semantics of kernels/evaluate_groups/evaluate_groups_pyccel.py:16-297: cells with
 * node_qabs == 0 skipped; a cell not containing the target is accepted when
 * 2h < theta |r - c| (coulomb, quadrupole term with rho = sqrt(D^2 + eps^2)) or
 * 4h^2 < theta^2 |r - c|^2 (log, order-p series, unsoftened); leaves sum their particles
 * directly (softened), skipping exclude[i] and zero charges.  order (M,) is the target
 * order of the wavefronts (td3_target_order); sorted_pos / sorted_charge / perm are the
 * sources in tree order.  target_root (M,) int32, by target index, is the node the walk
 * of each target starts from: NULL for node 0 (one tree); in a forest (td3_forest_build)
 * the root of the target's system, so that systems never see each other.  phi (M,),
 * field (2, M) written in target order.
 */
int td3_walk(int kernel, int precision, const double *targets, int64_t m, const int64_t *order,
             const int64_t *exclude, const double *sorted_pos, const double *sorted_charge,
             const int64_t *perm, int64_t n, const double *node_center, const double *node_half,
             const void *node_info, const double *node_qabs, const double *node_q,
             const double *node_p, const double *node_s, const double *moment_re,
             const double *moment_im, int64_t order_p, int64_t tree_depth,
             const int32_t *target_root, double theta, double softening, int with_potential,
             double *phi, double *field, int64_t *status, void *stream);

/*
 * td3_walk (same arguments, same outputs, the same walk) that also adds up its work into
 * counts[0..3] (uint64, added to: the caller zeroes them): cells accepted per target, source
 * terms per target, cells accepted per wavefront (a cell some lane of the wavefront accepts,
 * whose data the wavefront reads once) and sources read per wavefront (every source of every
 * leaf a wavefront visits).  A separate launch for measurement (the walk's flops and its
 * bytes' lower bound); the reference has no counterpart.
 */
int td3_walk_counts(int kernel, int precision, const double *targets, int64_t m,
                    const int64_t *order, const int64_t *exclude, const double *sorted_pos,
                    const double *sorted_charge, const int64_t *perm, int64_t n,
                    const double *node_center, const double *node_half, const void *node_info,
                    const double *node_qabs, const double *node_q, const double *node_p,
                    const double *node_s, const double *moment_re, const double *moment_im,
                    int64_t order_p, int64_t tree_depth, const int32_t *target_root,
                    double theta, double softening, int with_potential, double *phi,
                    double *field, uint64_t *counts, int64_t *status, void *stream);

/*
 * The walk's launch shape: wavefronts (64 targets each) per block, 1, 2, 4 (default) or 8,
 * for every later td3_walk of the process; 0 only asks.  Returns the previous value.  No
 * counterpart in the reference: the knob of this build's tuner (Ensemble.autotune times
 * the choices on the ensemble's own step and keeps the fastest).
 */
int td3_walk_waves(int waves);

/*
 * Dual-tree traversal of one tree with itself (the FMM's interaction lists).  Replaces
 * flat_tree.py:1318-1358 and kernels/fmm_traverse (fmm_traverse_cuda.cu:14-97) with the
 * rules of kernels/fmm_traverse/fmm_traverse_pyccel.py:15-106.  Writes the mutual M2L
 * pairs m2l (cap_m2l, 2) int32 and the P2P leaf pairs p2p (cap_p2p, 2) int32, in no
 * particular order (the pair SETS equal the reference's).  out_counts (HOST int64[4]) gets
 * {n_m2l, n_p2p, largest frontier, overflow}; TD3_ERR_CAPACITY when a buffer was too small
 * (retry with out_counts[0..2] or larger).  Synchronises the stream once.
 */
size_t td3_traverse_workspace_bytes(int64_t cap_front);
int td3_traverse(const double *node_center, const double *node_half, const void *node_info,
                 int64_t node_count, int64_t tree_depth, double theta, int64_t cap_m2l,
                 int64_t cap_p2p, int64_t cap_front, int32_t *m2l, int32_t *p2p, void *workspace,
                 size_t workspace_bytes, int64_t *out_counts, void *stream);

/*
 * The FMM's interaction rows by node, straight from the traversal rules of
 * kernels/fmm_traverse/fmm_traverse_pyccel.py:15-106 (flat_tree.py:1318-1358) without
 * the pair lists: level by level, every node (a wavefront) walks the pairs it inherits
 * from its parent (one lane per pair, depth first over the other cell) into slots of
 * `width` entries of its own (M2L row, P2P row, pairs handed to its children); one scan
 * per row kind and a packing copy then place the rows.  The rows hold the same pairs as
 * td3_traverse + td3_lists (M2L both ways, mode 0; P2P both ways with a self pair once,
 * mode 1), each row in a fixed walk order (not sorted); m2l_dst (may be NULL) gets the
 * row's node of every M2L entry.  levels_host as td3_moments.  out_counts (HOST
 * int64[4]): M2L entries, P2P entries, non-empty leaves (self pairs), overflow (0; else
 * the slot width a node needs, or 1 when only cap_m2l / cap_p2p were too small).
 * TD3_ERR_CAPACITY asks for larger cap_m2l / cap_p2p / width.  Any depth (the walks follow
 * child and parent links, no stack).  Synchronises the stream once, unless device_counts
 * (DEVICE int64[4], may be NULL) is given: then the same four counts are written on the
 * device, with no read-back and no synchronisation (out_counts may then be NULL), and
 * on overflow every row is left empty, so the rows can be evaluated safely before the
 * caller reads the counts and builds again with larger capacities.
 */
size_t td3_rows_workspace_bytes(const int64_t *levels_host, int64_t n_levels, int64_t width);
int td3_rows(const double *node_center, const double *node_half, const void *node_info,
             const int64_t *levels_host, int64_t n_levels, double theta, int64_t cap_m2l,
             int64_t cap_p2p, int64_t width, uint32_t *m2l_row, int32_t *m2l_col,
             int32_t *m2l_dst, uint32_t *p2p_row, int32_t *p2p_col, void *workspace,
             size_t workspace_bytes, int64_t *out_counts, int64_t *device_counts,
             void *stream);

/*
 * Rows by node of a pair list (kernels/pair_csr/pair_csr_pyccel.py:12-45): mode 0 lists
 * b under a and a under b; mode 1 the same with a self pair once (the P2P leaves); mode 2
 * one way, b under a (fmm_at's target rows).  row_start (node_count + 1) uint32, col
 * int32 sorted ascending within each row, dst (may be NULL) the row of each entry.
 * col_count: the number of nodes the columns index (node_count for modes 0 and 1, the
 * source tree's for mode 2).  The directed entries are radix-sorted as (row, column)
 * keys, so the rows and their order come with no atomics; the workspace
 * (td3_lists_workspace_bytes) holds the keys.  No read-back.
 */
size_t td3_lists_workspace_bytes(int64_t node_count, int64_t count, int mode);
int td3_lists(const int32_t *pairs, int64_t count, int mode, int64_t node_count,
              int64_t col_count, uint32_t *row_start, int32_t *col, int32_t *dst,
              void *workspace, size_t workspace_bytes, void *stream);

/*
 * Radius-scaled moments A_k = a_k / r^k, r = sqrt(2) node_half (FP64 -> double or float
 * by precision), for the M2L.
 */
int td3_scale_moments(int precision, const double *moment_re, const double *moment_im,
                      const double *node_half, int64_t node_count, int64_t order, void *out_re,
                      void *out_im, void *stream);

/*
 * M2L into the scaled locals B_l = b_l r^l of the destination rows [node_lo, node_hi).
 * Replaces kernels/fmm_m2l (fmm_m2l_cuda.cu:12-80) / kernels/fmm_m2l_host
 * (fmm_m2l_host_pyccel.py:16-105): b_0 += a_0 ln t - s_0, b_l += (-1)^(l+1) s_l w^l with
 * s_l = sum_k coeff[k, l] a_k w^k, w = 1/t, t = c_dst - c_src (ln t principal).  col/dst/
 * row_start are td3_lists rows of the M2L pairs; coeff (order+1)^2 device table
 * (flat_tree.py:234-245, either truncation); scaled_re/im from td3_scale_moments of the
 * same precision; dst_* are the destination tree's cells (the rows), src_* the source
 * tree's (the columns; the same arrays for fmm, the target tree for fmm_at).  Every row of
 * the range is WRITTEN (not accumulated): zeros for a row without pairs.  FP64 up to
 * order 15 with workspace NULL runs on the matrix cores, a wavefront per row; otherwise a
 * lane per pair, with the workspace (td3_m2l_workspace_bytes of at least the range's pair
 * count) holding the partial sums of rows that span several wavefronts.  The fast FP32
 * path up to order 15 is td3_m2l_f32.  No host read-back.
 */
size_t td3_m2l_workspace_bytes(int64_t pair_count, int64_t order, int precision);
int td3_m2l(int precision, int64_t order, const int32_t *col, const int32_t *dst,
            const uint32_t *row_start, int64_t node_lo, int64_t node_hi,
            const double *dst_center, const double *dst_half, const double *src_center,
            const double *src_half, const void *scaled_re, const void *scaled_im,
            const double *coeff, double *local_re, double *local_im, void *workspace,
            size_t workspace_bytes, void *stream);

/*
 * Source records of the FP32 M2L (order 1..15), one per cell of td3_m2l_record_bytes(order)
 * bytes (a multiple of 32): the cell's centre (x, y) as FP64, then A_k = a_k / r^k
 * (r = sqrt(2) h) as FP32 (re, im) for k <= order, except that the slot of Im A_0 holds r
 * (a_0 is a net charge: real), then zeros.  A lane of the M2L loads its four coefficients
 * in two 16-byte reads and the cell's geometry from the same record.  No counterpart in the
 * reference (its M2L reads the FP64 moments directly, fmm_m2l_cuda.cu:12-80); records is
 * DEVICE memory of node_count * td3_m2l_record_bytes(order) bytes.  No read-back.
 */
size_t td3_m2l_record_bytes(int64_t order);
int td3_m2l_records(int64_t order, const double *moment_re, const double *moment_im,
                    const double *node_center, const double *node_half, int64_t node_count,
                    void *records, void *stream);

/*
 * td3_m2l in FP32 (order 1..15) on the matrix cores from td3_m2l_records of the source tree:
 * the same rows (col, row_start), coefficients, destination cells and outputs (scaled FP64
 * locals, every row of [node_lo, node_hi) WRITTEN, zeros without pairs).  Replaces
 * kernels/fmm_m2l (fmm_m2l_cuda.cu:12-80) / kernels/fmm_m2l_host (fmm_m2l_host_pyccel.py:
 * 16-105) below the FP64 top levels of the mixed-precision FMM.  A wavefront takes eight
 * consecutive rows and runs their pairs through the matrix product 16 at a time; every row
 * is summed in one fixed order.  The imaginary part of a_0 ln t is dropped (it reaches
 * neither the potential -2 Re W, nor the field W', nor, through L2L, any other
 * coefficient).  No workspace, no read-back.
 */
int td3_m2l_f32(int64_t order, const int32_t *col, const uint32_t *row_start,
                int64_t node_lo, int64_t node_hi, const double *dst_center,
                const double *dst_half, const void *records, const double *coeff,
                double *local_re, double *local_im, void *stream);

/*
 * L2L down the tree on scaled locals, level by level from the root (flat_tree.py:1540-1552,
 * kernels/fmm_downward/fmm_downward_cuda.cu:10-51).  levels_host: HOST levels array.
 */
int td3_l2l(int64_t order, const double *node_center, const double *node_half,
            const void *node_info, const int64_t *levels_host, int64_t n_levels,
            const double *binomial, int64_t bsz, double *local_re, double *local_im,
            void *stream);

/* b_l = B_l / r^l: the reference's unscaled locals, in place. */
int td3_unscale_locals(const double *node_half, int64_t node_count, int64_t order,
                       double *local_re, double *local_im, void *stream);

/*
 * L2P + P2P at every source (flat_tree.py:1580-1600, kernels/fmm_evaluate
 * fmm_evaluate_cuda.cu:11-87; kernels/fmm_evaluate_host/fmm_evaluate_host_pyccel.py:14-100):
 * one lane per particle, owner-computes over its leaf's P2P row (td3_lists of the P2P
 * pairs, mode 1); leaf_of from td3_leaf_index.  Own charge excluded; softening on the near pairs only.  kernel TD3_LOG:
 * scaled complex locals in local_re/local_im; TD3_COULOMB: Cartesian locals (C, nc) in
 * local_re (cart_evaluate_host_pyccel.py:16-110).  phi (N,), field (2, N) written in input
 * order (phi zero when with_potential is 0).
 */
int td3_fmm_eval(int kernel, int precision, int64_t order, const double *sorted_pos,
                 const double *sorted_charge, const int64_t *perm, int64_t n,
                 const double *node_center, const double *node_half, const void *node_info,
                 int64_t node_count, const double *local_re, const double *local_im,
                 const uint32_t *p2p_row, const int32_t *p2p_col, const int32_t *leaf_of,
                 const void *sources4, double softening, int with_potential, double *phi,
                 double *field, int64_t *status, void *stream);

/*
 * td3_fmm_eval in two launches, so that the near field can run on a second stream while
 * the far field is translated: td3_fmm_near sums the P2P row of the sources [first, first
 * + count) of tree order (the same terms, order and status as td3_fmm_eval) into their
 * columns of near (3, N) float64 in tree order (potential, field x, field y), so that the
 * near field can be split into launches; td3_fmm_far then adds each particle's L2P and
 * writes phi (N,) and field (2, N) in input order, bit for bit td3_fmm_eval's.
 */
int td3_fmm_near(int kernel, int precision, const double *sorted_pos, const int64_t *perm,
                 int64_t n, int64_t first, int64_t count, const void *node_info,
                 const uint32_t *p2p_row, const int32_t *p2p_col, const int32_t *leaf_of,
                 const void *sources4, double softening, int with_potential, double *near,
                 int64_t *status, void *stream);
int td3_fmm_far(int kernel, int64_t order, const double *sorted_pos, const int64_t *perm,
                int64_t n, const double *node_center, const double *node_half,
                const double *local_re, const double *local_im, const int32_t *leaf_of,
                const double *near, int with_potential, double *phi, double *field,
                void *stream);

/*
 * Log-kernel moments of sources that carry dipoles as well as charges: td3_moments
 * (kernel TD3_LOG) plus a_n -= n D (z - c)^(n-1) per source, the moments of a dipole's
 * potential -2 D.(z - z_s) / |z - z_s|^2 (sorted_dipole (2, N): x components, then y, in
 * tree order).  No counterpart in the reference (which has no gradients): the source
 * side of the transposed FMM for losses on fields (their gradient is a field of dipoles).
 */
int td3_log_moments_dipole(int64_t order, const double *sorted_pos, const double *sorted_charge,
                           const double *sorted_dipole, int64_t n, const double *node_center,
                           const void *node_info, int64_t node_count, const int64_t *levels_host,
                           int64_t n_levels, const double *binomial, int64_t bsz,
                           double *node_qabs, double *moment_re, double *moment_im, void *stream);

/*
 * td3_fmm_eval for the log kernel in FP64 with the adjoint's extras; no counterpart in the
 * reference.  sorted_dipole (2, N, tree order, may be NULL): the sources' dipoles, near
 * field -2 D.d / (r^2 + eps^2) in phi and 2 D / (r^2 + eps^2) - 4 (D.d) d / (r^2 + eps^2)^2
 * in field (d = target - source; the far field comes with td3_log_moments_dipole's
 * moments).  jacobian (3, N, may be NULL; not with dipoles): rows xx, xy, yy of dE/dx at
 * every target, far field 2 W'' and near field 2 q (I / (r^2 + eps^2) - 2 d d^T /
 * (r^2 + eps^2)^2) (the second-derivative kernel of the adjoint of a loss on fields).
 */
int td3_log_fmm_eval_ext(int64_t order, const double *sorted_pos, const int64_t *perm, int64_t n,
                         const double *node_center, const double *node_half,
                         const void *node_info, const double *local_re, const double *local_im,
                         const uint32_t *p2p_row, const int32_t *p2p_col, const int32_t *leaf_of,
                         const void *sources4, const double *sorted_dipole, double softening,
                         double *phi, double *field, double *jacobian, int64_t *status,
                         void *stream);

/*
 * td3_fmm_at_eval for the log kernel with td3_log_fmm_eval_ext's dipoles and Jacobian
 * (jacobian (3, M) in target order); no counterpart in the reference.  t_info (t_count,)
 * int4 {start, end, child, depth} are the target tree's cells: a wavefront per target
 * leaf, whose targets split the leaf's P2P row between lanes.
 */
int td3_log_fmm_at_eval_ext(int64_t order, const double *sorted_targets,
                            const int64_t *target_perm, const int64_t *exclude, int64_t m,
                            const double *t_center, const double *t_half, const void *t_info,
                            int64_t t_count, const double *local_re, const double *local_im,
                            const uint32_t *p2p_row, const int32_t *p2p_col,
                            const int64_t *perm, int64_t n, const void *s_info,
                            const void *s_sources4, const double *sorted_dipole,
                            double softening, double *phi, double *field, double *jacobian,
                            int64_t *status, void *stream);

/*
 * (x, y, q, 0) float64 per sorted source (N x 32 bytes): the near field's source records
 * (td3_fmm_eval's sources4, td3_fmm_at_eval's s_sources4).  No counterpart in the
 * reference (it reads the (2, N) positions and the charges apart).
 */
int td3_pack_sources(const double *sorted_pos, const double *sorted_charge, int64_t n,
                     void *sources4, void *stream);

/*
 * leaf_of[k] (int32, N) = the leaf holding sorted particle k: the per-particle map the
 * evaluation kernels read (one lane per particle).  No counterpart in the reference (its
 * CPU loops run leaf by leaf).
 */
int td3_leaf_index(const void *node_info, int64_t node_count, int32_t *leaf_of, void *stream);

/*
 * Dual traversal of a target tree against a source tree (fmm_at's lists): replaces
 * flat_tree.py:1086-1122 (_dual_lists) and kernels/dual_traverse/dual_traverse_pyccel.py:14-86
 * (same pair sets; the reference walks depth first on the CPU).  m2l and p2p are (target
 * cell, source cell) pairs, one way; out_counts as td3_traverse.
 */
int td3_dual_traverse(const double *t_center, const double *t_half, const void *t_info,
                      int64_t t_count, int64_t t_depth, const double *s_center,
                      const double *s_half, const void *s_info, int64_t s_count, int64_t s_depth,
                      double theta, int64_t cap_m2l, int64_t cap_p2p, int64_t cap_front,
                      int32_t *m2l, int32_t *p2p, void *workspace, size_t workspace_bytes,
                      int64_t *out_counts, void *stream);

/*
 * fmm_at's evaluation (flat_tree.py:1267-1316, kernels/log_targets_host/log_targets_host_pyccel.py:14-94):
 * per target leaf of the target tree, L2P of its scaled local expansion plus exact sums
 * over its P2P source leaves (td3_lists mode 2), skipping uncharged sources and source
 * exclude[i]; a coincident charge reports the target.  sorted_targets (2, M) and
 * target_perm describe the target tree; sorted_pos / sorted_charge / perm / s_info the
 * sources.  phi (M,), field (2, M) written in target order.
 */
int td3_fmm_at_eval(int kernel, int precision, int64_t order, const double *sorted_targets,
                    const int64_t *target_perm, const int64_t *exclude, int64_t m,
                    const double *t_center, const double *t_half, const void *t_info,
                    int64_t t_count, const int32_t *t_leaf_of, const double *local_re,
                    const double *local_im, const uint32_t *p2p_row, const int32_t *p2p_col,
                    const double *sorted_pos, const double *sorted_charge, const int64_t *perm,
                    int64_t n, const void *s_info, const void *s_sources4, double softening,
                    int with_potential, double *phi, double *field, int64_t *status,
                    void *stream);

/*
 * The Coulomb (Cartesian) FMM of the softened kernel (flat_tree.py:1360-1428, _cartesian.py,
 * kernels/cart_*_pyccel.py; CPU-only in the reference).  td3_cart_moments: moments (C, nc)
 * with nc = (p+1)(p+2)/2, m_g = sum q (r - c)^g by P2M + M2M, and the primed moments
 * m'_a = (-1)^|a| m_a / a! the M2L reads.  td3_cart_m2l: T_b = sum_a (-1)^|a| C(a+b, a)
 * m_a A_{a+b}(c_dst - c_src) (A the Taylor coefficients of the softened kernel) into the
 * rows [node_lo, node_hi) of local (written).  td3_cart_l2l: shift down the tree.
 */
int td3_cart_moments(int64_t order, const double *sorted_pos, const double *sorted_charge,
                     int64_t n, const double *node_center, const void *node_info,
                     int64_t node_count, const int64_t *levels_host, int64_t n_levels,
                     const double *binomial, int64_t bsz, double *moments, double *mprime,
                     void *stream);
size_t td3_cart_m2l_workspace_bytes(int64_t pair_count, int64_t order, int64_t node_count);
int td3_cart_m2l(int64_t order, const int32_t *col, const int32_t *dst, const uint32_t *row_start,
                 int64_t node_lo, int64_t node_hi, const double *dst_center,
                 const double *src_center, const double *mprime, double softening, double *local,
                 void *workspace, size_t workspace_bytes, void *stream);

/*
 * The mixed-precision Cartesian M2L (order 1..12), no counterpart in the reference (its
 * Coulomb FMM is FP64 on the CPU): the translation of td3_cart_m2l on radius-scaled
 * quantities in FP32.  The softened kernel is homogeneous, A'_g(t; eps) = h^-(|g|+1)
 * A'_g(t/h; eps/h), so with h the destination's half size (dst_half) and mscaled_a =
 * m'_a / h_src^|a| (td3_cart_scale_prime of the source tree, FP32):
 * T_b = (1 / b!) h^-(|b|+1) sum_a mscaled_a (h_src / h)^|a| A'_{a+b}(t/h; eps/h), every
 * term bounded by the pair's separation.  The recurrence, products and per-pair sums in
 * FP32; sums over a destination's pairs, the scaling and local (written as td3_cart_m2l
 * writes it) in FP64.  Same rows, workspace and read-back as td3_cart_m2l.
 */
int td3_cart_scale_prime(int64_t order, const double *mprime, const double *node_half,
                         int64_t node_count, float *out, void *stream);
int td3_cart_m2l_f32(int64_t order, const int32_t *col, const int32_t *dst,
                     const uint32_t *row_start, int64_t node_lo, int64_t node_hi,
                     const double *dst_center, const double *dst_half,
                     const double *src_center, const double *src_half, const float *mscaled,
                     double softening, double *local, void *workspace,
                     size_t workspace_bytes, void *stream);
int td3_cart_l2l(int64_t order, const double *node_center, const void *node_info,
                 const int64_t *levels_host, int64_t n_levels, const double *binomial,
                 int64_t bsz, double *local, void *stream);

/*
 * Particle push of the simulation step (simulation.py:800-835).  first_half: the electric
 * half-kick v += dt/(2m) qE, then for mobile particles the exact drift
 * r += dt (sinc(a) v + (a/2) sinc^2(a/2) (v_y, -v_x)) and rotation by a = (q/m) B_z dt;
 * otherwise the second half-kick and forces = qE + q B_z (v_y, -v_x).  electric holds qE
 * (k_e included); positions, velocities, forces (2, N); fixed (N,) uint8 (not moved).
 */
int td3_push(int first_half, int64_t n, double *positions, double *velocities, double *forces,
             const double *electric, const double *masses, const double *charges,
             const uint8_t *fixed, double dt, double b_z, void *stream);

/*
 * The roots of an ensemble's systems (simulation.py:625-639 per system): a root square
 * that no longer holds its system's particles becomes the square around them and the old
 * root (centre lo/2 + hi/2, half max(hi/2 - lo/2) (1 + 1e-12)).  positions (2, N), the
 * systems one after another (system_start (S + 1) int64); roots (S, 3) updated in place;
 * *grown (device u32) counts the roots that grew.  Non-finite positions leave NaN in the
 * bounds: the caller checks them.  No counterpart in the reference.
 */
int td3_system_bounds(const double *positions, int64_t n, int64_t systems,
                      const int64_t *system_start, double *roots, uint32_t *grown, void *stream);

/*
 * Close-pair softening of line charges (the grounded strip's correction in the physical
 * plane, strip.py:168-193 and :288-323; replaces the host k-d tree pair search plus
 * kernels/soften_pairs/soften_pairs_cuda.cu:9-42).  For every target t of targets (2, M),
 * every source s of the tree (sorted_pos (2, N), sorted_charge, perm, node arrays of
 * td3_tree_build) with |t - s| <= radius and perm index != exclude[t] adds the softened
 * minus bare line-charge kernel: phi[t] += -q ln((r^2 + eps^2) / r^2) (when
 * with_potential) and field[:, t] += 2 q d (1 / (r^2 + eps^2) - 1 / r^2), d = t - s
 * (accumulated, by target index).  A zero separation reports 1 + t in status.  order (M,
 * may be NULL) is the order the targets are walked in; exclude (M,) may be NULL.
 * workspace: td3_soften_near_workspace_bytes.  No read-back.
 */
size_t td3_soften_near_workspace_bytes(int64_t node_count);
int td3_soften_near(const double *targets, int64_t m, const int64_t *order,
                    const int64_t *exclude, const double *sorted_pos, const double *sorted_charge,
                    const int64_t *perm, int64_t n, const double *node_center,
                    const double *node_half, const void *node_info, int64_t node_count,
                    double radius, double softening, int with_potential, double *phi,
                    double *field, int64_t *status, void *workspace, size_t workspace_bytes,
                    void *stream);

/*
 * Energies of an ensemble's systems (simulation.py:837-916 per system: kinetic
 * sum 1/2 m |v|^2 and potential sum 1/2 q phi), in FP64: out[s] the kinetic and
 * out[systems + s] the potential energy of system s (particles system_start[s] ..
 * system_start[s + 1] - 1; velocities (2, N); potential (N,) may be NULL: zeros).  One
 * workgroup per system adds its particles in a fixed order, so results do not change from
 * run to run.  No counterpart in the reference (its systems are summed one at a time on
 * the host).
 */
int td3_system_energies(const double *masses, const double *velocities, const double *charges,
                        const double *potential, int64_t n, int64_t systems,
                        const int64_t *system_start, double *out, void *stream);

/*
 * Static square cells of the smooth-membership FMM (smooth.py:83-121, BoxTree): the cells
 * td3_tree_build makes (same split rule, children, breadth-first numbering; the reference's
 * BoxTree passes max_depth 40), except that a point with y == cy goes to the lower
 * children (smooth.py:109-110; td3_tree_build sends it to the upper ones,
 * quadtree_quadrants_pyccel.py:38-41).  Same arguments, outputs and workspace as
 * td3_tree_build.
 */
int td3_box_tree_build(const double *pos, int64_t n, double xmid, double ymid, double half,
                       int64_t leaf_capacity, int64_t max_depth, int64_t capacity,
                       double *node_center, double *node_half, int64_t *node_start,
                       int64_t *node_end, int64_t *node_child, int64_t *node_depth,
                       int64_t *levels, int64_t *perm, double *sorted_pos, void *node_info,
                       void *workspace, size_t workspace_bytes, int64_t *out_sizes,
                       int64_t *levels_host, void *stream);

/*
 * Depth-first order of the leaves of a cell set (node_info of td3_tree_build, levels_host
 * its level starts): leaf_span[c] (C,) the leaves below cell c (1 for a leaf), leaf_rank[c]
 * (C,) the depth-first rank of its first leaf, so the leaves below c are the ranks
 * leaf_rank[c] .. leaf_rank[c] + leaf_span[c] - 1.  For td3_smooth_members.  No
 * counterpart in the reference (its memberships are grouped by leaf with a dictionary,
 * smooth.py:356-358).
 */
int td3_box_leaf_order(const void *node_info, int64_t node_count, const int64_t *levels_host,
                       int64_t n_levels, uint32_t *leaf_rank, uint32_t *leaf_span, void *stream);

/*
 * Memberships of N particles in the leaves of a cell set (smooth.py:122-193), and their
 * layout for the FMM (smooth.py:283-303).  band <= 0: the one leaf of the descent right
 * when x > cx, down when y <= cy, weight 1.  band > 0 (at most the smallest leaf
 * half-width; the caller checks): every leaf whose box grown by band strictly holds the
 * particle, weight w = product of the quintic steps across its four edges (edges on the
 * root (root_x, root_y, root_half) keep 1), kept when w > 0 - the reference's operations
 * in its order, so sets and weights agree bit for bit.  Outputs, with S = 4 slots per
 * particle: slot_leaf (N, S) int32 (-1: empty), slot_weight (3, N, S) (w, dw/dx, dw/dy),
 * slot_member (N, S) int32 (the slot's sorted membership, -1: empty); and per sorted
 * membership (capacity 4N, sorted by the depth-first rank of its leaf): member_pos (2, 4N)
 * (stride 4N), member_charge (4N,) q w, member_src (4N, 4) (x, y, q w, particle index as a
 * double), member_leaf (4N,) int32; member_info (C, 4) int32 {start, end, child, depth}:
 * each cell's range of sorted memberships (a cell is occupied when it is not empty), for
 * td3_moments (kernel TD3_LOG, n = 4N) and td3_rows; member_count (DEVICE int64[1]) the
 * number of memberships.  leaf_rank / leaf_span from td3_box_leaf_order, leaf_count the
 * number of leaves.  status (DEVICE int64[1], zeroed by the caller) gets 1 + a particle
 * that met more than four grown cells on one level.  No read-back.
 */
size_t td3_smooth_members_workspace_bytes(int64_t n, int64_t leaf_count);
int td3_smooth_members(const double *pos, const double *charge, int64_t n,
                       const double *node_center, const double *node_half, const void *node_info,
                       int64_t node_count, const uint32_t *leaf_rank, const uint32_t *leaf_span,
                       int64_t leaf_count, double root_x, double root_y, double root_half,
                       double band, int32_t *slot_leaf, double *slot_weight, int32_t *slot_member,
                       double *member_pos, double *member_charge, double *member_src,
                       int32_t *member_leaf, void *member_info, int64_t *member_count,
                       int64_t *status, void *workspace, size_t workspace_bytes, void *stream);

/*
 * Per-membership potential Phi_m and gradient G_m (smooth.py:344-379): the L2P of the
 * leaf's scaled local expansion (td3_l2l over td3_m2l of the td3_moments of the
 * memberships) plus the compact-softened kernel over the memberships of the leaf's P2P row
 * (td3_rows on member_info) that belong to other particles.  phi_m (4N,), g_m (2, 4N)
 * (stride cap = 4N), by sorted membership.  status: 1 + a particle that met another at
 * distance 0 with softening 0.  No read-back.
 */
int td3_smooth_eval(int64_t order, const double *member_src, const int32_t *member_leaf,
                    int64_t cap, const int64_t *member_count, const double *node_center,
                    const double *node_half, const void *member_info, const double *local_re,
                    const double *local_im, const uint32_t *p2p_row, const int32_t *p2p_col,
                    double softening, double *phi_m, double *g_m, int64_t *status,
                    void *stream);

/*
 * Per-particle potential phi = -2 sum w Phi_m, field E = 2 sum (w G_m + Phi_m grad w) over
 * its slots (smooth.py:415-422) in input order (either may be NULL), and the energy
 * U = -sum q w Phi_m (smooth.py:410-413) into energy (DEVICE double[1], may be NULL),
 * summed in a fixed order (workspace td3_smooth_combine_workspace_bytes).  No read-back.
 */
size_t td3_smooth_combine_workspace_bytes(int64_t n);
int td3_smooth_combine(int64_t n, int64_t cap, const double *charge, const int32_t *slot_member,
                       const double *slot_weight, const double *phi_m, const double *g_m,
                       double *phi, double *field, double *energy, void *workspace,
                       size_t workspace_bytes, void *stream);

/*
 * Direct sums of the compact-softened line-charge kernel (smooth.py:196-201 direct_energy,
 * :424-428 potential_at, :431-440 direct_fields): phi[t] = -2 sum q Phi(t - s) and
 * field[:, t] = 2 sum q grad Phi(t - s) (either may be NULL), Phi = ln r outside the
 * cloud of radius softening, ln a - 11/12 + (3u - 3u^2/2 + u^3/3)/2 inside (u = r^2/a^2).
 * self_pairs: targets are the sources and source t is skipped for target t; then energy
 * (DEVICE double[1], may be NULL) gets sum q phi / 2 in a fixed order (workspace
 * td3_compact_direct_workspace_bytes(m)).  A zero separation without softening reports
 * 1 + t in status.  No read-back.
 */
size_t td3_compact_direct_workspace_bytes(int64_t m);
int td3_compact_direct(const double *targets, int64_t m, int self_pairs, const double *sources,
                       const double *charge, int64_t n, double softening, double *phi,
                       double *field, double *energy, int64_t *status, void *workspace,
                       size_t workspace_bytes, void *stream);

/*
 * Kernel timing for measurement (no counterpart in the reference, whose profiling is
 * host-side regions): td3_timing_enable(1) clears the record and brackets every later
 * kernel launch with HIP events on its stream; td3_timing_get(name) synchronises and
 * returns the summed elapsed milliseconds and launch count of the kernels of that name
 * ("m2l", "eval", "traverse", "sort_scatter", ...).  Costs one branch per launch when off.
 */
int td3_timing_enable(int on);
int td3_timing_get(const char *name, double *total_ms, int64_t *launches);

#ifdef __cplusplus
}
#endif

#endif /* TD3_H */
