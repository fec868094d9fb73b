/* MI355X block-tensor contraction (libtensor btod_contract2 path) - C ABI.
 *
 * Plain C: opaque handles, pointers and sizes, no C++ or torch types.  Every
 * function returns LT_OK (0) or a negative error code; lt_last_error() gives
 * the message of the calling thread's last failure.  C++ exceptions never
 * cross this boundary.
 *
 * Index conventions are libtensor's: row-major blocks, last index fastest;
 * a permutation p maps position i of the result to position p[i] of the
 * source (libtensor/core/permutation.h); block indexes are absolute indexes
 * over the block-index dimensions.
 *
 * Each entry point names the libtensor interface it replaces.
 */
#ifndef LT_AMD_H
#define LT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LT_OK             0
#define LT_ERR_PARAM     -1   /* bad_parameter, out_of_bounds */
#define LT_ERR_DIMS      -2   /* bad_dimensions */
#define LT_ERR_SYMMETRY  -3   /* bad_symmetry */
#define LT_ERR_DEVICE    -4   /* HIP runtime failure or no GPU */
#define LT_ERR_NOMEM     -5   /* HBM budget exceeded (admission refused) */
#define LT_ERR_STATE     -6
#define LT_ERR_INTERNAL  -7

#define LT_MODE_PER_BLOCK 0   /* one GEMM launch per C block */
#define LT_MODE_GROUPED   1   /* one grouped GEMM launch for all C blocks */
#define LT_MODE_OZAKI     2   /* grouped, split-precision INT8 products */

#define LT_MAX_ORDER 8

typedef struct lt_ctx lt_ctx;
typedef struct lt_btensor lt_btensor;
typedef struct lt_symel lt_symel;
typedef struct lt_btod_contract2 lt_btod_contract2;
typedef struct lt_btod_symmetrize2 lt_btod_symmetrize2;
/* btod_symmetrize3 handles are used with the lt_btod_symmetrize2_* calls. */
typedef struct lt_btod_symmetrize2 lt_btod_symmetrize3;
typedef struct lt_btod_contract3 lt_btod_contract3;
typedef struct lt_server lt_server;

typedef struct lt_stats {
    double t_schedule_ms;   /* schedule build (constructor) */
    double t_plan_ms;       /* work-list build (first perform) */
    double t_gpu_ms;        /* device time of the last perform */
    double t_gemm_ms;       /* device time of its contraction phase */
    double flops;           /* sum over pairs of 2 |C block| K block */
    double bytes_perm;      /* permute traffic of one perform */
    uint64_t n_cblocks, n_pairs, n_tiles, n_unfold, n_launches;
    int32_t tile_cfg;
    int32_t ozaki_slices;
    double ozaki_max_rel_err;  /* split-precision mode: the stated a-priori bound
                                of max |error| / max |C| over the C blocks of the
                                last perform */
    double tune_ms[32];      /* autotuner: grouped GEMM time per kernel
                               configuration (0 = not measured) */
    int32_t graph;           /* 1: the last perform ran as a HIP graph
                                (t_gemm_ms is then 0: not separately timed) */
    int32_t tile_order;      /* grouped launch tile order: 0 longest C block
                                first, 1 tiles dealt to XCD groups */
    uint64_t n_plans;        /* execution plans built so far (a plan is reused
                                while the operands' structure, the result
                                tensor and the mode of addition stay the same) */
    uint64_t n_batches;      /* batches of C blocks of the last perform: 1, or
                                more when the plan of the whole schedule does
                                not fit the context's device memory budget
                                (gen_bto_contract2_batching_policy.h:48-70) */
} lt_stats;

/* Message of the calling thread's last failed call. */
const char *lt_last_error(void);
int lt_version(void);

/* ---- context (replaces the implicit host context of libtensor:
 *      allocator<double>::init, libtensor/core/allocator.h:84-85, and
 *      libutil::thread_pool, libutil/thread_pool/thread_pool.h:22) ---- */

/* device >= 0: GPU context with an HBM budget (0 = most of free memory).
 * device < 0: host-only context (schedules, symmetry, block storage in host
 * memory; device operations fail with LT_ERR_DEVICE). */
int lt_ctx_create(int device, size_t hbm_cap_bytes, lt_ctx **out);
int lt_ctx_destroy(lt_ctx *ctx);
int lt_ctx_sync(lt_ctx *ctx);
int lt_ctx_mem(lt_ctx *ctx, size_t *used, size_t *cap);
/* The context's device memory budget, the counterpart of the memory limit
 * of allocator<double>::init (libtensor/core/allocator.h:84-85) that the
 * contraction's batching policy reads (gen_bto_contract2_batching_policy.h:
 * 48-70): a contraction whose plan does not fit runs in batches of C blocks. */
int lt_ctx_set_cap(lt_ctx *ctx, size_t cap);
int lt_ctx_set_threads(lt_ctx *ctx, size_t nthreads);
/* Device time (HIP events) of the kernels of the last block operation
 * (lt_btod_copy, lt_btod_add, lt_btod_symmetrize2_perform) on this context.
 * No counterpart in the reference (its timings<> report host wall time). */
int lt_ctx_last_op_ms(lt_ctx *ctx, double *ms);

/* Device memory for the dense-tensor entry points. */
int lt_dev_alloc(lt_ctx *ctx, size_t bytes, void **dptr);
int lt_dev_free(lt_ctx *ctx, void *dptr);
int lt_dev_upload(lt_ctx *ctx, void *dst, const void *host, size_t bytes);
int lt_dev_download(lt_ctx *ctx, void *host, const void *src, size_t bytes);

/* ---- dense tensors on the device (replace dense_tensor ops) ---- */

/* tod_copy<N>(a, perm, c).perform(zero, b)
 * (libtensor/dense_tensor/tod_copy.h:95, :112): b = c P(a) (zero != 0) or
 * b += c P(a).  b has dimensions P(dims_a).  perm may be NULL (identity). */
int lt_tod_copy(lt_ctx *ctx, size_t n, const size_t *dims_a, const double *a,
    const size_t *perm, double c, int zero, double *b);

/* tod_add<N> (libtensor/dense_tensor/tod_add.h:54-100: the constructor
 * takes the first operand, add_op() the others, perform(zero, b)):
 * b = (zero ? 0 : b) + sum_i c[i] P_i(a[i]), operands added in order (one
 * accumulating permute per operand).  All a[i] have dimensions dims (before
 * their permutation); perms holds nops permutations of n entries each (NULL:
 * identities). */
int lt_tod_add(lt_ctx *ctx, size_t n, const size_t *dims, size_t nops,
    const double *const *a, const size_t *perms, const double *c, int zero, double *b);

/* tod_set<N>(v).perform(zero, a) (libtensor/dense_tensor/tod_set.h:37). */
int lt_tod_set(lt_ctx *ctx, size_t nelem, double v, double *a);

/* tod_contract2<N, M, K>(contr, a, b, d).perform(zero, c)
 * (libtensor/dense_tensor/tod_contract2.h:114, :171).  The contraction is
 * contraction2<N, M, K>(permc) with contract(ka[i], kb[i]) for i < K
 * (libtensor/core/contraction2.h).  permc may be NULL. */
int lt_tod_contract2(lt_ctx *ctx, size_t N, size_t M, size_t K,
    const size_t *permc, const size_t *ka, const size_t *kb,
    const size_t *dims_a, const double *a, const size_t *dims_b,
    const double *b, double d, int zero, double *c);

/* No counterpart in the reference: lt_tod_contract2 for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops and
 * harnesses that forbid hipMalloc).  Same arguments and results as
 * lt_tod_contract2 (same kernels); every temporary (aligned operand copies,
 * the aligned result, descriptors, tile list) lives in a fixed device region
 * of LT_NOALLOC_SCRATCH_BYTES per GPU, held from context creation until the
 * process ends: no hipMalloc, hipFree or allocating library call.  A call
 * whose temporaries do not fit fails with LT_ERR_PARAM before any device
 * work.  The region is shared by every call on one GPU and keeps the last
 * call's contents: calls that overlap on one GPU (several threads or
 * contexts) race. */
#define LT_NOALLOC_SCRATCH_BYTES ((size_t)32 << 20)
int lt_tod_contract2_noalloc(lt_ctx *ctx, size_t N, size_t M, size_t K,
    const size_t *permc, const size_t *ka, const size_t *kb,
    const size_t *dims_a, const double *a, const size_t *dims_b,
    const double *b, double d, int zero, double *c);

/* tod_dotprod<N>(a, perma, b, permb).calculate()
 * (libtensor/dense_tensor/tod_dotprod.h:77, :98) times c. */
int lt_tod_dotprod(lt_ctx *ctx, size_t n, const size_t *dims_a,
    const double *a, const size_t *perma, const size_t *dims_b,
    const double *b, const size_t *permb, double c, double *result);

/* ---- point-group product tables (point_group_table +
 *      product_table_container, libtensor/symmetry/point_group_table.h) ---- */
int lt_product_table_create(const char *id, size_t nlabels);
int lt_product_table_add_product(const char *id, size_t l1, size_t l2, size_t lr);

/* ---- block tensors (block_tensor<N, double>,
 *      libtensor/block_tensor/block_tensor.h:19; block_index_space
 *      libtensor/core/block_index_space.h:144) ---- */
int lt_btensor_create(lt_ctx *ctx, size_t n, const size_t *dims, lt_btensor **out);
int lt_btensor_destroy(lt_btensor *bt);
/* block_index_space::split(mask, pos); mask has n entries (0/1). */
int lt_btensor_split(lt_btensor *bt, const int *mask, size_t pos);
int lt_btensor_order(lt_btensor *bt, size_t *n);
int lt_btensor_bidims(lt_btensor *bt, size_t *bidims);
int lt_btensor_block_dims(lt_btensor *bt, size_t aidx, size_t *dims);
int lt_btensor_block_start(lt_btensor *bt, size_t aidx, size_t *start);
/* gen_block_tensor_ctrl::req_nonzero_blocks: sorted canonical non-zero blocks. */
int lt_btensor_nonzero(lt_btensor *bt, size_t *out, size_t cap, size_t *count);
/* Upload / download one canonical block (req_block / ret_block). */
int lt_btensor_set_block(lt_btensor *bt, size_t aidx, const double *host);
int lt_btensor_get_block(lt_btensor *bt, size_t aidx, double *host);
int lt_btensor_zero_block(lt_btensor *bt, size_t aidx);
int lt_btensor_zero_all(lt_btensor *bt);
/* Orbit of a block: canonical index, transformation canonical -> block
 * (perm[n], coefficient) and whether the orbit is allowed
 * (libtensor/core/orbit.h:46). */
int lt_btensor_orbit(lt_btensor *bt, size_t aidx, size_t *acidx,
    size_t *perm, double *coeff, int *allowed);
/* btod_random semantics with a fixed seed: every canonical allowed block is
 * filled (dist 0: U[-1,1], 1: N(0,1)) times scale, optionally divided by
 * 1 + |x_d0 - x_d1| (global element indexes; d0 < 0 disables), and
 * symmetrized over the block's stabilizer. */
int lt_btensor_random(lt_btensor *bt, uint64_t seed, int dist, double scale,
    int d0, int d1);
/* tod_btconv (libtensor/dense_tensor/tod_btconv.h:17): unfold into a dense
 * row-major host array. */
int lt_btensor_export(lt_btensor *bt, double *host);
/* btod_import_raw: take the canonical allowed blocks of a dense host array
 * (blocks that are exactly zero are left zero). */
int lt_btensor_import(lt_btensor *bt, const double *host);
/* block_tensor::set_immutable / is_immutable (libtensor/core/immutable.h):
 * an immutable tensor refuses every later write (LT_ERR_STATE); operations
 * may keep data derived from it (the split-precision mode keeps its int8
 * slices across performs). */
int lt_btensor_set_immutable(lt_btensor *bt);
int lt_btensor_is_immutable(lt_btensor *bt, int *out);
/* Copy the symmetry of src (so_copy). */
int lt_btensor_copy_symmetry(lt_btensor *dst, lt_btensor *src);
/* Lower bt's symmetry to the symmetry bt shares with `other`, as a sum of
 * the two would have it (gen_bto_aux_add::open, gen_block_tensor/impl/
 * gen_bto_aux_add_impl.h:40-66: so_merge(so_dirsum(...)); addition_schedule
 * touches the orbits it splits, :115-145): canonical blocks of the new
 * symmetry that bt held only through an orbit of its old one are written out
 * on the device.  A tensor without blocks may be on a host-only context. */
int lt_btensor_merge_symmetry(lt_btensor *bt, lt_btensor *other);

/* ---- symmetry elements (libtensor/symmetry/se_perm.h:28, se_part.h:32,
 *      se_label.h:37); build, then lt_btensor_insert (symmetry::insert) ---- */
int lt_se_perm_create(size_t n, const size_t *perm, double coeff, lt_symel **out);
int lt_se_part_create(lt_btensor *bt, const size_t *pdims, lt_symel **out);
int lt_se_part_add_map(lt_symel *e, const size_t *idx1, const size_t *idx2, double coeff);
int lt_se_part_mark_forbidden(lt_symel *e, const size_t *idx);
int lt_se_label_create(lt_btensor *bt, const char *table_id, lt_symel **out);
/* block_labeling::assign(mask, block, label) */
int lt_se_label_assign(lt_symel *e, const int *mask, size_t blk, size_t label);
/* se_label::set_rule(label_set) */
int lt_se_label_set_rule(lt_symel *e, const size_t *intrinsic, size_t nintr);
/* Evaluation rule term by term (libtensor/symmetry/evaluation_rule.h:56
 * new_product, product_rule.h add(seq, intr); evaluation_rule::clear): a
 * block is allowed if, for one product, every term's labels (dimension i
 * counted seqs[t * n + i] times) have the intrinsic label intr[t] in their
 * product; (size_t)-1 as intr[t] accepts any label. */
int lt_se_label_clear_rule(lt_symel *e);
int lt_se_label_add_product(lt_symel *e, size_t nterms, const size_t *seqs,
    const size_t *intr);
int lt_symel_destroy(lt_symel *e);
int lt_btensor_insert(lt_btensor *bt, const lt_symel *e);

/* ---- btod_contract2<N, M, K> (libtensor/block_tensor/btod_contract2.h:29;
 *      constructor :68, perform(btc) :115, perform(btc, d) :121,
 *      compute_block :124; called from
 *      libtensor/expr/btensor/impl/eval_btensor_double_contract.C:218) ---- */
int lt_btod_contract2_create(size_t N, size_t M, size_t K, const size_t *permc,
    const size_t *ka_idx, const size_t *kb_idx, lt_btensor *a, double ka,
    lt_btensor *b, double kb, double kc, lt_btod_contract2 **out);
int lt_btod_contract2_destroy(lt_btod_contract2 *op);
/* perform(btc): C := result.  mode: LT_MODE_*. */
int lt_btod_contract2_perform(lt_btod_contract2 *op, lt_btensor *c, int mode);
/* perform(btc, d): C += d * result.  C of another symmetry is first lowered
 * to the symmetry it shares with the result (lt_btensor_merge_symmetry), as
 * addition_schedule + gen_bto_aux_add do (block_tensor/impl/
 * btod_contract2_impl.h:70-90). */
int lt_btod_contract2_perform_add(lt_btod_contract2 *op, lt_btensor *c,
    double d, int mode);
/* get_schedule(): sorted non-zero canonical C blocks. */
int lt_btod_contract2_schedule(lt_btod_contract2 *op, size_t *out, size_t cap,
    size_t *count);
/* get_symmetry() (libtensor/gen_block_tensor/direct_gen_bto.h:38): dst (of
 * C's block index space, holding no blocks) takes the result symmetry. */
int lt_btod_contract2_get_symmetry(lt_btod_contract2 *op, lt_btensor *dst);
/* Result symmetry: canonical index / allowedness of a C block. */
int lt_btod_contract2_symc_orbit(lt_btod_contract2 *op, size_t aidx,
    size_t *acidx, size_t *perm, double *coeff, int *allowed);
/* Contraction list of a C block (gen_bto_contract2_clst_builder::build_list
 * after coalescing), 2 + 2n + 2 values per pair:
 * aia, acia, perm_a[NA], coeff_a, aib, acib, perm_b[NB], coeff_b. */
int lt_btod_contract2_clst(lt_btod_contract2 *op, size_t aidx_c, double *out,
    size_t cap_pairs, size_t *npairs);
/* compute_block(zero = true, ic, identity, blk) on the host (FP64 loops). */
int lt_btod_contract2_compute_block_host(lt_btod_contract2 *op, size_t aidx_c,
    double *host);
/* compute_block(zero, ic, tensor_transf(perm, coeff), blk)
 * (libtensor/block_tensor/btod_contract2.h:124,
 * gen_block_tensor/impl/gen_bto_contract2_impl.h:312-345): block aidx_c of
 * the result (any block of C's block index space, canonical or not),
 * computed on the device from the resident canonical blocks of A and B and
 * transformed into blk (device array, dims of the block permuted by perm;
 * zero: blk := ..., else blk += ...).  perm may be NULL (identity).  For
 * direct block tensors (gen_bto_contract2::compute_block). */
int lt_btod_contract2_compute_block(lt_btod_contract2 *op, int zero, size_t aidx_c,
    const size_t *perm, double coeff, double *blk);
int lt_btod_contract2_stats(lt_btod_contract2 *op, lt_stats *out);
/* Kernel configuration of the contraction GEMM: 0..27 fixed, -1 chosen by a
 * padding estimate, -2 (default) measured on the plan's own work list when
 * it is built (autotuner). */
int lt_btod_contract2_set_tile(lt_btod_contract2 *op, int cfg);
/* Grouped performs (LT_MODE_GROUPED) run as a HIP graph captured once per
 * execution plan (on != 0, the default) or as eager launches (on == 0).
 * No counterpart in the reference (gen_bto_contract2_batch_impl.h:273-276
 * runs the blocks as thread-pool tasks). */
int lt_btod_contract2_set_graph(lt_btod_contract2 *op, int on);
/* FP64 performs of this operation return once their work is enqueued on the
 * context's stream (on != 0) or when it is done (on == 0, the default).
 * Every later operation on the context is ordered after an enqueued perform:
 * performs, block operations, tensor uploads (they wait for it) and reads of
 * results (they wait for it and then copy); device times in lt_stats wait
 * for it when read.  Split-precision performs always wait (they read their
 * stated bounds).  No counterpart in the reference (its perform runs the
 * contraction on the calling threads, gen_bto_contract2_impl.h:42-308). */
int lt_btod_contract2_set_async(lt_btod_contract2 *op, int on);
/* Split-precision mode (LT_MODE_OZAKI): number of int8 slices per operand,
 * 1..8 (0 = 7).  No counterpart in the reference (all of its arithmetic is
 * FP64 BLAS, linalg_cblas_level3.C:10-60). */
int lt_btod_contract2_set_ozaki(lt_btod_contract2 *op, int slices);
/* Stated error of the last split-precision perform, per C block: the
 * canonical block, an a-priori bound of max |error| over its elements
 * (sum over its pairs of |d| K (s + 2) 2^(e_max + f_max - 7 s), e / f the
 * largest row / column exponents of the pair's operands, s the slices) and
 * the largest |C| of the block; count = number of blocks (cap 0: count
 * only).  lt_stats.ozaki_max_rel_err = max bound / max |C|.  A NaN or an
 * infinity in an operand row or column makes its outputs NaN and its bound
 * infinite.  No counterpart in the reference. */
int lt_btod_contract2_ozaki_bounds(lt_btod_contract2 *op, size_t *aidx, double *bound,
    double *cmax, size_t cap, size_t *count);
/* C tile shape of the split-precision product kernel: -1 (default) = each
 * shape timed once on the plan's work list, the fastest kept; 0 = 16x32,
 * 1 = 16x64, 2 = 32x32 (rows x columns).  No counterpart in the reference. */
int lt_btod_contract2_set_ozaki_tile(lt_btod_contract2 *op, int tile);

/* ---- btod_dotprod (libtensor/block_tensor/btod_dotprod.h:35-66:
 *      btod_dotprod(bta, tra, btb, trb).calculate()): sum over all elements
 *      of ca P_a(A) * cb P_b(B) for resident block tensors whose permuted
 *      block index spaces agree; perma / permb may be NULL. ---- */
int lt_btod_dotprod(lt_btensor *a, const size_t *perma, double ca, lt_btensor *b,
    const size_t *permb, double cb, double *result);

/* ---- btod_copy (libtensor/block_tensor/btod_copy.h:20-80:
 *      btod_copy(bta, perm, c).perform(btb) and .perform(btb, d);
 *      gen_block_tensor/impl/gen_bto_copy_impl.h:150-163, :302-330) ----
 * add == 0: B := c P(A); B (block index space of A permuted by perm) takes
 * the permuted symmetry of A.  add != 0: B += d c P(A); a B of another
 * symmetry is first lowered to the symmetry both share (through
 * gen_bto_aux_add in the reference).  perm may be NULL (identity). */
int lt_btod_copy(lt_btensor *a, const size_t *perm, double c, lt_btensor *b, int add,
    double d);

/* ---- btod_add (libtensor/block_tensor/btod_add.h: btod_add(bta, c) plus
 *      add_op(bta, perm, c), then perform(btb) / perform(btb, d);
 *      gen_block_tensor/impl/gen_bto_add_impl.h) ----
 * B := sum_i c[i] P_i(A_i) (add == 0) or B += d sum_i ... (add != 0).  perms
 * holds nops permutations of the tensor order (NULL: identities).  B takes
 * the symmetry of the sum (gen_bto_add::add_operand, gen_bto_add_impl.h:
 * 226-270: so_dirsum + so_merge operand by operand); with add != 0, B is
 * first lowered to the symmetry it shares with that.  Every block of the sum
 * is computed (the reference's schedule can miss blocks where the sum's
 * symmetry is lower than an operand's). */
int lt_btod_add(size_t nops, lt_btensor *const *a, const size_t *perms, const double *c,
    lt_btensor *b, int add, double d);

/* ---- btod_contract3 (libtensor/block_tensor/btod_contract3.h:30-150;
 *      gen_block_tensor/impl/gen_bto_contract3_impl.h) ----
 * D = kd (A B) C: contr1 contracts A (order N1+K1) with B (order N2+K1+K2)
 * into AB (order N1+N2+K2, index order permab, pairs ka1[i]-kb1[i]); contr2
 * contracts AB with C (order N3+K2) into D (order N1+N2+N3, index order
 * permd, pairs kab2[i]-kc2[i]), each as for lt_btod_contract2_create.  The
 * result symmetry and schedule are the reference's (the second
 * contraction's over AB's schedule).  AB is kept in a resident temporary
 * owned by the operation.  perform: D := result; perform_add: D += k *
 * result for D with the result symmetry; stats: device ms of the last
 * perform (both contractions) and the FLOPs of both. */
int lt_btod_contract3_create(size_t N1, size_t N2, size_t N3, size_t K1, size_t K2,
    const size_t *permab, const size_t *ka1, const size_t *kb1, const size_t *permd,
    const size_t *kab2, const size_t *kc2, lt_btensor *a, lt_btensor *b, lt_btensor *c,
    double kd, lt_btod_contract3 **out);
/* btod_symcontract3 (libtensor/block_tensor/btod_symcontract3.h;
 * gen_block_tensor/impl/gen_bto_symcontract3_impl.h): as btod_contract3 with
 * AB replaced by AB + s P(AB) before the second contraction (P =
 * sympermab, a pair permutation of AB's indexes; s = +1 if symmab), with
 * the symmetry and schedule of btod_symmetrize2.  Returns a btod_contract3
 * handle (the lt_btod_contract3_* calls apply). */
int lt_btod_symcontract3_create(size_t N1, size_t N2, size_t N3, size_t K1, size_t K2,
    const size_t *permab, const size_t *ka1, const size_t *kb1, const size_t *permd,
    const size_t *kab2, const size_t *kc2, lt_btensor *a, lt_btensor *b,
    const size_t *sympermab, int symmab, lt_btensor *c, double kd, lt_btod_contract3 **out);
int lt_btod_contract3_destroy(lt_btod_contract3 *op);
int lt_btod_contract3_perform(lt_btod_contract3 *op, lt_btensor *d, int mode);
int lt_btod_contract3_perform_add(lt_btod_contract3 *op, lt_btensor *d, double k, int mode);
int lt_btod_contract3_schedule(lt_btod_contract3 *op, size_t *out, size_t cap,
    size_t *count);
int lt_btod_contract3_sym_orbit(lt_btod_contract3 *op, size_t aidx, size_t *acidx,
    size_t *perm, double *coeff, int *allowed);
int lt_btod_contract3_stats(lt_btod_contract3 *op, double *gpu_ms, double *flops);
/* Execution plans built so far by the two contractions (see lt_stats.n_plans). */
int lt_btod_contract3_plans(lt_btod_contract3 *op, uint64_t *n);

/* ---- btod_symmetrize2 (libtensor/block_tensor/btod_symmetrize2.h:20-110;
 *      gen_block_tensor/impl/gen_bto_symmetrize2_impl.h) ----
 * B = X + s P(X), X the result of a btod_contract2 (create), of another
 * btod_symmetrize2 (create_nested) or of a btod_contract3 (create_c3), P a pair permutation (P P = 1, P not the
 * identity; btod_symmetrize2(op, i1, i2, symm) is the transposition of i1
 * and i2), s = +1 (symm != 0) or -1.  The result symmetry is libtensor's
 * so_symmetrize of the symmetry of X (sym_orbit queries it), the schedule
 * that of gen_bto_symmetrize2::make_schedule.  The operand must outlive the
 * operation.  perform: B := result (B takes the result symmetry);
 * perform_add: B += d * result for B with that symmetry (the reference's
 * perform(btb, d)); mode as for btod_contract2, used for X.  X is kept in a
 * resident temporary owned by the operation.  compute_block: blk (=, or +=
 * unless zero) P_blk c B(aidx_b) on the device (btod_symmetrize2::
 * compute_block, direct tensors); perm may be NULL. */
int lt_btod_symmetrize2_create(lt_btod_contract2 *op, const size_t *perm, int symm,
    lt_btod_symmetrize2 **out);
int lt_btod_symmetrize2_create_nested(lt_btod_symmetrize2 *op, const size_t *perm, int symm,
    lt_btod_symmetrize2 **out);
/* Operand a btod_contract3 (libtensor's symmetrization of any additive
 * operation; btod_symcontract3 is this with the symmetrized pair inside). */
int lt_btod_symmetrize2_create_c3(lt_btod_contract3 *op, const size_t *perm, int symm,
    lt_btod_symmetrize2 **out);
/* btod_symmetrize3 (libtensor/block_tensor/btod_symmetrize3.h;
 * gen_block_tensor/impl/gen_bto_symmetrize3_impl.h): B = sum over the six
 * exchanges g of three index groups, generated by the pair permutations
 * perm1 and perm2 (btod_symmetrize3(op, i1, i2, i3, symm) exchanges single
 * indexes: perm1 = (i1 i2), perm2 = (i1 i3)), of s(g) g(X), s(g) = s for
 * the three pair exchanges and 1 for the identity and the two cycles.
 * Result symmetry and schedule as in the reference; the handle is used with
 * the lt_btod_symmetrize2_* calls below (perform, perform_add, schedule,
 * sym_orbit, compute_block, stats, destroy); create_nested takes another
 * symmetrization as the operand. */
int lt_btod_symmetrize3_create(lt_btod_contract2 *op, const size_t *perm1,
    const size_t *perm2, int symm, lt_btod_symmetrize3 **out);
int lt_btod_symmetrize3_create_nested(lt_btod_symmetrize2 *op, const size_t *perm1,
    const size_t *perm2, int symm, lt_btod_symmetrize3 **out);
int lt_btod_symmetrize2_destroy(lt_btod_symmetrize2 *op);
int lt_btod_symmetrize2_perform(lt_btod_symmetrize2 *op, lt_btensor *b, int mode);
int lt_btod_symmetrize2_perform_add(lt_btod_symmetrize2 *op, lt_btensor *b, double d,
    int mode);
int lt_btod_symmetrize2_schedule(lt_btod_symmetrize2 *op, size_t *out, size_t cap,
    size_t *count);
int lt_btod_symmetrize2_sym_orbit(lt_btod_symmetrize2 *op, size_t aidx, size_t *acidx,
    size_t *perm, double *coeff, int *allowed);
int lt_btod_symmetrize2_compute_block(lt_btod_symmetrize2 *op, int zero, size_t aidx_b,
    const size_t *perm, double coeff, double *blk);
/* Last perform: device time in ms (operand and gather kernels), gather
 * kernels alone, bytes the gather moved (each source block read, each
 * destination written, and also read where added to), and the operand's
 * FLOPs.  Any pointer may be NULL.  No counterpart in the reference (its
 * timings go to libutil's timings store). */
int lt_btod_symmetrize2_stats(lt_btod_symmetrize2 *op, double *gpu_ms, double *gather_ms,
    double *gather_bytes, double *flops);

/* ---- request loop: many calculations sharing one GPU.  No counterpart in
 *      the reference (its operations run synchronously in the caller's
 *      thread, libutil/thread_pool/thread_pool.h:22).  Block tensors stay
 *      resident; requests are admitted by free device memory (estimate of
 *      a plan's footprint, hbm_reserve kept free), planned on a host thread
 *      while the GPU runs the previous batch, and all ready requests run as
 *      one batch (one grouped GEMM launch over the union of their work
 *      lists per operand layout and kernel configuration).  Requests that
 *      arrive meanwhile join the next batch. ---- */
typedef struct lt_server lt_server;
typedef struct lt_request_stats {
    double t_queue_ms;      /* submit -> start of its batch */
    double t_run_ms;        /* start -> end of its batch */
    double t_total_ms;      /* submit -> result ready (latency) */
    uint64_t batch;         /* batch number, 1-based */
    uint64_t batch_size;    /* requests in that batch */
    int32_t status;         /* LT_OK or the request's error code */
    int32_t pad;
} lt_request_stats;
int lt_server_create(lt_ctx *ctx, size_t max_batch, size_t hbm_reserve, lt_server **out);
/* Drains every submitted request, then stops. */
int lt_server_destroy(lt_server *srv);
/* perform(btc) (add == 0) or perform(btc, d) (add != 0), grouped mode.  One
 * request in flight per operation and per result tensor; neither may be
 * used by the caller until lt_server_wait returns. */
int lt_server_submit(lt_server *srv, lt_btod_contract2 *op, lt_btensor *c, int add,
    double d, uint64_t *id);
/* Blocks until the request is done; returns its status. */
int lt_server_wait(lt_server *srv, uint64_t id, lt_request_stats *st);
int lt_server_stats(lt_server *srv, uint64_t *batches, uint64_t *requests,
    double *mean_batch, uint64_t *max_batch);

#ifdef __cplusplus
}
#endif

#endif /* LT_AMD_H */
