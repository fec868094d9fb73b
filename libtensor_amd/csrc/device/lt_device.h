// Device work descriptors and kernel launchers for the block-contraction path
// on CDNA4 (gfx950).  Plain structs only: these are copied to HBM as work
// lists and read by the kernels.
#pragma once

#include <cstddef>
#include <cstdint>
#include <hip/hip_runtime_api.h>

namespace lta {
namespace dev {

static const int kPermMaxOrder = 8;

//! Fast unsigned division by a runtime constant for n < 2^31
//! (Granlund & Montgomery, "Division by invariant integers using
//! multiplication", PLDI 1994): q = (mulhi(n, mul) + n) >> shift.
struct fastdiv {
    uint32_t d, mul, shift, pad;
};
fastdiv make_fastdiv(uint32_t d);

//! dst = alpha * P(src) (+ dst if accumulate).  dims are the destination
//! dimensions after fusing; sstride[i] is the source stride of destination
//! dimension i; the destination is contiguous.
struct perm_desc {
    const double *src;
    double *dst;
    uint32_t n;                       //!< fused order, 1..8
    uint32_t total;                   //!< elements (< 2^31)
    fastdiv div[kPermMaxOrder];       //!< destination dimensions
    uint32_t sstride[kPermMaxOrder];  //!< source strides
    double alpha;
    int32_t accumulate;
    //! Transposing descriptors (the source's unit-stride dimension is
    //! destination dimension tile_p < n - 1): 32 x 32 tiles over
    //! (tile_p, n - 1) staged in LDS, read and written contiguously.
    //! -1: element-wise gather.
    int32_t tile_p;
    uint32_t tsq;                     //!< source stride of dimension n - 1
    fastdiv dps;                      //!< destination stride of tile_p
    fastdiv tdp, tdq;                 //!< extents of tile_p and n - 1
};
static const uint32_t kPermTile = 32;
static const uint32_t kPermTileItem = 0x80000000u;

//! One C block of a grouped GEMM: C(m x n, leading dimension ldc) is
//! overwritten (zero != 0) or accumulated with the sum over its pairs.
struct gemm_cblock {
    double *c;
    uint32_t m, n, ldc;
    uint32_t pair_begin, pair_end;
    int32_t zero;
};

//! One product d * A B contributing to a C block.  A is stored m x k
//! (A(m, k) at a[m * lda + k]) or, with kPairAKMajor, k x m (at
//! a[k * lda + m]); B is stored k x n (B(k, n) at b[k * ldb + n]) or, with
//! kPairBNMajor, n x k (at b[n * ldb + k]).
struct gemm_pair {
    const double *a;
    const double *b;
    uint32_t k, lda, ldb, flags;
    double d;
};
static const uint32_t kPairAKMajor = 1, kPairBNMajor = 2;
//! Stacked kernels: k-contiguous rows of A (kPairVecA) or of B
//! (kPairVecB, with kPairBNMajor) start 16-byte aligned and the extent is
//! even, so a lane loads two k at once.
static const uint32_t kPairVecA = 4, kPairVecB = 8;
//! Offset-table B whose k offsets come in adjacent even pairs and whose n
//! offsets are even (two k per load through the tables).
static const uint32_t kPairVecBT = 16;
//! Offset-table B whose tables are linear in at most two index groups:
//! koff[k] = (k / kd) ks0 + (k % kd) ks1 and noff[n] likewise, computed in
//! the kernel instead of read.
static const uint32_t kPairOffArith = 32;

//! Work item of a grouped GEMM launch: one output tile of one C block.
struct gemm_tile {
    uint32_t cblock;
    uint16_t tm, tn;
};

//! GEMM kernel configurations.  TILE_*: workgroup tiles with LDS-staged
//! operands; SPLITK_*_Wn: one workgroup of n waves per tile, the tile's
//! k chunks dealt to the waves, operands loaded straight into MFMA
//! fragments, partial tiles summed in LDS in a fixed order.
//! STACK_*: the split-K kernels over stacks (C blocks sharing their B
//! operands, rows concatenated; _On: compiled for n waves per SIMD; LDS: A
//! chunks staged in LDS for two column halves, two k groups; _P1:
//! the two-stage pipeline that loads a pair's row table when it starts
//! the pair's first chunk); used by grouped launches, per-block launches
//! run the SPLITK configuration of the same shape.
enum { TILE_64x64 = 0, TILE_32x64 = 1, TILE_64x32 = 2, TILE_32x32 = 3,
       SPLITK_16x64_W4 = 4, SPLITK_32x32_W4 = 5, SPLITK_32x64_W4 = 6,
       SPLITK_64x32_W4 = 7, SPLITK_16x64_W2 = 8, SPLITK_32x32_W2 = 9,
       SPLITK_16x32_W4 = 10, SPLITK_32x16_W4 = 11,
       STACK_16x64_W4 = 12, STACK_32x32_W4 = 13, STACK_32x64_W4 = 14,
       STACK_64x32_W4 = 15, STACK_48x64_W4 = 16, STACK_64x48_W4 = 17,
       STACK_48x32_W4 = 18, STACK_64x32_W4_P1 = 19, STACK_64x48_W2 = 20,
       STACK_48x64_W2 = 21, STACK_48x32_W4_O2 = 22, STACK_32x32_W4_O3 = 23,
       STACK_16x64_W4_O3 = 24, STACK_VAR_64x32_W4 = 25, STACK_VAR_64x48_W4 = 26,
       STACK_LDS_64x64 = 27, TILE_COUNT = 28 };
void tile_shape(int cfg, uint32_t &bm, uint32_t &bn);
inline bool is_stacked(int cfg) { return cfg >= STACK_16x64_W4 && cfg <= STACK_LDS_64x64; }
//! Stacked configurations over tiles of varying shape (var_tile lists).
inline bool is_var(int cfg) { return cfg == STACK_VAR_64x32_W4 || cfg == STACK_VAR_64x48_W4; }
//! The per-C-block configuration with the same summation order (same
//! split-K width; the tile shape only partitions the output).
inline int unstacked(int cfg) {
    switch(cfg) {
    case STACK_16x64_W4: return SPLITK_16x64_W4;
    case STACK_32x32_W4: return SPLITK_32x32_W4;
    case STACK_32x64_W4: case STACK_48x64_W4: return SPLITK_32x64_W4;
    case STACK_64x32_W4: case STACK_64x48_W4: case STACK_48x32_W4: case STACK_64x32_W4_P1:
    case STACK_48x32_W4_O2: case STACK_32x32_W4_O3: case STACK_16x64_W4_O3:
    case STACK_VAR_64x32_W4: case STACK_VAR_64x48_W4:
        return SPLITK_64x32_W4;
    case STACK_64x48_W2: case STACK_48x64_W2: case STACK_LDS_64x64: return SPLITK_32x32_W2;
    default: return cfg;
    }
}

//! A stack: C blocks whose contraction lists read the same B operands in the
//! same order (same layouts and extents); their rows are concatenated
//! (rows in all), so a tile reads each B chunk once for rows of several C
//! blocks.  Pair j of the stack is stack_pair[pair_begin + j]; row r of the
//! stack reads A through arow[arow_begin + j * rows + r] and writes through
//! crow[crow_begin + r].
struct stack_desc {
    uint32_t rows, n, npairs, pair_begin;
    uint32_t arow_begin, crow_begin;
    uint32_t nchunk;                  //!< sum over pairs of ceil(k / 16)
    uint32_t pad1;
};
struct stack_pair {
    const double *b;
    //! B read through offset tables (not null): B(k, n) = b[koff[k] + noff[n]]
    //! (a permuted operand read in place from its stored block)
    const int32_t *koff, *noff;
    uint32_t k, ldb, flags, pad;      //!< flags: kPairAKMajor / kPairBNMajor ...
    //! kPairOffArith: the tables as koff[k] = (k / kd) ks0 + (k % kd) ks1,
    //! noff[n] = (n / nd) ns0 + (n % nd) ns1
    fastdiv kd, nd;
    int32_t ks0, ks1, ns0, ns1;
};
//! A operand of one row for one pair: row r of A at p[k] (or p[k * lda]
//! with kPairAKMajor), coefficient d.
struct stack_arow {
    const double *p;
    double d;
    uint32_t lda, pad;
};
struct stack_crow {
    double *c;                        //!< start of the row in its C block
    uint32_t zero, pad;
};
//! Tile of a varying-shape stacked launch: rows 16 r16 .. + 16 fm,
//! columns 16 c16 .. + 16 fn of stack `stack`.
struct var_tile {
    uint32_t stack;
    uint16_t r16, c16;
    uint16_t fm, fn;
};
hipError_t launch_gemm_stacked_var(int cfg, const stack_desc *st, const stack_pair *sp,
    const stack_arow *ar, const stack_crow *cr, const var_tile *tiles, uint32_t ntiles,
    hipStream_t s);
hipError_t launch_gemm_stacked(int cfg, const stack_desc *st, const stack_pair *sp,
    const stack_arow *ar, const stack_crow *cr, const gemm_tile *tiles, uint32_t ntiles,
    hipStream_t s);

//! Batched permutation: work item j is chunk_desc[j] / chunk_off[j]: a
//! descriptor and the first destination element of a chunk of kPermChunk
//! elements, or (chunk_desc[j] | kPermTileItem) and the destination index
//! of a tile's corner for transposing descriptors.
hipError_t launch_permute(const perm_desc *descs, const uint32_t *chunk_desc,
    const uint32_t *chunk_off, uint32_t nchunks, hipStream_t s);
static const uint32_t kPermChunk = 2048;

//! Grouped FP64 MFMA GEMM over a tile list (operand layouts per pair).
//! nwg workgroups loop over the tiles (nwg == ntiles gives one workgroup
//! per tile).
hipError_t launch_gemm(int cfg, const gemm_cblock *cb,
    const gemm_pair *pairs, const gemm_tile *tiles, uint32_t ntiles,
    uint32_t nwg, hipStream_t s);

// ---- split-precision (Ozaki scheme) FP64 emulation on INT8 MFMA ----

//! Splitting of one operand matrix into int8 slices with one power-of-two
//! scale per row: row r = 2^exps[r] * sum_t slice_t[r][.] * 2^(-7 t),
//! t = 1..slices.  Element (r, k) of the source is src[r * rs + k * ks].
//! Slice t - 1 starts at dst + (t - 1) * sstride and is fragment-major: rows
//! form 16-row panels (the last one zero-padded), k forms 64-k chunks (zero
//! beyond k), panel p, chunk c is 1024 contiguous bytes at
//! (p * kp / 64 + c) * 1024, and within it byte j of lane l (row 16 p + l % 16,
//! k = 64 c + 16 (l / 16) + j) sits at l * 16 + j.  sstride = panels * 16 * kp.
struct oz_split_desc {
    const double *src;
    int8_t *dst;
    int32_t *exps;
    uint32_t rows, k, kp, slices;
    uint32_t rs, ks, sstride, pad;
};

//! One product of a C block in split form: A slices (rows of C) and B
//! slices (columns of C) in the fragment-major layout of oz_split_desc, kp
//! (multiple of 64) padded k.
struct oz_pair {
    const int8_t *a;
    const int8_t *b;
    const int32_t *ea;
    const int32_t *eb;
    uint32_t astride, bstride, kp, pad;
    double d;
    uint32_t ia, ib;                  //!< split descriptors of the two operands
};

//! C tile shapes of the split-precision product kernel (rows x columns).
enum oz_tile_cfg { OZ_TILE_16x32 = 0, OZ_TILE_16x64 = 1, OZ_TILE_32x32 = 2, OZ_TILE_COUNT = 3 };
void oz_tile_shape(int tile, uint32_t &bm, uint32_t &bn);

//! items: nitems pairs (descriptor, row).
hipError_t launch_oz_split(const oz_split_desc *descs, const uint32_t *items,
    uint32_t nitems, hipStream_t s);
//! C tiles of shape `tile` (oz_tile_cfg); products of slices t, u with
//! t + u <= slices + 1, exact in int32, recombined in FP64.
hipError_t launch_oz_gemm(const gemm_cblock *cb, const oz_pair *pairs,
    const gemm_tile *tiles, uint32_t ntiles, int slices, int tile, hipStream_t s);

//! Exponent of a split-precision row or column that holds a NaN or an
//! infinity (its products are NaN).
static const int32_t kOzNonFinite = 0x3fffffff;

//! Stated error of the split-precision products: out[2 b] = bound of
//! max |error| of C block b, out[2 b + 1] = max |C| of the block as written.
hipError_t launch_oz_bound(const gemm_cblock *cb, const gemm_pair *pairs, const oz_pair *ozp,
    const int32_t *opmax, uint32_t ncb, int slices, double *out, hipStream_t s);
//! Largest exponent of every split operand (opmax[i] over the rows of
//! descs[i]; kOzNonFinite when a row holds a NaN or an infinity), read by
//! the bound kernel; also zeroes the max |C| slots bound_out[2 b + 1] of
//! the ncb C blocks that the bound kernel fills next.
hipError_t launch_oz_opmax(const oz_split_desc *descs, uint32_t nops, int32_t *opmax,
    uint32_t ncb, double *bound_out, hipStream_t s);

//! Fill n doubles with v.
hipError_t launch_fill(double *p, size_t n, double v, hipStream_t s);

//! Dot product of permuted operands: sum_i a[i] * b(P(i)) over descriptors
//! whose `src` is b and `dst` is (const) a; partials[wg] receive one value
//! per workgroup, then a final pass sums them in a fixed order into *out.
hipError_t launch_dot(const perm_desc *descs, const uint32_t *chunk_desc,
    const uint32_t *chunk_off, uint32_t nchunks, double *partials,
    double *out, hipStream_t s);

//! Size of the fixed device scratch region of the no-allocation entry
//! points (one per GPU, held until the process ends).
static const size_t kNoallocScratchBytes = size_t(32) << 20;
static const int kMaxDevices = 64;
//! Device address of that region on `device`, which must be the current
//! device on the first call for it (resolved once, then cached).
hipError_t noalloc_scratch(int device, void **out);

} // namespace dev
} // namespace lta
