// CDNA4 (gfx950) kernels of the block-contraction path:
//   * batched N-d permute / scale / accumulate (tod_copy, tod_add, symmetry
//     unfolding, C write-back): gather chunks or LDS-tiled transposes,
//   * FP64 GEMM on v_mfma_f64_16x16x4f64 over the contraction lists of the
//     C blocks: LDS-staged workgroup tiles (gemm_kernel), split-K tiles with
//     operands loaded straight into MFMA fragments per C block
//     (gemm_direct_kernel) or over stacks of C blocks that share their B
//     operands (gemm_stacked_kernel, its varying-shape and LDS-staged
//     variants).  Every workgroup owns one C tile and runs the whole list of
//     that tile, the waves' partial tiles are added in LDS in wave order: each
//     C element is written once with a fixed summation order (no atomics),
//   * permuted dot product with a fixed-order two-pass reduction,
//   * fill.
#include <hip/hip_runtime.h>
#include "lt_device.h"

namespace lta {
namespace dev {

typedef double double4_t __attribute__((ext_vector_type(4)));

fastdiv make_fastdiv(uint32_t d) {
    fastdiv f;
    f.d = d;
    f.pad = 0;
    uint32_t l = 0;
    while(l < 32 && (uint64_t(1) << l) < d) l++;
    f.shift = l;
    uint64_t m = ((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1;
    f.mul = uint32_t(m);
    return f;
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const fastdiv &f) {
    return (__umulhi(n, f.mul) + n) >> f.shift;
}

void tile_shape(int cfg, uint32_t &bm, uint32_t &bn) {
    switch(cfg) {
    case TILE_32x64: bm = 32; bn = 64; break;
    case TILE_64x32: bm = 64; bn = 32; break;
    case TILE_32x32: bm = 32; bn = 32; break;
    case SPLITK_16x64_W4: case SPLITK_16x64_W2: case STACK_16x64_W4: bm = 16; bn = 64; break;
    case SPLITK_32x32_W4: case SPLITK_32x32_W2: case STACK_32x32_W4: bm = 32; bn = 32; break;
    case SPLITK_32x64_W4: case STACK_32x64_W4: bm = 32; bn = 64; break;
    case SPLITK_64x32_W4: case STACK_64x32_W4: case STACK_64x32_W4_P1: bm = 64; bn = 32; break;
    case STACK_LDS_64x64: bm = 64; bn = 64; break;
    case STACK_48x64_W4: bm = 48; bn = 64; break;
    case STACK_64x48_W4: bm = 64; bn = 48; break;
    case STACK_48x32_W4: bm = 48; bn = 32; break;
    case STACK_64x48_W2: bm = 64; bn = 48; break;
    case STACK_48x64_W2: bm = 48; bn = 64; break;
    case STACK_48x32_W4_O2: bm = 48; bn = 32; break;
    case STACK_32x32_W4_O3: bm = 32; bn = 32; break;
    case STACK_16x64_W4_O3: bm = 16; bn = 64; break;
    case SPLITK_16x32_W4: bm = 16; bn = 32; break;
    case SPLITK_32x16_W4: bm = 32; bn = 16; break;
    default: bm = 64; bn = 64; break;
    }
}

// ---------------------------------------------------------------------------
// Permute

__device__ __forceinline__ uint32_t src_offset(const perm_desc &d, uint32_t e) {
    uint32_t off = 0;
#pragma unroll
    for(int i = kPermMaxOrder - 1; i >= 0; i--) {
        if(i < (int)d.n) {
            uint32_t q = fdiv(e, d.div[i]);
            uint32_t r = e - q * d.div[i].d;
            off += r * d.sstride[i];
            e = q;
        }
    }
    return off;
}

__global__ void __launch_bounds__(256)
permute_kernel(const perm_desc *__restrict__ descs,
    const uint32_t *__restrict__ chunk_desc,
    const uint32_t *__restrict__ chunk_off) {

    __shared__ double tile[kPermTile][kPermTile + 1];
    const uint32_t ch = blockIdx.x;
    const uint32_t cd = chunk_desc[ch];
    const perm_desc d = descs[cd & ~kPermTileItem];
    const uint32_t base = chunk_off[ch];

    if(cd & kPermTileItem) {
        // 32 x 32 tile over destination dimensions (p, q = n - 1); the
        // source is contiguous along p, the destination along q
        // (extents and strides from dedicated fields: indexing the
        // descriptor's arrays with a runtime dimension would put it in
        // scratch memory)
        const uint32_t dq = d.tdq.d, dp = d.tdp.d;
        const uint32_t q0 = base - fdiv(base, d.tdq) * dq;
        const uint32_t e1 = fdiv(base, d.dps);
        const uint32_t p0 = e1 - fdiv(e1, d.tdp) * dp;
        const uint32_t np = min(kPermTile, dp - p0), nq = min(kPermTile, dq - q0);
        const uint32_t s0 = src_offset(d, base), sq = d.tsq, dps = d.dps.d;
        const uint32_t ti = threadIdx.x & 31, tj = threadIdx.x >> 5;
#pragma unroll
        for(uint32_t r = 0; r < kPermTile; r += 8) {
            const uint32_t iq = tj + r;
            if(ti < np && iq < nq) tile[iq][ti] = d.src[s0 + ti + iq * sq];
        }
        __syncthreads();
#pragma unroll
        for(uint32_t r = 0; r < kPermTile; r += 8) {
            const uint32_t ip = tj + r;
            if(ti < nq && ip < np) {
                double *o = d.dst + base + ip * dps + ti;
                const double v = d.alpha * tile[ti][ip];
                *o = d.accumulate ? v + *o : v;
            }
        }
        return;
    }

    // element-wise gather: each thread owns kPermChunk / 256 elements of the
    // chunk; all source offsets are computed and all loads issued before the
    // first store, so every thread has that many gathers in flight
    constexpr int E = kPermChunk / 256;
    const uint32_t end = min(base + kPermChunk, d.total);
    double v[E];
#pragma unroll
    for(int i = 0; i < E; i++) {
        const uint32_t e = min(base + threadIdx.x + 256u * i, end - 1);
        v[i] = d.src[src_offset(d, e)];
    }
    if(d.accumulate) {
#pragma unroll
        for(int i = 0; i < E; i++) {
            const uint32_t e = base + threadIdx.x + 256u * i;
            if(e < end) d.dst[e] = d.alpha * v[i] + d.dst[e];
        }
    } else {
#pragma unroll
        for(int i = 0; i < E; i++) {
            const uint32_t e = base + threadIdx.x + 256u * i;
            if(e < end) d.dst[e] = d.alpha * v[i];
        }
    }
}

hipError_t launch_permute(const perm_desc *descs, const uint32_t *chunk_desc,
    const uint32_t *chunk_off, uint32_t nchunks, hipStream_t s) {
    if(nchunks == 0) return hipSuccess;
    hipLaunchKernelGGL(permute_kernel, dim3(nchunks), dim3(256), 0, s,
        descs, chunk_desc, chunk_off);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Dot product

__global__ void __launch_bounds__(256)
dot_kernel(const perm_desc *__restrict__ descs,
    const uint32_t *__restrict__ chunk_desc,
    const uint32_t *__restrict__ chunk_off, double *__restrict__ partials) {

    __shared__ double red[256];
    const uint32_t ch = blockIdx.x;
    const perm_desc d = descs[chunk_desc[ch]];
    const uint32_t base = chunk_off[ch];
    const uint32_t end = min(base + kPermChunk, d.total);
    double acc = 0.0;
    for(uint32_t e = base + threadIdx.x; e < end; e += blockDim.x)
        acc += d.dst[e] * d.src[src_offset(d, e)];
    red[threadIdx.x] = acc * d.alpha;
    __syncthreads();
    for(uint32_t w = 128; w > 0; w >>= 1) {
        if(threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if(threadIdx.x == 0) partials[ch] = red[0];
}

__global__ void __launch_bounds__(256)
sum_kernel(const double *__restrict__ partials, uint32_t n, double *__restrict__ out) {
    __shared__ double red[256];
    double acc = 0.0;
    for(uint32_t i = threadIdx.x; i < n; i += 256) acc += partials[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for(uint32_t w = 128; w > 0; w >>= 1) {
        if(threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if(threadIdx.x == 0) *out = red[0];
}

hipError_t launch_dot(const perm_desc *descs, const uint32_t *chunk_desc,
    const uint32_t *chunk_off, uint32_t nchunks, double *partials,
    double *out, hipStream_t s) {
    if(nchunks > 0) {
        hipLaunchKernelGGL(dot_kernel, dim3(nchunks), dim3(256), 0, s,
            descs, chunk_desc, chunk_off, partials);
    }
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, s, partials, nchunks, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Fill

__global__ void __launch_bounds__(256) fill_kernel(double *p, size_t n, double v) {
    size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
    size_t stride = size_t(gridDim.x) * 256;
    for(; i < n; i += stride) p[i] = v;
}

hipError_t launch_fill(double *p, size_t n, double v, hipStream_t s) {
    if(n == 0) return hipSuccess;
    size_t nb = (n + 255) / 256;
    if(nb > 65536) nb = 65536;
    hipLaunchKernelGGL(fill_kernel, dim3((uint32_t)nb), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Grouped FP64 MFMA GEMM
//
// Workgroup: 256 threads = 4 waves in a 2 x 2 arrangement over a BM x BN
// tile; each wave owns (BM/2) x (BN/2) = FM x FN fragments of 16 x 16.
// Operand tiles of BK = 16 are staged in LDS as As[k][m], Bs[k][n] (row
// padding keeps the four k-rows read by one MFMA on different bank halves).
//
// v_mfma_f64_16x16x4f64 operand layout (lane l = 0..63):
//   A: A[m = l % 16][k = l / 16]      B: B[k = l / 16][n = l % 16]
//   D: D[m = (l / 16) + 4 * v][n = l % 16], v = 0..3 (measured on gfx950)

template<int BM, int BN>
__global__ void __launch_bounds__(256)
gemm_kernel(const gemm_cblock *__restrict__ cbs,
    const gemm_pair *__restrict__ pairs,
    const gemm_tile *__restrict__ tiles, uint32_t ntiles) {

    constexpr int BK = 16;
    constexpr int PAD = 16;
    constexpr int WM = BM / 2, WN = BN / 2;
    constexpr int FM = WM / 16, FN = WN / 16;
    __shared__ double As[BK][BM + PAD];
    __shared__ double Bs[BK][BN + PAD];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm0 = (wave >> 1) * WM;
    const int wn0 = (wave & 1) * WN;
    const int lrow = lane & 15;
    const int lk = lane >> 4;

    for(uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const gemm_tile tile = tiles[t];
        const gemm_cblock cb = cbs[tile.cblock];
        const int m0 = int(tile.tm) * BM, n0 = int(tile.tn) * BN;
        const int mrem = int(cb.m) - m0, nrem = int(cb.n) - n0;

        double4_t acc[FM][FN];
#pragma unroll
        for(int i = 0; i < FM; i++)
#pragma unroll
            for(int j = 0; j < FN; j++) acc[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};

        for(uint32_t p = cb.pair_begin; p < cb.pair_end; p++) {
            const gemm_pair pr = pairs[p];
            const int K = int(pr.k);
            const double d = pr.d;
            const bool TA = (pr.flags & kPairAKMajor) != 0;
            const bool TB = (pr.flags & kPairBNMajor) != 0;
            for(int k0 = 0; k0 < K; k0 += BK) {
                const int krem = K - k0;
                // A tile -> As[k][m]
                if(!TA) {
                    // A(m, k) at a[m * lda + k]: threads run along k
#pragma unroll
                    for(int e = tid; e < BM * BK; e += 256) {
                        const int kk = e % BK, mm = e / BK;
                        double v = 0.0;
                        if(mm < mrem && kk < krem)
                            v = pr.a[size_t(m0 + mm) * pr.lda + (k0 + kk)];
                        As[kk][mm] = v * d;
                    }
                } else {
                    // A(m, k) at a[k * lda + m]: threads run along m
#pragma unroll
                    for(int e = tid; e < BM * BK; e += 256) {
                        const int mm = e % BM, kk = e / BM;
                        double v = 0.0;
                        if(mm < mrem && kk < krem)
                            v = pr.a[size_t(k0 + kk) * pr.lda + (m0 + mm)];
                        As[kk][mm] = v * d;
                    }
                }
                // B tile -> Bs[k][n]
                if(!TB) {
                    // B(k, n) at b[k * ldb + n]
#pragma unroll
                    for(int e = tid; e < BN * BK; e += 256) {
                        const int nn = e % BN, kk = e / BN;
                        double v = 0.0;
                        if(nn < nrem && kk < krem)
                            v = pr.b[size_t(k0 + kk) * pr.ldb + (n0 + nn)];
                        Bs[kk][nn] = v;
                    }
                } else {
                    // B(k, n) at b[n * ldb + k]
#pragma unroll
                    for(int e = tid; e < BN * BK; e += 256) {
                        const int kk = e % BK, nn = e / BK;
                        double v = 0.0;
                        if(nn < nrem && kk < krem)
                            v = pr.b[size_t(n0 + nn) * pr.ldb + (k0 + kk)];
                        Bs[kk][nn] = v;
                    }
                }
                __syncthreads();
#pragma unroll
                for(int kb = 0; kb < BK; kb += 4) {
                    double af[FM], bf[FN];
#pragma unroll
                    for(int i = 0; i < FM; i++) af[i] = As[kb + lk][wm0 + i * 16 + lrow];
#pragma unroll
                    for(int j = 0; j < FN; j++) bf[j] = Bs[kb + lk][wn0 + j * 16 + lrow];
#pragma unroll
                    for(int i = 0; i < FM; i++)
#pragma unroll
                        for(int j = 0; j < FN; j++)
                            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                af[i], bf[j], acc[i][j], 0, 0, 0);
                }
                __syncthreads();
            }
        }

        // Epilogue: each C element of the tile is written exactly once.
#pragma unroll
        for(int i = 0; i < FM; i++) {
#pragma unroll
            for(int j = 0; j < FN; j++) {
                const int col = wn0 + j * 16 + lrow;
#pragma unroll
                for(int v = 0; v < 4; v++) {
                    const int row = wm0 + i * 16 + lk + 4 * v;
                    if(row < mrem && col < nrem) {
                        double *cp = cb.c + size_t(m0 + row) * cb.ldc + (n0 + col);
                        double x = acc[i][j][v];
                        if(!cb.zero) x += *cp;
                        *cp = x;
                    }
                }
            }
        }
    }
}

template<int BM, int BN>
static hipError_t launch_gemm_t(const gemm_cblock *cb,
    const gemm_pair *pairs, const gemm_tile *tiles, uint32_t ntiles,
    uint32_t nwg, hipStream_t s) {
    hipLaunchKernelGGL((gemm_kernel<BM, BN>), dim3(nwg), dim3(256), 0, s, cb, pairs, tiles, ntiles);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Direct split-K FP64 MFMA GEMM.  A workgroup of W waves owns one
// (16 FM) x (16 FN) C tile; the k chunks (16 k each) of the tile's whole
// pair list, taken in list order, are dealt round-robin to the W waves, so
// a C block with a long contraction list keeps W waves busy instead of one.
// Operands go from memory straight into MFMA fragments: per chunk, lane l
// (r = l % 16, g = l / 16) loads A[m0 + r][k0 + 4g .. 4g + 3] and
// B[k0 + 4g .. 4g + 3][n0 + r]; MFMA step j feeds element j, so fragment
// slot g of step j carries k = k0 + 4g + j for both operands and each k is
// summed once.  Loads are branch-free: rows or columns past the block edge
// read row 0 (their outputs are never stored), k past the pair's extent
// reads element K - 1 and is replaced by zero.  Each wave double-buffers
// its chunks (the next chunk is in flight while the current one is on the
// matrix cores).  The W partial tiles are summed in LDS in wave order
// 0, 1, ..., W - 1, so results do not depend on timing.

typedef const double __attribute__((address_space(1))) gdouble;

//! Advance (p, c) to the first chunk at or after chunk c of pair p.
__device__ __forceinline__ void seek_chunk(const gemm_pair *__restrict__ pairs,
    uint32_t end, uint32_t &p, int &c) {
    while(p < end) {
        const int nch = (int(pairs[p].k) + 15) >> 4;
        if(c < nch) break;
        c -= nch;
        p++;
    }
}

//! Element of a uniform base pointer at a 32-bit byte offset (one
//! address register per load, base in scalar registers).
__device__ __forceinline__ double ldg(const double *base, uint32_t off_bytes) {
    return *(gdouble *)((const __attribute__((address_space(1))) char *)base + off_bytes);
}

// Loads of one chunk.  k beyond the pair's extent reads element K - 1 (in
// bounds); mma_chunk multiplies those A values by zero, so the loaded values
// are not consumed here and the loads of the next chunk stay in flight
// while the current chunk is on the matrix cores.  Offsets are 32-bit
// byte offsets: the host refuses plans with an operand or C block of 2^29
// or more elements (4 GiB), so every offset fits.
template<int FM, int FN, bool TA, bool TB>
__device__ __forceinline__ void load_chunk(double (&af)[FM][4], double (&bf)[FN][4],
    const gemm_pair &pr, int c, const int (&arow)[FM], const int (&bcol)[FN], int lg) {
    const uint32_t K = pr.k;
    const uint32_t kk = 16 * c + 4 * lg;
    uint32_t kc[4];
#pragma unroll
    for(int j = 0; j < 4; j++) kc[j] = min(kk + j, K - 1);
#pragma unroll
    for(int i = 0; i < FM; i++) {
        const uint32_t r = uint32_t(arow[i]);
#pragma unroll
        for(int j = 0; j < 4; j++)
            af[i][j] = ldg(pr.a, 8u * (TA ? kc[j] * pr.lda + r : r * pr.lda + kc[j]));
    }
#pragma unroll
    for(int f = 0; f < FN; f++) {
        const uint32_t n = uint32_t(bcol[f]);
#pragma unroll
        for(int j = 0; j < 4; j++)
            bf[f][j] = ldg(pr.b, 8u * (TB ? n * pr.ldb + kc[j] : kc[j] * pr.ldb + n));
    }
}

//! A scaled by the pair's coefficient, by zero for k beyond its extent.
template<int FM, int FN>
__device__ __forceinline__ void mma_chunk(double4_t (&acc)[FM][FN],
    double (&af)[FM][4], const double (&bf)[FN][4], double d, uint32_t K, int c, int lg) {
    const uint32_t kk = 16 * c + 4 * lg;
#pragma unroll
    for(int q = 0; q < 4; q++) {
        const double dq = kk + q < K ? d : 0.0;
#pragma unroll
        for(int i = 0; i < FM; i++) af[i][q] *= dq;
    }
#pragma unroll
    for(int q = 0; q < 4; q++)
#pragma unroll
        for(int i = 0; i < FM; i++)
#pragma unroll
            for(int j = 0; j < FN; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                    af[i][q], bf[j][q], acc[i][j], 0, 0, 0);
}

//! Loads of one chunk in the pair's operand layout (wave-uniform branch).
template<int FM, int FN>
__device__ __forceinline__ void load_chunk_any(double (&af)[FM][4], double (&bf)[FN][4],
    const gemm_pair &pr, int c, const int (&arow)[FM], const int (&bcol)[FN], int lg) {
    switch(pr.flags & (kPairAKMajor | kPairBNMajor)) {
    case 0: load_chunk<FM, FN, false, false>(af, bf, pr, c, arow, bcol, lg); break;
    case kPairAKMajor: load_chunk<FM, FN, true, false>(af, bf, pr, c, arow, bcol, lg); break;
    case kPairBNMajor: load_chunk<FM, FN, false, true>(af, bf, pr, c, arow, bcol, lg); break;
    default: load_chunk<FM, FN, true, true>(af, bf, pr, c, arow, bcol, lg); break;
    }
}

template<int FM, int FN, int W>
__global__ void __launch_bounds__(64 * W)
gemm_direct_kernel(const gemm_cblock *__restrict__ cbs,
    const gemm_pair *__restrict__ pairs,
    const gemm_tile *__restrict__ tiles, uint32_t ntiles) {

    constexpr int TM = 16 * FM, TN = 16 * FN;
    // partial tiles of waves 1..W-1 in fragment order: [w][i][j][v][lane]
    __shared__ double red[(W > 1 ? W - 1 : 1) * FM * FN * 4 * 64];
    const int lane = threadIdx.x & 63;
    // wave-uniform: pair metadata and chunk bookkeeping stay in scalar
    // registers and scalar loads
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lr = lane & 15, lg = lane >> 4;

    for(uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const gemm_tile tile = tiles[t];
        const gemm_cblock cb = cbs[tile.cblock];
        const int m0 = int(tile.tm) * TM, n0 = int(tile.tn) * TN;
        const int mrem = int(cb.m) - m0, nrem = int(cb.n) - n0;

        double4_t acc[FM][FN];
#pragma unroll
        for(int i = 0; i < FM; i++)
#pragma unroll
            for(int j = 0; j < FN; j++) acc[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};

        int arow[FM], bcol[FN];
#pragma unroll
        for(int i = 0; i < FM; i++) arow[i] = (16 * i + lr) < mrem ? m0 + 16 * i + lr : m0;
#pragma unroll
        for(int j = 0; j < FN; j++) bcol[j] = (16 * j + lr) < nrem ? n0 + 16 * j + lr : n0;

        const uint32_t pend = cb.pair_end;
        uint32_t px = cb.pair_begin;
        int cx = wave;
        seek_chunk(pairs, pend, px, cx);
        if(px < pend) {
            double ax[FM][4], bx[FN][4], ay[FM][4], by[FN][4];
            gemm_pair prx = pairs[px];
            load_chunk_any<FM, FN>(ax, bx, prx, cx, arow, bcol, lg);
            while(true) {
                uint32_t py = px;
                int cy = cx + W;
                seek_chunk(pairs, pend, py, cy);
                const bool hy = py < pend;
                if(!hy) { py = px; cy = cx; }       // dummy in-bounds load
                const gemm_pair pry = pairs[py];
                load_chunk_any<FM, FN>(ay, by, pry, cy, arow, bcol, lg);
                mma_chunk<FM, FN>(acc, ax, bx, prx.d, prx.k, cx, lg);
                if(!hy) break;
                uint32_t pz = py;
                int cz = cy + W;
                seek_chunk(pairs, pend, pz, cz);
                const bool hz = pz < pend;
                if(!hz) { pz = py; cz = cy; }
                prx = pairs[pz];
                load_chunk_any<FM, FN>(ax, bx, prx, cz, arow, bcol, lg);
                mma_chunk<FM, FN>(acc, ay, by, pry.d, pry.k, cy, lg);
                if(!hz) break;
                px = pz;
                cx = cz;
            }
        }

        if(W > 1) {
            if(wave > 0) {
                double *r = red + (wave - 1) * (FM * FN * 4 * 64);
#pragma unroll
                for(int i = 0; i < FM; i++)
#pragma unroll
                    for(int j = 0; j < FN; j++)
#pragma unroll
                        for(int v = 0; v < 4; v++)
                            r[((i * FN + j) * 4 + v) * 64 + lane] = acc[i][j][v];
            }
            __syncthreads();
            if(wave == 0) {
                for(int w = 1; w < W; w++) {
                    const double *r = red + (w - 1) * (FM * FN * 4 * 64);
#pragma unroll
                    for(int i = 0; i < FM; i++)
#pragma unroll
                        for(int j = 0; j < FN; j++)
#pragma unroll
                            for(int v = 0; v < 4; v++)
                                acc[i][j][v] += r[((i * FN + j) * 4 + v) * 64 + lane];
                }
            }
        }

        if(wave == 0) {
#pragma unroll
            for(int i = 0; i < FM; i++) {
#pragma unroll
                for(int j = 0; j < FN; j++) {
                    const int col = 16 * j + lr;
#pragma unroll
                    for(int v = 0; v < 4; v++) {
                        const int row = 16 * i + lg + 4 * v;
                        if(row < mrem && col < nrem) {
                            double *cp = cb.c + size_t(m0 + row) * cb.ldc + (n0 + col);
                            double x = acc[i][j][v];
                            if(!cb.zero) x += *cp;
                            *cp = x;
                        }
                    }
                }
            }
        }
        if(W > 1) __syncthreads();   // red is reused by the next tile
    }
}

template<int FM, int FN, int W>
static hipError_t launch_direct_t(const gemm_cblock *cb,
    const gemm_pair *pairs, const gemm_tile *tiles, uint32_t ntiles,
    uint32_t nwg, hipStream_t s) {
    hipLaunchKernelGGL((gemm_direct_kernel<FM, FN, W>), dim3(nwg), dim3(64 * W), 0, s,
        cb, pairs, tiles, ntiles);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Stacked split-K FP64 MFMA GEMM: the direct kernel's scheme over a stack of
// C blocks (rows of several blocks in one tile, one read of each B chunk for
// all of them).  Per pair, each lane takes the A row of its fragment rows
// from the stack's row table (row start, coefficient, stride) and the C row
// from the stack's output-row table.  The chunks of the stack's pair list
// are dealt to the W waves as in the direct kernel, so a C element gets the
// same sums in the same order as there.

template<int FM, int FN>
struct st_rowsrc {
    const double *p[FM];
    double d[FM];
    uint32_t lda[FM];
    int32_t no[FN];           //!< B column offsets (offset-table pairs)
};

//! Offset of index k in a table linear in two index groups:
//! (k / f.d) s0 + (k % f.d) s1.
__device__ __forceinline__ int32_t off_arith(uint32_t k, const fastdiv &f, int32_t s0, int32_t s1) {
    const uint32_t q = fdiv(k, f);
    return int32_t(q) * s0 + int32_t(k - q * f.d) * s1;
}

//! B k offsets of one chunk: none (KT 0), read from the table (KT 1) or
//! computed (KT 2).
template<int KT>
__device__ __forceinline__ int32_t koff_at(const stack_pair &sp, uint32_t k) {
    if(KT == 1) return sp.koff[k];
    if(KT == 2) return off_arith(k, sp.kd, sp.ks0, sp.ks1);
    return 0;
}

typedef double double2_t __attribute__((ext_vector_type(2)));
typedef const double2_t __attribute__((address_space(1))) gdouble2;

// Operand loads of one chunk (16 k) of a stacked tile.  Lane (lr, lg) holds
// k = 16 c + 4 lg + q (q = 0..3) of its fragment rows (A) and columns (B).
// A k past the extent reads an in-bounds element and is zeroed through the
// A scale.  A and B are loaded independently, each in the form its layout
// allows:
//   A: k-contiguous rows, 16-byte aligned, even extent (kPairVecA): two
//      16-byte loads per row; k-major (kPairAKMajor): one element per k,
//      consecutive rows at consecutive addresses; else one element per k.
//   B: k-contiguous columns, aligned, even extent (kPairVecB); offset tables
//      with adjacent even k pairs (kPairVecBT): two 16-byte loads per
//      column; offset tables (computed or read), n-major or k-major: one
//      element per k.
// With an even extent the pairs (kk, kk + 1) and (kk + 2, kk + 3) (kk even)
// are each inside or outside the extent; an outside pair reads the last
// pair.
template<int FM, int FN>
__device__ __forceinline__ void load_a_st(double (&af)[FM][4], const st_rowsrc<FM, FN> &ar,
    const stack_pair &sp, int c, int lg) {
    const uint32_t K = sp.k;
    const uint32_t kk = 16 * c + 4 * lg;
    if(sp.flags & kPairVecA) {
        const uint32_t k0 = min(kk, K - 2), k2 = min(kk + 2, K - 2);
#pragma unroll
        for(int i = 0; i < FM; i++) {
            gdouble *pa = (gdouble *)ar.p[i];
            const double2_t v0 = *(gdouble2 *)(pa + k0);
            const double2_t v1 = *(gdouble2 *)(pa + k2);
            af[i][0] = v0.x; af[i][1] = v0.y; af[i][2] = v1.x; af[i][3] = v1.y;
        }
        return;
    }
    uint32_t kc[4];
#pragma unroll
    for(int j = 0; j < 4; j++) kc[j] = min(kk + j, K - 1);
    if(sp.flags & kPairAKMajor) {
#pragma unroll
        for(int i = 0; i < FM; i++) {
            gdouble *pa = (gdouble *)ar.p[i];
#pragma unroll
            for(int j = 0; j < 4; j++) af[i][j] = pa[kc[j] * ar.lda[i]];
        }
    } else {
#pragma unroll
        for(int i = 0; i < FM; i++) {
            gdouble *pa = (gdouble *)ar.p[i];
#pragma unroll
            for(int j = 0; j < 4; j++) af[i][j] = pa[kc[j]];
        }
    }
}

template<int FN, int KT>
__device__ __forceinline__ void load_b_vt(double (&bf)[FN][4], const int32_t (&no)[FN],
    const stack_pair &sp, uint32_t k0, uint32_t k2) {
    // table pairs: koff[k0 + 1] = koff[k0] + 1, all offsets even
    const int32_t o0 = koff_at<KT>(sp, k0), o2 = koff_at<KT>(sp, k2);
    gdouble *pb = (gdouble *)sp.b;
#pragma unroll
    for(int f = 0; f < FN; f++) {
        const double2_t v0 = *(gdouble2 *)(pb + uint32_t(o0 + no[f]));
        const double2_t v1 = *(gdouble2 *)(pb + uint32_t(o2 + no[f]));
        bf[f][0] = v0.x; bf[f][1] = v0.y; bf[f][2] = v1.x; bf[f][3] = v1.y;
    }
}

template<int FN, int KT>
__device__ __forceinline__ void load_b_el(double (&bf)[FN][4], const int32_t (&no)[FN],
    const stack_pair &sp, const uint32_t (&kc)[4], const int (&bcol)[FN]) {
    int32_t ko[4];
#pragma unroll
    for(int j = 0; j < 4; j++) ko[j] = koff_at<KT>(sp, kc[j]);
#pragma unroll
    for(int f = 0; f < FN; f++) {
        const uint32_t n = uint32_t(bcol[f]);
#pragma unroll
        for(int j = 0; j < 4; j++) {
            const uint32_t e = KT ? uint32_t(ko[j] + no[f])
                                  : ((sp.flags & kPairBNMajor) ? n * sp.ldb + kc[j] : kc[j] * sp.ldb + n);
            bf[f][j] = ldg(sp.b, 8u * e);
        }
    }
}

template<int FM, int FN>
__device__ __forceinline__ void load_b_st(double (&bf)[FN][4], const st_rowsrc<FM, FN> &ar,
    const stack_pair &sp, int c, const int (&bcol)[FN], int lg) {
    const uint32_t K = sp.k;
    const uint32_t kk = 16 * c + 4 * lg;
    const bool arith = (sp.flags & kPairOffArith) != 0;
    if(sp.flags & (kPairVecB | kPairVecBT)) {
        const uint32_t k0 = min(kk, K - 2), k2 = min(kk + 2, K - 2);
        if(sp.flags & kPairVecB) {
#pragma unroll
            for(int f = 0; f < FN; f++) {
                gdouble *pb = (gdouble *)sp.b + size_t(uint32_t(bcol[f]) * sp.ldb);
                const double2_t v0 = *(gdouble2 *)(pb + k0);
                const double2_t v1 = *(gdouble2 *)(pb + k2);
                bf[f][0] = v0.x; bf[f][1] = v0.y; bf[f][2] = v1.x; bf[f][3] = v1.y;
            }
        } else if(arith) load_b_vt<FN, 2>(bf, ar.no, sp, k0, k2);
        else load_b_vt<FN, 1>(bf, ar.no, sp, k0, k2);
        return;
    }
    uint32_t kc[4];
#pragma unroll
    for(int j = 0; j < 4; j++) kc[j] = min(kk + j, K - 1);
    if(arith) load_b_el<FN, 2>(bf, ar.no, sp, kc, bcol);
    else if(sp.koff) load_b_el<FN, 1>(bf, ar.no, sp, kc, bcol);
    else load_b_el<FN, 0>(bf, ar.no, sp, kc, bcol);
}

template<int FM, int FN>
__device__ __forceinline__ void load_chunk_st_any(double (&af)[FM][4], double (&bf)[FN][4],
    const st_rowsrc<FM, FN> &ar, const stack_pair &sp, int c, const int (&bcol)[FN], int lg) {
    load_a_st<FM, FN>(af, ar, sp, c, lg);
    load_b_st<FM, FN>(bf, ar, sp, c, bcol, lg);
}

template<int FM, int FN>
__device__ __forceinline__ void st_rows(st_rowsrc<FM, FN> &r, const stack_arow *__restrict__ arow,
    uint32_t base, const int (&srow)[FM], const stack_pair &sp, const int (&bcol)[FN]) {
#pragma unroll
    for(int i = 0; i < FM; i++) {
        const stack_arow e = arow[base + srow[i]];
        r.p[i] = e.p;
        r.d[i] = e.d;
        r.lda[i] = e.lda;
    }
    if(sp.flags & kPairOffArith) {
#pragma unroll
        for(int f = 0; f < FN; f++) r.no[f] = off_arith(uint32_t(bcol[f]), sp.nd, sp.ns0, sp.ns1);
    } else {
#pragma unroll
        for(int f = 0; f < FN; f++) r.no[f] = sp.noff ? sp.noff[bcol[f]] : 0;
    }
}

template<int FM, int FN>
__device__ __forceinline__ void mma_chunk_st(double4_t (&acc)[FM][FN],
    double (&af)[FM][4], const double (&bf)[FN][4], const double (&d)[FM],
    uint32_t K, int c, int lg) {
    const uint32_t kk = 16 * c + 4 * lg;
#pragma unroll
    for(int q = 0; q < 4; q++) {
        const bool kv = kk + q < K;
#pragma unroll
        for(int i = 0; i < FM; i++) af[i][q] *= kv ? d[i] : 0.0;
    }
#pragma unroll
    for(int q = 0; q < 4; q++)
#pragma unroll
        for(int i = 0; i < FM; i++)
#pragma unroll
            for(int j = 0; j < FN; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                    af[i][q], bf[j][q], acc[i][j], 0, 0, 0);
}

__device__ __forceinline__ void seek_chunk_st(const stack_pair *__restrict__ sp,
    uint32_t end, uint32_t &p, int &c) {
    while(p < end) {
        const int nch = (int(sp[p].k) + 15) >> 4;
        if(c < nch) break;
        c -= nch;
        p++;
    }
}

//! Operand forms of a stacked pair with a branch-free load path: A in
//! 16-byte row loads and B k-contiguous (VA_VB), row-major (VA_ROW), through
//! computed offsets element by element (VA_AR) or in 16-byte pairs
//! (VA_VTAR); everything else takes the general path (ANY).
enum { ST_FORM_ANY = 0, ST_FORM_VA_VB = 1, ST_FORM_VA_ROW = 2, ST_FORM_VA_AR = 3,
       ST_FORM_VA_VTAR = 4 };

__device__ __forceinline__ int st_form(const stack_pair &sp) {
    const uint32_t f = sp.flags;
    if(!(f & kPairVecA)) return ST_FORM_ANY;
    if(f & kPairVecB) return ST_FORM_VA_VB;
    if(f & kPairVecBT) return (f & kPairOffArith) ? ST_FORM_VA_VTAR : ST_FORM_ANY;
    if(f & kPairOffArith) return ST_FORM_VA_AR;
    if(!sp.koff && !(f & kPairBNMajor)) return ST_FORM_VA_ROW;
    return ST_FORM_ANY;
}

//! Row sources of one pair without the B column offsets (computed or read
//! when the operands are loaded, so the row table loads do not wait for the
//! pair's descriptor).
template<int FM, int FN>
__device__ __forceinline__ void st_rows_a(st_rowsrc<FM, FN> &r, const stack_arow *__restrict__ arow,
    uint32_t base, const int (&srow)[FM]) {
#pragma unroll
    for(int i = 0; i < FM; i++) {
        const stack_arow e = arow[base + srow[i]];
        r.p[i] = e.p;
        r.d[i] = e.d;
        r.lda[i] = e.lda;
    }
}

template<int FM, int FN>
__device__ __forceinline__ void st_cols(st_rowsrc<FM, FN> &r, const stack_pair &sp,
    const int (&bcol)[FN]) {
    if(sp.flags & kPairOffArith) {
#pragma unroll
        for(int f = 0; f < FN; f++) r.no[f] = off_arith(uint32_t(bcol[f]), sp.nd, sp.ns0, sp.ns1);
    } else {
#pragma unroll
        for(int f = 0; f < FN; f++) r.no[f] = sp.noff ? sp.noff[bcol[f]] : 0;
    }
}

//! B operand forms without a condition on A (for kernels that stage A
//! separately): the same forms as st_form.
__device__ __forceinline__ int st_bform(const stack_pair &sp) {
    const uint32_t f = sp.flags;
    if(f & kPairVecB) return ST_FORM_VA_VB;
    if(f & kPairVecBT) return (f & kPairOffArith) ? ST_FORM_VA_VTAR : ST_FORM_ANY;
    if(f & kPairOffArith) return ST_FORM_VA_AR;
    if(!sp.koff && !(f & kPairBNMajor)) return ST_FORM_VA_ROW;
    return ST_FORM_ANY;
}

//! B operand of one chunk in the pair's B form (FORM != ANY: no branches).
template<int FM, int FN, int FORM>
__device__ __forceinline__ void load_b_form(double (&bf)[FN][4], st_rowsrc<FM, FN> &ar,
    const stack_pair &sp, int c, const int (&bcol)[FN], int lg) {
    if(FORM == ST_FORM_ANY) {
        st_cols<FM, FN>(ar, sp, bcol);
        load_b_st<FM, FN>(bf, ar, sp, c, bcol, lg);
        return;
    }
    const uint32_t K = sp.k;
    const uint32_t kk = 16 * c + 4 * lg;
    const uint32_t k0 = min(kk, K - 2), k2 = min(kk + 2, K - 2);
    if(FORM == ST_FORM_VA_VB) {
#pragma unroll
        for(int f = 0; f < FN; f++) {
            gdouble *pb = (gdouble *)sp.b + size_t(uint32_t(bcol[f]) * sp.ldb);
            const double2_t v0 = *(gdouble2 *)(pb + k0);
            const double2_t v1 = *(gdouble2 *)(pb + k2);
            bf[f][0] = v0.x; bf[f][1] = v0.y; bf[f][2] = v1.x; bf[f][3] = v1.y;
        }
        return;
    }
    int32_t no[FN];
#pragma unroll
    for(int f = 0; f < FN; f++)
        no[f] = FORM == ST_FORM_VA_ROW ? bcol[f] : off_arith(uint32_t(bcol[f]), sp.nd, sp.ns0, sp.ns1);
    if(FORM == ST_FORM_VA_VTAR) {
        load_b_vt<FN, 2>(bf, no, sp, k0, k2);
        return;
    }
    uint32_t kc[4];
#pragma unroll
    for(int j = 0; j < 4; j++) kc[j] = min(kk + j, K - 1);
    int32_t ko[4];
#pragma unroll
    for(int j = 0; j < 4; j++)
        ko[j] = FORM == ST_FORM_VA_ROW ? int32_t(kc[j] * sp.ldb) : off_arith(kc[j], sp.kd, sp.ks0, sp.ks1);
#pragma unroll
    for(int f = 0; f < FN; f++)
#pragma unroll
        for(int j = 0; j < 4; j++) bf[f][j] = ldg(sp.b, 8u * uint32_t(ko[j] + no[f]));
}

//! Operands of one chunk in the pair's form (FORM != ANY: no branches).
template<int FM, int FN, int FORM>
__device__ __forceinline__ void load_form(double (&af)[FM][4], double (&bf)[FN][4],
    st_rowsrc<FM, FN> &ar, const stack_pair &sp, int c, const int (&bcol)[FN], int lg) {
    if(FORM == ST_FORM_ANY) {
        st_cols<FM, FN>(ar, sp, bcol);
        load_chunk_st_any<FM, FN>(af, bf, ar, sp, c, bcol, lg);
        return;
    }
    const uint32_t K = sp.k;
    const uint32_t kk = 16 * c + 4 * lg;
    const uint32_t k0 = min(kk, K - 2), k2 = min(kk + 2, K - 2);
#pragma unroll
    for(int i = 0; i < FM; i++) {
        gdouble *pa = (gdouble *)ar.p[i];
        const double2_t v0 = *(gdouble2 *)(pa + k0);
        const double2_t v1 = *(gdouble2 *)(pa + k2);
        af[i][0] = v0.x; af[i][1] = v0.y; af[i][2] = v1.x; af[i][3] = v1.y;
    }
    load_b_form<FM, FN, FORM>(bf, ar, sp, c, bcol, lg);
}

//! Cursor of a wave over its chunks of a stack's pair list (pair p, chunk
//! c of that pair) with the pair's descriptor and the lane's row sources.
//! After the last chunk it stays on it with valid = false, so a load issued
//! from it stays in bounds.
template<int FM, int FN>
struct st_cursor {
    uint32_t p;
    int c;
    bool valid;
    stack_pair sp;
    st_rowsrc<FM, FN> r;
};

//! Moves the cursor W chunks on.  A new pair's descriptor and row sources
//! are loaded here, one chunk before the chunk's operands.
template<int FM, int FN, int W>
__device__ __forceinline__ void st_advance(st_cursor<FM, FN> &u,
    const stack_pair *__restrict__ spairs, const stack_arow *__restrict__ arow,
    const stack_desc &st, const int (&srow)[FM], const int (&bcol)[FN]) {
    if(!u.valid) return;
    int c = u.c + W;
    const int nch = (int(u.sp.k) + 15) >> 4;
    if(c < nch) { u.c = c; return; }
    uint32_t p = u.p + 1;
    const uint32_t pend = st.pair_begin + st.npairs;
    c -= nch;
    seek_chunk_st(spairs, pend, p, c);
    if(p >= pend) { u.valid = false; return; }
    u.sp = spairs[p];
    st_rows<FM, FN>(u.r, arow, st.arow_begin + (p - st.pair_begin) * st.rows, srow, u.sp, bcol);
    u.p = p;
    u.c = c;
}

//! As st_advance, loading only the new pair's row sources (st_rows_a).
template<int FM, int FN, int W>
__device__ __forceinline__ void st_advance_a(st_cursor<FM, FN> &u,
    const stack_pair *__restrict__ spairs, const stack_arow *__restrict__ arow,
    const stack_desc &st, const int (&srow)[FM]) {
    if(!u.valid) return;
    int c = u.c + W;
    const int nch = (int(u.sp.k) + 15) >> 4;
    if(c < nch) { u.c = c; return; }
    uint32_t p = u.p + 1;
    const uint32_t pend = st.pair_begin + st.npairs;
    c -= nch;
    seek_chunk_st(spairs, pend, p, c);
    if(p >= pend) { u.valid = false; return; }
    u.sp = spairs[p];
    st_rows_a<FM, FN>(u.r, arow, st.arow_begin + (p - st.pair_begin) * st.rows, srow);
    u.p = p;
    u.c = c;
}

//! Loads chunk u (in its pair's form) and multiplies the chunk held in
//! (ac, bc): one basic block per form, so the scheduler can interleave the
//! loads with the products.
template<int FM, int FN>
__device__ __forceinline__ void st_step(double4_t (&acc)[FM][FN], double (&an)[FM][4],
    double (&bn)[FN][4], st_cursor<FM, FN> &u, double (&ac)[FM][4], const double (&bc)[FN][4],
    const double (&dc)[FM], uint32_t kc, int cc, const int (&bcol)[FN], int lg) {
    switch(st_form(u.sp)) {
    case ST_FORM_VA_VB:
        load_form<FM, FN, ST_FORM_VA_VB>(an, bn, u.r, u.sp, u.c, bcol, lg);
        mma_chunk_st<FM, FN>(acc, ac, bc, dc, kc, cc, lg);
        break;
    case ST_FORM_VA_ROW:
        load_form<FM, FN, ST_FORM_VA_ROW>(an, bn, u.r, u.sp, u.c, bcol, lg);
        mma_chunk_st<FM, FN>(acc, ac, bc, dc, kc, cc, lg);
        break;
    case ST_FORM_VA_AR:
        load_form<FM, FN, ST_FORM_VA_AR>(an, bn, u.r, u.sp, u.c, bcol, lg);
        mma_chunk_st<FM, FN>(acc, ac, bc, dc, kc, cc, lg);
        break;
    case ST_FORM_VA_VTAR:
        load_form<FM, FN, ST_FORM_VA_VTAR>(an, bn, u.r, u.sp, u.c, bcol, lg);
        mma_chunk_st<FM, FN>(acc, ac, bc, dc, kc, cc, lg);
        break;
    default:
        load_form<FM, FN, ST_FORM_ANY>(an, bn, u.r, u.sp, u.c, bcol, lg);
        mma_chunk_st<FM, FN>(acc, ac, bc, dc, kc, cc, lg);
        break;
    }
}

template<int FM, int FN, int W, int PIPE>
__device__ __forceinline__ void stacked_body(const stack_desc *__restrict__ stacks,
    const stack_pair *__restrict__ spairs, const stack_arow *__restrict__ arow,
    const stack_crow *__restrict__ crow, const gemm_tile *__restrict__ tiles,
    uint32_t ntiles, double *red);

//! PIPE 0: one chunk in registers at a time; 1: the next chunk's operands
//! load during the current chunk's products; 2: also the descriptor and row
//! sources of the chunk after that; 3: as 2, with one branch-free block of
//! loads and products per operand form.
template<int FM, int FN, int W, int PIPE>
__global__ void __launch_bounds__(64 * W)
gemm_stacked_kernel(const stack_desc *__restrict__ stacks,
    const stack_pair *__restrict__ spairs, const stack_arow *__restrict__ arow,
    const stack_crow *__restrict__ crow, const gemm_tile *__restrict__ tiles,
    uint32_t ntiles) {
    __shared__ double red[(W > 1 ? W - 1 : 1) * FM * FN * 4 * 64];
    stacked_body<FM, FN, W, PIPE>(stacks, spairs, arow, crow, tiles, ntiles, red);
}

//! The same kernel compiled for OCC waves per SIMD (register budget
//! 512 / OCC per lane, accumulators in architectural registers).
template<int FM, int FN, int W, int OCC>
__global__ void __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(OCC)))
gemm_stacked_occ_kernel(const stack_desc *__restrict__ stacks,
    const stack_pair *__restrict__ spairs, const stack_arow *__restrict__ arow,
    const stack_crow *__restrict__ crow, const gemm_tile *__restrict__ tiles,
    uint32_t ntiles) {
    __shared__ double red[(W > 1 ? W - 1 : 1) * FM * FN * 4 * 64];
    stacked_body<FM, FN, W, 2>(stacks, spairs, arow, crow, tiles, ntiles, red);
}

//! One (16 FM) x (16 FN) tile of a stack at rows m0, columns n0.
template<int FM, int FN, int W, int PIPE>
__device__ __forceinline__ void stacked_tile(const stack_desc &st,
    const stack_pair *__restrict__ spairs, const stack_arow *__restrict__ arow,
    const stack_crow *__restrict__ crow, int m0, int n0, double *red) {

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lr = lane & 15, lg = lane >> 4;
    const int mrem = int(st.rows) - m0, nrem = int(st.n) - n0;

    double4_t acc[FM][FN];
#pragma unroll
    for(int i = 0; i < FM; i++)
#pragma unroll
        for(int j = 0; j < FN; j++) acc[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};

    // lanes of rows or columns past the extent read the tile's first row or
    // column: their loads fall on cache lines the other lanes request anyway
    int srow[FM], bcol[FN];
#pragma unroll
    for(int i = 0; i < FM; i++) srow[i] = (16 * i + lr) < mrem ? m0 + 16 * i + lr : m0;
#pragma unroll
    for(int j = 0; j < FN; j++) bcol[j] = (16 * j + lr) < nrem ? n0 + 16 * j + lr : n0;

    const uint32_t pbeg = st.pair_begin, pend = st.pair_begin + st.npairs;
    uint32_t px = pbeg;
    int cx = wave;
    seek_chunk_st(spairs, pend, px, cx);
    if(PIPE == 0) {
        // single buffer: fewer registers, latency hidden by other waves
        double ax[FM][4], bx[FN][4];
        st_rowsrc<FM, FN> rx;
        uint32_t pr = 0xffffffffu;
        stack_pair sx;
        while(px < pend) {
            if(px != pr) {
                sx = spairs[px];
                st_rows<FM, FN>(rx, arow, st.arow_begin + (px - pbeg) * st.rows, srow, sx, bcol);
                pr = px;
            }
            load_chunk_st_any<FM, FN>(ax, bx, rx, sx, cx, bcol, lg);
            mma_chunk_st<FM, FN>(acc, ax, bx, rx.d, sx.k, cx, lg);
            cx += W;
            seek_chunk_st(spairs, pend, px, cx);
        }
    } else if(PIPE == 1 && px < pend) {
        double ax[FM][4], bx[FN][4], ay[FM][4], by[FN][4];
        st_rowsrc<FM, FN> rx, ry;
        stack_pair sx = spairs[px];
        st_rows<FM, FN>(rx, arow, st.arow_begin + (px - pbeg) * st.rows, srow, sx, bcol);
        load_chunk_st_any<FM, FN>(ax, bx, rx, sx, cx, bcol, lg);
        while(true) {
            uint32_t py = px;
            int cy = cx + W;
            seek_chunk_st(spairs, pend, py, cy);
            const bool hy = py < pend;
            if(!hy) { py = px; cy = cx; }       // dummy in-bounds load
            const stack_pair sy = spairs[py];
            if(py == px) ry = rx;
            else st_rows<FM, FN>(ry, arow, st.arow_begin + (py - pbeg) * st.rows, srow, sy, bcol);
            load_chunk_st_any<FM, FN>(ay, by, ry, sy, cy, bcol, lg);
            mma_chunk_st<FM, FN>(acc, ax, bx, rx.d, sx.k, cx, lg);
            if(!hy) break;
            uint32_t pz = py;
            int cz = cy + W;
            seek_chunk_st(spairs, pend, pz, cz);
            const bool hz = pz < pend;
            if(!hz) { pz = py; cz = cy; }
            sx = spairs[pz];
            if(pz == py) rx = ry;
            else st_rows<FM, FN>(rx, arow, st.arow_begin + (pz - pbeg) * st.rows, srow, sx, bcol);
            load_chunk_st_any<FM, FN>(ax, bx, rx, sx, cz, bcol, lg);
            mma_chunk_st<FM, FN>(acc, ay, by, ry.d, sy.k, cy, lg);
            if(!hz) break;
            px = pz;
            cx = cz;
        }
    } else if(PIPE == 2 && px < pend) {
        // three stages: operands of chunk i + 1 and the descriptor and row
        // sources of chunk i + 2 are in flight during the products of chunk i
        st_cursor<FM, FN> u;
        u.p = px;
        u.c = cx;
        u.valid = true;
        u.sp = spairs[px];
        st_rows<FM, FN>(u.r, arow, st.arow_begin + (px - pbeg) * st.rows, srow, u.sp, bcol);
        double ax[FM][4], bx[FN][4], ay[FM][4], by[FN][4];
        double dx[FM], dy[FM];
        uint32_t kx, ky;
        int ccx, ccy;
        load_chunk_st_any<FM, FN>(ax, bx, u.r, u.sp, u.c, bcol, lg);
#pragma unroll
        for(int i = 0; i < FM; i++) dx[i] = u.r.d[i];
        kx = u.sp.k;
        ccx = u.c;
        st_advance<FM, FN, W>(u, spairs, arow, st, srow, bcol);
        while(true) {
            const bool vy = u.valid;            // else a dummy in-bounds load
            load_chunk_st_any<FM, FN>(ay, by, u.r, u.sp, u.c, bcol, lg);
#pragma unroll
            for(int i = 0; i < FM; i++) dy[i] = u.r.d[i];
            ky = u.sp.k;
            ccy = u.c;
            st_advance<FM, FN, W>(u, spairs, arow, st, srow, bcol);
            mma_chunk_st<FM, FN>(acc, ax, bx, dx, kx, ccx, lg);
            if(!vy) break;
            const bool vx = u.valid;
            load_chunk_st_any<FM, FN>(ax, bx, u.r, u.sp, u.c, bcol, lg);
#pragma unroll
            for(int i = 0; i < FM; i++) dx[i] = u.r.d[i];
            kx = u.sp.k;
            ccx = u.c;
            st_advance<FM, FN, W>(u, spairs, arow, st, srow, bcol);
            mma_chunk_st<FM, FN>(acc, ay, by, dy, ky, ccy, lg);
            if(!vx) break;
        }
    } else if(PIPE == 3 && px < pend) {
        // u1: the chunk loaded during the current products; u2: the chunk
        // after it, whose descriptor and row sources load one step earlier
        st_cursor<FM, FN> u1, u2;
        u1.p = px;
        u1.c = cx;
        u1.valid = true;
        u1.sp = spairs[px];
        st_rows_a<FM, FN>(u1.r, arow, st.arow_begin + (px - pbeg) * st.rows, srow);
        double ax[FM][4], bx[FN][4], ay[FM][4], by[FN][4];
        double dx[FM], dy[FM];
        load_form<FM, FN, ST_FORM_ANY>(ax, bx, u1.r, u1.sp, u1.c, bcol, lg);
#pragma unroll
        for(int i = 0; i < FM; i++) dx[i] = u1.r.d[i];
        uint32_t kx = u1.sp.k, ky;
        int ccx = u1.c, ccy;
        u2 = u1;
        st_advance_a<FM, FN, W>(u2, spairs, arow, st, srow);
        while(true) {
            u1 = u2;
            st_advance_a<FM, FN, W>(u2, spairs, arow, st, srow);
            const bool vy = u1.valid;           // else a dummy in-bounds load
            st_step<FM, FN>(acc, ay, by, u1, ax, bx, dx, kx, ccx, bcol, lg);
#pragma unroll
            for(int i = 0; i < FM; i++) dy[i] = u1.r.d[i];
            ky = u1.sp.k;
            ccy = u1.c;
            if(!vy) break;
            u1 = u2;
            st_advance_a<FM, FN, W>(u2, spairs, arow, st, srow);
            const bool vx = u1.valid;
            st_step<FM, FN>(acc, ax, bx, u1, ay, by, dy, ky, ccy, bcol, lg);
#pragma unroll
            for(int i = 0; i < FM; i++) dx[i] = u1.r.d[i];
            kx = u1.sp.k;
            ccx = u1.c;
            if(!vx) break;
        }
    }

    if(W > 1) {

        if(wave > 0) {
            double *r = red + (wave - 1) * (FM * FN * 4 * 64);
#pragma unroll
            for(int i = 0; i < FM; i++)
#pragma unroll
                for(int j = 0; j < FN; j++)
#pragma unroll
                    for(int v = 0; v < 4; v++)
                        r[((i * FN + j) * 4 + v) * 64 + lane] = acc[i][j][v];
        }
        __syncthreads();
        if(wave == 0) {
            for(int w = 1; w < W; w++) {
                const double *r = red + (w - 1) * (FM * FN * 4 * 64);
#pragma unroll
                for(int i = 0; i < FM; i++)
#pragma unroll
                    for(int j = 0; j < FN; j++)
#pragma unroll
                        for(int v = 0; v < 4; v++)
                            acc[i][j][v] += r[((i * FN + j) * 4 + v) * 64 + lane];
            }
        }
    }

    if(wave == 0) {
#pragma unroll
        for(int i = 0; i < FM; i++) {
#pragma unroll
            for(int v = 0; v < 4; v++) {
                const int row = 16 * i + lg + 4 * v;
                if(row >= mrem) continue;
                const stack_crow cr = crow[st.crow_begin + m0 + row];
#pragma unroll
                for(int j = 0; j < FN; j++) {
                    const int col = 16 * j + lr;
                    if(col < nrem) {
                        double *cp = cr.c + (n0 + col);
                        double x = acc[i][j][v];
                        if(!cr.zero) x += *cp;
                        *cp = x;
                    }
                }
            }
        }
    }
    if(W > 1) __syncthreads();   // red is reused by the next tile
}

template<int FM, int FN, int W, int PIPE>
__device__ __forceinline__ void stacked_body(const stack_desc *__restrict__ stacks,
    const stack_pair *__restrict__ spairs, const stack_arow *__restrict__ arow,
    const stack_crow *__restrict__ crow, const gemm_tile *__restrict__ tiles,
    uint32_t ntiles, double *red) {
    for(uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const gemm_tile tile = tiles[t];
        const stack_desc st = stacks[tile.cblock];
        stacked_tile<FM, FN, W, PIPE>(st, spairs, arow, crow, int(tile.tm) * 16 * FM,
            int(tile.tn) * 16 * FN, red);
    }
}

// LDS-staged stacked kernel.  A workgroup of four waves computes a
// (16 FM) x (32 FN) tile: waves g = 0, 1 (k group) x h = 0, 1 (column half).
// The chunks of the stack's pair list are dealt to the two k groups in turn
// (as to the two waves of the two-wave split-K kernels); the two waves of a
// group share the group's A chunk, staged in LDS once per chunk (each lane
// loads 16 bytes of one row, scaled by the row coefficient and zeroed past
// the extent), and load their own B columns into registers.  Group 1's
// partial tile is added to group 0's in LDS at the end, so every C element
// gets the same sums in the same order as in the two-wave split-K kernels.
static const int kLdsRowStride = 18;    // doubles per staged A row (144 bytes)

//! Staging of one A chunk: lane (row srw + 16 j, segment seg) holds k =
//! 16 c + 2 seg and 2 seg + 1 of its rows.
template<int FM, int FN>
__device__ __forceinline__ void lds_stage_load(double2_t (&sa)[FM], const st_rowsrc<FM, FN> &r,
    const stack_pair &sp, int c, int seg) {
    const uint32_t K = sp.k;
    const uint32_t k = 16 * c + 2 * seg;
    if(sp.flags & kPairVecA) {
        const uint32_t k0 = min(k, K - 2);
#pragma unroll
        for(int j = 0; j < FM; j++) sa[j] = *(gdouble2 *)((gdouble *)r.p[j] + k0);
        return;
    }
    const uint32_t k0 = min(k, K - 1), k1 = min(k + 1, K - 1);
    if(sp.flags & kPairAKMajor) {
#pragma unroll
        for(int j = 0; j < FM; j++) {
            gdouble *pa = (gdouble *)r.p[j];
            sa[j].x = pa[k0 * r.lda[j]];
            sa[j].y = pa[k1 * r.lda[j]];
        }
    } else {
#pragma unroll
        for(int j = 0; j < FM; j++) {
            gdouble *pa = (gdouble *)r.p[j];
            sa[j].x = pa[k0];
            sa[j].y = pa[k1];
        }
    }
}

template<int FM, int FN>
__device__ __forceinline__ void lds_stage_store(double *buf, const double2_t (&sa)[FM],
    const st_rowsrc<FM, FN> &r, uint32_t K, int c, int srw, int seg) {
    const uint32_t k = 16 * c + 2 * seg;
    const bool v0 = k < K, v1 = k + 1 < K;
#pragma unroll
    for(int j = 0; j < FM; j++) {
        double2_t w;
        w.x = sa[j].x * (v0 ? r.d[j] : 0.0);
        w.y = sa[j].y * (v1 ? r.d[j] : 0.0);
        *(double2_t *)(buf + (srw + 16 * j) * kLdsRowStride + 2 * seg) = w;
    }
}

//! One step of the LDS-staged tile: stage chunk un (A through registers into
//! wbuf, zeros when un is past the group's last chunk; B into bn) while the
//! chunk staged in rbuf (with B in bc) is multiplied.  One basic block per
//! B form.
template<int FM, int FN, int FORM>
__device__ __forceinline__ void lds_step(double4_t (&acc)[FM][FN], double (&bn)[FN][4],
    const double (&bc)[FN][4], st_cursor<FM, FN> &un, const double *rbuf, double *wbuf,
    const int (&bcol)[FN], int lr, int lg, int srw, int seg) {
    double2_t sa[FM];
    lds_stage_load<FM, FN>(sa, un.r, un.sp, un.c, seg);
    load_b_form<FM, FN, FORM>(bn, un.r, un.sp, un.c, bcol, lg);
    double af[FM][4];
#pragma unroll
    for(int i = 0; i < FM; i++) {
        const double *rp = rbuf + (16 * i + lr) * kLdsRowStride + 4 * lg;
        const double2_t v0 = *(const double2_t *)rp;
        const double2_t v1 = *(const double2_t *)(rp + 2);
        af[i][0] = v0.x; af[i][1] = v0.y; af[i][2] = v1.x; af[i][3] = v1.y;
    }
#pragma unroll
    for(int q = 0; q < 4; q++)
#pragma unroll
        for(int i = 0; i < FM; i++)
#pragma unroll
            for(int j = 0; j < FN; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i][q], bc[j][q], acc[i][j],
                    0, 0, 0);
    lds_stage_store<FM, FN>(wbuf, sa, un.r, un.valid ? un.sp.k : 0u, un.c, srw, seg);
}

template<int FM, int FN>
__device__ __forceinline__ void lds_step_any(double4_t (&acc)[FM][FN], double (&bn)[FN][4],
    const double (&bc)[FN][4], st_cursor<FM, FN> &un, const double *rbuf, double *wbuf,
    const int (&bcol)[FN], int lr, int lg, int srw, int seg) {
    switch(st_bform(un.sp)) {
    case ST_FORM_VA_VB:
        lds_step<FM, FN, ST_FORM_VA_VB>(acc, bn, bc, un, rbuf, wbuf, bcol, lr, lg, srw, seg); break;
    case ST_FORM_VA_ROW:
        lds_step<FM, FN, ST_FORM_VA_ROW>(acc, bn, bc, un, rbuf, wbuf, bcol, lr, lg, srw, seg); break;
    case ST_FORM_VA_AR:
        lds_step<FM, FN, ST_FORM_VA_AR>(acc, bn, bc, un, rbuf, wbuf, bcol, lr, lg, srw, seg); break;
    case ST_FORM_VA_VTAR:
        lds_step<FM, FN, ST_FORM_VA_VTAR>(acc, bn, bc, un, rbuf, wbuf, bcol, lr, lg, srw, seg); break;
    default:
        lds_step<FM, FN, ST_FORM_ANY>(acc, bn, bc, un, rbuf, wbuf, bcol, lr, lg, srw, seg); break;
    }
}

template<int FM, int FN>
__device__ __forceinline__ void lds_tile(const stack_desc &st,
    const stack_pair *__restrict__ spairs, const stack_arow *__restrict__ arow,
    const stack_crow *__restrict__ crow, int m0, int n0, double *lds) {
    constexpr int ABUF = 16 * FM * kLdsRowStride;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = wave & 1, h = wave >> 1;
    const int lr = lane & 15, lg = lane >> 4;
    const int lig = lane + 64 * h;
    const int seg = lig & 7, srw = lig >> 3;
    const int nw = n0 + 16 * FN * h;
    const int mrem = int(st.rows) - m0, nrem = int(st.n) - nw;
    double *abuf = lds + g * 2 * ABUF;

    double4_t acc[FM][FN];
#pragma unroll
    for(int i = 0; i < FM; i++)
#pragma unroll
        for(int j = 0; j < FN; j++) acc[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};

    int srow[FM], bcol[FN];
#pragma unroll
    for(int j = 0; j < FM; j++) srow[j] = (srw + 16 * j) < mrem ? m0 + srw + 16 * j : m0;
#pragma unroll
    for(int f = 0; f < FN; f++) bcol[f] = (16 * f + lr) < nrem ? nw + 16 * f + lr : n0;

    const int nchunk = int(st.nchunk);
    if(nchunk > 0) {
        const uint32_t pbeg = st.pair_begin, pend = st.pair_begin + st.npairs;
        // the group's first chunk; a group without chunks keeps an in-bounds
        // cursor on chunk 0, stages zeros and only joins the barriers
        st_cursor<FM, FN> u, u2;
        u.p = pbeg;
        u.c = g;
        seek_chunk_st(spairs, pend, u.p, u.c);
        u.valid = u.p < pend;
        if(!u.valid) { u.p = pbeg; u.c = 0; }
        u.sp = spairs[u.p];
        st_rows_a<FM, FN>(u.r, arow, st.arow_begin + (u.p - pbeg) * st.rows, srow);
        st_cols<FM, FN>(u.r, u.sp, bcol);
        double bx[FN][4], by[FN][4];
        {
            double2_t sa[FM];
            lds_stage_load<FM, FN>(sa, u.r, u.sp, u.c, seg);
            load_b_st<FM, FN>(bx, u.r, u.sp, u.c, bcol, lg);
            lds_stage_store<FM, FN>(abuf, sa, u.r, u.valid ? u.sp.k : 0u, u.c, srw, seg);
        }
        u2 = u;
        st_advance_a<FM, FN, 2>(u2, spairs, arow, st, srow);
        __syncthreads();
        // iteration it multiplies chunk it (LDS buffer it & 1, B in bx or by)
        // and stages chunk it + 1 in the other buffer
        const int niter = (nchunk + 1) >> 1;
        for(int it = 0; it < niter; it += 2) {
            st_cursor<FM, FN> un = u2;
            st_advance_a<FM, FN, 2>(u2, spairs, arow, st, srow);
            lds_step_any<FM, FN>(acc, by, bx, un, abuf, abuf + ABUF, bcol, lr, lg, srw, seg);
            __syncthreads();
            if(it + 1 >= niter) break;
            un = u2;
            st_advance_a<FM, FN, 2>(u2, spairs, arow, st, srow);
            lds_step_any<FM, FN>(acc, bx, by, un, abuf + ABUF, abuf, bcol, lr, lg, srw, seg);
            __syncthreads();
        }
    }

    // group 1's partial tile onto group 0's, then the C rows
    double *red = lds;
    if(g == 1) {
        double *r = red + h * (FM * FN * 4 * 64);
#pragma unroll
        for(int i = 0; i < FM; i++)
#pragma unroll
            for(int j = 0; j < FN; j++)
#pragma unroll
                for(int v = 0; v < 4; v++)
                    r[((i * FN + j) * 4 + v) * 64 + lane] = acc[i][j][v];
    }
    __syncthreads();
    if(g == 0) {
        const double *r = red + h * (FM * FN * 4 * 64);
#pragma unroll
        for(int i = 0; i < FM; i++)
#pragma unroll
            for(int j = 0; j < FN; j++)
#pragma unroll
                for(int v = 0; v < 4; v++)
                    acc[i][j][v] += r[((i * FN + j) * 4 + v) * 64 + lane];
#pragma unroll
        for(int i = 0; i < FM; i++) {
#pragma unroll
            for(int v = 0; v < 4; v++) {
                const int row = 16 * i + lg + 4 * v;
                if(row >= mrem) continue;
                const stack_crow cr = crow[st.crow_begin + m0 + row];
#pragma unroll
                for(int j = 0; j < FN; j++) {
                    const int col = 16 * j + lr;
                    if(col < nrem) {
                        double *cp = cr.c + (nw + col);
                        double x = acc[i][j][v];
                        if(!cr.zero) x += *cp;
                        *cp = x;
                    }
                }
            }
        }
    }
    __syncthreads();   // LDS is reused by the next tile
}

template<int FM, int FN>
__global__ void __launch_bounds__(256)
gemm_stacked_lds_kernel(const stack_desc *__restrict__ stacks,
    const stack_pair *__restrict__ spairs, const stack_arow *__restrict__ arow,
    const stack_crow *__restrict__ crow, const gemm_tile *__restrict__ tiles,
    uint32_t ntiles) {
    constexpr int NA = 4 * 16 * FM * kLdsRowStride, NR = 2 * FM * FN * 4 * 64;
    __shared__ double lds[NA > NR ? NA : NR];
    for(uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const gemm_tile tile = tiles[t];
        const stack_desc st = stacks[tile.cblock];
        lds_tile<FM, FN>(st, spairs, arow, crow, int(tile.tm) * 16 * FM,
            int(tile.tn) * 32 * FN, lds);
    }
}

//! Stacked kernel over tiles of varying height and width (16 fm x 16 fn,
//! fm <= 4, fn <= FNX): every stack is covered by full tiles and one
//! narrower tile per remainder, so rows and columns are padded to 16 only.
template<int W, int FNX>
__global__ void __launch_bounds__(64 * W)
gemm_stacked_var_kernel(const stack_desc *__restrict__ stacks,
    const stack_pair *__restrict__ spairs, const stack_arow *__restrict__ arow,
    const stack_crow *__restrict__ crow, const var_tile *__restrict__ tiles,
    uint32_t ntiles) {
    __shared__ double red[(W > 1 ? W - 1 : 1) * 4 * FNX * 4 * 64];
    for(uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const var_tile vt = tiles[t];
        const stack_desc st = stacks[vt.stack];
        const int m0 = 16 * int(vt.r16), n0 = 16 * int(vt.c16);
        switch(vt.fm * 4 + vt.fn) {
        case 1 * 4 + 1: stacked_tile<1, 1, W, 2>(st, spairs, arow, crow, m0, n0, red); break;
        case 2 * 4 + 1: stacked_tile<2, 1, W, 2>(st, spairs, arow, crow, m0, n0, red); break;
        case 3 * 4 + 1: stacked_tile<3, 1, W, 2>(st, spairs, arow, crow, m0, n0, red); break;
        case 4 * 4 + 1: stacked_tile<4, 1, W, 2>(st, spairs, arow, crow, m0, n0, red); break;
        case 1 * 4 + 2: stacked_tile<1, 2, W, 2>(st, spairs, arow, crow, m0, n0, red); break;
        case 2 * 4 + 2: stacked_tile<2, 2, W, 2>(st, spairs, arow, crow, m0, n0, red); break;
        case 3 * 4 + 2: stacked_tile<3, 2, W, 2>(st, spairs, arow, crow, m0, n0, red); break;
        case 4 * 4 + 2: stacked_tile<4, 2, W, 2>(st, spairs, arow, crow, m0, n0, red); break;
        default:
            if(FNX >= 3) {
                switch(vt.fm) {
                case 1: stacked_tile<1, (FNX >= 3 ? 3 : 1), W, 2>(st, spairs, arow, crow, m0, n0, red); break;
                case 2: stacked_tile<2, (FNX >= 3 ? 3 : 1), W, 2>(st, spairs, arow, crow, m0, n0, red); break;
                case 3: stacked_tile<3, (FNX >= 3 ? 3 : 1), W, 2>(st, spairs, arow, crow, m0, n0, red); break;
                default: stacked_tile<4, (FNX >= 3 ? 3 : 1), W, 2>(st, spairs, arow, crow, m0, n0, red); break;
                }
            }
            break;
        }
    }
}

hipError_t launch_gemm_stacked_var(int cfg, const stack_desc *st, const stack_pair *sp,
    const stack_arow *ar, const stack_crow *cr, const var_tile *tiles, uint32_t ntiles,
    hipStream_t s) {
    if(ntiles == 0) return hipSuccess;
    if(cfg == STACK_VAR_64x32_W4)
        hipLaunchKernelGGL((gemm_stacked_var_kernel<4, 2>), dim3(ntiles), dim3(256), 0, s,
            st, sp, ar, cr, tiles, ntiles);
    else if(cfg == STACK_VAR_64x48_W4)
        hipLaunchKernelGGL((gemm_stacked_var_kernel<4, 3>), dim3(ntiles), dim3(256), 0, s,
            st, sp, ar, cr, tiles, ntiles);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

template<int FM, int FN, int W, int OCC>
static hipError_t launch_stacked_occ_t(const stack_desc *st, const stack_pair *sp,
    const stack_arow *ar, const stack_crow *cr, const gemm_tile *tiles, uint32_t ntiles,
    hipStream_t s) {
    hipLaunchKernelGGL((gemm_stacked_occ_kernel<FM, FN, W, OCC>), dim3(ntiles), dim3(64 * W), 0, s,
        st, sp, ar, cr, tiles, ntiles);
    return hipGetLastError();
}

template<int FM, int FN, int W, int PIPE = 2>
static hipError_t launch_stacked_t(const stack_desc *st, const stack_pair *sp,
    const stack_arow *ar, const stack_crow *cr, const gemm_tile *tiles, uint32_t ntiles,
    hipStream_t s) {
    hipLaunchKernelGGL((gemm_stacked_kernel<FM, FN, W, PIPE>), dim3(ntiles), dim3(64 * W), 0, s,
        st, sp, ar, cr, tiles, ntiles);
    return hipGetLastError();
}

hipError_t launch_gemm_stacked(int cfg, const stack_desc *st, const stack_pair *sp,
    const stack_arow *ar, const stack_crow *cr, const gemm_tile *tiles, uint32_t ntiles,
    hipStream_t s) {
    if(ntiles == 0) return hipSuccess;
    switch(cfg) {
    case STACK_16x64_W4: return launch_stacked_t<1, 4, 4>(st, sp, ar, cr, tiles, ntiles, s);
    case STACK_32x32_W4: return launch_stacked_t<2, 2, 4>(st, sp, ar, cr, tiles, ntiles, s);
    case STACK_32x64_W4: return launch_stacked_t<2, 4, 4>(st, sp, ar, cr, tiles, ntiles, s);
    case STACK_64x32_W4: return launch_stacked_t<4, 2, 4, 3>(st, sp, ar, cr, tiles, ntiles, s);
    case STACK_LDS_64x64:
        hipLaunchKernelGGL((gemm_stacked_lds_kernel<4, 2>), dim3(ntiles), dim3(256), 0, s,
            st, sp, ar, cr, tiles, ntiles);
        return hipGetLastError();
    case STACK_48x64_W4: return launch_stacked_t<3, 4, 4>(st, sp, ar, cr, tiles, ntiles, s);
    case STACK_64x48_W4: return launch_stacked_t<4, 3, 4, 3>(st, sp, ar, cr, tiles, ntiles, s);
    case STACK_48x32_W4: return launch_stacked_t<3, 2, 4>(st, sp, ar, cr, tiles, ntiles, s);
    case STACK_64x32_W4_P1: return launch_stacked_t<4, 2, 4, 1>(st, sp, ar, cr, tiles, ntiles, s);
    case STACK_64x48_W2: return launch_stacked_t<4, 3, 2>(st, sp, ar, cr, tiles, ntiles, s);
    case STACK_48x64_W2: return launch_stacked_t<3, 4, 2>(st, sp, ar, cr, tiles, ntiles, s);
    case STACK_48x32_W4_O2: return launch_stacked_occ_t<3, 2, 4, 2>(st, sp, ar, cr, tiles, ntiles, s);
    case STACK_32x32_W4_O3: return launch_stacked_occ_t<2, 2, 4, 3>(st, sp, ar, cr, tiles, ntiles, s);
    case STACK_16x64_W4_O3: return launch_stacked_occ_t<1, 4, 4, 3>(st, sp, ar, cr, tiles, ntiles, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_gemm(int cfg, const gemm_cblock *cb,
    const gemm_pair *pairs, const gemm_tile *tiles, uint32_t ntiles,
    uint32_t nwg, hipStream_t s) {
    if(ntiles == 0) return hipSuccess;
    if(nwg == 0 || nwg > ntiles) nwg = ntiles;
    cfg = unstacked(cfg);
    switch(cfg) {
    case TILE_32x64: return launch_gemm_t<32, 64>(cb, pairs, tiles, ntiles, nwg, s);
    case TILE_64x32: return launch_gemm_t<64, 32>(cb, pairs, tiles, ntiles, nwg, s);
    case TILE_32x32: return launch_gemm_t<32, 32>(cb, pairs, tiles, ntiles, nwg, s);
    case SPLITK_16x64_W4: return launch_direct_t<1, 4, 4>(cb, pairs, tiles, ntiles, nwg, s);
    case SPLITK_32x32_W4: return launch_direct_t<2, 2, 4>(cb, pairs, tiles, ntiles, nwg, s);
    case SPLITK_32x64_W4: return launch_direct_t<2, 4, 4>(cb, pairs, tiles, ntiles, nwg, s);
    case SPLITK_64x32_W4: return launch_direct_t<4, 2, 4>(cb, pairs, tiles, ntiles, nwg, s);
    case SPLITK_16x64_W2: return launch_direct_t<1, 4, 2>(cb, pairs, tiles, ntiles, nwg, s);
    case SPLITK_32x32_W2: return launch_direct_t<2, 2, 2>(cb, pairs, tiles, ntiles, nwg, s);
    case SPLITK_16x32_W4: return launch_direct_t<1, 2, 4>(cb, pairs, tiles, ntiles, nwg, s);
    case SPLITK_32x16_W4: return launch_direct_t<2, 1, 4>(cb, pairs, tiles, ntiles, nwg, s);
    default: return launch_gemm_t<64, 64>(cb, pairs, tiles, ntiles, nwg, s);
    }
}

} // namespace dev
} // namespace lta
