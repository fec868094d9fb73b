// CDNA4 (gfx950) kernels of the block-contraction path:
//   * batched N-d permute / scale / accumulate (tod_copy, tod_add, symmetry
//     unfolding, C write-back),
//   * grouped FP64 GEMM on v_mfma_f64_16x16x4f64 with LDS-staged operand
//     tiles; every workgroup owns one C tile and runs the whole contraction
//     list of that C block, so each C element is written once and the sum
//     order is fixed (no atomics),
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
    case DIRECT_16x64: bm = 16; bn = 64; break;
    case DIRECT_32x32: bm = 32; bn = 32; break;
    case DIRECT_32x64: bm = 32; bn = 64; break;
    case DIRECT_16x128: bm = 16; bn = 128; break;
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

    const uint32_t ch = blockIdx.x;
    const perm_desc d = descs[chunk_desc[ch]];
    const uint32_t base = chunk_off[ch];
    const uint32_t end = min(base + kPermChunk, d.total);
    for(uint32_t e = base + threadIdx.x; e < end; e += blockDim.x) {
        double v = d.alpha * d.src[src_offset(d, e)];
        if(d.accumulate) v += d.dst[e];
        d.dst[e] = v;
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

template<int BM, int BN, bool TA, bool TB>
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
static hipError_t launch_gemm_t(bool ta, bool tb, const gemm_cblock *cb,
    const gemm_pair *pairs, const gemm_tile *tiles, uint32_t ntiles,
    uint32_t nwg, hipStream_t s) {
    dim3 g(nwg), b(256);
    if(!ta && !tb) hipLaunchKernelGGL((gemm_kernel<BM, BN, false, false>), g, b, 0, s, cb, pairs, tiles, ntiles);
    else if(!ta && tb) hipLaunchKernelGGL((gemm_kernel<BM, BN, false, true>), g, b, 0, s, cb, pairs, tiles, ntiles);
    else if(ta && !tb) hipLaunchKernelGGL((gemm_kernel<BM, BN, true, false>), g, b, 0, s, cb, pairs, tiles, ntiles);
    else hipLaunchKernelGGL((gemm_kernel<BM, BN, true, true>), g, b, 0, s, cb, pairs, tiles, ntiles);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Direct FP64 MFMA GEMM: one wave per (16 FM) x (16 FN) tile, four waves per
// workgroup working on four tiles.  Operands go from memory straight into
// MFMA fragments.  Per chunk of 16 k, lane l (r = l % 16, g = l / 16) loads
// A[m0 + r][k0 + 4g .. 4g + 3] and B[k0 + 4g .. 4g + 3][n0 + r]; MFMA step j
// then feeds element j, so fragment slot g of step j carries k = k0 + 4g + j
// for both operands and the 16 k of the chunk are each summed once.

template<bool T>
__device__ __forceinline__ void load_frag(double (&x)[4], const double *base,
    bool rowok, int kk, int krem_full, int ld) {
    // !T: element (row, k) at base[k] (row-major, k contiguous);
    //  T: element (row, k) at base[k * ld] (k-major).
    if(krem_full >= 4) {
#pragma unroll
        for(int j = 0; j < 4; j++)
            x[j] = rowok ? (T ? base[size_t(kk + j) * ld] : base[kk + j]) : 0.0;
    } else {
#pragma unroll
        for(int j = 0; j < 4; j++)
            x[j] = (rowok && j < krem_full) ? (T ? base[size_t(kk + j) * ld] : base[kk + j]) : 0.0;
    }
}

template<int FM, int FN, bool TA, bool TB>
__global__ void __launch_bounds__(256)
gemm_direct_kernel(const gemm_cblock *__restrict__ cbs,
    const gemm_pair *__restrict__ pairs,
    const gemm_tile *__restrict__ tiles, uint32_t ntiles) {

    constexpr int TM = 16 * FM, TN = 16 * FN;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int lr = lane & 15, lg = lane >> 4;

    for(uint32_t t = blockIdx.x * 4 + wave; t < ntiles; t += gridDim.x * 4) {
        const gemm_tile tile = tiles[t];
        const gemm_cblock cb = cbs[tile.cblock];
        const int m0 = int(tile.tm) * TM, n0 = int(tile.tn) * TN;
        const int mrem = int(cb.m) - m0, nrem = int(cb.n) - n0;

        double4_t acc[FM][FN];
#pragma unroll
        for(int i = 0; i < FM; i++)
#pragma unroll
            for(int j = 0; j < FN; j++) acc[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};

        bool aok[FM], bok[FN];
#pragma unroll
        for(int i = 0; i < FM; i++) aok[i] = (16 * i + lr) < mrem;
#pragma unroll
        for(int j = 0; j < FN; j++) bok[j] = (16 * j + lr) < nrem;

        for(uint32_t p = cb.pair_begin; p < cb.pair_end; p++) {
            const gemm_pair pr = pairs[p];
            const int K = int(pr.k);
            const double d = pr.d;
            const double *ab[FM];
            const double *bb[FN];
#pragma unroll
            for(int i = 0; i < FM; i++) {
                const int m = aok[i] ? (m0 + 16 * i + lr) : 0;
                ab[i] = TA ? (pr.a + m) : (pr.a + size_t(m) * pr.lda);
            }
#pragma unroll
            for(int j = 0; j < FN; j++) {
                const int n = bok[j] ? (n0 + 16 * j + lr) : 0;
                bb[j] = TB ? (pr.b + size_t(n) * pr.ldb) : (pr.b + n);
            }
            for(int k0 = 0; k0 < K; k0 += 16) {
                const int kk = k0 + 4 * lg;
                const int kr = K - kk;
                double af[FM][4], bf[FN][4];
#pragma unroll
                for(int i = 0; i < FM; i++) load_frag<TA>(af[i], ab[i], aok[i], kk, kr, pr.lda);
#pragma unroll
                for(int j = 0; j < FN; j++) load_frag<!TB>(bf[j], bb[j], bok[j], kk, kr, pr.ldb);
#pragma unroll
                for(int i = 0; i < FM; i++)
#pragma unroll
                    for(int q = 0; q < 4; q++) af[i][q] *= d;
#pragma unroll
                for(int q = 0; q < 4; q++)
#pragma unroll
                    for(int i = 0; i < FM; i++)
#pragma unroll
                        for(int j = 0; j < FN; j++)
                            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                af[i][q], bf[j][q], acc[i][j], 0, 0, 0);
            }
        }

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
}

template<int FM, int FN>
static hipError_t launch_direct_t(bool ta, bool tb, const gemm_cblock *cb,
    const gemm_pair *pairs, const gemm_tile *tiles, uint32_t ntiles,
    uint32_t nwg, hipStream_t s) {
    dim3 g((nwg + 3) / 4), b(256);
    if(!ta && !tb) hipLaunchKernelGGL((gemm_direct_kernel<FM, FN, false, false>), g, b, 0, s, cb, pairs, tiles, ntiles);
    else if(!ta && tb) hipLaunchKernelGGL((gemm_direct_kernel<FM, FN, false, true>), g, b, 0, s, cb, pairs, tiles, ntiles);
    else if(ta && !tb) hipLaunchKernelGGL((gemm_direct_kernel<FM, FN, true, false>), g, b, 0, s, cb, pairs, tiles, ntiles);
    else hipLaunchKernelGGL((gemm_direct_kernel<FM, FN, true, true>), g, b, 0, s, cb, pairs, tiles, ntiles);
    return hipGetLastError();
}

hipError_t launch_gemm(int cfg, bool ta, bool tb, const gemm_cblock *cb,
    const gemm_pair *pairs, const gemm_tile *tiles, uint32_t ntiles,
    uint32_t nwg, hipStream_t s) {
    if(ntiles == 0) return hipSuccess;
    if(nwg == 0 || nwg > ntiles) nwg = ntiles;
    switch(cfg) {
    case TILE_32x64: return launch_gemm_t<32, 64>(ta, tb, cb, pairs, tiles, ntiles, nwg, s);
    case TILE_64x32: return launch_gemm_t<64, 32>(ta, tb, cb, pairs, tiles, ntiles, nwg, s);
    case TILE_32x32: return launch_gemm_t<32, 32>(ta, tb, cb, pairs, tiles, ntiles, nwg, s);
    case DIRECT_16x64: return launch_direct_t<1, 4>(ta, tb, cb, pairs, tiles, ntiles, nwg, s);
    case DIRECT_32x32: return launch_direct_t<2, 2>(ta, tb, cb, pairs, tiles, ntiles, nwg, s);
    case DIRECT_32x64: return launch_direct_t<2, 4>(ta, tb, cb, pairs, tiles, ntiles, nwg, s);
    case DIRECT_16x128: return launch_direct_t<1, 8>(ta, tb, cb, pairs, tiles, ntiles, nwg, s);
    default: return launch_gemm_t<64, 64>(ta, tb, cb, pairs, tiles, ntiles, nwg, s);
    }
}

} // namespace dev
} // namespace lta
