// Split-precision FP64 emulation (Ozaki scheme with integer slices) on
// CDNA4 INT8 MFMA: every operand row (A) / column (B) gets one power-of-two
// scale 2^e with max |x| < 2^e, the scaled values are cut into 7-bit signed
// digits (int8 slices, exact), slice products run on v_mfma_i32_16x16x64_i8
// with exact int32 accumulation, and the products of equal weight
// 2^(-7 (t + u)) are scaled and summed in FP64.  Products with
// t + u > slices + 1 are dropped (triangular scheme).  Error per element of
// one product of extent K (s = slices): each operand keeps a remainder
// below 2^(e - 7 s), and a dropped slice product t + u >= s + 2 is below
// 127^2 2^(e + f - 7 (t + u)); summed, |error| < K (s + 2) 2^(e_i + f_j - 7 s)
// (oz_bound_kernel states it per C block).  A row or column holding a NaN
// or an infinity gets the exponent kNonFinite and its outputs are NaN.
// References: K. Ozaki, T. Ogita, S. Oishi, S. M. Rump, "Error-free
// transformations of matrix multiplication by using fast routines of
// matrix multiplication and its applications", Numer. Algorithms 59 (2012);
// H. Ootomo, K. Ozaki, R. Yokota, "DGEMM on integer matrix multiplication
// unit", IJHPCA (2024).
#include <hip/hip_runtime.h>
#include <algorithm>
#include "lt_device.h"

namespace lta {
namespace dev {

typedef int int4_t __attribute__((ext_vector_type(4)));

static const uint32_t kOzBoundSlices = 8;

void oz_tile_shape(int tile, uint32_t &bm, uint32_t &bn) {
    switch(tile) {
    case OZ_TILE_16x32: bm = 16; bn = 32; break;
    case OZ_TILE_16x64: bm = 16; bn = 64; break;
    default: bm = 32; bn = 32; break;
    }
}

// One wave per row of a 16-row panel (rows past the operand's last row are
// written as zeros).  Lane l cuts element k = 64 c + l of the row (loads
// coalesced over k) and stores its slice bytes in the panel's
// fragment-major layout (see oz_split_desc): within chunk c the byte sits at
// (r % 16 + 16 (l / 16)) * 16 + l % 16, so a wave's store is four runs of
// 16 bytes.
__global__ void __launch_bounds__(256)
oz_split_kernel(const oz_split_desc *__restrict__ descs,
    const uint32_t *__restrict__ items, uint32_t nitems) {

    const uint32_t it = blockIdx.x * 4 + (threadIdx.x >> 6);
    if(it >= nitems) return;          // whole wave: no barrier in this kernel
    const int lane = threadIdx.x & 63;
    const oz_split_desc d = descs[items[2 * it]];
    const uint32_t r = items[2 * it + 1];
    const bool live = r < d.rows;
    const double *row = d.src + size_t(live ? r : 0) * d.rs;

    int e = 0;
    if(live) {
        double mx = 0.0;
        int nonfinite = 0;
        for(uint32_t k = lane; k < d.k; k += 64) {
            const double v = fabs(row[size_t(k) * d.ks]);
            nonfinite |= !(v <= 1.7976931348623157e308);    // NaN or infinity
            mx = fmax(mx, v);
        }
#pragma unroll
        for(int off = 32; off > 0; off >>= 1) {
            mx = fmax(mx, __shfl_xor(mx, off));
            nonfinite |= __shfl_xor(nonfinite, off);
        }
        if(nonfinite) e = kOzNonFinite;            // slices all zero, outputs NaN
        else if(mx > 0.0) (void)frexp(mx, &e);     // mx = f 2^e, 0.5 <= f < 1
        if(lane == 0) d.exps[r] = e;
    }

    const uint32_t nchunks = d.kp / 64, panel = r / 16;
    int8_t *pbase = d.dst + size_t(panel) * nchunks * 1024 + size_t(r % 16) * 16
        + size_t(lane / 16) * 256 + size_t(lane % 16);
    for(uint32_t c = 0; c < nchunks; c++) {
        const uint32_t k = 64 * c + lane;
        double x = (live && k < d.k) ? ldexp(row[size_t(k) * d.ks], -e) : 0.0;   // |x| < 1
        int8_t *out = pbase + size_t(c) * 1024;
        for(uint32_t t = 0; t < d.slices; t++) {
            const double y = x * 128.0;
            const double q = trunc(y);           // |q| <= 127
            x = y - q;                           // exact
            out[size_t(t) * d.sstride] = (int8_t)(int)q;
        }
    }
}

hipError_t launch_oz_split(const oz_split_desc *descs, const uint32_t *items,
    uint32_t nitems, hipStream_t s) {
    if(nitems == 0) return hipSuccess;
    hipLaunchKernelGGL(oz_split_kernel, dim3((nitems + 3) / 4), dim3(256), 0, s,
        descs, items, nitems);
    return hipGetLastError();
}

// One wave per 16 FM x 16 FN C tile, four waves (tiles) per workgroup.
// v_mfma_i32_16x16x64_i8 takes 16 bytes of A and of B per lane and returns
// D[4 (l / 16) + v][l % 16], v = 0..3.  Lane l supplies row (A) / column (B)
// l % 16 at k = 16 (l / 16) + j of the 64-k chunk; A and B use the same
// assignment, and the integer sum over k does not depend on the order in
// which the instruction adds the k terms, so the products are exact whatever
// its internal k order.  A fragment is one 16-byte load of a 1 KB
// fragment-major panel chunk (a wave's load covers 8 whole cache lines).  Per
// chunk the S slices of A are loaded once, each slice of B is loaded and
// used at once: every product t + u = L <= S + 1 runs from registers into
// the int32 accumulator of its weight L; at the end of a pair each weight is
// scaled by 2^(e_i + f_j - 7 L) and added in FP64 in the order
// L = 2, 3, ..., S + 1.
template<int S, int FM, int FN>
__global__ void __launch_bounds__(256)
oz_gemm_kernel(const gemm_cblock *__restrict__ cbs,
    const oz_pair *__restrict__ pairs,
    const gemm_tile *__restrict__ tiles, uint32_t ntiles) {

    const uint32_t t = blockIdx.x * 4 + (threadIdx.x >> 6);
    if(t >= ntiles) return;           // whole wave: no barrier in this kernel
    const int lane = threadIdx.x & 63;
    const int lr = lane & 15, lg = lane >> 4;
    const gemm_tile tile = tiles[t];
    const gemm_cblock cb = cbs[tile.cblock];
    const int m0 = int(tile.tm) * 16 * FM, n0 = int(tile.tn) * 16 * FN;
    const int mrem = int(cb.m) - m0, nrem = int(cb.n) - n0;
    // panels past the last row / column of C are replaced by the last panel
    // (read in bounds, those rows and columns are not written)
    const int mpan = (int(cb.m) + 15) / 16, npan = (int(cb.n) + 15) / 16;
    int apan[FM], bpan[FN];
#pragma unroll
    for(int i = 0; i < FM; i++) apan[i] = min(m0 / 16 + i, mpan - 1);
#pragma unroll
    for(int f = 0; f < FN; f++) bpan[f] = min(n0 / 16 + f, npan - 1);
    int bcol[FN];
#pragma unroll
    for(int f = 0; f < FN; f++) bcol[f] = (16 * f + lr) < nrem ? n0 + 16 * f + lr : 0;
    int orow[FM][4];
#pragma unroll
    for(int i = 0; i < FM; i++)
#pragma unroll
        for(int v = 0; v < 4; v++)
            orow[i][v] = (16 * i + 4 * lg + v) < mrem ? m0 + 16 * i + 4 * lg + v : 0;

    double cacc[FM][FN][4];
#pragma unroll
    for(int i = 0; i < FM; i++)
#pragma unroll
        for(int f = 0; f < FN; f++)
#pragma unroll
            for(int v = 0; v < 4; v++) cacc[i][f][v] = 0.0;

    for(uint32_t p = cb.pair_begin; p < cb.pair_end; p++) {
        const oz_pair pr = pairs[p];
        int4_t acc[S][FM][FN];
#pragma unroll
        for(int l = 0; l < S; l++)
#pragma unroll
            for(int i = 0; i < FM; i++)
#pragma unroll
                for(int f = 0; f < FN; f++) acc[l][i][f] = int4_t{0, 0, 0, 0};
        const uint32_t nch = pr.kp / 64;
        const int8_t *pa[FM], *pb[FN];
#pragma unroll
        for(int i = 0; i < FM; i++) pa[i] = pr.a + (size_t(apan[i]) * nch) * 1024 + lane * 16;
#pragma unroll
        for(int f = 0; f < FN; f++) pb[f] = pr.b + (size_t(bpan[f]) * nch) * 1024 + lane * 16;
        for(uint32_t kc = 0; kc < nch; kc++) {
            int4_t av[S][FM];
#pragma unroll
            for(int ts = 0; ts < S; ts++)
#pragma unroll
                for(int i = 0; i < FM; i++)
                    av[ts][i] = *(const int4_t *)(pa[i] + size_t(ts) * pr.astride + size_t(kc) * 1024);
#pragma unroll
            for(int us = 0; us < S; us++) {
                int4_t bv[FN];
#pragma unroll
                for(int f = 0; f < FN; f++)
                    bv[f] = *(const int4_t *)(pb[f] + size_t(us) * pr.bstride + size_t(kc) * 1024);
#pragma unroll
                for(int ts = 0; ts + us < S; ts++)       // t + u <= S + 1 (1-based)
#pragma unroll
                    for(int i = 0; i < FM; i++)
#pragma unroll
                        for(int f = 0; f < FN; f++)
                            acc[ts + us][i][f] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                                av[ts][i], bv[f], acc[ts + us][i][f], 0, 0, 0);
            }
        }
        int eb[FN];
#pragma unroll
        for(int f = 0; f < FN; f++) eb[f] = pr.eb[bcol[f]];
#pragma unroll
        for(int i = 0; i < FM; i++) {
            int ea[4];
#pragma unroll
            for(int v = 0; v < 4; v++) ea[v] = pr.ea[orow[i][v]];
#pragma unroll
            for(int l = 0; l < S; l++)               // weight L = l + 2
#pragma unroll
                for(int f = 0; f < FN; f++)
#pragma unroll
                    for(int v = 0; v < 4; v++) {
                        const bool nf = ea[v] == kOzNonFinite || eb[f] == kOzNonFinite;
                        cacc[i][f][v] += nf ? __builtin_nan("") : pr.d * ldexp(
                            double(acc[l][i][f][v]), ea[v] + eb[f] - 7 * (l + 2));
                    }
        }
    }

#pragma unroll
    for(int i = 0; i < FM; i++)
#pragma unroll
        for(int f = 0; f < FN; f++) {
            const int col = 16 * f + lr;
#pragma unroll
            for(int v = 0; v < 4; v++) {
                const int row = 16 * i + 4 * lg + v;
                if(row < mrem && col < nrem) {
                    double *cp = cb.c + size_t(m0 + row) * cb.ldc + (n0 + col);
                    double x = cacc[i][f][v];
                    if(!cb.zero) x += *cp;
                    *cp = x;
                }
            }
        }
}

// Largest exponent of every split operand: one wave per operand, its rows'
// exponents (as the split kernel wrote them) reduced with max (a row with a
// NaN or an infinity holds kOzNonFinite, larger than any finite exponent).
__global__ void __launch_bounds__(256)
oz_opmax_kernel(const oz_split_desc *__restrict__ descs, uint32_t nops,
    int32_t *__restrict__ opmax, uint32_t ncb, double *__restrict__ bound_out) {
    // the max |C| slots of the bound kernel (which runs next) start at zero
    for(uint32_t b = blockIdx.x * 256 + threadIdx.x; b < ncb; b += gridDim.x * 256)
        bound_out[2 * b + 1] = 0.0;
    const uint32_t i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if(i >= nops) return;             // whole wave: no barrier in this kernel
    const int lane = threadIdx.x & 63;
    const oz_split_desc d = descs[i];
    int m = -2147483647;
    for(uint32_t r = lane; r < d.rows; r += 64) m = max(m, d.exps[r]);
#pragma unroll
    for(int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
    if(lane == 0) opmax[i] = m;
}

// Stated error of every C block of a split-precision product: workgroups
// (C block b, row slice y of gridDim.y); slice 0 states the bound = sum over
// the block's pairs (in list order) of |d| K (s + 2) 2^(max_i e_i + max_j
// f_j - 7 s) (the per-element bound above with the largest row and column
// exponents of the pair's operands), and every slice folds the largest |C|
// of its rows (as the product kernel wrote them) into the block's slot with
// an integer atomic max: |C| >= 0, and non-negative doubles order as their
// bit patterns, with infinity and then NaN above every finite value, so
// NaN propagates and the result does not depend on the order.
// out[2 b] = bound, out[2 b + 1] = max |C| (zeroed by oz_opmax_kernel).
__global__ void __launch_bounds__(256)
oz_bound_kernel(const gemm_cblock *__restrict__ cbs, const gemm_pair *__restrict__ pairs,
    const oz_pair *__restrict__ ozp, const int32_t *__restrict__ opmax, int slices,
    double *__restrict__ out) {

    __shared__ double rc[256];
    __shared__ double rt[256];
    const gemm_cblock cb = cbs[blockIdx.x];
    const int tid = threadIdx.x;
    const uint32_t ns = gridDim.y, sl = blockIdx.y;
    // the terms of up to 256 pairs at a time, one per thread, then added by
    // thread 0 in list order (slice 0 only: uniform over the workgroup)
    double bound = 0.0;
    for(uint32_t p0 = cb.pair_begin; sl == 0 && p0 < cb.pair_end; p0 += 256) {
        const uint32_t p = p0 + tid;
        double term = 0.0;
        if(p < cb.pair_end) {
            const oz_pair o = ozp[p];
            const int ma = opmax[o.ia], mb = opmax[o.ib];
            if(ma > -1000000 && mb > -1000000) {             // rows and columns present
                const bool nf = ma == kOzNonFinite || mb == kOzNonFinite;
                term = nf ? __builtin_inf() : fabs(pairs[p].d) * double(pairs[p].k) *
                    double(slices + 2) * ldexp(1.0, ma + mb - 7 * slices);
            }
        }
        rt[tid] = term;
        __syncthreads();
        if(tid == 0) {
            const uint32_t n = min(256u, cb.pair_end - p0);
            for(uint32_t q = 0; q < n; q++) bound += rt[q];
        }
        __syncthreads();
    }
    // max |C|: wave w of slice y rows 4 y + w, 4 y + w + 4 ns, ...; lanes
    // columns, four loads in flight
    double cm = 0.0;
    for(uint32_t r = 4 * sl + (tid >> 6); r < cb.m; r += 4 * ns) {
        const double *row = cb.c + size_t(r) * cb.ldc;
        for(uint32_t x = tid & 63; x < cb.n; x += 256) {
            double v[4];
#pragma unroll
            for(int u = 0; u < 4; u++) v[u] = x + 64 * u < cb.n ? fabs(row[x + 64 * u]) : 0.0;
#pragma unroll
            for(int u = 0; u < 4; u++) cm = (v[u] > cm || v[u] != v[u]) ? v[u] : cm;
        }
    }
    rc[tid] = cm;
    __syncthreads();
    for(int w = 128; w > 0; w >>= 1) {
        if(tid < w) {
            const double v = rc[tid + w];
            rc[tid] = (v > rc[tid] || v != v) ? v : rc[tid];
        }
        __syncthreads();
    }
    if(tid == 0) {
        if(sl == 0) out[2 * blockIdx.x] = bound;
        atomicMax((unsigned long long *)&out[2 * blockIdx.x + 1],
            (unsigned long long)__double_as_longlong(rc[0]));
    }
}

hipError_t launch_oz_opmax(const oz_split_desc *descs, uint32_t nops, int32_t *opmax,
    uint32_t ncb, double *bound_out, hipStream_t s) {
    if(nops == 0 && ncb == 0) return hipSuccess;
    const uint32_t g = std::max((nops + 3) / 4, (ncb + 255) / 256);
    hipLaunchKernelGGL(oz_opmax_kernel, dim3(g), dim3(256), 0, s, descs, nops, opmax, ncb,
        bound_out);
    return hipGetLastError();
}

hipError_t launch_oz_bound(const gemm_cblock *cb, const gemm_pair *pairs, const oz_pair *ozp,
    const int32_t *opmax, uint32_t ncb, int slices, double *out, hipStream_t s) {
    if(ncb == 0) return hipSuccess;
    // rows of a C block in up to kOzBoundSlices workgroups
    hipLaunchKernelGGL(oz_bound_kernel, dim3(ncb, kOzBoundSlices), dim3(256), 0, s, cb, pairs,
        ozp, opmax,
        slices, out);
    return hipGetLastError();
}

template<int S>
static void launch_oz_s(const gemm_cblock *cb, const oz_pair *pairs,
    const gemm_tile *tiles, uint32_t ntiles, int tile, hipStream_t s) {
    const dim3 grid((ntiles + 3) / 4), block(256);
    switch(tile) {
    case OZ_TILE_16x32:
        hipLaunchKernelGGL((oz_gemm_kernel<S, 1, 2>), grid, block, 0, s, cb, pairs, tiles, ntiles);
        break;
    case OZ_TILE_16x64:
        hipLaunchKernelGGL((oz_gemm_kernel<S, 1, 4>), grid, block, 0, s, cb, pairs, tiles, ntiles);
        break;
    default:
        hipLaunchKernelGGL((oz_gemm_kernel<S, 2, 2>), grid, block, 0, s, cb, pairs, tiles, ntiles);
        break;
    }
}

hipError_t launch_oz_gemm(const gemm_cblock *cb, const oz_pair *pairs,
    const gemm_tile *tiles, uint32_t ntiles, int slices, int tile, hipStream_t s) {
    if(ntiles == 0) return hipSuccess;
    if(tile < 0 || tile >= OZ_TILE_COUNT) return hipErrorInvalidValue;
    switch(slices) {
    case 1: launch_oz_s<1>(cb, pairs, tiles, ntiles, tile, s); break;
    case 2: launch_oz_s<2>(cb, pairs, tiles, ntiles, tile, s); break;
    case 3: launch_oz_s<3>(cb, pairs, tiles, ntiles, tile, s); break;
    case 4: launch_oz_s<4>(cb, pairs, tiles, ntiles, tile, s); break;
    case 5: launch_oz_s<5>(cb, pairs, tiles, ntiles, tile, s); break;
    case 6: launch_oz_s<6>(cb, pairs, tiles, ntiles, tile, s); break;
    case 7: launch_oz_s<7>(cb, pairs, tiles, ntiles, tile, s); break;
    case 8: launch_oz_s<8>(cb, pairs, tiles, ntiles, tile, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

} // namespace dev
} // namespace lta
