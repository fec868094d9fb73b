// Split-precision FP64 emulation (Ozaki scheme with integer slices) on
// CDNA4 INT8 MFMA: every operand row (A) / column (B) gets one power-of-two
// scale 2^e with max |x| < 2^e, the scaled values are cut into 7-bit signed
// digits (int8 slices, exact), slice products run on v_mfma_i32_16x16x32_i8
// with exact int32 accumulation, and the products of equal weight
// 2^(-7 (t + u)) are scaled and summed in FP64.  Products with
// t + u > slices + 1 are dropped (triangular scheme); the truncation error
// per element is below 2^(e_i + f_j) * K * (slices + 1) * 2^(-7 (slices + 1)).
// References: K. Ozaki, T. Ogita, S. Oishi, S. M. Rump, "Error-free
// transformations of matrix multiplication by using fast routines of
// matrix multiplication and its applications", Numer. Algorithms 59 (2012);
// H. Ootomo, K. Ozaki, R. Yokota, "DGEMM on integer matrix multiplication
// unit", IJHPCA (2024).
#include <hip/hip_runtime.h>
#include "lt_device.h"

namespace lta {
namespace dev {

typedef int int4_t __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256)
oz_split_kernel(const oz_split_desc *__restrict__ descs,
    const uint32_t *__restrict__ items, uint32_t nitems) {

    const uint32_t it = blockIdx.x * 4 + (threadIdx.x >> 6);
    if(it >= nitems) return;          // whole wave: no barrier in this kernel
    const int lane = threadIdx.x & 63;
    const oz_split_desc d = descs[items[2 * it]];
    const uint32_t r = items[2 * it + 1];
    const double *row = d.src + size_t(r) * d.rs;

    double mx = 0.0;
    for(uint32_t k = lane; k < d.k; k += 64) mx = fmax(mx, fabs(row[size_t(k) * d.ks]));
#pragma unroll
    for(int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
    int e = 0;
    if(mx > 0.0) (void)frexp(mx, &e);  // mx = f 2^e, 0.5 <= f < 1
    if(lane == 0) d.exps[r] = e;

    int8_t *out = d.dst + size_t(r) * d.kp;
    for(uint32_t k = lane; k < d.kp; k += 64) {
        double x = k < d.k ? ldexp(row[size_t(k) * d.ks], -e) : 0.0;   // |x| < 1
        for(uint32_t t = 0; t < d.slices; t++) {
            const double y = x * 128.0;
            const double q = trunc(y);           // |q| <= 127
            x = y - q;                           // exact
            out[size_t(t) * d.sstride + k] = (int8_t)(int)q;
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

// One wave per 16 x 16 FN C tile, four waves (tiles) per workgroup.
// v_mfma_i32_16x16x32_i8: lane l supplies A[l % 16][8 (l / 16) + j] and
// B[8 (l / 16) + j][l % 16], j = 0..7, as 8 packed bytes; it receives
// D[4 (l / 16) + v][l % 16], v = 0..3.  Per chunk of 32 k the S slices of A
// and of B are loaded once and every product t + u = L <= S + 1 runs from
// registers into the int32 accumulator of its weight L; at the end of a
// pair each weight is scaled by 2^(e_i + f_j - 7 L) and added in FP64 in
// the order L = 2, 3, ..., S + 1.
template<int S, int FN>
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
    const int m0 = int(tile.tm) * 16, n0 = int(tile.tn) * 16 * FN;
    const int mrem = int(cb.m) - m0, nrem = int(cb.n) - n0;

    const int arow = lr < mrem ? m0 + lr : 0;
    int bcol[FN];
#pragma unroll
    for(int f = 0; f < FN; f++) bcol[f] = (16 * f + lr) < nrem ? n0 + 16 * f + lr : 0;
    int orow[4];
#pragma unroll
    for(int v = 0; v < 4; v++) orow[v] = (4 * lg + v) < mrem ? m0 + 4 * lg + v : 0;

    double cacc[FN][4];
#pragma unroll
    for(int f = 0; f < FN; f++)
#pragma unroll
        for(int v = 0; v < 4; v++) cacc[f][v] = 0.0;

    for(uint32_t p = cb.pair_begin; p < cb.pair_end; p++) {
        const oz_pair pr = pairs[p];
        int4_t acc[S][FN];
#pragma unroll
        for(int l = 0; l < S; l++)
#pragma unroll
            for(int f = 0; f < FN; f++) acc[l][f] = int4_t{0, 0, 0, 0};
        const int8_t *pa = pr.a + size_t(arow) * pr.kp + 8 * lg;
        const int8_t *pb[FN];
#pragma unroll
        for(int f = 0; f < FN; f++) pb[f] = pr.b + size_t(bcol[f]) * pr.kp + 8 * lg;
        for(uint32_t kc = 0; kc < pr.kp; kc += 32) {
            long av[S], bv[S][FN];
#pragma unroll
            for(int ts = 0; ts < S; ts++) av[ts] = *(const long *)(pa + size_t(ts) * pr.astride + kc);
#pragma unroll
            for(int us = 0; us < S; us++)
#pragma unroll
                for(int f = 0; f < FN; f++)
                    bv[us][f] = *(const long *)(pb[f] + size_t(us) * pr.bstride + kc);
#pragma unroll
            for(int ts = 0; ts < S; ts++)
#pragma unroll
                for(int us = 0; us + ts < S; us++)       // t + u <= S + 1 (1-based)
#pragma unroll
                    for(int f = 0; f < FN; f++)
                        acc[ts + us][f] = __builtin_amdgcn_mfma_i32_16x16x32_i8(
                            av[ts], bv[us][f], acc[ts + us][f], 0, 0, 0);
        }
        int ea[4], eb[FN];
#pragma unroll
        for(int v = 0; v < 4; v++) ea[v] = pr.ea[orow[v]];
#pragma unroll
        for(int f = 0; f < FN; f++) eb[f] = pr.eb[bcol[f]];
#pragma unroll
        for(int l = 0; l < S; l++)               // weight L = l + 2
#pragma unroll
            for(int f = 0; f < FN; f++)
#pragma unroll
                for(int v = 0; v < 4; v++)
                    cacc[f][v] += pr.d * ldexp(double(acc[l][f][v]), ea[v] + eb[f] - 7 * (l + 2));
    }

#pragma unroll
    for(int f = 0; f < FN; f++) {
        const int col = 16 * f + lr;
#pragma unroll
        for(int v = 0; v < 4; v++) {
            const int row = 4 * lg + v;
            if(row < mrem && col < nrem) {
                double *cp = cb.c + size_t(m0 + row) * cb.ldc + (n0 + col);
                double x = cacc[f][v];
                if(!cb.zero) x += *cp;
                *cp = x;
            }
        }
    }
}

template<int S>
static void launch_oz_s(const gemm_cblock *cb, const oz_pair *pairs,
    const gemm_tile *tiles, uint32_t ntiles, hipStream_t s) {
    hipLaunchKernelGGL((oz_gemm_kernel<S, kOzTileN / 16>), dim3((ntiles + 3) / 4), dim3(256), 0, s,
        cb, pairs, tiles, ntiles);
}

hipError_t launch_oz_gemm(const gemm_cblock *cb, const oz_pair *pairs,
    const gemm_tile *tiles, uint32_t ntiles, int slices, hipStream_t s) {
    if(ntiles == 0) return hipSuccess;
    static_assert(kOzTileM == 16 && kOzTileN % 16 == 0, "tile shape");
    switch(slices) {
    case 1: launch_oz_s<1>(cb, pairs, tiles, ntiles, s); break;
    case 2: launch_oz_s<2>(cb, pairs, tiles, ntiles, s); break;
    case 3: launch_oz_s<3>(cb, pairs, tiles, ntiles, s); break;
    case 4: launch_oz_s<4>(cb, pairs, tiles, ntiles, s); break;
    case 5: launch_oz_s<5>(cb, pairs, tiles, ntiles, s); break;
    case 6: launch_oz_s<6>(cb, pairs, tiles, ntiles, s); break;
    case 7: launch_oz_s<7>(cb, pairs, tiles, ntiles, s); break;
    case 8: launch_oz_s<8>(cb, pairs, tiles, ntiles, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

} // namespace dev
} // namespace lta
