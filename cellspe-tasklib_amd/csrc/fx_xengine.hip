// fx_xengine.hip -- X-engine: per-channel complex Hermitian GEMM on MFMA.
//
// V_ij(c) += sum_t X_i(c,t) conj(X_j(c,t)), i <= j, the baseline MAC of
// genericAddProduct_32fc_conj (architecture.h:427-434) that the minicorrelator
// runs per baseline (three_baseline_naive_MACauto, minicorrelator_spu.c:745-786,
// with its missing conjugation fixed).
//
// Complex arithmetic as two real GEMMs sharing the left operand: with
// u_i = (re_i(t0), im_i(t0), re_i(t1), ...) along K and the rotated copy
// v_j = (-im_j(t0), re_j(t0), ...), Re V = U U^T and Im V = U V^T.  Both
// operands are read straight from the channelized layout X[c][t][i] (input
// fastest), so each MFMA fragment load is a 128-byte coalesced row segment.
//
// A workgroup of 4 waves owns one 64x64 input tile (bi64 <= bj64) of one
// channel; wave w computes the 32x32 block (2*bi64 + w/2, 2*bj64 + w%2) and
// skips the strictly-lower block of diagonal tiles.  Accumulators live in
// registers for the chunk and are added into an engine-layout buffer
// acc[c][tile][wave][reg][lane] (fully coalesced), which the dump kernel
// converts to the canonical [p][c] complex64 visibilities.
#include "fx_internal.h"

#include <hip/hip_fp8.h>

typedef _Float16 fx_half8 __attribute__((ext_vector_type(8)));
typedef float fx_float16 __attribute__((ext_vector_type(16)));
typedef int fx_int8 __attribute__((ext_vector_type(8)));

namespace {

// number of 64x64 upper-triangle tiles for ninp inputs
__host__ __device__ inline int ntiles64(int ninp)
{
    int nb = (ninp + 63) / 64;
    return nb * (nb + 1) / 2;
}

// tile index -> (bi, bj), row-major over the upper triangle
__device__ inline void tile_coords(int tile, int nb, int& bi, int& bj)
{
    int r = 0, rem = tile;
    while (rem >= nb - r) {
        rem -= nb - r;
        r++;
    }
    bi = r;
    bj = r + rem;
}

__device__ inline void acc_store(float* dst, int lane, const fx_float16& re, const fx_float16& im)
{
#pragma unroll
    for (int r = 0; r < 16; r++) {
        dst[r * 64 + lane] += re[r];
        dst[(16 + r) * 64 + lane] += im[r];
    }
}

// ---------------------------------------------------------------------------
// FP32: v_mfma_f32_32x32x2_f32, K = 2 = one complex sample.
//   lanes 0-31 carry k=0 (real parts), lanes 32-63 k=1 (imaginary parts).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_xengine_f32(const float2* __restrict__ X, int nchan,
                                                     long nt, int ninp, float* __restrict__ acc)
{
    const int c = blockIdx.y;
    const int tile = blockIdx.x;
    const int nb = (ninp + 63) / 64;
    const int ntl = nb * (nb + 1) / 2;
    int bi64, bj64;
    tile_coords(tile, nb, bi64, bj64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int bi = 2 * bi64 + (w >> 1), bj = 2 * bj64 + (w & 1);
    if (bi > bj || bi * 32 >= ninp || bj * 32 >= ninp) return;
    const int ra = bi * 32 + (lane & 31), rb = bj * 32 + (lane & 31);
    const bool va = ra < ninp, vb = rb < ninp;
    const bool hi = lane >= 32;
    const float2* base = X + (size_t)c * (size_t)nt * (size_t)ninp;
    fx_float16 dre = {}, dim = {};
    const float2 z = make_float2(0.f, 0.f);
    long t = 0;
    for (; t + 4 <= nt; t += 4) {
        float2 xa[4], xb[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            xa[q] = va ? base[(t + q) * ninp + ra] : z;
            xb[q] = vb ? base[(t + q) * ninp + rb] : z;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            float a = hi ? xa[q].y : xa[q].x;
            float b = hi ? xb[q].y : xb[q].x;
            float br = hi ? xb[q].x : -xb[q].y;
            dre = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, dre, 0, 0, 0);
            dim = __builtin_amdgcn_mfma_f32_32x32x2f32(a, br, dim, 0, 0, 0);
        }
    }
    for (; t < nt; t++) {
        float2 xa = va ? base[t * ninp + ra] : z;
        float2 xb = vb ? base[t * ninp + rb] : z;
        float a = hi ? xa.y : xa.x;
        float b = hi ? xb.y : xb.x;
        float br = hi ? xb.x : -xb.y;
        dre = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, dre, 0, 0, 0);
        dim = __builtin_amdgcn_mfma_f32_32x32x2f32(a, br, dim, 0, 0, 0);
    }
    float* dst = acc + (((size_t)c * ntl + tile) * 4 + w) * (32 * 64);
    acc_store(dst, lane, dre, dim);
}

// ---------------------------------------------------------------------------
// FP16: v_mfma_f32_32x32x16_f16, K = 16 = eight complex samples; lane half h
// carries samples t + 4h .. t + 4h + 3 as (re, im) pairs.  nt % 8 == 0.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t rot_h2(uint32_t v)
{
    // (re, im) -> (-im, re) on a packed half2: swap halves, flip the new low sign
    return ((v >> 16) | (v << 16)) ^ 0x00008000u;
}

__global__ __launch_bounds__(256) void k_xengine_f16(const uint32_t* __restrict__ X, int nchan,
                                                     long nt, int ninp, float* __restrict__ acc)
{
    const int c = blockIdx.y;
    const int tile = blockIdx.x;
    const int nb = (ninp + 63) / 64;
    const int ntl = nb * (nb + 1) / 2;
    int bi64, bj64;
    tile_coords(tile, nb, bi64, bj64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int bi = 2 * bi64 + (w >> 1), bj = 2 * bj64 + (w & 1);
    if (bi > bj || bi * 32 >= ninp || bj * 32 >= ninp) return;
    const int ra = bi * 32 + (lane & 31), rb = bj * 32 + (lane & 31);
    const bool va = ra < ninp, vb = rb < ninp;
    const int h = lane >> 5;
    const uint32_t* base = X + (size_t)c * (size_t)nt * (size_t)ninp;
    fx_float16 dre = {}, dim = {};
    for (long t = 0; t < nt; t += 8) {
        uint32_t ua[4], ub[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            long tt = t + 4 * h + q;
            bool vt = tt < nt;
            ua[q] = (va && vt) ? base[tt * ninp + ra] : 0u;
            ub[q] = (vb && vt) ? base[tt * ninp + rb] : 0u;
        }
        fx_int8 ia, ib, ir;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            ia[q] = (int)ua[q];
            ib[q] = (int)ub[q];
            ir[q] = (int)rot_h2(ub[q]);
        }
        // only the low 4 dwords carry data: repack as 8 halfs
        fx_half8 a, b, br;
        {
            typedef int int4v __attribute__((ext_vector_type(4)));
            int4v a4 = {ia[0], ia[1], ia[2], ia[3]};
            int4v b4 = {ib[0], ib[1], ib[2], ib[3]};
            int4v r4 = {ir[0], ir[1], ir[2], ir[3]};
            a = __builtin_bit_cast(fx_half8, a4);
            b = __builtin_bit_cast(fx_half8, b4);
            br = __builtin_bit_cast(fx_half8, r4);
        }
        dre = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, dre, 0, 0, 0);
        dim = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, br, dim, 0, 0, 0);
    }
    float* dst = acc + (((size_t)c * ntl + tile) * 4 + w) * (32 * 64);
    acc_store(dst, lane, dre, dim);
}

// ---------------------------------------------------------------------------
// dump: engine layout -> vis[p][c] = sum_s coef[s] * acc_s / (sc_c^2 s_i s_j)
// MFMA 32x32 output layout: lane l, register r holds row 8(r/4) + 4(l/32) + r%4,
// column l%32.
// ---------------------------------------------------------------------------
__global__ void k_xengine_dump(const float* __restrict__ acc, int nchan, int ninp, int nsets,
                               size_t set_stride, const float* __restrict__ coef,
                               const float* __restrict__ sc_chan, const float* __restrict__ sc_inp,
                               const int2* __restrict__ pij, long nprod, int accumulate,
                               float2* __restrict__ vis)
{
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nprod * nchan) return;
    const int c = (int)(idx % nchan);
    const long p = idx / nchan;
    const int2 ij = pij[p];
    const int i = ij.x, j = ij.y;
    const int nb = (ninp + 63) / 64;
    const int ntl = nb * (nb + 1) / 2;
    const int bi64 = i >> 6, bj64 = j >> 6;
    const int tile = bi64 * nb - bi64 * (bi64 - 1) / 2 + (bj64 - bi64);
    const int w = (((i >> 5) & 1) << 1) | ((j >> 5) & 1);
    const int ii = i & 31, jj = j & 31;
    const int r = 4 * (ii >> 3) + (ii & 3);
    const int l = jj + 32 * ((ii >> 2) & 1);
    const size_t off = (((size_t)c * ntl + tile) * 4 + w) * (32 * 64);
    float re = 0.f, im = 0.f;
    for (int s = 0; s < nsets; s++) {
        const float* a = acc + s * set_stride + off;
        float k = coef ? coef[s] : 1.0f;
        re += k * a[r * 64 + l];
        im += k * a[(16 + r) * 64 + l];
    }
    if (sc_chan) {
        float d = sc_chan[c] * sc_chan[c] * sc_inp[i] * sc_inp[j];
        re /= d;
        im /= d;
    }
    if (accumulate) {
        float2 o = vis[idx];
        re += o.x;
        im += o.y;
    }
    vis[idx] = make_float2(re, im);
}

// ---------------------------------------------------------------------------
// MFMA layout probe: A[i][k] = (i+1) [k==0], B[k][j] = (j+1) [k==0]
// ---------------------------------------------------------------------------
__global__ void k_mfma_probe(int which, float* out)
{
    const int lane = threadIdx.x;
    fx_float16 d = {};
    const int row = lane & 31;
    const bool k0 = lane < 32;
    if (which == 0) {
        float a = k0 ? (float)(row + 1) : 0.f;
        d = __builtin_amdgcn_mfma_f32_32x32x2f32(a, a, d, 0, 0, 0);
    } else if (which == 1) {
        fx_half8 a = {};
        if (k0) a[0] = (_Float16)(row + 1);
        d = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, a, d, 0, 0, 0);
    } else {
        fx_int8 a = {};
        if (k0) {
            __hip_fp8_e4m3 v((float)(row + 1));
            a[0] = (int)(uint32_t)v.__x;
        }
        d = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, a, d, 0, 0, 0, 127, 0, 127);
    }
#pragma unroll
    for (int r = 0; r < 16; r++) {
        int i = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        int j = lane & 31;
        out[i * 32 + j] = d[r];
    }
}

} // namespace

size_t fxk_xengine_acc_floats(int nchan, int ninp, int nsets)
{
    return (size_t)nsets * (size_t)nchan * (size_t)ntiles64(ninp) * 4u * 32u * 64u;
}

int fxk_xengine(hipStream_t s, int in_type, const void* X, int nchan, long nt, int ninp, float* acc)
{
    dim3 grid((unsigned)ntiles64(ninp), (unsigned)nchan);
    if (in_type == FX_PREC_F32) {
        hipLaunchKernelGGL(k_xengine_f32, grid, dim3(256), 0, s,
                           reinterpret_cast<const float2*>(X), nchan, nt, ninp, acc);
    } else if (in_type == FX_PREC_F16) {
        hipLaunchKernelGGL(k_xengine_f16, grid, dim3(256), 0, s,
                           reinterpret_cast<const uint32_t*>(X), nchan, nt, ninp, acc);
    } else {
        return FX_ERR_UNSUPPORTED;
    }
    return hipGetLastError() == hipSuccess ? FX_OK : FX_ERR_HIP;
}

int fxk_xengine_dump_pij(hipStream_t s, const float* acc, int nchan, int ninp, int nsets,
                         const float* set_coef, const float* scale_chan, const float* scale_inp,
                         const int2* pij, int accumulate, float2* vis)
{
    long nprod = (long)ninp * (ninp + 1) / 2;
    long total = nprod * nchan;
    size_t set_stride = fxk_xengine_acc_floats(nchan, ninp, 1);
    hipLaunchKernelGGL(k_xengine_dump, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, acc,
                       nchan, ninp, nsets, set_stride, set_coef, scale_chan, scale_inp, pij, nprod,
                       accumulate, vis);
    return hipGetLastError() == hipSuccess ? FX_OK : FX_ERR_HIP;
}

int fxk_mfma_probe(hipStream_t s, int which, float* out)
{
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, s, which, out);
    return hipGetLastError() == hipSuccess ? FX_OK : FX_ERR_HIP;
}
