// fx_xtile.h -- one tile of the FP32/FP16 X-engine, shared by the X-engine
// kernels (fx_xengine.hip) and the persistent batch kernel (fx_persist.hip).
//
// V_ij(c) += sum_t X_i(c,t) conj(X_j(c,t)) over one TSZ x TSZ input tile of
// one channel, the baseline MAC of genericAddProduct_32fc_conj
// (architecture.h:427-434), as two real GEMMs sharing the left operand: with
// u_i = (re_i(t0), im_i(t0), re_i(t1), ...) along K and the rotated copy
// v_j = (-im_j(t0), re_j(t0), ...), Re V = U U^T and Im V = U V^T.  A
// workgroup of 4 waves; per stage the KT-sample slabs X[t0..t0+KT][rows] of
// its two row ranges (one for diagonal tiles) go through LDS with coalesced
// 16-byte loads, the next stage prefetched into registers while the MFMAs
// run.  Wave w accumulates the (TSZ/2) x (TSZ/2) quadrant (w/2, w%2) as
// NBW x NBW blocks of 32 x 32 (re and im); the strictly-lower quadrant of a
// diagonal tile idles.
#pragma once

#include "fx_internal.h"

typedef _Float16 fx_half8 __attribute__((ext_vector_type(8)));
typedef float fx_float16 __attribute__((ext_vector_type(16)));
typedef int fx_int4v __attribute__((ext_vector_type(4)));
typedef int fx_int8 __attribute__((ext_vector_type(8)));

namespace {

constexpr int TS = 128;     // tile rows of the 8-bit engine (and default of the others)
constexpr int XNT = 256;    // threads per workgroup

__host__ __device__ inline int ntiles_ts(int ninp, int ts)
{
    int nb = (ninp + ts - 1) / ts;
    return nb * (nb + 1) / 2;
}
__host__ __device__ inline int ntiles128(int ninp) { return ntiles_ts(ninp, TS); }

// tile index -> (ti, tj), row-major over the upper triangle
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

__device__ __forceinline__ uint32_t rot_h2(uint32_t v)
{
    // (re, im) -> (-im, re) on a packed half2 in one packed multiply: the low
    // result is the high half times -1, the high result the low half times +1
    // (op_sel picks the halves) -- exact for every finite value and signed
    // zero, the same bits as swapping the halves and flipping the new low sign
    uint32_t r;
    asm("v_pk_mul_f16 %0, %1, %2 op_sel:[1,0] op_sel_hi:[0,1]" : "=v"(r) : "v"(v), "s"(0x3C00BC00u));
    return r;
}

template <int PREC> struct XTraits;
template <> struct XTraits<FX_PREC_F32> {
    typedef float2 E;                // one complex sample
    static constexpr int KT = 16;    // samples per LDS stage (16 KiB per slab at TSZ = 128)
};
template <> struct XTraits<FX_PREC_F16> {
    typedef uint32_t E;
    static constexpr int KT = 32;
};

// Accumulate the tile (ti, tj) of the channel whose samples start at base
// ([t][i], nt spectra of ninp inputs) and hand each active wave's
// accumulators to out(dre, dim, w, qi, qj, diag, lane): dre[a][b][r] and
// dim[a][b][r] hold rows 8 (r/4) + 4 (lane/32) + r%4 and column lane%32 of
// block (a, b) of the wave's quadrant.  slab: 2 x KT x TSZ samples of LDS.
template <int PREC, int TSZ, typename Out>
__device__ __forceinline__ void xengine_tile(const typename XTraits<PREC>::E* __restrict__ base, long nt, int ninp,
                                             int ti, int tj,
                                             typename XTraits<PREC>::E (*slab)[XTraits<PREC>::KT][TSZ], Out&& out)
{
    typedef typename XTraits<PREC>::E E;
    constexpr int KT = XTraits<PREC>::KT;
    constexpr int NBW = TSZ / 64;                         // blocks per wave per dimension
    constexpr int CH = 16 / sizeof(E);                    // samples per 16-byte chunk
    constexpr int SLAB_CHUNKS = KT * TSZ / CH;            // chunks per slab
    constexpr int LPT = SLAB_CHUNKS / XNT;                // chunks per thread per slab
    static_assert(LPT >= 1, "slab smaller than the workgroup");
    const bool diag = ti == tj;
    const int nslab = diag ? 1 : 2;
    const int tid = threadIdx.x;
    const int w = tid >> 6, lane = tid & 63;
    const int qi = w >> 1, qj = w & 1;
    const bool active = !(diag && qi > qj);
    const int row0[2] = {ti * TSZ, tj * TSZ};

    fx_float16 dre[NBW][NBW], dim[NBW][NBW];
#pragma unroll
    for (int a = 0; a < NBW; a++)
#pragma unroll
        for (int b = 0; b < NBW; b++) {
            dre[a][b] = fx_float16{};
            dim[a][b] = fx_float16{};
        }

    // global -> register prefetch of one stage: chunk q of slab s covers
    // sample t0 + q / (TSZ/CH), rows (q % (TSZ/CH)) * CH .. + CH
    uint4 pre[2][LPT];
    auto gload = [&](long t0) {
#pragma unroll
        for (int s = 0; s < 2; s++) {
            if (s >= nslab) break;
#pragma unroll
            for (int k = 0; k < LPT; k++) {
                const int q = tid + k * XNT;
                const int tt = q / (TSZ / CH);
                const int rr = (q - tt * (TSZ / CH)) * CH;
                const long t = t0 + tt;
                const int row = row0[s] + rr;
                uint4 v = make_uint4(0u, 0u, 0u, 0u);
                if (t < nt) {
                    if (row + CH <= ninp && (ninp % CH) == 0) {
                        v = *reinterpret_cast<const uint4*>(base + (size_t)t * ninp + row);
                    } else {
                        E tmp[CH];
#pragma unroll
                        for (int e = 0; e < CH; e++) {
                            if (row + e < ninp) tmp[e] = base[(size_t)t * ninp + row + e];
                            else __builtin_memset(&tmp[e], 0, sizeof(E));
                        }
                        v = *reinterpret_cast<uint4*>(tmp);
                    }
                }
                pre[s][k] = v;
            }
        }
    };
    auto lstore = [&]() {
#pragma unroll
        for (int s = 0; s < 2; s++) {
            if (s >= nslab) break;
#pragma unroll
            for (int k = 0; k < LPT; k++) {
                const int q = tid + k * XNT;
                const int tt = q / (TSZ / CH);
                const int rr = (q - tt * (TSZ / CH)) * CH;
                *reinterpret_cast<uint4*>(&slab[s][tt][rr]) = pre[s][k];
            }
        }
    };

    const int sb = diag ? 0 : 1;          // LDS slab holding the j rows
    const int r32 = lane & 31, h = lane >> 5;
    const long nstage = (nt + KT - 1) / KT;
    gload(0);
    for (long st = 0; st < nstage; st++) {
        __syncthreads();
        lstore();
        __syncthreads();
        if (st + 1 < nstage) gload((st + 1) * KT);
        if (!active) continue;
        if constexpr (PREC == FX_PREC_F16) {
#pragma unroll
            for (int ks = 0; ks < KT / 8; ks++) {
                fx_half8 fa[NBW], fb[NBW], fr[NBW];
#pragma unroll
                for (int a = 0; a < NBW; a++) {
                    fx_int4v ia;
                    const int row = (NBW * qi + a) * 32 + r32;
#pragma unroll
                    for (int q = 0; q < 4; q++) ia[q] = (int)slab[0][ks * 8 + 4 * h + q][row];
                    fa[a] = __builtin_bit_cast(fx_half8, ia);
                }
#pragma unroll
                for (int b = 0; b < NBW; b++) {
                    fx_int4v ib, ir;
                    const int row = (NBW * qj + b) * 32 + r32;
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        uint32_t v = slab[sb][ks * 8 + 4 * h + q][row];
                        ib[q] = (int)v;
                        ir[q] = (int)rot_h2(v);
                    }
                    fb[b] = __builtin_bit_cast(fx_half8, ib);
                    fr[b] = __builtin_bit_cast(fx_half8, ir);
                }
#pragma unroll
                for (int a = 0; a < NBW; a++)
#pragma unroll
                    for (int b = 0; b < NBW; b++) {
                        if (diag && qi == qj && a > b) continue;
                        dre[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fa[a], fb[b], dre[a][b], 0, 0, 0);
                        dim[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fa[a], fr[b], dim[a][b], 0, 0, 0);
                    }
            }
        } else {
            // FP32: K = 2 reals = one complex sample per MFMA; lanes 0-31
            // carry the real parts, lanes 32-63 the imaginary parts
#pragma unroll
            for (int ks = 0; ks < KT; ks++) {
                float fa[NBW], fb[NBW], fr[NBW];
#pragma unroll
                for (int a = 0; a < NBW; a++) {
                    const float2 v = slab[0][ks][(NBW * qi + a) * 32 + r32];
                    fa[a] = h ? v.y : v.x;
                }
#pragma unroll
                for (int b = 0; b < NBW; b++) {
                    const float2 v = slab[sb][ks][(NBW * qj + b) * 32 + r32];
                    fb[b] = h ? v.y : v.x;
                    fr[b] = h ? v.x : -v.y;
                }
#pragma unroll
                for (int a = 0; a < NBW; a++)
#pragma unroll
                    for (int b = 0; b < NBW; b++) {
                        if (diag && qi == qj && a > b) continue;
                        dre[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[a], fb[b], dre[a][b], 0, 0, 0);
                        dim[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[a], fr[b], dim[a][b], 0, 0, 0);
                    }
            }
        }
    }
    if (!active) return;
    out(dre, dim, w, qi, qj, diag, lane);
}

} // namespace
