// fx_xengine.hip -- X-engine: per-channel complex Hermitian GEMM on MFMA.
//
// V_ij(c) += sum_t X_i(c,t) conj(X_j(c,t)), i <= j, the baseline MAC of
// genericAddProduct_32fc_conj (architecture.h:427-434) that the minicorrelator
// runs per baseline (three_baseline_naive_MACauto, minicorrelator_spu.c:745-786,
// with its missing conjugation fixed).
//
// Complex arithmetic as two real GEMMs sharing the left operand: with
// u_i = (re_i(t0), im_i(t0), re_i(t1), ...) along K and the rotated copy
// v_j = (-im_j(t0), re_j(t0), ...), Re V = U U^T and Im V = U V^T.
//
// Tiling: a workgroup of 4 waves owns one 128 x 128 input tile (ti <= tj) of
// one channel.  Per stage it copies the KT-sample slabs X[c][t0..t0+KT][rows]
// of its two row ranges (one for diagonal tiles) from HBM into LDS with
// coalesced 16-byte loads, prefetching the next stage into registers while the
// MFMAs run, so every channelized sample is read from HBM once per tile row.
// Wave w accumulates the 64 x 64 quadrant (w/2, w%2) as 2 x 2 blocks of
// 32 x 32 (re and im: 128 accumulators); the strictly-lower quadrant of a
// diagonal tile idles.  Accumulators are added into an engine-layout buffer
// acc[c][tile][wave][block][reg][lane] (coalesced), converted to the
// canonical [p][c] complex64 visibilities by k_xengine_dump.
#include "fx_internal.h"
#include "fx_xtile.h"

#include <hip/hip_fp8.h>

#include <cstdlib>

namespace {

// TSZ x TSZ input tiles (TSZ = 128 or 64); wave w = (qi, qj) owns the
// (TSZ/2) x (TSZ/2) quadrant as NBW x NBW blocks of 32 x 32, NBW = TSZ/64.
// Engine image: acc[c][tile][block][reg][lane] with block = w NBW^2 + a NBW + b.
template <int PREC, int TSZ>
__global__ __launch_bounds__(XNT, 2) void k_xengine(const typename XTraits<PREC>::E* __restrict__ X,
                                                    int nchan, long nt, int ninp,
                                                    float* __restrict__ acc)
{
    typedef typename XTraits<PREC>::E E;
    constexpr int KT = XTraits<PREC>::KT;
    constexpr int NBW = TSZ / 64;                         // blocks per wave per dimension
    constexpr int NBT = (TSZ / 32) * (TSZ / 32);          // blocks per tile
    __shared__ E slab[2][KT][TSZ];

    const int nb = (ninp + TSZ - 1) / TSZ;
    const int ntl = nb * (nb + 1) / 2;
    const long lb = xcd_block((long)ntl * nchan);
    if (lb >= (long)ntl * nchan) return;
    const int c = (int)(lb / ntl);
    const int tile = (int)(lb - (long)c * ntl);
    int ti, tj;
    tile_coords(tile, nb, ti, tj);
    xengine_tile<PREC, TSZ>(X + (size_t)c * (size_t)nt * (size_t)ninp, nt, ninp, ti, tj, slab,
                            [&](const auto& dre, const auto& dim, int w, int qi, int qj, bool diag, int lane) {
        float* dst = acc + (((size_t)c * ntl + tile) * NBT + w * NBW * NBW) * (32 * 64);
#pragma unroll
        for (int a = 0; a < NBW; a++)
#pragma unroll
            for (int b = 0; b < NBW; b++) {
                if (diag && qi == qj && a > b) continue;
                float* d = dst + (a * NBW + b) * (32 * 64);
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    d[r * 64 + lane] += dre[a][b][r];
                    d[(16 + r) * 64 + lane] += dim[a][b][r];
                }
            }
    });
}

// ---------------------------------------------------------------------------
// One full 128 x 128 tile (ninp == 128, nt a multiple of KT): the stages go
// through a ring of RING slabs in LDS filled by LDS-DMA loads
// (global_load_lds_dwordx4: lane l's 16 bytes land at the wave's LDS base +
// 16 l), so RING - 1 stages are in flight while one is multiplied and no
// registers hold prefetched data.  Same products, order and engine image as
// k_xengine<PREC, 128>.
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ void wait_vmcnt()
{
    // s_waitcnt simm16 on gfx9: vmcnt[3:0] | expcnt[6:4] | lgkmcnt[11:8] | vmcnt[5:4] << 14
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 0xF) | (0x7 << 4) | (0xF << 8) | (((N >> 4) & 3) << 14));
}

template <int PREC>
__global__ __launch_bounds__(XNT, 2) void k_xengine_ring(const typename XTraits<PREC>::E* __restrict__ X,
                                                         int nchan, long nt, float* __restrict__ acc)
{
    typedef typename XTraits<PREC>::E E;
    constexpr int TSZ = 128, NBW = 2, NBT = 16;
    constexpr int KT = XTraits<PREC>::KT;
    constexpr int RING = 4;
    constexpr int STAGE_BYTES = KT * TSZ * (int)sizeof(E);
    constexpr int NPER = STAGE_BYTES / 1024 / (XNT / 64);   // 1 KB DMA loads per wave per stage
    static_assert(NPER * 1024 * (XNT / 64) == STAGE_BYTES, "stage is not whole 1 KB loads per wave");
    __shared__ E slab[RING][KT][TSZ];

    const long lb = xcd_block((long)nchan);
    if (lb >= (long)nchan) return;
    const int c = (int)lb;
    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int qi = w >> 1, qj = w & 1;
    const bool active = qi <= qj;
    const char* base = reinterpret_cast<const char*>(X + (size_t)c * (size_t)nt * TSZ);
    const long nstage = nt / KT;

    auto issue = [&](long st) {
        const char* src = base + (size_t)st * STAGE_BYTES;
        char* dst = reinterpret_cast<char*>(&slab[st % RING][0][0]);
#pragma unroll
        for (int i = 0; i < NPER; i++) {
            const int blk = w * NPER + i;
            __builtin_amdgcn_global_load_lds(src + blk * 1024 + lane * 16,
                                             (__attribute__((address_space(3))) void*)(dst + blk * 1024), 16, 0, 0);
        }
    };

    fx_float16 dre[NBW][NBW], dim[NBW][NBW];
#pragma unroll
    for (int a = 0; a < NBW; a++)
#pragma unroll
        for (int b = 0; b < NBW; b++) {
            dre[a][b] = fx_float16{};
            dim[a][b] = fx_float16{};
        }
    const int r32 = lane & 31, h = lane >> 5;
    for (long st = 0; st < RING - 1 && st < nstage; st++) issue(st);
    for (long st = 0; st < nstage; st++) {
        // this wave's loads of stage st are done once at most the later
        // stages' loads are outstanding; the barrier makes every wave's
        // stage visible and frees the slab of stage st - 1
        const long ahead = nstage - 1 - st;
        if (ahead >= RING - 2) wait_vmcnt<(RING - 2) * NPER>();
        else if (ahead == 1) wait_vmcnt<NPER>();
        else wait_vmcnt<0>();
        // a plain barrier: __syncthreads() would also wait for the later
        // stages' loads (its release fence drains vmcnt)
        asm volatile("s_barrier" ::: "memory");
        if (st + RING - 1 < nstage) issue(st + RING - 1);
        if (!active) continue;
        const E (*sl)[TSZ] = slab[st % RING];
        if constexpr (PREC == FX_PREC_F16) {
#pragma unroll
            for (int ks = 0; ks < KT / 8; ks++) {
                fx_half8 fa[NBW], fb[NBW], fr[NBW];
#pragma unroll
                for (int a = 0; a < NBW; a++) {
                    fx_int4v ia;
                    const int row = (NBW * qi + a) * 32 + r32;
#pragma unroll
                    for (int q = 0; q < 4; q++) ia[q] = (int)sl[ks * 8 + 4 * h + q][row];
                    fa[a] = __builtin_bit_cast(fx_half8, ia);
                }
#pragma unroll
                for (int b = 0; b < NBW; b++) {
                    fx_int4v ib, ir;
                    const int row = (NBW * qj + b) * 32 + r32;
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        uint32_t v = sl[ks * 8 + 4 * h + q][row];
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
                        if (qi == qj && a > b) continue;
                        dre[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fa[a], fb[b], dre[a][b], 0, 0, 0);
                        dim[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fa[a], fr[b], dim[a][b], 0, 0, 0);
                    }
            }
        } else {
#pragma unroll
            for (int ks = 0; ks < KT; ks++) {
                float fa[NBW], fb[NBW], fr[NBW];
#pragma unroll
                for (int a = 0; a < NBW; a++) {
                    const float2 v = sl[ks][(NBW * qi + a) * 32 + r32];
                    fa[a] = h ? v.y : v.x;
                }
#pragma unroll
                for (int b = 0; b < NBW; b++) {
                    const float2 v = sl[ks][(NBW * qj + b) * 32 + r32];
                    fb[b] = h ? v.y : v.x;
                    fr[b] = h ? v.x : -v.y;
                }
#pragma unroll
                for (int a = 0; a < NBW; a++)
#pragma unroll
                    for (int b = 0; b < NBW; b++) {
                        if (qi == qj && a > b) continue;
                        dre[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[a], fb[b], dre[a][b], 0, 0, 0);
                        dim[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[a], fr[b], dim[a][b], 0, 0, 0);
                    }
            }
        }
    }
    if (!active) return;
    float* dst = acc + ((size_t)c * NBT + w * NBW * NBW) * (32 * 64);
#pragma unroll
    for (int a = 0; a < NBW; a++)
#pragma unroll
        for (int b = 0; b < NBW; b++) {
            if (qi == qj && a > b) continue;
            float* d = dst + (a * NBW + b) * (32 * 64);
#pragma unroll
            for (int r = 0; r < 16; r++) {
                d[r * 64 + lane] += dre[a][b][r];
                d[(16 + r) * 64 + lane] += dim[a][b][r];
            }
        }
}

// ---------------------------------------------------------------------------
// Full 128 x 128 tiles of more inputs (ninp a multiple of 128): a stage holds
// the KT/2-sample slabs of the tile's row block and column block, ring of 4
// in LDS, filled by LDS-DMA loads as in k_xengine_ring; a slab's time rows
// are ninp apart in HBM, so each lane computes its own source address.  The
// waves split the tile into quadrants as k_xengine<PREC, 128> does (same
// products in the same order, same engine image).
//
// LOCK: the workgroups form groups of ntl (the tiles of one channel), all of
// a group on one XCD (blocks b = x + 8 (g ntl + tile) of XCD x); a group
// walks the channels grp, grp + 8 ngx, ... together.  Each 128-input block of
// a channel is read by 5 tile slabs; kept within a few stages of each other,
// the group's workgroups find the blocks their siblings fetched still in the
// XCD's L2, where tiles scheduled independently drift apart by more than the
// L2 holds.  Every XLP stages wave 0 publishes the stages it has consumed in
// prog[grp ntl + tile] and, if a sibling was more than lag stages behind at
// the previous round, polls until it has caught up; the other waves wait at
// the stage barrier.  The waits only order the fetches -- results do not
// depend on them -- and are bounded: after XSPIN polls the workgroup stops
// waiting for the rest of the launch.  The host sizes the grid so that every
// group is resident at once (its tiles' workgroups fit one XCD's CUs).
// ---------------------------------------------------------------------------
constexpr int XLP = 4;             // stages per progress round
constexpr int XSPIN = 1 << 14;     // polls of a wait before lockstep is dropped

template <int PREC, bool LOCK>
__global__ __launch_bounds__(XNT, 2) void k_xengine_ring2(const typename XTraits<PREC>::E* __restrict__ X,
                                                          int nchan, long nt, int ninp, float* __restrict__ acc,
                                                          unsigned* __restrict__ prog, int ngx, unsigned lag)
{
    typedef typename XTraits<PREC>::E E;
    constexpr int TSZ = 128, NBW = 2, NBT = 16;
    constexpr int KT = XTraits<PREC>::KT / 2;            // samples per stage and slab
    constexpr int RING = 4;
    constexpr int ROWB = TSZ * (int)sizeof(E);           // bytes of one slab time row
    constexpr int RPB = 1024 / ROWB;                     // time rows per 1 KB load
    constexpr int LPR = 64 / RPB;                        // lanes per time row
    constexpr int CHE = 16 / (int)sizeof(E);             // samples per lane
    constexpr int BPS = KT * ROWB / 1024;                // 1 KB loads per slab
    constexpr int NPER = 2 * BPS / (XNT / 64);           // per wave per stage
    static_assert(NPER * (XNT / 64) == 2 * BPS && RPB * ROWB == 1024, "stage is not whole 1 KB loads per wave");
    __shared__ E slab[RING][2][KT][TSZ];

    const int nb = ninp / TSZ;
    const int ntl = nb * (nb + 1) / 2;
    int c0, tile, cstep, grp = 0;
    if constexpr (LOCK) {
        const int b = blockIdx.x, l = b >> 3;
        const int gx = l / ntl;
        tile = l - gx * ntl;
        grp = (b & 7) + 8 * gx;
        c0 = grp;
        cstep = 8 * ngx;
    } else {
        const long lb = xcd_block((long)ntl * nchan);
        if (lb >= (long)ntl * nchan) return;
        c0 = (int)(lb / ntl);
        tile = (int)(lb - (long)c0 * ntl);
        cstep = nchan;
    }
    int ti, tj;
    tile_coords(tile, nb, ti, tj);
    const bool diag = ti == tj;
    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int qi = w >> 1, qj = w & 1;
    const bool active = !(diag && qi > qj);
    const long nstage = nt / KT;
    // this lane's part of every 1 KB load: time row within the block, column
    const int trow = lane / LPR, col = (lane % LPR) * CHE;
    unsigned* const sib = prog + (size_t)grp * ntl;
    bool lock = LOCK;
    unsigned seen = 0;      // (wave 0, lanes < ntl) the siblings' progress at the last round
    auto peek = [&]() { return __hip_atomic_load(sib + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); };
    unsigned sg = 0;        // stages this workgroup has consumed, all channels
    for (int c = c0; c < nchan; c += cstep) {
        if (c != c0) __syncthreads();   // every wave is done with the previous channel's ring
        // this lane's source of each of its NPER loads in stage 0; stage s is
        // s KT time rows further
        const E* lsrc[NPER];
#pragma unroll
        for (int i = 0; i < NPER; i++) {
            const int blk = w * NPER + i;                    // 0 .. 2 BPS - 1
            const int sl = blk / BPS, b = blk - sl * BPS;
            lsrc[i] = X + ((size_t)c * (size_t)nt + (size_t)(b * RPB + trow)) * (size_t)ninp + (sl ? tj : ti) * TSZ +
                      col;
        }
        const size_t sstep = (size_t)KT * (size_t)ninp;     // samples per stage of the channel

        // the loads of stage st into its ring slot; past the last stage the
        // last stage's samples again (every stage issues NPER loads per wave,
        // so one wait count fits every stage, and the compiler can count them)
        auto issue = [&](long st) {
            const size_t so = (size_t)(st < nstage ? st : nstage - 1) * sstep;
#pragma unroll
            for (int i = 0; i < NPER; i++) {
                const int blk = w * NPER + i;
                const int sl = blk / BPS, b = blk - sl * BPS;
                char* dst = reinterpret_cast<char*>(&slab[st % RING][sl][0][0]) + b * 1024;
                const char* src = reinterpret_cast<const char*>(lsrc[i] + so);
                __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
            }
        };

        fx_float16 dre[NBW][NBW], dim[NBW][NBW];
#pragma unroll
        for (int a = 0; a < NBW; a++)
#pragma unroll
            for (int b = 0; b < NBW; b++) {
                dre[a][b] = fx_float16{};
                dim[a][b] = fx_float16{};
            }
        const int r32 = lane & 31, h = lane >> 5;
        for (long st = 0; st < RING - 1; st++) issue(st);
        // LOCK: rounds of XLP stages unrolled (the host ensures nstage % XLP
        // == 0), so the round's progress load sits at one place in straight
        // code and the compiler can count the loads issued after it
        constexpr int SPR = LOCK ? XLP : 1;
        for (long st0 = 0; st0 < nstage; st0 += SPR) {
#pragma unroll
            for (int k = 0; k < SPR; k++, sg++) {
                const long st = st0 + k;
                // this wave's loads of stage st are done once at most the two
                // later stages' loads are outstanding
                wait_vmcnt<(RING - 2) * NPER>();
                if constexpr (LOCK) {
                    if (k == 0 && lock && w == 0 && sg) {
                        const unsigned need = sg > lag ? sg - lag : 0u;
                        // seen was loaded a round ago (the wait above has retired it:
                        // three stages' loads followed it); if a sibling trails, poll
                        // until it catches up
                        if (!__all(lane >= ntl || seen >= need)) {
                            for (int n = 0;; n++) {
                                if (n == XSPIN) {
                                    lock = false;
                                    break;
                                }
                                __builtin_amdgcn_s_sleep(1);
                                const unsigned v = lane < ntl ? peek() : ~0u;
                                if (__all(v >= need)) break;
                            }
                        }
                        if (lane == 0) __hip_atomic_store(sib + tile, sg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (lane < ntl) seen = peek();
                    }
                }
                asm volatile("s_barrier" ::: "memory");
                issue(st + RING - 1);
                if (!active) continue;
                const E (*sa)[TSZ] = slab[st % RING][0];
                const E (*sb)[TSZ] = slab[st % RING][1];
                if constexpr (PREC == FX_PREC_F16) {
#pragma unroll
                    for (int ks = 0; ks < KT / 8; ks++) {
                        fx_half8 fa[NBW], fb[NBW], fr[NBW];
#pragma unroll
                        for (int a = 0; a < NBW; a++) {
                            fx_int4v ia;
                            const int row = (NBW * qi + a) * 32 + r32;
#pragma unroll
                            for (int q = 0; q < 4; q++) ia[q] = (int)sa[ks * 8 + 4 * h + q][row];
                            fa[a] = __builtin_bit_cast(fx_half8, ia);
                        }
#pragma unroll
                        for (int b = 0; b < NBW; b++) {
                            fx_int4v ib, ir;
                            const int row = (NBW * qj + b) * 32 + r32;
#pragma unroll
                            for (int q = 0; q < 4; q++) {
                                uint32_t v = sb[ks * 8 + 4 * h + q][row];
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
#pragma unroll
                    for (int ks = 0; ks < KT; ks++) {
                        float fa[NBW], fb[NBW], fr[NBW];
#pragma unroll
                        for (int a = 0; a < NBW; a++) {
                            const float2 v = sa[ks][(NBW * qi + a) * 32 + r32];
                            fa[a] = h ? v.y : v.x;
                        }
#pragma unroll
                        for (int b = 0; b < NBW; b++) {
                            const float2 v = sb[ks][(NBW * qj + b) * 32 + r32];
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
        }
        wait_vmcnt<0>();   // the loads past the last stage have landed before the ring is reused or freed
        if (!active) continue;
        float* dst = acc + (((size_t)c * ntl + tile) * NBT + w * NBW * NBW) * (32 * 64);
#pragma unroll
        for (int a = 0; a < NBW; a++)
#pragma unroll
            for (int b = 0; b < NBW; b++) {
                if (diag && qi == qj && a > b) continue;
                float* d = dst + (a * NBW + b) * (32 * 64);
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    d[r * 64 + lane] += dre[a][b][r];
                    d[(16 + r) * 64 + lane] += dim[a][b][r];
                }
            }
    }
}

// ---------------------------------------------------------------------------
// 8-bit engine: v_mfma_scale_f32_32x32x64_f8f6f4 with unit (E8M0 = 127)
// scales, K = 64 e4m3 values = 32 complex samples per instruction.
//
// Input is input-major, X[c][i][t], so one row of a stage (KT samples of one
// input) is a contiguous HBM run and lands in LDS as the contiguous K-run a
// lane feeds to the MFMA: A and B operands are two 16-byte LDS reads per lane.
// (Which K index each byte carries is irrelevant: A and B are packed alike.)
//
// NS = 1: e4m3 (re, im) pairs, one accumulator set.
// NS = 3: 2-bit codes expanded while staging into the three exact operands of
// X = S + (H-1) T: S = sign (+-1), U = S + T (0, +-1, +-2) and T (0, +-1),
// each an e4m3 table lookup per code (v_perm_b32), so every product and every
// FP32 partial sum is exact.
//
// Workgroup: 8 waves own one 128 x 128 tile of one channel; wave w computes
// rows 32 (w/2) .. +32 against columns 64 (w%2) .. +64 (two 32 x 32 blocks).
// Re V = A B^T with A = B = (re, im) pairs; Im V = A' B^T with the rotated
// A' = (im, -re).  Accumulators go to the same engine layout as k_xengine,
// one set after another (set stride = one engine image).
// ---------------------------------------------------------------------------
constexpr int X8NT = 512;                // threads (8 waves)
constexpr int X8KT = 64;                 // samples per stage

// e4m3 images indexed by the 2-bit code of {-H, -1, +1, +H} (byte c = code c)
constexpr uint32_t kTblS = 0x3838B8B8u;  // S:  -1 -1 +1 +1
constexpr uint32_t kTblU = 0x4038B8C0u;  // U:  -2 -1 +1 +2
constexpr uint32_t kTblT = 0x380000B8u;  // T:  -1  0  0 +1

// four code bytes (re | im << 2) -> four e4m3 (re, im) pairs
__device__ __forceinline__ uint2 expand4(uint32_t w, uint32_t tbl)
{
    const uint32_t re = __builtin_amdgcn_perm(0u, tbl, w & 0x03030303u);
    const uint32_t im = __builtin_amdgcn_perm(0u, tbl, (w >> 2) & 0x03030303u);
    return make_uint2(__builtin_amdgcn_perm(im, re, 0x05010400u), __builtin_amdgcn_perm(im, re, 0x07030602u));
}

// two e4m3 pairs (re0, im0, re1, im1) -> (im0, -re0, im1, -re1)
__device__ __forceinline__ int rot_p8(int v)
{
    const uint32_t u = (uint32_t)v;
    return (int)(__builtin_amdgcn_perm(u, u, 0x02030001u) ^ 0x80008000u);
}

template <int NS>
__global__ __launch_bounds__(X8NT, (NS == 1 ? 2 : 1)) void k_xengine8(const uint8_t* __restrict__ X, int nchan,
                                                                      long nt, int ninp, float* __restrict__ acc,
                                                                      size_t set_stride)
{
    constexpr int BPS = NS == 1 ? 2 : 1;          // HBM bytes per complex sample
    constexpr int CPR = X8KT * BPS / 16;          // 16-byte chunks per row per stage (8 or 4)
    constexpr int RCH = TS * CPR;                 // chunks per row range
    constexpr int LPT = 2 * RCH / X8NT;           // chunks per thread (4 or 2)
    constexpr int SPC = 16 / BPS;                 // samples per chunk (8 or 16)
    constexpr int RS = X8KT * 2 + 16;             // LDS row stride in bytes (padded)
    __shared__ __attribute__((aligned(16))) uint8_t op[2][NS][TS][RS];

    const int nb = (ninp + TS - 1) / TS;
    const int ntl = nb * (nb + 1) / 2;
    const long lb = xcd_block((long)ntl * nchan);
    if (lb >= (long)ntl * nchan) return;
    const int c = (int)(lb / ntl);
    const int tile = (int)(lb - (long)c * ntl);
    int ti, tj;
    tile_coords(tile, nb, ti, tj);
    const bool diag = ti == tj;
    const int nslab = diag ? 1 : 2;
    const int tid = threadIdx.x;
    const int w = tid >> 6, lane = tid & 63;
    const int qa = w >> 1, qb = w & 1;
    bool act[2];
    act[0] = !(diag && qa > 2 * qb);
    act[1] = !(diag && qa > 2 * qb + 1);
    const uint8_t* base = X + (size_t)c * (size_t)ninp * (size_t)nt * BPS;
    const int row0[2] = {ti * TS, tj * TS};

    fx_float16 are[NS][2], aim[NS][2];
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            are[s][b] = fx_float16{};
            aim[s][b] = fx_float16{};
        }

    // chunk q = tid + k * X8NT: range q / RCH, row (q % RCH) / CPR, part q % CPR
    uint4 pre[LPT];
    bool val[LPT];
    auto gload = [&](long t0) {
#pragma unroll
        for (int k = 0; k < LPT; k++) {
            const int q = tid + k * X8NT;
            const int s = q / RCH;
            const int row = (q - s * RCH) / CPR;
            const int part = q % CPR;
            const int r = row0[s] + row;
            const long t = t0 + (long)part * SPC;
            val[k] = s < nslab && r < ninp && t < nt;
            pre[k] = val[k] ? *reinterpret_cast<const uint4*>(base + ((size_t)r * (size_t)nt + (size_t)t) * BPS)
                            : make_uint4(0u, 0u, 0u, 0u);
        }
    };
    auto lstore = [&]() {
#pragma unroll
        for (int k = 0; k < LPT; k++) {
            const int q = tid + k * X8NT;
            const int s = q / RCH;
            if (s >= nslab) continue;
            const int row = (q - s * RCH) / CPR;
            const int part = q % CPR;
            if constexpr (NS == 1) {
                *reinterpret_cast<uint4*>(&op[s][0][row][part * 16]) = pre[k];
            } else {
                const uint32_t wv[4] = {pre[k].x, pre[k].y, pre[k].z, pre[k].w};
                const uint32_t tbl[3] = {kTblS, kTblU, kTblT};
#pragma unroll
                for (int set = 0; set < 3; set++) {
                    uint2 e[4];
#pragma unroll
                    for (int d = 0; d < 4; d++) e[d] = val[k] ? expand4(wv[d], tbl[set]) : make_uint2(0u, 0u);
                    uint4* dst = reinterpret_cast<uint4*>(&op[s][set][row][part * 32]);
                    dst[0] = make_uint4(e[0].x, e[0].y, e[1].x, e[1].y);
                    dst[1] = make_uint4(e[2].x, e[2].y, e[3].x, e[3].y);
                }
            }
        }
    };

    const int sb = diag ? 0 : 1;
    const int r32 = lane & 31, h = lane >> 5;
    const long nstage = (nt + X8KT - 1) / X8KT;
    gload(0);
    for (long st = 0; st < nstage; st++) {
        __syncthreads();
        lstore();
        __syncthreads();
        if (st + 1 < nstage) gload((st + 1) * X8KT);
        if (!act[0] && !act[1]) continue;
#pragma unroll
        for (int ks = 0; ks < X8KT / 32; ks++) {
            const int kb = ks * 64 + h * 32;      // byte offset of this lane's 16 samples
#pragma unroll
            for (int set = 0; set < NS; set++) {
                const uint4* pa = reinterpret_cast<const uint4*>(&op[0][set][qa * 32 + r32][kb]);
                const uint4 a0 = pa[0], a1 = pa[1];
                fx_int8 fa = {(int)a0.x, (int)a0.y, (int)a0.z, (int)a0.w, (int)a1.x, (int)a1.y, (int)a1.z, (int)a1.w};
                fx_int8 fr;
#pragma unroll
                for (int d = 0; d < 8; d++) fr[d] = rot_p8(fa[d]);
#pragma unroll
                for (int b = 0; b < 2; b++) {
                    if (!act[b]) continue;
                    const uint4* pb = reinterpret_cast<const uint4*>(&op[sb][set][(2 * qb + b) * 32 + r32][kb]);
                    const uint4 b0 = pb[0], b1 = pb[1];
                    fx_int8 fb = {(int)b0.x, (int)b0.y, (int)b0.z, (int)b0.w, (int)b1.x, (int)b1.y, (int)b1.z, (int)b1.w};
                    are[set][b] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(fa, fb, are[set][b], 0, 0, 0, 127, 0, 127);
                    aim[set][b] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(fr, fb, aim[set][b], 0, 0, 0, 127, 0, 127);
                }
            }
        }
    }
#pragma unroll
    for (int b = 0; b < 2; b++) {
        if (!act[b]) continue;
        const int bi = qa, bj = 2 * qb + b;
        const int w4 = (bi >> 1) * 2 + (bj >> 1), s4 = (bi & 1) * 2 + (bj & 1);
#pragma unroll
        for (int set = 0; set < NS; set++) {
            float* d = acc + set * set_stride + (((size_t)c * ntl + tile) * 16 + w4 * 4 + s4) * (32 * 64);
#pragma unroll
            for (int r = 0; r < 16; r++) {
                d[r * 64 + lane] += are[set][b][r];
                d[(16 + r) * 64 + lane] += aim[set][b][r];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Exact 2-bit engine on FP4 MFMA: v_mfma_scale_f32_32x32x64_f8f6f4 with
// cbsz = blgp = 4 (e2m1: A and B are 32 values = four VGPRs a lane, twice the
// e4m3 rate) and unit scales.  The operands are the same three exact parts of
// the 2-bit levels as the FP8 form -- S (sign), U = S + T, T (sign on the
// outer levels, 0 inside), all of them e2m1 values (0, +-1, +-2) -- so every
// product and every FP32 partial sum is the same exact integer.
//
// Operand layout, re-planar: a lane's 16 bytes are 16 complex samples, words
// 0 and 1 their re parts as nibbles and words 2 and 3 their im parts in the
// same sample order, so the rotated operand (im, -re) of Im V is words
// (2, 3, 0 ^ sign, 1 ^ sign): two XORs a 16-sample operand.  Within a word
// nibble 2k is sample k and nibble 2k+1 sample k+4 of an 8-sample group
// (A and B are packed alike, so the K order is free).
//
// The code bits map to e2m1 nibbles bitwise, eight samples a 32-bit word:
// code (hi, lo) = (0,0) -H, (0,1) -1, (1,0) +1, (1,1) +H; sign = !hi,
// eq = (hi == lo) marks |x| = H; S = 0x2 | sign, U = S + (eq << 1) (0x4 = 2),
// T = eq ? S : 0.
//
// Workgroup and tile as k_xengine8 (8 waves on one 128 x 128 tile of one
// channel, wave w on rows 32 (w/2) and columns 64 (w%2), three accumulator
// sets), with the stage loop double-buffered in LDS and one barrier a stage:
// while the MFMAs read stage st from one buffer, the same waves expand stage
// st+1 from registers into the other and fetch stage st+2 from HBM.
// ---------------------------------------------------------------------------
constexpr int X4KT = 64;                 // samples per stage
constexpr int X4RS = X4KT + 16;          // LDS row stride in bytes: 64 samples x 2 nibbles + pad

// 16 code bytes (re | im << 2 each) -> the re-planar e2m1 operands of S, U, T
__device__ __forceinline__ void expand16_e2m1(uint4 w, uint4& vs, uint4& vu, uint4& vt)
{
    constexpr uint32_t N = 0x0F0F0F0Fu;
    // P: byte k = code of sample k | code of sample k + 4 << 4 (upper code bits ignored)
    const uint32_t P[2] = {(w.x & N) | ((w.y << 4) & ~N), (w.z & N) | ((w.w << 4) & ~N)};
    uint32_t sr[2], ur[2], tr[2], si[2], ui[2], ti[2];
#pragma unroll
    for (int g = 0; g < 2; g++) {
        const uint32_t p = P[g];
        const uint32_t eq = ~(p ^ (p >> 1)) & 0x55555555u;       // bit 0: re lo == hi, bit 2: im lo == hi
        const uint32_t er = eq & 0x11111111u, ei = (eq >> 2) & 0x11111111u;
        sr[g] = ((~p & 0x22222222u) << 2) | 0x22222222u;         // sign = !re hi (bit 1 -> bit 3)
        si[g] = (~p & 0x88888888u) | 0x22222222u;                // sign = !im hi (bit 3)
        ur[g] = sr[g] + (er << 1);
        ui[g] = si[g] + (ei << 1);
        tr[g] = (er << 1) | ((er << 3) & sr[g]);
        ti[g] = (ei << 1) | ((ei << 3) & si[g]);
    }
    vs = make_uint4(sr[0], sr[1], si[0], si[1]);
    vu = make_uint4(ur[0], ur[1], ui[0], ui[1]);
    vt = make_uint4(tr[0], tr[1], ti[0], ti[1]);
}

__global__ __launch_bounds__(X8NT, 1) void k_xengine4(const uint8_t* __restrict__ X, int nchan, long nt, int ninp,
                                                      float* __restrict__ acc, size_t set_stride)
{
    __shared__ __attribute__((aligned(16))) uint8_t op[2][2][3][TS][X4RS];   // [buffer][slab][set][row][byte]

    const int nb = (ninp + TS - 1) / TS;
    const int ntl = nb * (nb + 1) / 2;
    const long lb = xcd_block((long)ntl * nchan);
    if (lb >= (long)ntl * nchan) return;
    const int c = (int)(lb / ntl);
    const int tile = (int)(lb - (long)c * ntl);
    int ti, tj;
    tile_coords(tile, nb, ti, tj);
    const bool diag = ti == tj;
    const int nslab = diag ? 1 : 2;
    const int tid = threadIdx.x;
    const int w = tid >> 6, lane = tid & 63;
    const int qa = w >> 1, qb = w & 1;
    // blocks (qa, 2 qb) and (qa, 2 qb + 1); in a diagonal tile the strictly lower ones are not stored,
    // and a wave with both strictly lower idles.  A wave with one lower block still computes both:
    // it shares its SIMD with an idle or a one-block wave, so the stage time is unchanged.
    const int mode = __builtin_amdgcn_readfirstlane(!(diag && qa > 2 * qb) ? 2 : !(diag && qa > 2 * qb + 1) ? 1 : 0);

    // staging: lane l of wave w moves the 16-sample chunk (row 16 w + l % 16, part l / 16) of each slab:
    // a quarter-wave stores one part of 16 rows (80-byte stride: on distinct LDS banks), and a wave
    // still loads 16 whole 64-byte row segments from HBM
    static_assert(X8NT == 512 && TS == 128, "8 waves x 16 rows cover the tile's rows");
    const int row = ((tid >> 6) << 4) | (tid & 15), part = (tid >> 4) & 3;
    const uint8_t* src[2];
    bool rok[2];
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const int r = (s == 0 ? ti : tj) * TS + row;
        rok[s] = s < nslab && r < ninp;
        src[s] = X + ((size_t)c * (size_t)ninp + (size_t)(rok[s] ? r : 0)) * (size_t)nt + part * 16;
    }

    fx_float16 are[3][2], aim[3][2];
#pragma unroll
    for (int set = 0; set < 3; set++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            are[set][b] = fx_float16{};
            aim[set][b] = fx_float16{};
        }

    uint4 pre[2];
    bool val[2];
    auto gload = [&](long t0) {
#pragma unroll
        for (int s = 0; s < 2; s++) {
            val[s] = rok[s] && t0 + part * 16 < nt;   // nt is a multiple of 16: chunks are whole
            pre[s] = val[s] ? *reinterpret_cast<const uint4*>(src[s] + t0) : make_uint4(0u, 0u, 0u, 0u);
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int s = 0; s < 2; s++) {
            if (s >= nslab) break;
            uint4* dst[3] = {reinterpret_cast<uint4*>(&op[buf][s][0][row][part * 16]),
                             reinterpret_cast<uint4*>(&op[buf][s][1][row][part * 16]),
                             reinterpret_cast<uint4*>(&op[buf][s][2][row][part * 16])};
            if (val[s]) {
                uint4 vs, vu, vt;
                expand16_e2m1(pre[s], vs, vu, vt);
                *dst[0] = vs;
                *dst[1] = vu;
                *dst[2] = vt;
            } else {
                // rows past the last input and samples past the end contribute zeros
                *dst[0] = *dst[1] = *dst[2] = make_uint4(0u, 0u, 0u, 0u);
            }
        }
    };

    const int sb = diag ? 0 : 1;
    const int r32 = lane & 31, h = lane >> 5;
    // the MFMAs of one stage
    auto stage = [&](int buf) {
#pragma unroll
        for (int ks = 0; ks < X4KT / 32; ks++) {
            const int kb = ks * 32 + h * 16;      // byte offset of this lane's 16 samples
#pragma unroll
            for (int set = 0; set < 3; set++) {
                const uint4 a = *reinterpret_cast<const uint4*>(&op[buf][0][set][qa * 32 + r32][kb]);
                const fx_int8 fa = {(int)a.x, (int)a.y, (int)a.z, (int)a.w, 0, 0, 0, 0};
                const fx_int8 fr = {(int)a.z, (int)a.w, (int)(a.x ^ 0x88888888u), (int)(a.y ^ 0x88888888u), 0, 0, 0, 0};
#pragma unroll
                for (int b = 0; b < 2; b++) {
                    const uint4 v =
                        *reinterpret_cast<const uint4*>(&op[buf][sb][set][(2 * qb + b) * 32 + r32][kb]);
                    const fx_int8 fb = {(int)v.x, (int)v.y, (int)v.z, (int)v.w, 0, 0, 0, 0};
                    are[set][b] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(fa, fb, are[set][b], 4, 4, 0, 127, 0,
                                                                                  127);
                    aim[set][b] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(fr, fb, aim[set][b], 4, 4, 0, 127, 0,
                                                                                  127);
                }
            }
        }
    };

    const long nstage = (nt + X4KT - 1) / X4KT;
    gload(0);
    lstore(0);
    if (nstage > 1) gload(X4KT);
    __syncthreads();
    for (long st = 0; st < nstage; st++) {
        const int buf = (int)(st & 1);
        if (mode) stage(buf);
        if (st + 1 < nstage) {
            // the other buffer was last read in stage st - 1, before the barrier that ended it
            lstore(buf ^ 1);
            if (st + 2 < nstage) gload((st + 2) * X4KT);
        }
        __syncthreads();
    }
    if (mode == 0) return;
#pragma unroll
    for (int b = 0; b < 2; b++) {
        if (b == 0 && mode == 1) continue;
        const int bi = qa, bj = 2 * qb + b;
        const int w4 = (bi >> 1) * 2 + (bj >> 1), s4 = (bi & 1) * 2 + (bj & 1);
#pragma unroll
        for (int set = 0; set < 3; set++) {
            float* d = acc + set * set_stride + (((size_t)c * ntl + tile) * 16 + w4 * 4 + s4) * (32 * 64);
#pragma unroll
            for (int r = 0; r < 16; r++) {
                d[r * 64 + lane] += are[set][b][r];
                d[(16 + r) * 64 + lane] += aim[set][b][r];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// dump: engine layout -> vis[p][c] = sum_s acc_s / (sc_c^2 s_i s_j), or for
// the exact 2-bit decomposition (three sets of exact integer sums A0, A1, A2)
// the visibility (2-H) A0 + (H-1) A1 + (H-1)(H-2) A2 rounded once to FP32.
// With H = h 2^-22 (h an integer: the FP32 level value) it is
//   K 2^-22 + L 2^-44,  K = e0 A0 + e1 A1 + e2 A2,  L = e3 A2,
// e = {2^23 - h, h - 2^22, hi and lo parts (c2 = e2 2^22 + e3) of
// (h - 2^22)(h - 2^23)}: K and L are exact 64-bit integers and exact doubles,
// their two-sum is the exact value as an unevaluated pair, and the FP32
// rounding of the pair is corrected where the high part sits on a midpoint.
__device__ __forceinline__ float round_pair_f32(double hi, double lo)
{
    float f = (float)hi;
    const double fd = (double)f;
    if (lo != 0.0 && fd != hi) {
        // g: the float on the other side of hi; a pair on the midpoint of f
        // and g goes to the side lo points to
        const float g = nextafterf(f, hi > fd ? INFINITY : -INFINITY);
        if (hi == 0.5 * (fd + (double)g) && ((lo > 0.0) == ((double)g > fd))) f = g;
    }
    return f;
}

__device__ __forceinline__ float exact2bit(const long long* __restrict__ e, float a0, float a1, float a2)
{
    const long long A0 = (long long)a0, A1 = (long long)a1, A2 = (long long)a2;
    const long long K = e[0] * A0 + e[1] * A1 + e[2] * A2;
    const long long L = e[3] * A2;
    const double x = ldexp((double)K, -22), y = ldexp((double)L, -44);
    const double s = x + y;                      // two-sum: s + err = x + y exactly
    const double bb = s - x;
    const double err = (x - (s - bb)) + (y - bb);
    return round_pair_f32(s, err);
}
// MFMA 32x32 output layout: lane l, register r holds row 8(r/4) + 4(l/32) + r%4,
// column l%32.
// ---------------------------------------------------------------------------
__global__ void k_xengine_dump(const float* __restrict__ acc, int nchan, int ninp, int ts, int nsets,
                               size_t set_stride, const long long* __restrict__ ecoef,
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
    const int nb = (ninp + ts - 1) / ts;
    const int ntl = nb * (nb + 1) / 2;
    const int ti = i / ts, tj = j / ts;
    const int tile = ti * nb - ti * (ti - 1) / 2 + (tj - ti);
    // block of (i, j) in the tile: wave (qi, qj) quadrant, sub-block (a, b)
    const int nbw = ts / 64;
    const int bi = (i % ts) >> 5, bj = (j % ts) >> 5;
    const int w = (bi / nbw) * 2 + (bj / nbw);
    const int blk = w * nbw * nbw + (bi % nbw) * nbw + (bj % nbw);
    const int ii = i & 31, jj = j & 31;
    const int r = 4 * (ii >> 3) + (ii & 3);
    const int l = jj + 32 * ((ii >> 2) & 1);
    const size_t off = (((size_t)c * ntl + tile) * (size_t)((ts / 32) * (ts / 32)) + blk) * (32 * 64);
    float re = 0.f, im = 0.f;
    if (ecoef) {
        const float* a0 = acc + off;
        const float* a1 = a0 + set_stride;
        const float* a2 = a1 + set_stride;
        re = exact2bit(ecoef, a0[r * 64 + l], a1[r * 64 + l], a2[r * 64 + l]);
        im = exact2bit(ecoef, a0[(16 + r) * 64 + l], a1[(16 + r) * 64 + l], a2[(16 + r) * 64 + l]);
    } else {
        for (int q = 0; q < nsets; q++) {
            const float* a = acc + q * set_stride + off;
            re += a[r * 64 + l];
            im += a[(16 + r) * 64 + l];
        }
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

size_t fxk_xengine_acc_floats(int nchan, int ninp, int nsets, int ts)
{
    return (size_t)nsets * (size_t)nchan * (size_t)ntiles_ts(ninp, ts) * (size_t)((ts / 32) * (ts / 32)) * 32u * 64u;
}

// FX_XENGINE_RING=0 selects the register-prefetch engine for every shape
// (the default of a new correlator; fx_corr_set_xengine overrides it)
bool fxk_xengine_ring_default()
{
    static const bool on = [] {
        const char* e = getenv("FX_XENGINE_RING");
        return !(e && e[0] == '0');
    }();
    return on;
}

namespace {

enum XKind { XK_NONE, XK_RING2_F16, XK_RING2_F32, XK_RING_F16, XK_RING_F32, XK_F32_128, XK_F32_64, XK_F16_128,
             XK_F16_64 };

XKind xengine_select(int in_type, int ts, const void* X, long nt, int ninp, bool ring)
{
    const bool aligned = ((uintptr_t)X % 16) == 0;
    // whole 128-input tiles of more inputs and whole stages: the two-slab LDS ring
    if (ring && ts == 128 && ninp > 128 && ninp % 128 == 0 && aligned) {
        if (in_type == FX_PREC_F16 && nt % (XTraits<FX_PREC_F16>::KT / 2) == 0) return XK_RING2_F16;
        if (in_type == FX_PREC_F32 && nt % (XTraits<FX_PREC_F32>::KT / 2) == 0) return XK_RING2_F32;
    }
    // one full tile and whole stages: the LDS-ring engine
    if (ring && ts == 128 && ninp == 128 && aligned) {
        if (in_type == FX_PREC_F16 && nt % XTraits<FX_PREC_F16>::KT == 0) return XK_RING_F16;
        if (in_type == FX_PREC_F32 && nt % XTraits<FX_PREC_F32>::KT == 0) return XK_RING_F32;
    }
    if (in_type == FX_PREC_F32 && ts == 128) return XK_F32_128;
    if (in_type == FX_PREC_F32 && ts == 64) return XK_F32_64;
    if (in_type == FX_PREC_F16 && ts == 128) return XK_F16_128;
    if (in_type == FX_PREC_F16 && ts == 64) return XK_F16_64;
    return XK_NONE;
}

// Lockstep groups per XCD for the large-input ring engine (0: run it
// ungrouped): every group of ntl workgroups must be resident at once, so the
// groups of an XCD fill at most its CUs times the engine's workgroups per CU;
// a round is XLP whole stages.
template <int PREC>
int ring2_groups_per_xcd(int ntl, int nchan, long nt)
{
    static const int slots = [] {
        int dev = 0, ncu = 0, occ = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_xengine_ring2<PREC, true>, XNT, 0) != hipSuccess) {
            (void)hipGetLastError();
            return 0;
        }
        return ncu / 8 * occ;
    }();
    if (ntl < 2 || ntl > 64 || nchan < 8 || nt % ((long)XTraits<PREC>::KT / 2 * XLP) != 0) return 0;
    return std::min(slots / ntl, nchan / 8);
}

int ring2_groups(XKind kind, int ninp, int ts, int nchan, long nt, bool lock)
{
    if (!lock || (kind != XK_RING2_F16 && kind != XK_RING2_F32)) return 0;
    const int ntl = ntiles_ts(ninp, ts);
    return kind == XK_RING2_F16 ? ring2_groups_per_xcd<FX_PREC_F16>(ntl, nchan, nt)
                                : ring2_groups_per_xcd<FX_PREC_F32>(ntl, nchan, nt);
}

} // namespace

// FX_XENGINE_LOCK=1 runs the large-input ring engine in lockstep groups (the
// default of a new correlator; fx_corr_set_xengine overrides it)
bool fxk_xengine_lock_default()
{
    static const bool on = [] {
        const char* e = getenv("FX_XENGINE_LOCK");
        return e && e[0] == '1';
    }();
    return on;
}

int fxk_xengine_prog_words() { return 8 * 64; }

// stages a lockstep sibling may trail (FX_XENGINE_LAG overrides the 8)
static unsigned xlag()
{
    static const unsigned v = [] {
        const char* e = getenv("FX_XENGINE_LAG");
        return e && atoi(e) > 0 ? (unsigned)atoi(e) : 8u;
    }();
    return v;
}

const char* fxk_xengine_name(int in_type, int ts, const void* X, int nchan, long nt, int ninp, bool ring,
                             bool lock)
{
    const XKind kind = xengine_select(in_type, ts, X, nt, ninp, ring);
    const bool grouped = ring2_groups(kind, ninp, ts, nchan, nt, lock) > 0;
    switch (kind) {
    case XK_RING2_F16: return grouped ? "k_xengine_ring2<F16> (lockstep)" : "k_xengine_ring2<F16>";
    case XK_RING2_F32: return grouped ? "k_xengine_ring2<F32> (lockstep)" : "k_xengine_ring2<F32>";
    case XK_RING_F16: return "k_xengine_ring<F16>";
    case XK_RING_F32: return "k_xengine_ring<F32>";
    case XK_F32_128: return "k_xengine<F32, 128>";
    case XK_F32_64: return "k_xengine<F32, 64>";
    case XK_F16_128: return "k_xengine<F16, 128>";
    case XK_F16_64: return "k_xengine<F16, 64>";
    default: return "";
    }
}

int fxk_xengine(hipStream_t s, int in_type, int ts, const void* X, int nchan, long nt, int ninp, float* acc,
                bool ring, unsigned* prog)
{
    const long nblk = xcd_grid((long)ntiles_ts(ninp, ts) * nchan);
    const dim3 g((unsigned)nblk), b(XNT);
    const uint32_t* xh = reinterpret_cast<const uint32_t*>(X);
    const float2* xf = reinterpret_cast<const float2*>(X);
    const XKind kind = xengine_select(in_type, ts, X, nt, ninp, ring);
    const int ngx = ring2_groups(kind, ninp, ts, nchan, nt, prog != nullptr);
    if (ngx > 0) {
        const int ntl = ntiles_ts(ninp, ts);
        const int rc = fxk_fill32(s, prog, 0u, (size_t)8 * ngx * ntl);
        if (rc) return rc;
        const dim3 gl((unsigned)(8 * ngx * ntl));
        if (kind == XK_RING2_F16)
            hipLaunchKernelGGL((k_xengine_ring2<FX_PREC_F16, true>), gl, b, 0, s, xh, nchan, nt, ninp, acc, prog, ngx,
                               xlag());
        else
            hipLaunchKernelGGL((k_xengine_ring2<FX_PREC_F32, true>), gl, b, 0, s, xf, nchan, nt, ninp, acc, prog, ngx,
                               xlag());
        return hipGetLastError() == hipSuccess ? FX_OK : FX_ERR_HIP;
    }
    switch (kind) {
    case XK_RING2_F16:
        hipLaunchKernelGGL((k_xengine_ring2<FX_PREC_F16, false>), g, b, 0, s, xh, nchan, nt, ninp, acc, nullptr, 0,
                           0u);
        break;
    case XK_RING2_F32:
        hipLaunchKernelGGL((k_xengine_ring2<FX_PREC_F32, false>), g, b, 0, s, xf, nchan, nt, ninp, acc, nullptr, 0,
                           0u);
        break;
    case XK_RING_F16: hipLaunchKernelGGL(k_xengine_ring<FX_PREC_F16>, g, b, 0, s, xh, nchan, nt, acc); break;
    case XK_RING_F32: hipLaunchKernelGGL(k_xengine_ring<FX_PREC_F32>, g, b, 0, s, xf, nchan, nt, acc); break;
    case XK_F32_128: hipLaunchKernelGGL((k_xengine<FX_PREC_F32, 128>), g, b, 0, s, xf, nchan, nt, ninp, acc); break;
    case XK_F32_64: hipLaunchKernelGGL((k_xengine<FX_PREC_F32, 64>), g, b, 0, s, xf, nchan, nt, ninp, acc); break;
    case XK_F16_128: hipLaunchKernelGGL((k_xengine<FX_PREC_F16, 128>), g, b, 0, s, xh, nchan, nt, ninp, acc); break;
    case XK_F16_64: hipLaunchKernelGGL((k_xengine<FX_PREC_F16, 64>), g, b, 0, s, xh, nchan, nt, ninp, acc); break;
    default: return FX_ERR_UNSUPPORTED;
    }
    return hipGetLastError() == hipSuccess ? FX_OK : FX_ERR_HIP;
}

int fxk_xengine8(hipStream_t s, int prec, const void* X, int nchan, long nt, int ninp, float* acc)
{
    long nblk = xcd_grid((long)ntiles128(ninp) * nchan);
    const size_t stride = fxk_xengine_acc_floats(nchan, ninp, 1, TS);
    const uint8_t* x = reinterpret_cast<const uint8_t*>(X);
    if (prec == FX_PREC_FP8) {
        if (nt % 8) return FX_ERR_ARG;
        hipLaunchKernelGGL(k_xengine8<1>, dim3((unsigned)nblk), dim3(X8NT), 0, s, x, nchan, nt, ninp, acc, stride);
    } else if (prec == FX_PREC_FP8_2BIT) {
        if (nt % 16) return FX_ERR_ARG;
        hipLaunchKernelGGL(k_xengine8<3>, dim3((unsigned)nblk), dim3(X8NT), 0, s, x, nchan, nt, ninp, acc, stride);
    } else if (prec == FX_PREC_FP4_2BIT) {
        if (nt % 16) return FX_ERR_ARG;
        hipLaunchKernelGGL(k_xengine4, dim3((unsigned)nblk), dim3(X8NT), 0, s, x, nchan, nt, ninp, acc, stride);
    } else {
        return FX_ERR_UNSUPPORTED;
    }
    return hipGetLastError() == hipSuccess ? FX_OK : FX_ERR_HIP;
}

int fxk_xengine_dump_pij(hipStream_t s, const float* acc, int nchan, int ninp, int ts, int nsets,
                         const long long* exact_coef, const float* scale_chan, const float* scale_inp,
                         const int2* pij, int accumulate, float2* vis)
{
    long nprod = (long)ninp * (ninp + 1) / 2;
    long total = nprod * nchan;
    size_t set_stride = fxk_xengine_acc_floats(nchan, ninp, 1, ts);
    hipLaunchKernelGGL(k_xengine_dump, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, acc,
                       nchan, ninp, ts, nsets, set_stride, exact_coef, scale_chan, scale_inp, pij, nprod,
                       accumulate, vis);
    return hipGetLastError() == hipSuccess ? FX_OK : FX_ERR_HIP;
}

int fxk_mfma_probe(hipStream_t s, int which, float* out)
{
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, s, which, out);
    return hipGetLastError() == hipSuccess ? FX_OK : FX_ERR_HIP;
}
