// fx_frames.hip -- the real-file front end on the GPU: VDIF and Mark5B
// frames in HBM -> validated, header-stripped payload placed per thread and
// time slot.
//
//   k_frames_place   one workgroup per frame: header decode (format_vdif.c:
//                    135-206, mark5_format_mark5b.c:134-186), validity
//                    (format_vdif.c:2299-2345 invalid bit; blanker_vdif and
//                    blanker_mark5, blanker_mark5.c:54-200,282-320: Mark5
//                    fill pattern at the ends of the payload), placement of
//                    the payload words by (second, frame number, thread).
//
// The payload of a 2-bit VDIF frame is already this library's packed sample
// format (16 samples per little-endian word, sample j at bits 2j+1:2j; nchan
// channels interleaved per time step), and a Mark5B payload is after one bit
// swap per sample, so the placed rows feed the correlator (one row per input)
// and the spectrometer directly.
#include "fx_internal.h"

#include <algorithm>

namespace {

constexpr uint32_t kMark5bSync = 0xABADDEEDu;
constexpr unsigned long long kFill64 = 0x1122334411223344ull;
constexpr uint32_t kFill32 = 0x11223344u;

// BCD digits d7..d0 of a word, most significant nibble first
__device__ __forceinline__ int bcd(uint32_t w, int first, int count)
{
    int v = 0;
    for (int i = first; i < first + count; i++) v = 10 * v + (int)((w >> (28 - 4 * i)) & 0xFu);
    return v;
}

// Start of a trailing fill run, in bytes, of a payload of n 64-bit words
// whose first word is data and last word fill (findfirstinvalid,
// blanker_mark5.c:54-85).  The mark5access search starts at n/2 with steps
// n/4, n/8, ..., 1 and never probes the index its last step lands on, so it
// can report the run up to a few 64-bit words late (up to 36 bytes for a
// Mark5B payload); exact = true bisects to the first fill word instead.  Both
// then move back 4 bytes when the 32-bit word before the run is fill.
__device__ int first_invalid_byte(const unsigned long long* d, int n, bool exact)
{
    const uint32_t* d32 = reinterpret_cast<const uint32_t*>(d);
    int r;
    if (exact) {
        int lo = 0, hi = n - 1;       // d[lo] data, d[hi] fill
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (d[mid] == kFill64) hi = mid;
            else lo = mid;
        }
        r = hi;
    } else {
        int i = n / 2;
        r = n - 1;
        for (int s = n / 4; s > 0; s >>= 1) {
            if (d[i] == kFill64) {
                r = i;
                i -= s;
            } else {
                i += s;
            }
        }
    }
    r *= 8;
    if (r >= 4 && d32[r / 4 - 1] == kFill32) r -= 4;
    return r;
}

// End of a leading fill run, in bytes, of a payload whose first 64-bit word
// is fill and last word data (findfirstvalid, blanker_mark5.c:87-118).  The
// same search as above for the first data word; then 4 bytes more when that
// word's first 32-bit half is still fill.  mark5access tests the low half of
// the word 8 bytes before instead (d32[r/4 - 2]), a fill word whenever the
// search is exact, so its boundary then lands 4 bytes late; the reference
// form keeps that test, exact = true tests d32[r/4].
__device__ int first_valid_byte(const unsigned long long* d, int n, bool exact)
{
    const uint32_t* d32 = reinterpret_cast<const uint32_t*>(d);
    int r;
    if (exact) {
        int lo = 0, hi = n - 1;       // d[lo] fill, d[hi] data
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (d[mid] != kFill64) hi = mid;
            else lo = mid;
        }
        r = hi;
    } else {
        int i = n / 2;
        r = n - 1;
        for (int s = n / 4; s > 0; s >>= 1) {
            if (d[i] != kFill64) {
                r = i;
                i -= s;
            } else {
                i += s;
            }
        }
    }
    r *= 8;
    if (exact) {
        if (d32[r / 4] == kFill32) r += 4;
    } else if (r >= 8 && d32[r / 4 - 2] == kFill32) {
        r += 4;
    }
    return r;
}

// Mark5B 2-bit samples hold the sign in the low and the magnitude in the high
// bit of each pair (codes -H, +1, -1, +H; mark5_format_mark5b.c:59-96);
// swapping the two bits gives this library's codes -H, -1, +1, +H.  Channel
// interleave and sample order are the same.
__device__ __forceinline__ uint32_t swap_pairs(uint32_t v)
{
    return ((v & 0x55555555u) << 1) | ((v >> 1) & 0x55555555u);
}
__device__ __forceinline__ uint2 swap_pairs(uint2 v) { return make_uint2(swap_pairs(v.x), swap_pairs(v.y)); }
__device__ __forceinline__ uint4 swap_pairs(uint4 v)
{
    return make_uint4(swap_pairs(v.x), swap_pairs(v.y), swap_pairs(v.z), swap_pairs(v.w));
}

struct Placement {
    long dst;           // word offset of the slot in out, or -1
    long vslot;         // index of the slot's range in valid
    int vs, ve;         // valid word range of the payload
    int outcome;        // counter index: 0 placed, 2 out of range, 3 unmapped
};

// Header decode, slot and blanking of one frame (wave-uniform).
__device__ Placement decide(const uint8_t* f, int pbytes, int payload_offset, int format,
                            const int* __restrict__ thread_row, int map_size, int nrows, long sec0,
                            long frame0, int fps, long nslots, size_t out_stride, bool exact)
{
    const uint32_t* w = reinterpret_cast<const uint32_t*>(f);
    long sec, fnum;
    int thread;
    bool ok;
    if (format == FX_FMT_VDIF) {
        sec = (long)(w[0] & 0x3FFFFFFFu);
        fnum = (long)(w[1] & 0x00FFFFFFu);
        thread = (int)((w[3] >> 16) & 0x3FFu);
        ok = ((w[0] >> 31) & 1u) == 0u;
    } else {
        // Mark5B: day (MJD mod 1000) and second of day in BCD, 15-bit frame number
        sec = (long)bcd(w[2], 0, 3) * 86400L + (long)bcd(w[2], 3, 5);
        fnum = (long)(w[1] & 0x7FFFu);
        thread = 0;
        ok = w[0] == kMark5bSync;
    }
    int row = thread < map_size ? thread_row[thread] : -1;
    if (row >= nrows) row = -1;
    const long slot = (sec - sec0) * (long)fps + fnum - frame0;
    Placement pl{-1, 0, 0, 0, 3};
    if (row < 0) return pl;
    if (slot < 0 || slot >= nslots) {
        pl.outcome = 2;
        return pl;
    }
    const int pwords = pbytes / 4;
    pl.outcome = 0;
    pl.dst = (long)row * (long)out_stride + slot * (long)pwords;
    pl.vslot = ((long)row * nslots + slot) * 2;
    const unsigned long long* d = reinterpret_cast<const unsigned long long*>(f + payload_offset);
    const int n64 = pbytes / 8;
    const bool start_ok = d[0] != kFill64, end_ok = d[n64 - 1] != kFill64;
    pl.ve = pwords;
    if (!ok) {
        pl.ve = 0;
    } else if (format == FX_FMT_VDIF) {
        // one zone: a packet is good or bad
        if (!start_ok || !end_ok) pl.ve = 0;
    } else if (!start_ok && !end_ok) {
        pl.ve = 0;
    } else if (!end_ok) {
        pl.ve = first_invalid_byte(d, n64, exact) / 4;
    } else if (!start_ok) {
        pl.vs = first_valid_byte(d, n64, exact) / 4;
    }
    return pl;
}

// One wavefront per frame, frames strided over the grid's waves.  The first
// K vectors of each lane's share of the payload are loaded before the header
// is decoded, so the payload and header loads are in flight together; the
// counters are summed per wave and added once at the end.
template <typename V, int K>
__global__ __launch_bounds__(256) void k_frames_place(const uint8_t* __restrict__ raw, long nframes,
                                                      int framebytes, int payload_offset, int format,
                                                      const int* __restrict__ thread_row, int map_size,
                                                      int nrows, long sec0, long frame0, int fps, long nslots,
                                                      uint32_t* __restrict__ out, size_t out_stride,
                                                      uint32_t* __restrict__ valid,
                                                      unsigned long long* __restrict__ counts, bool exact)
{
    constexpr int VB = (int)sizeof(V);
    const int lane = threadIdx.x & 63;
    const long wave = (long)blockIdx.x * (blockDim.x >> 6) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long nwaves = (long)gridDim.x * (blockDim.x >> 6);
    const int pbytes = framebytes - payload_offset;
    const int nv = pbytes / VB;
    // placed, placed but empty, out of range, unmapped
    unsigned long long c0 = 0, c1 = 0, c2 = 0, c3 = 0;
    for (long fr = wave; fr < nframes; fr += nwaves) {
        const uint8_t* f = raw + (size_t)fr * framebytes;
        const V* pay = reinterpret_cast<const V*>(f + payload_offset);
        V r[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int i = k * 64 + lane;
            if (i < nv) r[k] = pay[i];
        }
        const Placement pl = decide(f, pbytes, payload_offset, format, thread_row, map_size, nrows, sec0, frame0,
                                    fps, nslots, out_stride, exact);
        c0 += pl.outcome == 0;
        c2 += pl.outcome == 2;
        c3 += pl.outcome == 3;
        if (pl.dst < 0) continue;
        c1 += pl.ve <= pl.vs;
        if (lane == 0) {
            valid[pl.vslot] = (uint32_t)pl.vs;
            valid[pl.vslot + 1] = (uint32_t)pl.ve;
        }
        V* o = reinterpret_cast<V*>(out + pl.dst);
        const bool m5b = format != FX_FMT_VDIF;
        for (int base = 0; base < nv; base += 64 * K) {
            if (base > 0) {
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const int i = base + k * 64 + lane;
                    if (i < nv) r[k] = pay[i];
                }
            }
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int i = base + k * 64 + lane;
                if (i < nv) o[i] = m5b ? swap_pairs(r[k]) : r[k];
            }
        }
    }
    if (lane == 0) {
        if (c0) atomicAdd(&counts[0], c0);
        if (c1) atomicAdd(&counts[1], c1);
        if (c2) atomicAdd(&counts[2], c2);
        if (c3) atomicAdd(&counts[3], c3);
    }
}

__global__ __launch_bounds__(256) void k_valid_words(const uint32_t* __restrict__ valid, long nslots, int pwords,
                                                     uint8_t* __restrict__ flags)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nslots * (size_t)pwords) return;
    const size_t slot = i / (size_t)pwords;
    const uint32_t j = (uint32_t)(i - slot * (size_t)pwords);
    flags[i] = (j >= valid[2 * slot] && j < valid[2 * slot + 1]) ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_any_zero(const uint8_t* __restrict__ flags, size_t n, int* __restrict__ any)
{
    bool z = false;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        z = z || flags[i] == 0;
    if (__any(z) && (threadIdx.x & 63) == 0) any[0] = 1;
}

} // namespace

int fxk_valid_words(hipStream_t s, const uint32_t* valid, long nslots, int pwords, uint8_t* flags)
{
    const size_t n = (size_t)nslots * (size_t)pwords;
    if (n == 0) return FX_OK;
    hipLaunchKernelGGL(k_valid_words, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, valid, nslots, pwords,
                       flags);
    return hipGetLastError() == hipSuccess ? FX_OK : FX_ERR_HIP;
}

int fxk_any_zero(hipStream_t s, const uint8_t* flags, size_t n, int* any)
{
    if (n == 0) return FX_OK;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_any_zero, dim3((unsigned)blocks), dim3(256), 0, s, flags, n, any);
    return hipGetLastError() == hipSuccess ? FX_OK : FX_ERR_HIP;
}

int fxk_frames_place(hipStream_t s, int format, const uint8_t* raw, long nframes, int framebytes,
                     int payload_offset, const int* thread_row, int map_size, int nrows, long sec0, long frame0,
                     int fps, long nslots, uint32_t* out, size_t out_stride_words, uint32_t* valid,
                     unsigned long long* counts, bool exact)
{
    if (nframes <= 0) return FX_OK;
    // 16-byte vectors when frames, payloads and slots are 16-byte aligned
    // (8032-byte VDIF and 10016-byte Mark5B frames are), else 8-byte ones
    const int pwords = (framebytes - payload_offset) / 4;
    const bool v16 = (uintptr_t)raw % 16 == 0 && framebytes % 16 == 0 && payload_offset % 16 == 0 &&
                     (uintptr_t)out % 16 == 0 && out_stride_words % 4 == 0 && pwords % 4 == 0;
    // waves of 64 lanes, 4 per workgroup; 8 workgroups per CU on 256 CUs fill the chip
    const long wgs = std::min<long>((nframes + 3) / 4, 2048);
    if (v16) {
        hipLaunchKernelGGL((k_frames_place<uint4, 8>), dim3((unsigned)wgs), dim3(256), 0, s, raw, nframes,
                           framebytes, payload_offset, format, thread_row, map_size, nrows, sec0, frame0, fps,
                           nslots, out, out_stride_words, valid, counts, exact);
    } else {
        hipLaunchKernelGGL((k_frames_place<uint2, 8>), dim3((unsigned)wgs), dim3(256), 0, s, raw, nframes,
                           framebytes, payload_offset, format, thread_row, map_size, nrows, sec0, frame0, fps,
                           nslots, out, out_stride_words, valid, counts, exact);
    }
    return hipGetLastError() == hipSuccess ? FX_OK : FX_ERR_HIP;
}
