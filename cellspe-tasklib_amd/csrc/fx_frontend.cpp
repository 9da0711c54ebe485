// fx_frontend.cpp -- host side of the real-file front end: VDIF and Mark5B
// header decoding, the frame placement entry point over k_frames_place, and
// the double-buffered file-to-HBM ingest stream.
#include "fx_internal.h"

#include <fcntl.h>
#include <unistd.h>

#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace {

// Modified Julian Date of a Gregorian calendar date (Fliegel & Van Flandern
// day number, minus 2400001)
int mjd_of(int y, int m, int d)
{
    const int a = (14 - m) / 12;
    const int yy = y + 4800 - a;
    const int mm = m + 12 * a - 3;
    const long jdn = d + (153 * mm + 2) / 5 + 365L * yy + yy / 4 - yy / 100 + yy / 400 - 32045;
    return (int)(jdn - 2400001);
}

uint32_t le32(const unsigned char* b) { return (uint32_t)b[0] | (uint32_t)b[1] << 8 | (uint32_t)b[2] << 16 | (uint32_t)b[3] << 24; }

int bcd_digits(uint32_t w, int first, int count)
{
    int v = 0;
    for (int i = first; i < first + count; i++) v = 10 * v + (int)((w >> (28 - 4 * i)) & 0xFu);
    return v;
}

} // namespace

extern "C" int fx_vdif_decode(const void* header, fx_vdif_header* h)
{
    FX_REQUIRE(header && h, FX_ERR_ARG, "fx_vdif_decode: null argument");
    const unsigned char* b = static_cast<const unsigned char*>(header);
    const uint32_t w0 = le32(b), w1 = le32(b + 4), w2 = le32(b + 8), w3 = le32(b + 12);
    h->seconds = w0 & 0x3FFFFFFFu;
    h->legacy = (int)((w0 >> 30) & 1u);
    h->invalid = (int)((w0 >> 31) & 1u);
    h->frame = w1 & 0x00FFFFFFu;
    h->epoch = (int)((w1 >> 24) & 0x3Fu);
    h->framebytes = (int)(w2 & 0x00FFFFFFu) * 8;
    h->headerbytes = h->legacy ? 16 : 32;
    h->nchan = 1u << ((w2 >> 24) & 0x1Fu);
    h->version = (int)((w2 >> 29) & 0x7u);
    h->station = (int)(w3 & 0xFFFFu);
    h->thread = (int)((w3 >> 16) & 0x3FFu);
    h->nbits = (int)((w3 >> 26) & 0x1Fu) + 1;
    h->is_complex = (int)((w3 >> 31) & 1u);
    // reference epoch e: 1 January (e even) or 1 July (e odd) of 2000 + e/2
    const int epoch_mjd = mjd_of(2000 + h->epoch / 2, (h->epoch & 1) ? 7 : 1, 1);
    h->mjd = epoch_mjd + (int)(h->seconds / 86400u);
    h->second_of_day = (int)(h->seconds % 86400u);
    return FX_OK;
}

extern "C" int fx_mark5b_decode(const void* header, fx_mark5b_header* h)
{
    FX_REQUIRE(header && h, FX_ERR_ARG, "fx_mark5b_decode: null argument");
    const unsigned char* b = static_cast<const unsigned char*>(header);
    const uint32_t w0 = le32(b), w1 = le32(b + 4), w2 = le32(b + 8), w3 = le32(b + 12);
    h->sync_ok = w0 == 0xABADDEEDu;
    h->frame = (int)(w1 & 0x7FFFu);
    h->tvg = (int)((w1 >> 15) & 1u);
    h->day = bcd_digits(w2, 0, 3);
    h->second_of_day = bcd_digits(w2, 3, 5);
    // top four BCD digits of word 3: 0.1 ms units, "unrounded" up to the
    // 156.25 us grid of the frames (mark5_format_mark5b.c:176-182)
    const long ns = (long)bcd_digits(w3, 0, 4) * 100000L;
    h->frac_ns = (double)(156250L * ((ns + 156249L) / 156250L));
    return FX_OK;
}

extern "C" int fx_frames_place(fx_ctx* ctx, int format, const void* d_raw, size_t nframes, int framebytes,
                               int payload_offset, const int* d_thread_row, int map_size, int nrows, long sec0,
                               long frame0, int frames_per_second, long nslots, uint32_t* d_out,
                               size_t out_stride_words, uint32_t* d_valid, uint64_t* d_counts, int flags)
{
    FX_REQUIRE(ctx && d_raw && d_thread_row && d_out && d_valid && d_counts, FX_ERR_ARG,
               "fx_frames_place: null argument");
    FX_HIP_TRY(hipSetDevice(fx_ctx_device_of(ctx)));
    FX_REQUIRE(format == FX_FMT_VDIF || format == FX_FMT_MARK5B, FX_ERR_ARG, "fx_frames_place: format %d", format);
    FX_REQUIRE(format != FX_FMT_MARK5B || (framebytes == 10016 && payload_offset == 16), FX_ERR_ARG,
               "fx_frames_place: Mark5B frames are 10016 bytes with a 16-byte header");
    FX_REQUIRE(format != FX_FMT_VDIF || ((payload_offset == 16 || payload_offset == 32) && framebytes % 8 == 0 &&
                                        framebytes > payload_offset),
               FX_ERR_ARG, "fx_frames_place: VDIF frames of %d bytes with a %d-byte header", framebytes,
               payload_offset);
    FX_REQUIRE(((uintptr_t)d_raw % 8) == 0, FX_ERR_ARG, "fx_frames_place: frames must be 8-byte aligned");
    FX_REQUIRE(map_size > 0 && map_size <= 1024 && nrows > 0 && nslots > 0 && frames_per_second > 0,
               FX_ERR_ARG, "fx_frames_place: map_size %d, nrows %d, nslots %ld, fps %d", map_size, nrows,
               nslots, frames_per_second);
    const size_t pwords = (size_t)(framebytes - payload_offset) / 4;
    FX_REQUIRE(out_stride_words % 2 == 0 && ((uintptr_t)d_out % 8) == 0, FX_ERR_ARG,
               "fx_frames_place: rows must be 8-byte aligned (even out_stride_words)");
    FX_REQUIRE(out_stride_words >= pwords * (size_t)nslots, FX_ERR_ARG,
               "fx_frames_place: rows of %zu words hold fewer than %ld slots of %zu", out_stride_words, nslots,
               pwords);
    FX_REQUIRE((flags & ~(FX_PLACE_CLEAR | FX_PLACE_EXACT_BLANK)) == 0, FX_ERR_ARG,
               "fx_frames_place: unknown flags 0x%x", flags);
    FX_REQUIRE(nframes <= 0x7FFFFFFFu, FX_ERR_ARG, "fx_frames_place: %zu frames in one call", nframes);
    hipStream_t s = fx_ctx_hip_stream(ctx);
    if (flags & FX_PLACE_CLEAR) {
        if (fxk_fill32(s, d_valid, 0u, 2 * (size_t)nrows * (size_t)nslots) != FX_OK ||
            fxk_fill32(s, d_counts, 0u, 8) != FX_OK) {
            fx_set_error("fx_frames_place: clear failed");
            return FX_ERR_HIP;
        }
    }
    if (nframes == 0) return FX_OK;
    if (fxk_frames_place(s, format, static_cast<const uint8_t*>(d_raw), (long)nframes, framebytes, payload_offset,
                         d_thread_row, map_size, nrows, sec0, frame0, frames_per_second, nslots, d_out,
                         out_stride_words,
                         d_valid, reinterpret_cast<unsigned long long*>(d_counts),
                         (flags & FX_PLACE_EXACT_BLANK) != 0) != FX_OK) {
        fx_set_error("fx_frames_place: launch failed: %s", hipGetErrorString(hipGetLastError()));
        return FX_ERR_HIP;
    }
    return FX_OK;
}

// ---------------------------------------------------------------------------
// double-buffered ingest
// ---------------------------------------------------------------------------
struct fx_ingest {
    fx_ctx* ctx = nullptr;
    int device = 0;
    int fd = -1;
    size_t offset = 0, chunk = 0, fsize = 0;
    int nbuf = 2;
    hipStream_t cs = nullptr;               // copy stream
    std::vector<void*> host, dev;
    std::vector<hipEvent_t> copied, consumed;
    std::vector<size_t> bytes;
    long nchunks = 0;                       // chunks in the file
    long ready = 0;                         // chunks whose copy is enqueued
    long given = 0;                         // chunks handed to the consumer
    bool quit = false, failed = false;
    char err[256] = "";
    std::mutex m;
    std::condition_variable cv;
    std::thread reader;
};

namespace {

void ingest_reader(fx_ingest* g)
{
    hipSetDevice(g->device);
    for (long k = 0; k < g->nchunks; k++) {
        const int b = (int)(k % g->nbuf);
        {
            // buffer b held chunk k - nbuf.  Its consumed event is recorded
            // when the consumer asks for chunk k - nbuf + 1, so the copy may
            // be queued once given >= k - nbuf + 2.
            std::unique_lock<std::mutex> lk(g->m);
            g->cv.wait(lk, [&] { return g->quit || k - g->given <= g->nbuf - 2; });
            if (g->quit) return;
        }
        hipEventSynchronize(g->copied[(size_t)b]);   // the host buffer's previous copy is done
        const size_t pos = g->offset + (size_t)k * g->chunk;
        const size_t want = std::min(g->chunk, g->fsize - pos);
        size_t got = 0;
        while (got < want) {
            const ssize_t r = pread(g->fd, static_cast<char*>(g->host[(size_t)b]) + got, want - got,
                                    (off_t)(pos + got));
            if (r <= 0) break;
            got += (size_t)r;
        }
        bool ok = got == want;
        if (ok) ok = hipStreamWaitEvent(g->cs, g->consumed[(size_t)b], 0) == hipSuccess &&
                     hipMemcpyAsync(g->dev[(size_t)b], g->host[(size_t)b], want, hipMemcpyHostToDevice, g->cs) ==
                         hipSuccess &&
                     hipEventRecord(g->copied[(size_t)b], g->cs) == hipSuccess;
        std::lock_guard<std::mutex> lk(g->m);
        if (!ok) {
            g->failed = true;
            snprintf(g->err, sizeof(g->err), "fx_ingest: chunk %ld: %s", k,
                     got != want ? "short read" : "copy failed");
            g->cv.notify_all();
            return;
        }
        g->bytes[(size_t)b] = want;
        g->ready = k + 1;
        g->cv.notify_all();
    }
}

void ingest_free(fx_ingest* g)
{
    for (void* p : g->host)
        if (p) hipHostFree(p);
    for (void* p : g->dev)
        if (p) hipFree(p);
    for (hipEvent_t e : g->copied)
        if (e) hipEventDestroy(e);
    for (hipEvent_t e : g->consumed)
        if (e) hipEventDestroy(e);
    if (g->cs) hipStreamDestroy(g->cs);
    if (g->fd >= 0) close(g->fd);
    delete g;
}

} // namespace

extern "C" int fx_ingest_open(fx_ctx* ctx, const char* path, size_t offset, size_t chunk_bytes, int nbuf,
                              fx_ingest** out)
{
    FX_REQUIRE(ctx && path && out && chunk_bytes > 0 && nbuf >= 2 && nbuf <= 16, FX_ERR_ARG,
               "fx_ingest_open: bad argument");
    *out = nullptr;
    fx_ingest* g = new fx_ingest;
    g->ctx = ctx;
    g->device = fx_ctx_device_of(ctx);
    g->offset = offset;
    g->chunk = chunk_bytes;
    g->nbuf = nbuf;
    g->fd = open(path, O_RDONLY);
    if (g->fd < 0) {
        delete g;
        fx_set_error("fx_ingest_open: cannot open %s", path);
        return FX_ERR_ARG;
    }
    const off_t end = lseek(g->fd, 0, SEEK_END);
    g->fsize = end > 0 ? (size_t)end : 0;
    const size_t body = g->fsize > offset ? g->fsize - offset : 0;
    g->nchunks = (long)((body + chunk_bytes - 1) / chunk_bytes);
    if (hipSetDevice(g->device) != hipSuccess ||
        hipStreamCreateWithFlags(&g->cs, hipStreamNonBlocking) != hipSuccess) {
        ingest_free(g);
        fx_set_error("fx_ingest_open: stream creation failed");
        return FX_ERR_HIP;
    }
    g->host.assign((size_t)nbuf, nullptr);
    g->dev.assign((size_t)nbuf, nullptr);
    g->copied.assign((size_t)nbuf, nullptr);
    g->consumed.assign((size_t)nbuf, nullptr);
    g->bytes.assign((size_t)nbuf, 0);
    for (int b = 0; b < nbuf; b++) {
        if (hipHostMalloc(&g->host[(size_t)b], chunk_bytes, hipHostMallocDefault) != hipSuccess ||
            hipMalloc(&g->dev[(size_t)b], chunk_bytes) != hipSuccess ||
            hipEventCreateWithFlags(&g->copied[(size_t)b], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&g->consumed[(size_t)b], hipEventDisableTiming) != hipSuccess) {
            hipGetLastError();
            ingest_free(g);
            fx_set_error("fx_ingest_open: buffer allocation failed (%d x %zu bytes)", nbuf, chunk_bytes);
            return FX_ERR_NOMEM;
        }
    }
    g->reader = std::thread(ingest_reader, g);
    *out = g;
    return FX_OK;
}

extern "C" int fx_ingest_next(fx_ingest* g, const void** d_chunk, size_t* bytes)
{
    FX_REQUIRE(g && d_chunk && bytes, FX_ERR_ARG, "fx_ingest_next: null argument");
    FX_HIP_TRY(hipSetDevice(g->device));
    hipStream_t s = fx_ctx_hip_stream(g->ctx);
    std::unique_lock<std::mutex> lk(g->m);
    const long k = g->given;
    if (k > 0) {
        // the previous chunk's buffer may be refilled after the work queued on it
        FX_HIP_TRY(hipEventRecord(g->consumed[(size_t)((k - 1) % g->nbuf)], s));
    }
    if (k >= g->nchunks) {
        g->given = k + (k < g->nchunks + 1 ? 1 : 0);
        g->cv.notify_all();
        *d_chunk = nullptr;
        *bytes = 0;
        return FX_OK;
    }
    g->cv.wait(lk, [&] { return g->failed || g->ready > k; });
    if (g->failed && g->ready <= k) {
        fx_set_error("%s", g->err);
        return FX_ERR_STATE;
    }
    const int b = (int)(k % g->nbuf);
    FX_HIP_TRY(hipStreamWaitEvent(s, g->copied[(size_t)b], 0));
    g->given = k + 1;
    g->cv.notify_all();
    *d_chunk = g->dev[(size_t)b];
    *bytes = g->bytes[(size_t)b];
    return FX_OK;
}

extern "C" int fx_ingest_close(fx_ingest* g)
{
    if (!g) return FX_OK;
    {
        std::lock_guard<std::mutex> lk(g->m);
        g->quit = true;
    }
    g->cv.notify_all();
    if (g->reader.joinable()) g->reader.join();
    hipSetDevice(g->device);
    hipStreamSynchronize(fx_ctx_hip_stream(g->ctx));
    hipStreamSynchronize(g->cs);
    ingest_free(g);
    return FX_OK;
}
