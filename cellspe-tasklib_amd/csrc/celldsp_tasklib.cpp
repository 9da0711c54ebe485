// celldsp_tasklib.cpp -- the cellspe-tasklib task runtime on MI355X units.
//
// A unit is a host worker thread with its own HIP stream (an fx_ctx created on
// the unit's first device task) and device staging buffers.  The PPU-side
// bookkeeping (reservation table under one mutex, blocking wait for a free
// unit, continuable tasks, streamed next_task chains, result codes) follows
// cellspe-tasklib.cpp:110-412; the per-command work follows the SPU dispatcher
// (cellspe-tasklib.cpp:477-580) and the DSP kernels of cellspe-tasklib-dsp.cpp,
// executed by the gfx950 kernels instead of SPU code.  Without a HIP device the
// bookkeeping still works and every command that needs the device completes
// with CELLDSP_TASK_GENERICERROR.
#include "fx_internal.h"

#include "celldsp_dsp.h"
#include "celldsp_tasklib.h"

#include <algorithm>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace {

// A repeated next_task chain replayed as one captured hipGraph
// (CDSP_BEHAV_GRAPH).  The host vectors of the chain are merged into
// disjoint address regions, each mirrored by a pinned staging region: the
// graph's copies read and write the staging, so copies between tasks that
// pass data through a host vector keep their order and aliasing, and the
// host copies user <-> staging once per execution.
struct ChainGraph {
    struct Region {
        char* user;
        size_t len, off;
        bool written;
    };
    std::vector<unsigned char> sig;   // commands, vectors, sub-contexts, FFT size
    int seen = 0;                     // eager executions with this signature
    bool broken = false;              // capture failed: stay eager
    hipGraphExec_t exec = nullptr;
    void* pinned = nullptr;
    size_t pinned_cap = 0;
    std::vector<Region> regions;
    unsigned fft_n_after = 0;         // host FFT state after the chain
    unsigned long slot_gen = 0;       // the unit's slot generation the graph was captured with
};

struct Unit {
    std::thread th;
    std::mutex m;
    std::condition_variable cv;
    CellDSPTask_t* posted = nullptr;
    bool busy = false;
    bool have_result = false;
    bool quit = false;
    int result = CELLDSP_TASK_OK;
    // device side, owned by the worker thread
    fx_ctx* ctx = nullptr;
    bool ctx_failed = false;
    void* slot[2 * CELLDSP_MAX_NUM_IOVECTORS + 2] = {};
    size_t cap[2 * CELLDSP_MAX_NUM_IOVECTORS + 2] = {};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    unsigned int fft_n = 0;   // continuable FFT size (0 = none)
    // bumped whenever a slot is freed or reallocated: a captured graph holds
    // the slot addresses, so it is valid only for the generation it saw
    unsigned long slot_gen = 0;
    ChainGraph cg;
};

struct State {
    std::mutex m;
    std::condition_variable freecv;
    int clients = 0;
    unsigned behaviour = CDSP_BEHAV_DEFAULT;
    unsigned flags[CELLDSP_MAX_SPE_NR] = {};
    int free_count = CELLDSP_MAX_SPE_NR;
    int waiters = 0;
    long graph_replays = 0;   // chains run from a captured graph
    Unit units[CELLDSP_MAX_SPE_NR];
};

State* g = nullptr;
std::mutex g_life;

// ---------------------------------------------------------------------------
// device helpers (worker thread only)
// ---------------------------------------------------------------------------
bool unit_device(Unit& u)
{
    if (u.ctx) return true;
    if (u.ctx_failed) return false;
    if (fx_ctx_create(0, &u.ctx) != FX_OK) {
        fprintf(stderr, "celldsp unit: no HIP device (%s)\n", fx_last_error());
        u.ctx_failed = true;
        return false;
    }
    if (hipEventCreate(&u.ev0) != hipSuccess || hipEventCreate(&u.ev1) != hipSuccess) {
        u.ctx_failed = true;
        return false;
    }
    return true;
}

void* unit_slot(Unit& u, int s, size_t bytes)
{
    if (u.cap[s] >= bytes && u.slot[s]) return u.slot[s];
    u.slot_gen++;
    if (u.slot[s]) hipFree(u.slot[s]);
    u.slot[s] = nullptr;
    u.cap[s] = 0;
    size_t b = bytes < 256 ? 256 : bytes;
    if (hipMalloc(&u.slot[s], b) != hipSuccess) {
        hipGetLastError();
        return nullptr;
    }
    u.cap[s] = b;
    return u.slot[s];
}

bool h2d(Unit& u, void* d, const void* h, size_t bytes)
{
    return !bytes || hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, fx_ctx_hip_stream(u.ctx)) == hipSuccess;
}

bool d2h(Unit& u, void* h, const void* d, size_t bytes)
{
    return !bytes || hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, fx_ctx_hip_stream(u.ctx)) == hipSuccess;
}

bool is_pow2u(unsigned n) { return n && !(n & (n - 1)); }

int fail(const char* what, unsigned cmd)
{
    fprintf(stderr, "celldsp task (cmd=%u): %s\n", cmd, what);
    return CELLDSP_TASK_GENERICERROR;
}

// stage input vectors 0..nin-1 into slots 0..nin-1, outputs into slots 8..
bool stage_in(Unit& u, CellDSPTask_t* t, int nin, int nout_fetch)
{
    for (int i = 0; i < nin; i++) {
        void* d = unit_slot(u, i, t->inputvector_lengths[i]);
        if (!d || !h2d(u, d, t->inputvector_ptrs[i].p, t->inputvector_lengths[i])) return false;
    }
    for (int o = 0; o < nout_fetch; o++) {
        void* d = unit_slot(u, CELLDSP_MAX_NUM_IOVECTORS + o, t->outputvector_lengths[o]);
        if (!d || !h2d(u, d, t->outputvector_ptrs[o].p, t->outputvector_lengths[o])) return false;
    }
    return true;
}

bool write_out(Unit& u, CellDSPTask_t* t, int o, size_t bytes)
{
    return d2h(u, t->outputvector_ptrs[o].p, u.slot[CELLDSP_MAX_NUM_IOVECTORS + o], bytes);
}

void* out_slot(Unit& u, CellDSPTask_t* t, int o)
{
    return unit_slot(u, CELLDSP_MAX_NUM_IOVECTORS + o, t->outputvector_lengths[o]);
}

// ---------------------------------------------------------------------------
// one task on one unit
// ---------------------------------------------------------------------------
int run_device_task(Unit& u, CellDSPTask_t* t)
{
    const unsigned cmd = t->command;
    hipStream_t s = fx_ctx_hip_stream(u.ctx);
    const float2* tw = fx_ctx_twiddles(u.ctx);
    switch (cmd) {
    case CELLDSP_TASK_IOTEST: {
        if (t->num_output_vectors < 1 || t->outputvector_lengths[0] < 16) return fail("IOTEST needs a 16-byte output", cmd);
        void* d = out_slot(u, t, 0);
        if (!d || fxk_iotest(s, (float*)d) != FX_OK || !write_out(u, t, 0, 16)) return fail("device error", cmd);
        return CELLDSP_TASK_OK;
    }
    case CELLDSP_TASK_MAC:
    case CELLDSP_TASK_MULC:
    case CELLDSP_TASK_ADDC: {
        if (t->num_input_vectors < 2 || t->num_output_vectors < 1) return fail("needs 2 inputs, 1 output", cmd);
        size_t bytes = t->inputvector_lengths[0];
        if (t->inputvector_lengths[1] < bytes || t->outputvector_lengths[0] < bytes || bytes % 8)
            return fail("vector lengths disagree", cmd);
        long n = (long)(bytes / 8);
        if (!stage_in(u, t, 2, cmd == CELLDSP_TASK_MAC ? 1 : 0)) return fail("staging failed", cmd);
        void* o = out_slot(u, t, 0);
        if (!o) return fail("allocation failed", cmd);
        const float2* a = (const float2*)u.slot[0];
        const float2* b = (const float2*)u.slot[1];
        int rc = cmd == CELLDSP_TASK_MAC ? fxk_complex_mac(s, a, b, (float2*)o, n, 0)
               : cmd == CELLDSP_TASK_MULC ? fxk_complex_mul(s, a, b, (float2*)o, n)
                                           : fxk_complex_add(s, a, b, (float2*)o, n);
        if (rc != FX_OK || !write_out(u, t, 0, bytes)) return fail("device error", cmd);
        return CELLDSP_TASK_OK;
    }
    case CELLDSP_TASK_FFT_EXEC:
    case CELLDSP_TASK_FFT_EXEC_DYN: {
        unsigned n = cmd == CELLDSP_TASK_FFT_EXEC ? u.fft_n : t->subcontext.FFT.nfft;
        if (cmd == CELLDSP_TASK_FFT_EXEC && n == 0) return fail("FFT_EXEC without FFT_INIT", cmd);
        if (!is_pow2u(n) || n < 2 || n > 16384) return fail("unsupported FFT length", cmd);
        if (cmd == CELLDSP_TASK_FFT_EXEC_DYN) u.fft_n = n;
        size_t bytes = (size_t)n * 8;
        if (t->num_input_vectors < 1 || t->num_output_vectors < 1 || t->inputvector_lengths[0] < bytes ||
            t->outputvector_lengths[0] < bytes)
            return fail("FFT vectors shorter than n complex", cmd);
        if (!stage_in(u, t, 1, 0)) return fail("staging failed", cmd);
        void* o = out_slot(u, t, 0);
        if (!o || fxk_fft_c2c(s, tw, (const float2*)u.slot[0], (float2*)o, (int)n, 1) != FX_OK ||
            !write_out(u, t, 0, bytes))
            return fail("device error", cmd);
        return CELLDSP_TASK_OK;
    }
    case CELLDSP_TASK_SINCOS:
    case CELLDSP_TASK_SINCOS_NORM: {
        long n = (long)t->subcontext.Phase.argcount;
        if (t->num_input_vectors < 1 || t->num_output_vectors < 1 ||
            t->inputvector_lengths[0] < (size_t)n * 4 || t->outputvector_lengths[0] < (size_t)n * 8)
            return fail("vector lengths disagree with argcount", cmd);
        if (!stage_in(u, t, 1, 0)) return fail("staging failed", cmd);
        void* o = out_slot(u, t, 0);
        if (!o || fxk_sincos(s, (const float*)u.slot[0], (float2*)o, n, cmd == CELLDSP_TASK_SINCOS_NORM) != FX_OK ||
            !write_out(u, t, 0, (size_t)n * 8))
            return fail("device error", cmd);
        return CELLDSP_TASK_OK;
    }
    case CELLDSP_TASK_SINCOS_GENERTR: {
        long n = (long)t->subcontext.Phase.argcount;
        if (t->num_output_vectors < 1 || t->outputvector_lengths[0] < (size_t)n * 8)
            return fail("output shorter than argcount", cmd);
        void* o = out_slot(u, t, 0);
        if (!o || fxk_sincos_gen(s, (double)t->subcontext.Phase.phase_start,
                                 (double)t->subcontext.Phase.phase_delta, (float2*)o, n) != FX_OK ||
            !write_out(u, t, 0, (size_t)n * 8))
            return fail("device error", cmd);
        return CELLDSP_TASK_OK;
    }
    case CELLDSP_TASK_UNPACK: {
        int nchan = t->subcontext.Unpack.nchan ? (int)t->subcontext.Unpack.nchan : 1;
        if (!(nchan == 1 || nchan == 2 || nchan == 4 || nchan == 8 || nchan == 16)) return fail("nchan", cmd);
        if (t->num_input_vectors < 1 || t->num_output_vectors < 1) return fail("needs 1 input, 1 output", cmd);
        size_t ntime = t->outputvector_lengths[0] / 4 / (size_t)nchan;
        if ((ntime * (size_t)nchan + 15) / 16 * 4 > t->inputvector_lengths[0]) return fail("input too short", cmd);
        if (!stage_in(u, t, 1, 0)) return fail("staging failed", cmd);
        void* o = out_slot(u, t, 0);
        if (!o || fxk_unpack2b(s, (const uint32_t*)u.slot[0], ntime, nchan, (float*)o) != FX_OK ||
            !write_out(u, t, 0, ntime * (size_t)nchan * 4))
            return fail("device error", cmd);
        return CELLDSP_TASK_OK;
    }
    case CELLDSP_TASK_DIFXPROCESS: {
        const CellDSPParams_Difx_t& d = t->subcontext.Difx;
        unsigned nfft = d.nfft ? d.nfft : d.twicenumchannels;
        unsigned nch = d.numchannels ? d.numchannels : d.twicenumchannels / 2;
        if (!is_pow2u(nfft) || nfft < 4 || nfft > 16384 || nch == 0 || 2 * nch > nfft)
            return fail("DIFXPROCESS sizes", cmd);
        size_t nproc = (size_t)nch + 1;
        if (t->num_input_vectors < 2 || t->num_output_vectors < 3 || t->inputvector_lengths[0] < nfft * 8u ||
            t->inputvector_lengths[1] < nproc * 8 || t->outputvector_lengths[0] < nproc * 8 ||
            t->outputvector_lengths[1] < nproc * 8 || t->outputvector_lengths[2] < nproc * 8)
            return fail("DIFXPROCESS vector lengths", cmd);
        if (!stage_in(u, t, 2, 0)) return fail("staging failed", cmd);
        // autocorrelation accumulator is read-modify-write
        void* acc = out_slot(u, t, 2);
        if (!acc || !h2d(u, acc, t->outputvector_ptrs[2].p, nproc * 8)) return fail("staging failed", cmd);
        void* fftout = unit_slot(u, 2 * CELLDSP_MAX_NUM_IOVECTORS, (size_t)nfft * 8);
        void* o0 = out_slot(u, t, 0);
        void* o1 = out_slot(u, t, 1);
        if (!fftout || !o0 || !o1) return fail("allocation failed", cmd);
        int band0 = d.useUpperOrLower ? (int)nch : 0;
        if (fxk_fft_c2c(s, tw, (const float2*)u.slot[0], (float2*)fftout, (int)nfft, 1) != FX_OK ||
            fxk_difx_post(s, (const float2*)fftout, (int)nfft, band0, (int)nproc, (const float2*)u.slot[1],
                          (float2*)o0, (float2*)o1, (float2*)acc) != FX_OK ||
            !write_out(u, t, 0, nproc * 8) || !write_out(u, t, 1, nproc * 8) || !write_out(u, t, 2, nproc * 8))
            return fail("device error", cmd);
        return CELLDSP_TASK_OK;
    }
    case CELLDSP_TASK_IOBENCH:
    case CELLDSP_TASK_DMASIMULATION: {
        // every output vector makes a round trip through HBM; DMASIMULATION
        // fills the outputs from input 0 first
        for (unsigned o = 0; o < t->num_output_vectors; o++) {
            size_t b = t->outputvector_lengths[o];
            void* dd = out_slot(u, t, (int)o);
            const void* src = (cmd == CELLDSP_TASK_DMASIMULATION && t->num_input_vectors > 0)
                                  ? t->inputvector_ptrs[0].p : t->outputvector_ptrs[o].p;
            size_t sb = (cmd == CELLDSP_TASK_DMASIMULATION && t->num_input_vectors > 0)
                            ? std::min<size_t>(b, t->inputvector_lengths[0]) : b;
            if (!dd || !h2d(u, dd, src, sb) || !write_out(u, t, (int)o, sb)) return fail("device error", cmd);
        }
        return CELLDSP_TASK_OK;
    }
    default:
        return CELLDSP_TASK_UNKOWNCOMMAND;
    }
}

int run_task(Unit& u, CellDSPTask_t* t)
{
    switch (t->command) {
    case CELLDSP_TASK_NOP:
    case CELLDSP_TASK_RUNALL:
    case CELLDSP_TASK_SIGTERM:
        t->tick_count = 0;
        return CELLDSP_TASK_OK;
    case CELLDSP_TASK_FFT_INIT: {
        unsigned n = t->subcontext.FFT.nfft;
        if (!is_pow2u(n) || n < 2 || n > 16384) return fail("unsupported FFT length", t->command);
        u.fft_n = n;
        t->subcontext.FFT.log2_nfft = (unsigned)__builtin_ctz(n);
        t->tick_count = 0;
        return CELLDSP_TASK_OK;
    }
    case CELLDSP_TASK_FFT_FREE:
        u.fft_n = 0;
        t->tick_count = 0;
        return CELLDSP_TASK_OK;
    default:
        break;
    }
    if (t->command > CELLDSP_TASK_SIGTERM) return CELLDSP_TASK_UNKOWNCOMMAND;
    if (!unit_device(u)) return fail("no HIP device for this unit", t->command);
    hipStream_t s = fx_ctx_hip_stream(u.ctx);
    hipEventRecord(u.ev0, s);
    int rc = run_device_task(u, t);
    hipEventRecord(u.ev1, s);
    if (hipStreamSynchronize(s) != hipSuccess) rc = fail("stream error", t->command);
    float ms = 0.f;
    hipEventElapsedTime(&ms, u.ev0, u.ev1);
    t->tick_count = (unsigned int)(ms * 1e-3 * CELLDSP_TIMEBASE);
    return rc;
}

// ---------------------------------------------------------------------------
// chains as graphs
// ---------------------------------------------------------------------------
bool graphable(unsigned cmd)
{
    switch (cmd) {
    case CELLDSP_TASK_IOTEST:
    case CELLDSP_TASK_MAC:
    case CELLDSP_TASK_MULC:
    case CELLDSP_TASK_ADDC:
    case CELLDSP_TASK_FFT_EXEC:
    case CELLDSP_TASK_FFT_EXEC_DYN:
    case CELLDSP_TASK_SINCOS:
    case CELLDSP_TASK_SINCOS_NORM:
    case CELLDSP_TASK_SINCOS_GENERTR:
    case CELLDSP_TASK_UNPACK:
    case CELLDSP_TASK_DIFXPROCESS:
        return true;
    default:
        return false;
    }
}

template <typename T>
void put(std::vector<unsigned char>& v, const T& x)
{
    const unsigned char* b = reinterpret_cast<const unsigned char*>(&x);
    v.insert(v.end(), b, b + sizeof(T));
}

// Signature of a chain; false if some task cannot be captured.
bool chain_signature(const Unit& u, CellDSPTask_t* head, std::vector<unsigned char>& sig, int& ntasks)
{
    sig.clear();
    ntasks = 0;
    // the unit's FFT size matters only to an FFT_EXEC before any FFT_EXEC_DYN
    bool dyn = false, uses_fft_n = false;
    for (CellDSPTask_t* t = head; t; t = (CellDSPTask_t*)t->next_task.p) {
        if (t->command == CELLDSP_TASK_FFT_EXEC && !dyn) uses_fft_n = true;
        if (t->command == CELLDSP_TASK_FFT_EXEC_DYN) dyn = true;
    }
    put(sig, uses_fft_n ? u.fft_n : 0u);
    for (CellDSPTask_t* t = head; t; t = (CellDSPTask_t*)t->next_task.p) {
        if (!graphable(t->command) || t->num_input_vectors > CELLDSP_MAX_NUM_IOVECTORS ||
            t->num_output_vectors > CELLDSP_MAX_NUM_IOVECTORS)
            return false;
        put(sig, t->command);
        put(sig, t->num_input_vectors);
        put(sig, t->num_output_vectors);
        for (unsigned i = 0; i < t->num_input_vectors; i++) {
            put(sig, t->inputvector_ptrs[i].p);
            put(sig, t->inputvector_lengths[i]);
        }
        for (unsigned o = 0; o < t->num_output_vectors; o++) {
            put(sig, t->outputvector_ptrs[o].p);
            put(sig, t->outputvector_lengths[o]);
        }
        put(sig, t->subcontext);
        ntasks++;
    }
    return ntasks >= 2;
}

void chain_graph_reset(ChainGraph& cg)
{
    if (cg.exec) hipGraphExecDestroy(cg.exec);
    cg.exec = nullptr;
    cg.sig.clear();
    cg.regions.clear();
    cg.seen = 0;
    cg.broken = false;
}

void chain_graph_free(ChainGraph& cg)
{
    chain_graph_reset(cg);
    if (cg.pinned) hipHostFree(cg.pinned);
    cg.pinned = nullptr;
    cg.pinned_cap = 0;
}

// Merge the chain's host vectors into disjoint regions with staging offsets.
bool chain_regions(ChainGraph& cg, CellDSPTask_t* head)
{
    struct Iv {
        char* p;
        size_t len;
        bool out;
    };
    std::vector<Iv> iv;
    for (CellDSPTask_t* t = head; t; t = (CellDSPTask_t*)t->next_task.p) {
        for (unsigned i = 0; i < t->num_input_vectors; i++)
            if (t->inputvector_lengths[i]) iv.push_back({(char*)t->inputvector_ptrs[i].p, t->inputvector_lengths[i], false});
        for (unsigned o = 0; o < t->num_output_vectors; o++)
            if (t->outputvector_lengths[o]) iv.push_back({(char*)t->outputvector_ptrs[o].p, t->outputvector_lengths[o], true});
    }
    for (const Iv& x : iv)
        if (!x.p) return false;
    std::sort(iv.begin(), iv.end(), [](const Iv& a, const Iv& b) { return a.p < b.p; });
    cg.regions.clear();
    size_t off = 0;
    for (const Iv& x : iv) {
        if (!cg.regions.empty()) {
            ChainGraph::Region& r = cg.regions.back();
            if (x.p < r.user + r.len) {
                const size_t end = std::max<size_t>(r.len, (size_t)(x.p - r.user) + x.len);
                off += end - r.len;
                r.len = end;
                r.written = r.written || x.out;
                continue;
            }
            off = (off + 255) & ~(size_t)255;
        }
        cg.regions.push_back({x.p, x.len, off, x.out});
        off += x.len;
    }
    if (off > cg.pinned_cap) {
        if (cg.pinned) hipHostFree(cg.pinned);
        cg.pinned = nullptr;
        cg.pinned_cap = 0;
        if (hipHostMalloc(&cg.pinned, off, hipHostMallocDefault) != hipSuccess) {
            hipGetLastError();
            cg.pinned = nullptr;
            return false;
        }
        cg.pinned_cap = off;
    }
    return true;
}

void* staging_of(const ChainGraph& cg, const void* p)
{
    const char* c = (const char*)p;
    for (const ChainGraph::Region& r : cg.regions)
        if (c >= r.user && c < r.user + r.len) return (char*)cg.pinned + r.off + (c - r.user);
    return nullptr;
}

// Capture the chain (its host vectors redirected to the staging) into cg.exec.
bool chain_capture(Unit& u, CellDSPTask_t* head)
{
    ChainGraph& cg = u.cg;
    if (!chain_regions(cg, head)) return false;
    std::vector<CellDSPTask_t> shadow;
    for (CellDSPTask_t* t = head; t; t = (CellDSPTask_t*)t->next_task.p) {
        CellDSPTask_t c;
        memcpy((void*)&c, (const void*)t, sizeof(c));
        for (unsigned i = 0; i < c.num_input_vectors; i++)
            if (c.inputvector_lengths[i]) c.inputvector_ptrs[i].p = staging_of(cg, t->inputvector_ptrs[i].p);
        for (unsigned o = 0; o < c.num_output_vectors; o++)
            if (c.outputvector_lengths[o]) c.outputvector_ptrs[o].p = staging_of(cg, t->outputvector_ptrs[o].p);
        c.next_task.p = nullptr;
        shadow.push_back(c);
    }
    hipStream_t s = fx_ctx_hip_stream(u.ctx);
    const unsigned fft_before = u.fft_n;
    const unsigned long gen = u.slot_gen;
    // the slots were sized by the eager run of this chain, so nothing below allocates
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        hipGetLastError();
        return false;
    }
    bool ok = true;
    for (CellDSPTask_t& c : shadow)
        if (run_device_task(u, &c) != CELLDSP_TASK_OK) ok = false;
    hipGraph_t graph = nullptr;
    if (hipStreamEndCapture(s, &graph) != hipSuccess) ok = false;
    cg.fft_n_after = u.fft_n;
    u.fft_n = fft_before;
    if (u.slot_gen != gen) ok = false;   // a slot moved under the capture
    cg.slot_gen = u.slot_gen;
    if (ok && graph) ok = hipGraphInstantiate(&cg.exec, graph, nullptr, nullptr, 0) == hipSuccess;
    if (graph) hipGraphDestroy(graph);
    if (!ok) {
        hipGetLastError();
        if (cg.exec) hipGraphExecDestroy(cg.exec);
        cg.exec = nullptr;
    }
    return ok;
}

int chain_replay(Unit& u, CellDSPTask_t* head, int ntasks)
{
    ChainGraph& cg = u.cg;
    for (const ChainGraph::Region& r : cg.regions) memcpy((char*)cg.pinned + r.off, r.user, r.len);
    hipStream_t s = fx_ctx_hip_stream(u.ctx);
    int rc = CELLDSP_TASK_OK;
    hipEventRecord(u.ev0, s);
    if (hipGraphLaunch(cg.exec, s) != hipSuccess) rc = fail("graph launch failed", head->command);
    hipEventRecord(u.ev1, s);
    if (hipStreamSynchronize(s) != hipSuccess) rc = fail("stream error", head->command);
    float ms = 0.f;
    hipEventElapsedTime(&ms, u.ev0, u.ev1);
    if (rc == CELLDSP_TASK_OK) {
        for (const ChainGraph::Region& r : cg.regions)
            if (r.written) memcpy(r.user, (char*)cg.pinned + r.off, r.len);
        u.fft_n = cg.fft_n_after;
    }
    // device time of the whole chain, shared evenly among its tasks
    const unsigned ticks = (unsigned)(ms * 1e-3 * CELLDSP_TIMEBASE / ntasks);
    for (CellDSPTask_t* t = head; t; t = (CellDSPTask_t*)t->next_task.p) {
        t->tick_count = ticks;
        t->completed_flag = (unsigned short)rc;
    }
    return rc;
}

// Runs the chain from its graph when it can; false: run it eagerly.
bool run_chain_graph(Unit& u, CellDSPTask_t* head, int* rc)
{
    std::vector<unsigned char> sig;
    int ntasks = 0;
    if (!chain_signature(u, head, sig, ntasks) || !unit_device(u)) return false;
    ChainGraph& cg = u.cg;
    if (sig != cg.sig) {
        chain_graph_reset(cg);
        cg.sig.swap(sig);
    }
    if (cg.exec && cg.slot_gen != u.slot_gen) {
        // another task on this unit grew a slot the graph reads or writes:
        // drop the graph, run this execution eagerly, capture again next time
        hipGraphExecDestroy(cg.exec);
        cg.exec = nullptr;
        cg.seen = 0;
    }
    if (!cg.exec) {
        if (cg.broken || cg.seen++ == 0) return false;     // first run eager: sizes the slots
        if (!chain_capture(u, head)) {
            cg.broken = true;
            return false;
        }
    }
    *rc = chain_replay(u, head, ntasks);
    {
        std::lock_guard<std::mutex> lk(g->m);
        g->graph_replays++;
    }
    return true;
}

void worker(Unit* u)
{
    for (;;) {
        CellDSPTask_t* head;
        {
            std::unique_lock<std::mutex> lk(u->m);
            u->cv.wait(lk, [u] { return u->quit || u->posted != nullptr; });
            if (u->posted == nullptr && u->quit) break;
            head = u->posted;
        }
        int rc = CELLDSP_TASK_OK;
        unsigned behaviour;
        {
            std::lock_guard<std::mutex> lk(g->m);
            behaviour = g->behaviour;
        }
        if (!(behaviour & CDSP_BEHAV_GRAPH) || !run_chain_graph(*u, head, &rc)) {
            // follow the streamed chain (cellspe-tasklib.cpp:594-602)
            for (CellDSPTask_t* t = head; t; t = (CellDSPTask_t*)t->next_task.p) {
                rc = run_task(*u, t);
                t->completed_flag = (unsigned short)rc;
            }
        }
        {
            std::lock_guard<std::mutex> lk(u->m);
            u->posted = nullptr;
            u->busy = false;
            u->result = rc;
            u->have_result = true;
        }
        u->cv.notify_all();
    }
    // release device resources on the thread that created them
    chain_graph_free(u->cg);
    for (auto& p : u->slot)
        if (p) hipFree(p);
    if (u->ev0) hipEventDestroy(u->ev0);
    if (u->ev1) hipEventDestroy(u->ev1);
    if (u->ctx) fx_ctx_destroy(u->ctx);
}

bool valid_unit(int i) { return g && i >= 0 && i < CELLDSP_MAX_SPE_NR; }

} // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" int celldsp_start(void)
{
    std::lock_guard<std::mutex> life(g_life);
    if (g) {
        std::lock_guard<std::mutex> lk(g->m);
        g->clients++;
        return 0;
    }
    g = new State;
    g->clients = 1;
    for (int i = 0; i < CELLDSP_MAX_SPE_NR; i++) g->units[i].th = std::thread(worker, &g->units[i]);
    return 0;
}

extern "C" int celldsp_stop(void)
{
    std::lock_guard<std::mutex> life(g_life);
    if (!g) return -1;
    {
        std::lock_guard<std::mutex> lk(g->m);
        if (--g->clients > 0) return 0;
    }
    for (auto& u : g->units) {
        {
            std::lock_guard<std::mutex> lk(u.m);
            u.quit = true;
        }
        u.cv.notify_all();
    }
    for (auto& u : g->units)
        if (u.th.joinable()) u.th.join();
    delete g;
    g = nullptr;
    return 0;
}

extern "C" void celldsp_setBehaviour(unsigned int flags)
{
    if (!g) return;
    std::lock_guard<std::mutex> lk(g->m);
    g->behaviour = flags;
    g->freecv.notify_all();
}

extern "C" unsigned int celldsp_getBehaviour(void)
{
    if (!g) return 0;
    std::lock_guard<std::mutex> lk(g->m);
    return g->behaviour;
}

extern "C" int celldsp_getFreeSPE(void)
{
    if (!g) return -1;
    std::unique_lock<std::mutex> lk(g->m);
    for (;;) {
        for (int i = 0; i < CELLDSP_MAX_SPE_NR; i++) {
            if (!(g->flags[i] & CDSP_EF_RESERVED)) {
                g->flags[i] |= CDSP_EF_RESERVED;
                g->free_count--;
                return i;
            }
        }
        if (!(g->behaviour & CDSP_BEHAV_BLOCKING)) return -1;
        g->waiters++;
        g->freecv.wait(lk);
        g->waiters--;
    }
}

extern "C" int celldsp_setFreeSPE(int target)
{
    if (!valid_unit(target)) return -1;
    std::lock_guard<std::mutex> lk(g->m);
    if (g->flags[target] & CDSP_EF_RESERVED) {
        g->flags[target] = CDSP_EF_DEFAULT;
        g->free_count++;
        if (g->waiters) g->freecv.notify_all();
    } else {
        fprintf(stderr, "warning: celldsp_setFreeSPE(unit %d) was already free\n", target);
    }
    return 0;
}

extern "C" int celldsp_startTask(int target, CellDSPTask_t* task)
{
    if (!valid_unit(target) || !task) return -1;
    {
        std::lock_guard<std::mutex> lk(g->m);
        if (!(g->flags[target] & CDSP_EF_RESERVED)) {
            fprintf(stderr, "celldsp_startTask: unit %d is not reserved\n", target);
            return -1;
        }
    }
    for (CellDSPTask_t* t = task; t; t = (CellDSPTask_t*)t->next_task.p) t->completed_flag = CELLDSP_TASK_INCOMPLETE;
    Unit& u = g->units[target];
    {
        std::unique_lock<std::mutex> lk(u.m);
        // the inbound mailbox holds one chain: wait for the unit to drain it
        u.cv.wait(lk, [&u] { return !u.busy; });
        u.posted = task;
        u.busy = true;
        u.have_result = false;
    }
    u.cv.notify_all();
    return 0;
}

extern "C" int celldsp_waitTask(int target)
{
    if (!valid_unit(target)) return -1;
    Unit& u = g->units[target];
    std::unique_lock<std::mutex> lk(u.m);
    u.cv.wait(lk, [&u] { return !u.busy && u.have_result; });
    u.have_result = false;
    return u.result;
}

extern "C" int celldsp_executeTask(CellDSPTask_t* task)
{
    int unit = celldsp_getFreeSPE();
    if (unit == -1) {
        fprintf(stderr, "celldsp_executeTask(cmd=%u): could not get any free unit\n", task ? task->command : 0u);
        return !vecNoErr;
    }
    if (celldsp_startTask(unit, task) == -1) {
        celldsp_setFreeSPE(unit);
        return !vecNoErr;
    }
    int result = celldsp_waitTask(unit);
    celldsp_setFreeSPE(unit);
    if (result != CELLDSP_TASK_OK) {
        fprintf(stderr, "celldsp_executeTask: task returned error code %d\n", result);
        return !vecNoErr;
    }
    return vecNoErr;
}

extern "C" int celldsp_startContinuableTask(CellDSPTask_t* task, int* speID)
{
    *speID = celldsp_getFreeSPE();
    if (*speID == -1) {
        fprintf(stderr, "celldsp_startContinuableTask(cmd=%u): could not get any free unit\n", task->command);
        return !vecNoErr;
    }
    {
        std::lock_guard<std::mutex> lk(g->m);
        g->flags[*speID] |= CDSP_EF_CONTINUABLE;
    }
    if (celldsp_startTask(*speID, task) == -1) return !vecNoErr;
    int result = celldsp_waitTask(*speID);
    if (result != CELLDSP_TASK_OK) {
        celldsp_setFreeSPE(*speID);
        return !vecNoErr;
    }
    return vecNoErr;
}

extern "C" int celldsp_continueTask(CellDSPTask_t* task, int* speID)
{
    if (celldsp_startTask(*speID, task) == -1) return !vecNoErr;
    int result = celldsp_waitTask(*speID);
    return result == CELLDSP_TASK_OK ? vecNoErr : !vecNoErr;
}

extern "C" int celldsp_stopContinuableTask(int* speID)
{
    celldsp_setFreeSPE(*speID);
    *speID = -1;
    return vecNoErr;
}

extern "C" long celldsp_graphReplays(void)
{
    if (!g) return 0;
    std::lock_guard<std::mutex> lk(g->m);
    return g->graph_replays;
}

extern "C" int celldsp_freeCount(void)
{
    if (!g) return 0;
    std::lock_guard<std::mutex> lk(g->m);
    return g->free_count;
}

extern "C" int celldsp_clientCount(void)
{
    std::lock_guard<std::mutex> life(g_life);
    if (!g) return 0;
    std::lock_guard<std::mutex> lk(g->m);
    return g->clients;
}

// ---------------------------------------------------------------------------
// spe* wrappers (cellspe-tasklib-dsp.cpp:95-359)
// ---------------------------------------------------------------------------
extern "C" int speNOP(void)
{
    CellDSPTaskClass nop(CELLDSP_TASK_NOP);
    return celldsp_executeTask(nop.getTask());
}

extern "C" int speIOCheck_out(float* out4)
{
    CellDSPTaskClass io(CELLDSP_TASK_IOTEST);
    io.addOutVec((void*)out4, 4 * sizeof(float));
    return celldsp_executeTask(io.getTask());
}

// cellspe-tasklib-dsp.cpp:111-124: the IOTEST task into a static 8-float
// buffer, then the values it wrote back are printed
extern "C" int speIOCheck(void)
{
    alignas(128) static float output[8];
    const int status = speIOCheck_out(output);
    printf("speIOCheck wrote back : vect={%g,%g,%g,%g}   expected would be {0.1,2.3,4.5,6.7}\n", output[0],
           output[1], output[2], output[3]);
    fflush(stdout);
    return status;
}

extern "C" int speAddProduct_32fc(fc32* src, fc32* src2, fc32* acc, int length)
{
    CellDSPTaskClass mac(CELLDSP_TASK_MAC);
    mac.addInVec((void*)src, (size_t)length * sizeof(fc32));
    mac.addInVec((void*)src2, (size_t)length * sizeof(fc32));
    mac.addOutVec((void*)acc, (size_t)length * sizeof(fc32));
    return celldsp_executeTask(mac.getTask());
}

extern "C" int speMul_32fc(fc32* src, fc32* src2, fc32* dest, int length)
{
    CellDSPTaskClass t(CELLDSP_TASK_MULC);
    t.addInVec((void*)src, (size_t)length * sizeof(fc32));
    t.addInVec((void*)src2, (size_t)length * sizeof(fc32));
    t.addOutVec((void*)dest, (size_t)length * sizeof(fc32));
    return celldsp_executeTask(t.getTask());
}

extern "C" int speAdd_32fc(fc32* src, fc32* src2, fc32* dest, int length)
{
    CellDSPTaskClass t(CELLDSP_TASK_ADDC);
    t.addInVec((void*)src, (size_t)length * sizeof(fc32));
    t.addInVec((void*)src2, (size_t)length * sizeof(fc32));
    t.addOutVec((void*)dest, (size_t)length * sizeof(fc32));
    return celldsp_executeTask(t.getTask());
}

extern "C" int speFFT_init(int fft_len, int* speID)
{
    CellDSPTaskClass t(CELLDSP_TASK_FFT_INIT);
    t.setFFTparams((unsigned)fft_len);
    return celldsp_startContinuableTask(t.getTask(), speID);
}

extern "C" int speFFT_exec(void* in, void* out, int fft_len, int* speID)
{
    CellDSPTaskClass t(CELLDSP_TASK_FFT_EXEC);
    t.addInVec(in, (size_t)fft_len * sizeof(fc32));
    t.addOutVec(out, (size_t)fft_len * sizeof(fc32));
    t.setFFTparams((unsigned)fft_len);
    return celldsp_continueTask(t.getTask(), speID);
}

extern "C" int speFFT_destroy(int* speID)
{
    CellDSPTaskClass t(CELLDSP_TASK_FFT_FREE);
    int status = celldsp_continueTask(t.getTask(), speID);
    celldsp_stopContinuableTask(speID);
    return status;
}

extern "C" int speFFT_singleexec(void* in, void* out, int fft_len)
{
    CellDSPTaskClass t(CELLDSP_TASK_FFT_EXEC_DYN);
    t.addInVec(in, (size_t)fft_len * sizeof(fc32));
    t.addOutVec(out, (size_t)fft_len * sizeof(fc32));
    t.setFFTparams((unsigned)fft_len);
    return celldsp_executeTask(t.getTask());
}

extern "C" int speSinCosGenerator(float startphase, float increment, fc32* out, int numiter)
{
    CellDSPTaskClass t(CELLDSP_TASK_SINCOS_GENERTR);
    t.addOutVec((void*)out, (size_t)numiter * sizeof(cf32));
    CellDSPParams_Phase_t* p = &t.getSubcontext()->Phase;
    p->argcount = (unsigned)numiter;
    p->phase_start = startphase;
    p->phase_delta = increment;
    return celldsp_executeTask(t.getTask());
}

extern "C" int speSinCosAccurate(float* arg, cf32* out, int n)
{
    CellDSPTaskClass t(CELLDSP_TASK_SINCOS);
    t.addInVec((void*)arg, (size_t)n * sizeof(float));
    t.addOutVec((void*)out, (size_t)n * sizeof(cf32));
    t.getSubcontext()->Phase.argcount = (unsigned)n;
    return celldsp_executeTask(t.getTask());
}

extern "C" int speSinCosAccurateNorm(float* arg, cf32* out, int n)
{
    CellDSPTaskClass t(CELLDSP_TASK_SINCOS_NORM);
    t.addInVec((void*)arg, (size_t)n * sizeof(float));
    t.addOutVec((void*)out, (size_t)n * sizeof(cf32));
    t.getSubcontext()->Phase.argcount = (unsigned)n;
    return celldsp_executeTask(t.getTask());
}

extern "C" int speUnpack2bit(const unsigned int* words, float* out, int nsamp, int nchan)
{
    CellDSPTaskClass t(CELLDSP_TASK_UNPACK);
    size_t nw = ((size_t)nsamp * (size_t)nchan + 15) / 16;
    t.addInVec((void*)words, nw * 4);
    t.addOutVec((void*)out, (size_t)nsamp * (size_t)nchan * sizeof(float));
    t.getSubcontext()->Unpack.nchan = (unsigned)nchan;
    return celldsp_executeTask(t.getTask());
}

extern "C" int speDIFX_process(void* sincosrotdone, void* complexfracmult, void* fftoutJ, void* conjfftoutJ,
                               void* autocorr, int nfft, int twicenumchannels, int useUpperOrLower)
{
    size_t dual = (size_t)twicenumchannels * sizeof(cf32);
    size_t normal = (size_t)(twicenumchannels / 2 + 1 + 3) * sizeof(cf32);
    CellDSPTaskClass t(CELLDSP_TASK_DIFXPROCESS);
    t.addInVec(sincosrotdone, dual);
    t.addInVec(complexfracmult, normal);
    t.addOutVec(fftoutJ, normal);
    t.addOutVec(conjfftoutJ, normal);
    t.addOutVec(autocorr, normal);
    CellDSPParams_Difx_t* d = &t.getSubcontext()->Difx;
    d->nfft = (unsigned)nfft;
    d->useUpperOrLower = (unsigned)useUpperOrLower;
    d->twicenumchannels = (unsigned)twicenumchannels;
    d->numchannels = (unsigned)(twicenumchannels / 2);
    return celldsp_executeTask(t.getTask());
}
