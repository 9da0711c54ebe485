// fx_noalloc.hip -- the spectrometer entry point that allocates nothing.
//
// fx_spectrometer_ex_noalloc is fx_spectrometer_ex for callers that cannot
// call hipMalloc (code under HIP graph capture, tight serving loops): the
// window table and the partial spectra live in one file-scope device array
// of FX_SPEC_NOALLOC_BYTES instead of per-context allocations.  Each GPU the
// library runs on holds its own copy of the array until the process ends;
// its address on a device is looked up once, with that device current, on
// the first call there.  The kernels are the ones fx_spectrometer_ex runs,
// with the same launch shape, so the two entry points give identical spectra.
//
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU
// (other streams, other contexts, other host threads) race.  Nothing in the
// array has to start at zero: each call writes the window table and every
// partial spectrum before they are read.
#include "fx_internal.h"

#include <mutex>

static __device__ __align__(256) unsigned char scratch[FX_SPEC_NOALLOC_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};   // the array's device address on each device (null: not looked up yet)

// window table at the head of the array, partial spectra after it (256-byte aligned)
constexpr size_t kWindowBytes = (size_t)FX_MAX_N * sizeof(float);

int scratch_ptr(int device, unsigned char** out)
{
    FX_REQUIRE(device >= 0 && device < kMaxDevices, FX_ERR_ARG, "fx_spectrometer_ex_noalloc: device %d", device);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        // the caller made `device` current; the array has a separate address on every device
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            hipGetLastError();
            fx_set_error("fx_spectrometer_ex_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", device,
                         hipGetErrorString(e));
            return FX_ERR_HIP;
        }
        g_ptr[device] = p;
    }
    *out = static_cast<unsigned char*>(g_ptr[device]);
    return FX_OK;
}

} // namespace

extern "C" int fx_spectrometer_ex_noalloc(fx_ctx* ctx, const uint32_t* d_words, size_t ntime, int nchan, int n,
                                          int overlap, int integrate_ffts, int window, int continuous,
                                          int keep_nyquist, float* d_out, long* nspec_out)
{
    FX_REQUIRE(ctx != nullptr, FX_ERR_ARG, "fx_spectrometer_ex_noalloc: null context");
    const int device = fx_ctx_device_of(ctx);
    FX_HIP_TRY(hipSetDevice(device));
    fx_spec_plan pl;
    int rc = fx_spec_make_plan(ntime, nchan, n, overlap, integrate_ffts, window, continuous, &pl);
    if (rc) return rc;
    FX_REQUIRE(kWindowBytes + pl.part_bytes <= (size_t)FX_SPEC_NOALLOC_BYTES, FX_ERR_ARG,
               "fx_spectrometer_ex_noalloc: %ld spectra x %d channels of %d points need %zu bytes of partial "
               "spectra; the static scratch holds %zu (use fx_spectrometer_ex, or fewer spectra per call)",
               pl.nspec, nchan, n, pl.part_bytes, (size_t)FX_SPEC_NOALLOC_BYTES - kWindowBytes);
    if (nspec_out) *nspec_out = pl.nspec;
    if (pl.nspec == 0) return FX_OK;
    unsigned char* base = nullptr;
    rc = scratch_ptr(device, &base);
    if (rc) return rc;
    float* win = reinterpret_cast<float*>(base);
    float* partial = reinterpret_cast<float*>(base + kWindowBytes);
    hipStream_t s = fx_ctx_hip_stream(ctx);
    if (fxk_window(s, window, n, win) != FX_OK ||
        fxk_spectrometer(s, fx_ctx_twiddles(ctx), win, d_words, pl.hop, pl.spec_stride, nchan, n, integrate_ffts,
                         (int)pl.nspec, partial, pl.groups, keep_nyquist ? 1 : 0, d_out) != FX_OK) {
        fx_set_error("fx_spectrometer_ex_noalloc: kernel launch failed: %s", hipGetErrorString(hipGetLastError()));
        return FX_ERR_HIP;
    }
    return FX_OK;
}
