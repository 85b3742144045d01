// esp_noalloc.hip -- the allocation-free window EA-Myers batch.
//
// esp_window_myers_batch grows context buffers on demand (hipMalloc).  Callers
// that cannot allocate -- code under HIP graph capture, tight serving loops,
// harnesses that forbid it -- call esp_window_myers_batch_noalloc instead.  Its
// scratch (the job texts, the Myers descriptors and the work-queue counter)
// lives in one file-scope device array, `scratch`, of kScratchBytes per GPU,
// which the code object holds from load until the process ends.
//
// Host code never passes the array's name to a kernel or a copy: its device
// address comes from hipGetSymbolAddress, taken once per device with that
// device current (noalloc_setup, called by esp_create), and every use goes
// through that address.  All calls on one GPU share the array, and the last
// call's contents stay in it: calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/esp_align.h"
#include "esp_ctx.hpp"

constexpr size_t kScratchBytes = (size_t)32 << 20;

// Referenced by host code (hipGetSymbolAddress below), so the compiler keeps
// it as an exported symbol of this file's code object.
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;

std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};
hipError_t g_err[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};

void *scratch_of(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_ptr[device];
}

}  // namespace

// Looks the array's device address up once per device; the caller has made
// `device` current.  A failure is kept, and the entry point reports it.
void esph::noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_done[device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_err[device] = e;
    g_ptr[device] = e == hipSuccess ? p : nullptr;
    g_done[device] = true;
}

extern "C" {

int esp_window_myers_batch_noalloc(esp_ctx *c, const esp_ref *r, const esp_window_job *jobs, size_t njobs,
                                   const uint8_t *reads, uint32_t max_pattern, uint32_t max_text, int mode,
                                   int32_t *dist, uint64_t *pos, int32_t *status) {
    using namespace esph;
    int rc = window_myers_check(c, r, jobs, njobs, reads, max_text, mode, dist, pos);
    if (rc || njobs == 0) return rc;
    // scratch layout: counter (256 B), descriptors, texts (256-byte aligned)
    const size_t desc = align_up(window_myers_desc_bytes(njobs), 256);
    const size_t need = 256 + desc + window_myers_arena_bytes(njobs, max_text);
    if (need > kScratchBytes)
        return fail(c, ESP_E_INVAL,
                    "window Myers batch needs " + std::to_string(need) +
                        " bytes of scratch; the allocation-free entry point holds " +
                        std::to_string(kScratchBytes) + " (split the batch or lower max_text)");
    uint8_t *base = (uint8_t *)scratch_of(c->device);
    if (!base) {
        std::lock_guard<std::mutex> lk(g_mu);
        const hipError_t e = c->device < kMaxDevices ? g_err[c->device] : hipErrorInvalidDevice;
        return fail(c, ESP_E_DEVICE, std::string("device scratch address unavailable: ") +
                                         hipGetErrorString(e == hipSuccess ? hipErrorNotInitialized : e));
    }
    HIP_TRY(c, hipSetDevice(c->device));
    return window_myers_core(c, r, jobs, njobs, reads, max_pattern, max_text, mode, dist, pos, status,
                             base + 256, base + 256 + desc, (uint32_t *)base);
}

}  // extern "C"
