// litho_socs_image_batch_noalloc: the mask -> image batch with its workspaces
// in a fixed device array of the library instead of context memory, for
// callers that cannot allocate (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).  It runs the same launches as
// litho_socs_image_batch and gives the same images bit for bit.
//
// The array lives once per GPU for the life of the process.  Every call on a
// GPU uses that GPU's copy, so calls that overlap in time on one GPU (from two
// contexts or streams) race on it; the library does not serialise them.  The
// kernels overwrite each workspace in full before reading it, so nothing has
// to be cleared between calls.
#include <hip/hip_runtime.h>

#include <mutex>

#include "batch.hpp"
#include "common.hpp"

static __device__ __align__(256) unsigned char scratch[LITHO_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};    // device address of `scratch` on each GPU
int g_state[kMaxDevices] = {};    // 0: not resolved, 1: resolved, < 0: LITHO_E* of the lookup

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

}  // namespace

namespace litho {

int noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return LITHO_EARG;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_state[device] == 1) return LITHO_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        set_last_hip_error(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        g_state[device] = LITHO_EHIP;
        return LITHO_EHIP;
    }
    g_ptr[device] = p;
    g_state[device] = 1;
    return LITHO_OK;
}

}  // namespace litho

extern "C" int litho_socs_image_batch_noalloc(litho_ctx* c, const litho_bank* bank,
                                              const float* d_mask, int B, int N,
                                              float* d_image) {
    using namespace litho;
    if (!c || !bank || !d_mask || !d_image || B <= 0 || B > 65535) return LITHO_EARG;
    size_t sz[6];
    int band_path = 0, grid = 0;
    LITHO_TRY(image_batch_sizes(ctx_path(c), bank, B, N, sz, &band_path, &grid));
    size_t off[6], total = 0;
    for (int i = 0; i < 6; i++) {
        off[i] = total;
        total += align256(sz[i]);
    }
    if (total > (size_t)LITHO_NOALLOC_SCRATCH_BYTES) return LITHO_ESCRATCH;
    const int dev = ctx_device(c);
    if (dev < 0 || dev >= kMaxDevices) return LITHO_EARG;
    char* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (g_state[dev] != 1) return g_state[dev] < 0 ? g_state[dev] : LITHO_EHIP;
        base = (char*)g_ptr[dev];
    }
    void* ws[6] = {};
    for (int i = 0; i < 6; i++)
        if (sz[i]) ws[i] = base + off[i];
    return image_batch_enqueue(c, bank, d_mask, B, N, d_image, band_path, grid, ws, false, 0);
}
