// Internal interface of the mask -> image batch (api.cpp) shared with the
// no-allocation entry point (noalloc.hip).
#pragma once

#include "common.hpp"

namespace litho {

// workspace bytes of litho_socs_image_batch (six slots, unused ones 0),
// whether the band pipeline runs and, if not, the conv grid G of the other
// pipeline; LITHO_EARG for a shape it does not take
int image_batch_sizes(int path, const litho_bank* bank, int B, int N, size_t sz[6],
                      int* band_path, int* grid);
// the batch's launches on the context stream into the given workspaces; the
// kernel bank read as bank_fmt (0 float2, 1 bf16, 2 fp16, 3 bf16 hi + lo)
int image_batch_enqueue(litho_ctx* c, const litho_bank* bank, const float* d_mask, int B, int N,
                        float* d_image, int band_path, int grid, void* const ws[6], bool timing,
                        int bank_fmt);
int ctx_path(litho_ctx* c);
int ctx_device(litho_ctx* c);
// resolves this device's copy of the no-allocation scratch array (noalloc.hip);
// called by litho_ctx_create with the device current
int noalloc_init(int device);

}  // namespace litho
