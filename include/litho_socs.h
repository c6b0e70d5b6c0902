/*
 * litho_socs.h -- C-ABI of the MI355X SOCS aerial-image engine.
 *
 * Plain C, no HIP or torch types: every buffer is a pointer plus sizes.  The
 * entry points replace the reference's calcTCC / calcKernels / calcSOCS /
 * FT_r2c_padded / FI (validation/golden/src/litho_thread.cpp) and the HLS
 * accelerator top calc_socs_2048_hls_stream_refactored_v18
 * (source/SOCS_HLS/src/socs_2048_stream_refactored_v18_hls_fft.h L86-97).
 *
 * Conventions (identical to the reference):
 *   - row-major arrays; complex data as separate real / imaginary planes on
 *     the host side (the BIN layout), interleaved float2 on the device;
 *   - kernel index ID(nx,ny) = (ny+Ny)*(2Nx+1) + (nx+Nx), kernels stored
 *     index-flipped K[ID] = v[n-1-ID] (litho_thread.cpp L643-652);
 *   - mask spectrum centred, DC at [N/2][N/2], divided by N^2;
 *   - SOCS inverse FFT unnormalised, FI divides its R2C by G^2.
 *
 * Errors: every function returns an int status, LITHO_OK (0) or a negative
 * LITHO_E* code; nothing exits or throws across the boundary.  Host-pointer
 * entry points are synchronous and run on a default context of the calling
 * thread, created on first use and kept (stream, workspaces) until the
 * process ends: thread pools that come and go should use explicit contexts.
 * A litho_ctx owns one HIP stream and its workspaces: use one context per
 * host thread.
 */
#ifndef LITHO_SOCS_H
#define LITHO_SOCS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LITHO_OK 0
#define LITHO_EARG -1      /* bad size / null pointer / unsupported shape */
#define LITHO_ENOMEM -2    /* device or host allocation failed */
#define LITHO_EHIP -3      /* HIP runtime error */
#define LITHO_ENOCONV -4   /* eigensolver did not converge */
#define LITHO_ENODEV -5    /* no usable GPU */
#define LITHO_ERANGE -6    /* pupil support leaves the (2Nx+1)x(2Ny+1) window */
#define LITHO_ESCRATCH -7  /* request exceeds the no-allocation scratch */

/* flag bits returned through `int* flags` */
#define LITHO_FLAG_CLUSTER_SPLIT 1 /* nk cuts a degenerate eigenvalue cluster */
#define LITHO_FLAG_GRAM_ROUTE 2    /* kernels came from the source-space Gram route */

/* Optics and grid of one simulation (the reference's FullConfig subset used
 * by calcTCC, litho_thread.cpp L441-442).  Lx/Ly are the mask periods in nm
 * (the reference conflates period and pixel count; the two are separate here). */
typedef struct {
    float lambda;      /* nm */
    float NA;
    float defocus;     /* nm */
    float outerSigma;  /* getOuterSigma clamped to 1 (litho_thread.cpp L1343-1344) */
    int Lx, Ly;        /* period, nm */
    int Nx, Ny;        /* floor(NA*L*(1+sigma)/lambda) (litho_thread.cpp L1346-1347) */
} LithoOptics;

typedef struct litho_ctx litho_ctx;   /* device, stream, workspaces, twiddle tables */
typedef struct litho_bank litho_bank; /* one kernel set resident in HBM */

const char* litho_status_string(int status);
int litho_version(void);

/* ---- context and device memory (for callers without a HIP toolchain) ---- */
int litho_ctx_create(litho_ctx** ctx, int device);
int litho_ctx_destroy(litho_ctx* ctx);
int litho_ctx_sync(litho_ctx* ctx);
int litho_dev_alloc(litho_ctx* ctx, void** dptr, size_t bytes);
int litho_dev_free(litho_ctx* ctx, void* dptr);
int litho_memcpy_h2d(litho_ctx* ctx, void* dst, const void* src, size_t bytes);
int litho_memcpy_d2h(litho_ctx* ctx, void* dst, const void* src, size_t bytes);
int litho_memset_dev(litho_ctx* ctx, void* dst, int value, size_t bytes);
/* device-to-device copy on the context stream (synchronous) */
int litho_memcpy_d2d(litho_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---- Module 1 helpers (host) --------------------------------------------- */
/* createSource (common/source.cpp L105-141): annular L5-23, cross quadrupole
 * L25-55, dipole L58-87 (note its size/2 centre), point L122-129, each
 * followed by normalizeMatrix L89-103 (float sum).  src is srcSize^2 f32. */
int litho_source_annular(float* src, int srcSize, float inner, float outer);
int litho_source_quadrupole(float* src, int srcSize, float radius, float offset);
int litho_source_dipole(float* src, int srcSize, float radius, float offset, int onXAxis);
int litho_source_point(float* src, int srcSize, float px, float py);
/* getOuterSigma, litho_thread.cpp L427-439, clamped to 1 as main L1344. */
float litho_outer_sigma(const float* src, int srcSize);
/* Nx = floor(NA*period*(1+sigma)/lambda) in float, litho_thread.cpp L1346. */
int litho_compute_nx(float NA, int period, float outerSigma, float lambda);
/* Seeded synthetic Manhattan mask (no counterpart in the reference, which
 * only generates LineSpace, common/mask.cpp L5-29): axis-aligned rectangles
 * with sides U[wmin,wmax] pixels added until the 0/1 fill reaches `fill`.
 * RNG: PCG32 (O'Neill), state seeded from `seed`. */
int litho_manhattan_mask(float* mask, int N, unsigned long long seed, int wmin, int wmax,
                         float fill);

/* ---- Module 3: FT_r2c_padded, litho_thread.cpp L255-331 ------------------- */
/* Full centred spectrum of an N x N real mask (N even, <= 4096). */
int litho_mask_spectrum(const float* mask, int N, float* mskf_r, float* mskf_i);
/* The same for a rectangular Ly x Lx clip (the reference's fftLx x fftLy,
 * litho_thread.cpp L1269-1278): mask and spectrum [Ly][Lx], divided by Lx Ly,
 * Lx and Ly powers of two in 32..4096. */
int litho_mask_spectrum_rect(const float* mask, int Lx, int Ly, float* mskf_r, float* mskf_i);

/* ---- Module 4: calcTCC, litho_thread.cpp L441-532 ------------------------- */
/* TCC = A diag(S) A^H with the reference's float pupil arithmetic, built on
 * the GPU (FP64 MFMA contraction).  tcc_* are n x n, n = (2Nx+1)(2Ny+1). */
int litho_tcc_build(const LithoOptics* op, const float* src, int srcSize,
                    double* tcc_re, double* tcc_im);

/* ---- Module 6: calcKernels, litho_thread.cpp L597-663 ---------------------- */
/* Top-nk eigenpairs of a Hermitian TCC (upper triangle read, as zheevr 'U'),
 * FP64 on the GPU: Jacobi below n = 1024, block subspace iteration on the
 * matrix cores with a Rayleigh-Ritz step above (Jacobi again if the block
 * cannot hold the top clusters or H is not positive semi-definite).  scales descending (FP64; the reference casts to
 * f32 at L645), kernels flipped (nk x n f32 planes), degenerate clusters in
 * the canonical mirror basis, every vector phase-fixed.  flags may be NULL. */
int litho_kernels_topk(const double* tcc_re, const double* tcc_im, int n, int Nx, int Ny,
                       int nk, double* scales, float* krn_r, float* krn_i, int* flags);
/* Host wall-clock breakdown of the calling thread's last litho_tcc_build*,
 * litho_kernels_from_source* or litho_kernels_topk call (no counterpart in
 * the reference): out[0] pupil matrix ms, out[1] TCC / Gram contraction ms,
 * out[2] eigensolve and kernel extraction ms, out[3] eigensolver (0: Jacobi,
 * 1: block subspace iteration, used for n >= 1024), out[4] its iterations.
 * Writes min(n, 5) values and returns 5. */
int litho_setup_timing(double* out, int n);
/* Source -> kernels without forming the TCC on the host: explicit TCC route
 * when n <= nsrc, source-space Gram route (nsrc x nsrc eigenproblem, lifted)
 * otherwise.  Same outputs as litho_kernels_topk(litho_tcc_build(...)). */
int litho_kernels_from_source(const LithoOptics* op, const float* src, int srcSize, int nk,
                              double* scales, float* krn_r, float* krn_i, int* flags);

/* Precision form of the two calls above (no counterpart in the reference):
 * the TCC / Gram contraction in split BF16 on the matrix cores (every operand
 * as bf16 hi + lo, three products per real product, FP32 accumulate, FP64
 * sum over K slices); the eigensolve, the Gram eigenproblem and the lift stay
 * FP64.  tcc_prec LITHO_TCC_FP64 gives the plain results. */
#define LITHO_TCC_FP64 0
#define LITHO_TCC_SPLIT_BF16 1
int litho_tcc_build_prec(const LithoOptics* op, const float* src, int srcSize, int tcc_prec,
                         double* tcc_re, double* tcc_im);
int litho_kernels_from_source_prec(const LithoOptics* op, const float* src, int srcSize, int nk,
                                   int tcc_prec, double* scales, float* krn_r, float* krn_i,
                                   int* flags);

/* ---- Module 7: calcSOCS, litho_thread.cpp L813-1040 ------------------------ */
/* Conv grid of the accelerator-shaped entry along an axis of window
 * half-width n: the v18 top's fixed 128 (socs_2048_stream_refactored_v18_hls_fft.h
 * L34-35) for every n it accepts (n <= 8, MAX_KERNEL_SIZE = 17 at L43), and
 * above that max(128, nextPow2(4n+1)) (litho_thread.cpp L833-834), where v18
 * itself would drop the kernel rows and columns beyond 17.  n <= 255. */
int litho_tmpimg_grid(int n);
/* Drop-in for the HLS top calc_socs_2048_hls_stream_refactored_v18
 * (socs_2048_stream_refactored_v18_hls_fft.h L86-97, .cpp L287-388), same
 * pointer order and meaning: full Lx x Ly centred spectrum planes, kernels
 * nk x (2ny+1)(2nx+1) concatenated ([ky][kx], row length 2nx+1), scales[nk];
 * writes the fftshifted intensity tmpImgp (before FI) on the Gy x Gx grid,
 * Gx = litho_tmpimg_grid(nx), Gy = litho_tmpimg_grid(ny) -- 128 x 128 for
 * nx, ny <= 8, as v18 -- to tmpImg and to output (either may be NULL, they
 * may alias).  Lx/Ly are the spectrum sizes in pixels.  nx != ny is accepted.
 * FP32 arithmetic (v18 runs a block-floating-point IFFT). */
int litho_socs_tmpimg(const float* mskf_r, const float* mskf_i, const float* krn_r,
                      const float* krn_i, const float* scales, float* tmpImg, float* output,
                      int nk, int nx, int ny, int Lx, int Ly);
/* The same on a caller-chosen grid: Gx, Gy powers of two, Gx >= 2nx+1,
 * Gy >= 2ny+1 (the intensity sampled there; litho_thread.cpp's calcSOCS uses
 * nextPow2(4n+1) per axis, L833-834, which FI needs to be alias-free). */
int litho_socs_tmpimg_on_grid(const float* mskf_r, const float* mskf_i, const float* krn_r,
                              const float* krn_i, const float* scales, float* tmpImg,
                              float* output, int nk, int nx, int ny, int Lx, int Ly, int Gx,
                              int Gy);
/* Device-resident form for repeated calls (no counterpart in the reference:
 * the v18 testbench re-sends the kernels with every call): kernels uploaded
 * once as a bank by litho_bank_create with nx, ny, spectrum planes and
 * outputs in device memory, asynchronous on the context stream.  Gx, Gy as
 * litho_socs_tmpimg_on_grid (litho_tmpimg_grid gives the v18 grid). */
int litho_socs_tmpimg_dev(litho_ctx* ctx, const litho_bank* bank, const float* d_mskf_r,
                          const float* d_mskf_i, int Lx, int Ly, int Gx, int Gy, float* d_tmpImg,
                          float* d_output);
/* Mask -> aerial image: mask spectrum, SOCS loop and FI (Modules 3 + 7). */
int litho_socs_image(const float* mask, int N, const float* krn_r, const float* krn_i,
                     const float* scales, int nk, int Nx, int Ny, float* image);
/* The same on a rectangular Ly x Lx clip (calcSOCS with fftLx != fftLy,
 * L813-1040): mask and image [Ly][Lx], Lx and Ly powers of two in 32..4096, the
 * SOCS loop on one square grid G = max(32, nextPow2(4 max(Nx, Ny) + 1)) <=
 * min(1024, Lx, Ly) (the image does not depend on the grid once it holds the
 * band). */
int litho_socs_image_rect(const float* mask, int Lx, int Ly, const float* krn_r,
                          const float* krn_i, const float* scales, int nk, int Nx, int Ny,
                          float* image);
/* FI alone, litho_thread.cpp L692-772: G x G (zero frequency centred, as
 * tmpImgp) -> N x N band-limited interpolation. */
int litho_fi(const float* tmpImgp, int G, int N, int band, float* image);

/* ---- Module 5 (verification): calcImage + FT_c2r_padded ------------------- */
/* TCC-direct aerial image (litho_thread.cpp L539-583, L371-425): the full-rank
 * reference image SOCS converges to as nk -> n.  N a power of two, 32..4096,
 * 4 Nx + 1 <= N and 4 Ny + 1 <= N, tcc_* n x n with n = (2Nx+1)(2Ny+1) (both
 * triangles read; non-finite entries are rejected). */
int litho_image_tcc(const float* mask, int N, const double* tcc_re, const double* tcc_im, int Nx,
                    int Ny, float* image);
/* The same on a rectangular Ly x Lx clip (calcImage's fftLx x fftLy): mask and
 * image [Ly][Lx], Lx and Ly powers of two in 32..4096. */
int litho_image_tcc_rect(const float* mask, int Lx, int Ly, const double* tcc_re,
                         const double* tcc_im, int Nx, int Ny, float* image);

/* ---- device-resident batched path (the bench path) ------------------------- */
/* Upload one kernel set (f32 planes as in kernels/krn_k_{r,i}.bin, each
 * (2Ny+1) x (2Nx+1); Nx != Ny is stored zero-padded to the square window of
 * half-width max(Nx, Ny), which leaves every product and field unchanged).  The
 * mask -> image paths below run a second form of the set: a kernel K with
 * K(-f) = e^{ib} conj(K(f)) about the window centre (zero defocus with a
 * point-symmetric source gives such kernels) has a real field on a real
 * mask, so two of them share one transform as sqrt(la) Ka + i sqrt(lb) Kb;
 * the other kernels keep their own entries.  litho_socs_tmpimg (a given
 * spectrum) uses the set as given. */
int litho_bank_create(litho_ctx* ctx, const float* krn_r, const float* krn_i,
                      const float* scales, int nk, int Nx, int Ny, litho_bank** bank);
int litho_bank_destroy(litho_bank* bank);
/* Entries of the mask -> image form of the set and how many are kernel pairs
 * (entries = nk - pairs).  No counterpart in the reference. */
int litho_bank_info(const litho_bank* bank, int* entries, int* pairs);
/* B masks (device, B x N x N f32) -> B images (device), asynchronous on the
 * context stream.  Workspaces are grown by the context as needed. */
int litho_socs_image_batch(litho_ctx* ctx, const litho_bank* bank, const float* d_mask, int B,
                           int N, float* d_image);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  litho_socs_image_batch with its workspaces in a fixed
 * device array of LITHO_NOALLOC_SCRATCH_BYTES (64 MiB) that the library holds
 * once per GPU until the process ends: same launches, same images bit for
 * bit, nothing allocated or freed.  LITHO_ESCRATCH when the batch's
 * workspaces do not fit (about 0.99 MiB per clip at 1024^2, Nx = 34, so
 * B <= 64; 63 KiB per clip at 256^2, Nx = 8).  All calls on one GPU share
 * the array: calls that overlap in time on one GPU (two contexts or streams)
 * race and must be serialised by the caller.  Records no per-stage timing. */
#define LITHO_NOALLOC_SCRATCH_BYTES (64u << 20)
int litho_socs_image_batch_noalloc(litho_ctx* ctx, const litho_bank* bank, const float* d_mask,
                                   int B, int N, float* d_image);
/* Conv-grid choice of the image path.  AUTO: smallest square grid of
 * 2^a 3^b points (36, 81, 144) >= 4Nx+1 with the intensity's band spectrum
 * taken inside the SOCS kernel, else the reference's nextPow2(4Nx+1) grid
 * (litho_thread.cpp L833).  FI makes the image independent of the grid. */
#define LITHO_PATH_AUTO 0
#define LITHO_PATH_POW2 1
int litho_ctx_set_path(litho_ctx* ctx, int path);
/* Per-stage HIP events around the launches of litho_socs_image_batch (on by
 * default; off, a batch records none and litho_last_timing reports no
 * timing).  No counterpart in the reference. */
int litho_ctx_set_timing(litho_ctx* ctx, int on);
/* Timing on the context stream of the last litho_socs_image_batch call:
 * total and per-stage device time (ms).  Returns the number of stages (>= 0)
 * or a negative status. */
int litho_last_timing(litho_ctx* ctx, float* total_ms, float* stage_ms, int nstage);
/* Name of stage i of the last batch call (static string, "" if none). */
const char* litho_stage_name(litho_ctx* ctx, int i);

/* Batched form (no counterpart in the reference): the batch is split into
 * `chunks` sub-batches spread over up to four streams (sub-batch k on stream
 * k mod 4, its five launches in order) so that the compute-bound SOCS of one
 * sub-batch shares CUs with the memory-bound FFTs of another (smaller blocks
 * are used when chunks > 1), and the whole step is captured once as a HIP
 * graph (per shape and buffers) and replayed.
 * `precision` is a set of flags:
 *  - INTER_FP16 stores the row-DFT window intermediate as fp16 (and the FI
 *    column output too, unless DFT_MFMA is set: the image-side product splits
 *    that FP32 output itself);
 *  - BANK_BF16 / BANK_FP16 read the kernel bank as 16-bit pairs with one FP32
 *    scale per entry for the real and one for the imaginary parts; BANK_BF16X2
 *    as bf16 hi + lo pairs (hi = bf16(x), lo = bf16(x - hi)) with the same
 *    scales, 16 significant bits (wave-synced SOCS blocks only);
 *  - DFT_MFMA computes the mask-row and image-row band DFTs as fp16 matrix
 *    products with FP32 accumulation on the matrix cores (the image side
 *    splits its input into hi + lo fp16 parts); the other stages stay FP32;
 *  - MASK_U8 (requires DFT_MFMA): d_mask holds uint8 levels, mask = level/255
 *    (the PNG -> BIN convention); MASK_1BIT (requires DFT_MFMA, not with
 *    MASK_U8): d_mask holds binary masks one bit per pixel, N / 8 bytes per
 *    row, pixel 8j + i at bit i of byte j (numpy packbits, bitorder little);
 *    IMAGE_FP16 (requires DFT_MFMA): d_image receives fp16 values.
 * LITHO_PREC_MIXED is the storage-only set that meets 1e-3 max relative error
 * against FP64; LITHO_PREC_COMPACT the matrix-core set with 8-bit masks, an
 * fp16 row-DFT window and fp16 images (6-7.5e-4 at 1024^2, K=24).
 * Uses the band conv grid (N in {256, 512, 768, 1024}, 4Nx+1 <= 144); other
 * clips (up to 4096^2) run the reference's nextPow2 grid in one sub-batch,
 * where the bank formats (BANK_BF16 / BANK_FP16 / BANK_BF16X2, on conv grids
 * of 512 points and up) are the only flags taken. */
#define LITHO_PREC_FP32 0
#define LITHO_PREC_INTER_FP16 1
#define LITHO_PREC_BANK_BF16 2
#define LITHO_PREC_BANK_FP16 4
#define LITHO_PREC_MASK_U8 8
#define LITHO_PREC_IMAGE_FP16 16
#define LITHO_PREC_DFT_MFMA 32
#define LITHO_PREC_BANK_BF16X2 64
#define LITHO_PREC_MASK_1BIT 128
#define LITHO_PREC_MIXED (LITHO_PREC_INTER_FP16 | LITHO_PREC_BANK_FP16)
#define LITHO_PREC_COMPACT                                                              \
    (LITHO_PREC_INTER_FP16 | LITHO_PREC_BANK_FP16 | LITHO_PREC_MASK_U8 | LITHO_PREC_DFT_MFMA | \
     LITHO_PREC_IMAGE_FP16)
int litho_socs_image_batch_pipelined(litho_ctx* ctx, const litho_bank* bank, const void* d_mask,
                                     int B, int N, void* d_image, int precision, int chunks,
                                     int use_graph);

/* Mixed clip sizes (no counterpart in the reference): one bucket per clip
 * size, each with its own kernel bank (optics of that period), all buckets
 * in one step -- their launches spread over up to four streams and, with
 * use_graph, captured once as one HIP graph and replayed.  Precision flags as
 * litho_socs_image_batch_pipelined; N in {256, 512, 768, 1024}.  Each bucket
 * runs under the plan litho_autotune chose for its shape (default: the
 * smallest band grid, wave-synced SOCS blocks, 256-thread FFT blocks). */
typedef struct {
    const litho_bank* bank;  /* kernel set of this clip size */
    int N;                   /* clip size in pixels */
    int count;               /* clips in the bucket (0 = skip) */
    const void* d_masks;     /* count x N x N: f32, uint8 levels (MASK_U8) or bits (MASK_1BIT) */
    void* d_images;          /* count x N x N: f32, or fp16 with IMAGE_FP16 */
} LithoBucket;
int litho_socs_image_buckets(litho_ctx* ctx, const LithoBucket* buckets, int nbuckets,
                             int precision, int use_graph);
/* Device time (ms) from the start of the last eager (use_graph = 0) bucket
 * step to the end of each bucket's last launch; -1 for a skipped bucket.
 * Returns the number of buckets written. */
int litho_bucket_times(litho_ctx* ctx, float* ms, int nbuckets);
/* Autotuner: for every bucket shape (N, Nx, nk, precision) not tuned yet on
 * this context, time each candidate plan -- conv grid 36/81/144 (SOCS FFT
 * radix 6/9/12), SOCS block shape (0: 48 transform slots, 1: 24 slots at
 * grid 144, 2: transforms aligned to wavefronts, wave-level syncs), FFT
 * row/column blocks of 256 or 128 threads -- on up to 32 of the bucket's clips
 * (their images are written) and keep the fastest.  Returns the number of
 * shapes tuned. */
int litho_autotune(litho_ctx* ctx, const LithoBucket* buckets, int nbuckets, int precision,
                   int reps);
/* Pin the plan of a shape without tuning (the candidate set, plus SOCS
 * block shape 3: shape 2 with each clip's kernel entries split over two
 * blocks and their band spectra added by the FI column pass, which halves the
 * SOCS latency of batches much smaller than the GPU's 256 CUs; LITHO_EARG for
 * a combination that is not built). */
int litho_plan_set(litho_ctx* ctx, int N, int Nx, int nk, int precision, int G, int socs_shape,
                   int lean_mem);
/* The plan used for a shape: returns 1 if tuned or set, 0 for the default. */
int litho_plan_get(litho_ctx* ctx, int N, int Nx, int nk, int precision, int* G, int* socs_shape,
                   int* lean_mem);

/* ---- learned surrogate (no counterpart in the reference) ------------------
 * A Fourier neural operator mask -> aerial image (width 32, 16 x 16 retained
 * modes per sign of ky, GELU, 4 Fourier layers, 128-channel projection),
 * trained on the GPU by its own forward / backward kernels.  Weights are
 * FP32 masters read through a bf16 copy.  N: power of two, 32..256.  Masks
 * and images are device f32 [B][N][N] on the context stream. */
typedef struct litho_fno litho_fno;
int litho_fno_create(litho_ctx* ctx, int N, unsigned long long seed, litho_fno** out);
int litho_fno_destroy(litho_fno* f);
long long litho_fno_param_count(const litho_fno* f);
/* which = 0: FP32 weights, 1: gradients of the last litho_fno_loss_grad */
int litho_fno_get_params(litho_fno* f, float* host, long long n, int which);
int litho_fno_set_params(litho_fno* f, const float* host, long long n);
/* Inference on the matrix cores: bf16 weights (exact) times activations and
 * cos / sin tables split into bf16 hi + lo parts, FP32 accumulate, one fused
 * kernel per Fourier layer (B <= 65535). */
int litho_fno_forward(litho_fno* f, const float* d_masks, int B, float* d_images);
/* The same map on the vector ALUs in FP32 arithmetic: the training forward. */
int litho_fno_forward_fp32(litho_fno* f, const float* d_masks, int B, float* d_images);
/* SOCS as judge and fallback of served clips, on the device: clip j of the
 * check is d_images + j * clip_stride * N^2 against d_socs + j * N^2 (n
 * clips).  errs[j] (host) = max |image - socs| / max |socs| (0 when both are
 * all zero); every clip with errs[j] > tol, or NaN, is overwritten with its
 * SOCS image.  Returns the number of clips replaced (>= 0) or a status. */
int litho_fno_verify(litho_fno* f, float* d_images, int clip_stride, const float* d_socs, int n,
                     float tol, float* errs);
/* MSE loss over the batch and the gradients of every weight (synchronous
 * when loss is non-NULL) */
int litho_fno_loss_grad(litho_fno* f, const float* d_masks, const float* d_targets, int B,
                        float* loss);
int litho_fno_adam(litho_fno* f, float lr, float beta1, float beta2, float eps);

#ifdef __cplusplus
}
#endif
#endif
