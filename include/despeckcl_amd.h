/*
 * despeckcl_amd.h -- C ABI of the MI355X NL-SAR / NL-InSAR library (libdespeckcl_amd.so).
 *
 * This is the drop-in boundary for despeckCL's nonlocal filters.  Every entry point takes
 * plain pointers and sizes; nothing here depends on C++ or on PyTorch.  The C++ API of
 * include/despeckcl.h and the Python module despeckcl_amd are thin layers over it.
 *
 * Conventions
 *  - images are row-major height x width float32 planes, as in the reference
 *    (swig/python/despeckcl_typemaps.i:15-24); covariance matrices are 2*dim*dim planes,
 *    real and imaginary parts interlaced per matrix element in row-major (r, c) order
 *    (swig/python/nlsar.i:146-153);
 *  - host-pointer entry points copy in, filter on the current HIP device and copy out,
 *    returning when the result is in the caller's arrays;
 *  - *_device entry points take device pointers and a hipStream_t (passed as void*),
 *    enqueue the work on that stream and return without synchronising;
 *  - return value 0 on success, a negative DKX_E* code otherwise; dkx_last_error() gives
 *    the message of the calling thread's last failure.  (The reference terminates the
 *    process on device errors, src/compute/nlsar/nlsar_filter_sub_image.cpp:74-78.)
 */
#ifndef DESPECKCL_AMD_H
#define DESPECKCL_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define DKX_OK 0
#define DKX_EINVAL (-1)       /* invalid sizes or parameters */
#define DKX_ERESOURCE (-2)    /* parameters exceed a kernel resource (LDS) or memory */
#define DKX_ESTATS (-3)       /* statistics keys do not match the patch x scale set */
#define DKX_EUNSUPPORTED (-4) /* valid in the reference, not supported here */
#define DKX_EHIP (-5)         /* HIP runtime error */

/* One entry of std::map<nlsar::params, nlsar::stats>
 * (include/compute/nlsar/parameters.h:24-56, include/compute/nlsar/stats.h:27-55). */
typedef struct {
    int patch_size;
    int scale_size;
    int lut_size;
    float dissims_min;
    float dissims_max;
    const float* quantilles;  /* lut_size floats */
    const float* chi2cdf_inv; /* lut_size floats */
} dkx_nlsar_stats;

const char* dkx_last_error(void);
int dkx_version(void);

/* despeckcl::nlsar, InSAR overload (include/despeckcl.h:45-59; src/compute/nlsar/nlsar.cpp:139-167).
 * The statistics must hold exactly one entry per (patch size, scale size) pair. */
int dkx_nlsar_insar(const float* ampl_master, const float* ampl_slave, const float* phase,
                    float* ref_filt, float* phase_filt, float* coh_filt, int height, int width,
                    int search_window_size, const int* patch_sizes, int n_patch_sizes,
                    const int* scale_sizes, int n_scale_sizes, const dkx_nlsar_stats* stats,
                    int n_stats, float h_param, float c_param);

/* Same filter on device buffers, enqueued on `stream` (a hipStream_t; NULL = default).
 * enl_maps / alpha_maps: optional device buffers of n_patch*n_scale*height*width floats that
 * receive the bias-corrected ENL and alpha of every (patch, scale) pair, parameter index
 * = scale_index*n_patch_sizes + patch_index.  ub_pixels: optional host pointer; when given
 * the call synchronises and stores the number of pixels for which the reference's best
 * parameter selection is undefined (src/compute/nlsar/best_params.cpp:27-53). */
int dkx_nlsar_insar_device(const float* d_ampl_master, const float* d_ampl_slave,
                           const float* d_phase, float* d_ref_filt, float* d_phase_filt,
                           float* d_coh_filt, int height, int width, int search_window_size,
                           const int* patch_sizes, int n_patch_sizes, const int* scale_sizes,
                           int n_scale_sizes, const dkx_nlsar_stats* stats, int n_stats,
                           float h_param, float c_param, float* d_enl_maps, float* d_alpha_maps,
                           unsigned long long* ub_pixels, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc): the same filter
 * and outputs as dkx_nlsar_insar_device, with every scratch buffer (padded covariance and
 * scale-average grids, flags, the undefined-selection count, the weight LUTs) in one
 * 64 MiB device array of the library, one per device, held until the process ends; the call
 * allocates nothing.  Images up to about 1200 x 1200 pixels fit (search window 21, patch
 * sizes up to 7, scale sizes up to 3); larger ones return DKX_ERESOURCE.  The array keeps the
 * last call's contents and is shared by every call on the device: calls that overlap on one
 * device race, so the caller orders them (one stream, or events). */
int dkx_nlsar_insar_device_noalloc(const float* d_ampl_master, const float* d_ampl_slave,
                                   const float* d_phase, float* d_ref_filt, float* d_phase_filt,
                                   float* d_coh_filt, int height, int width,
                                   int search_window_size, const int* patch_sizes,
                                   int n_patch_sizes, const int* scale_sizes, int n_scale_sizes,
                                   const dkx_nlsar_stats* stats, int n_stats, float h_param,
                                   float c_param, float* d_enl_maps, float* d_alpha_maps,
                                   unsigned long long* ub_pixels, void* stream);

/* despeckcl::nlsar, covariance overload (include/despeckcl.h:61-72;
 * src/compute/nlsar/nlsar.cpp:170-193).  covmat / covmat_filt: 2*dim*dim planes of
 * height x width, real and imaginary parts of each matrix element interlaced in row-major
 * element order (swig/python/nlsar.i:196-201); every plane of the result is the
 * alpha-blended weighted mean of the selected (patch, scale) pair.  dim 2 or 3 (the
 * reference's dissimilarities exist for those two, src/compute/nlsar/sim_measures.h:37-158). */
int dkx_nlsar_covmat(const float* covmat, float* covmat_filt, int height, int width, int dim,
                     int search_window_size, const int* patch_sizes, int n_patch_sizes,
                     const int* scale_sizes, int n_scale_sizes, const dkx_nlsar_stats* stats,
                     int n_stats, float h_param, float c_param);

/* Same on device buffers, enqueued on `stream`; the optional outputs as for
 * dkx_nlsar_insar_device. */
int dkx_nlsar_covmat_device(const float* d_covmat, float* d_covmat_filt, int height, int width,
                            int dim, int search_window_size, const int* patch_sizes,
                            int n_patch_sizes, const int* scale_sizes, int n_scale_sizes,
                            const dkx_nlsar_stats* stats, int n_stats, float h_param,
                            float c_param, float* d_enl_maps, float* d_alpha_maps,
                            unsigned long long* ub_pixels, void* stream);

/* despeckcl::nlinsar (include/despeckcl.h:97-109; src/compute/nlinsar/nlinsar.cpp:42-146).
 * ref_filt / phase_filt / coh_filt are read as the first iteration's estimates and
 * overwritten with the result, like the reference (nlinsar.cpp:59-61).  Odd patch sizes up
 * to 21.  h: 0.58 for patch 3 and 0.244 for patch 7 as tabulated by the reference; for the
 * others the reference's Monte-Carlo quantile (insarsim_simu.cpp:62-82), which it seeds from
 * std::random_device, here with a fixed seed (dkx_nlinsar_simulate_h with seed 5489). */
int dkx_nlinsar(const float* ampl_master, const float* ampl_slave, const float* phase,
                float* ref_filt, float* phase_filt, float* coh_filt, int height, int width,
                int search_window_size, int patch_size, int niter, int lmin);

/* nlinsar::simu::quantile_insar(patch_size, 0.92) (src/compute/nlinsar/insarsim_simu.cpp:62-82):
 * the h_theo the reference uses for patch sizes other than 3 and 7, simulated with a
 * std::mt19937 seeded with `seed` in place of the reference's std::random_device.  Host
 * only. */
int dkx_nlinsar_simulate_h(int patch_size, unsigned int seed, float* h_theo);

/* Same on device buffers.  n_smoothed: optional host pointer (synchronises) receiving the
 * number of pixel-iterations that went through the lmin smoothing. */
int dkx_nlinsar_device(const float* d_ampl_master, const float* d_ampl_slave,
                       const float* d_phase, float* d_ref_filt, float* d_phase_filt,
                       float* d_coh_filt, int height, int width, int search_window_size,
                       int patch_size, int niter, int lmin, unsigned long long* n_smoothed,
                       void* stream);

/* despeckcl::boxcar (include/despeckcl.h:34-43; src/compute/boxcar/boxcar.cpp:35-92,
 * cl_kernels/boxcar_kernel.cl:19-69): w x w window average of the interferogram.  The last
 * row / column carry the reference's tile-edge values (the value at the last interior row /
 * column of its last 512-pixel tile, computed on the edge-replicated image).  Odd window
 * widths up to 15, the reference's range. */
int dkx_boxcar(const float* ampl_master, const float* ampl_slave, const float* phase,
               float* ref_filt, float* phase_filt, float* coh_filt, int height, int width,
               int window_width);

/* Same on device buffers, enqueued on `stream`. */
int dkx_boxcar_device(const float* d_ampl_master, const float* d_ampl_slave,
                      const float* d_phase, float* d_ref_filt, float* d_phase_filt,
                      float* d_coh_filt, int height, int width, int window_width, void* stream);

/* despeckcl::goldstein (include/despeckcl.h:111-122; src/compute/goldstein/goldstein.cpp:39-106,
 * goldstein_filter_sub_image.cpp:73-187).  ref_filt / phase_filt receive |z| and arg z of the
 * filtered interferogram; coh_filt is not written, as in the reference.  Reproduces the
 * reference: the filtered patches are added onto the raw interferogram (so alpha = 0 gives
 * |z| = a1 + a2), the tile is the reference's single-device tile, and the last row / column
 * carry its tile-edge values.  Patch sizes 16, 32 and 64; 0 <= overlap <= patch_size / 4 (at
 * most two patches overlap, the case the reference's packing passes handle). */
int dkx_goldstein(const float* ampl_master, const float* ampl_slave, const float* phase,
                  float* ref_filt, float* phase_filt, float* coh_filt, int height, int width,
                  int patch_size, int overlap, float alpha);

/* Same on device buffers, enqueued on `stream`. */
int dkx_goldstein_device(const float* d_ampl_master, const float* d_ampl_slave,
                         const float* d_phase, float* d_ref_filt, float* d_phase_filt,
                         float* d_coh_filt, int height, int width, int patch_size, int overlap,
                         float alpha, void* stream);

/* despeckcl::nlsar_training, InSAR overload (include/despeckcl.h:75-83;
 * src/compute/nlsar/training/nlsar_training.cpp:57-79).  Writes n_patch*n_scale entries to
 * `out` in std::map order (patch size, then scale size); their LUT arrays point into
 * `lut_storage`, which must hold n_patch*n_scale*2*256 floats. */
int dkx_nlsar_training_insar(const float* ampl_master, const float* ampl_slave,
                             const float* phase, int height, int width, const int* patch_sizes,
                             int n_patch_sizes, const int* scale_sizes, int n_scale_sizes,
                             dkx_nlsar_stats* out, float* lut_storage);

/* despeckcl::nlsar_training, covariance overload (include/despeckcl.h:85-92).  Both
 * training entry points run on the current HIP device. */
int dkx_nlsar_training_covmat(const float* covmat, int height, int width, int dim,
                              const int* patch_sizes, int n_patch_sizes, const int* scale_sizes,
                              int n_scale_sizes, dkx_nlsar_stats* out, float* lut_storage);

/* Same on device covariance planes (2*dim*dim interlaced planes of height x width), on
 * `stream`; the statistics are written to host memory, so the call synchronises the stream. */
int dkx_nlsar_training_covmat_device(const float* d_covmat, int height, int width, int dim,
                                     const int* patch_sizes, int n_patch_sizes,
                                     const int* scale_sizes, int n_scale_sizes,
                                     dkx_nlsar_stats* out, float* lut_storage, void* stream);

/* nlsar::stats::stats(dissims, lut_size) (src/compute/nlsar/stats.cpp:28-80): quantile LUT
 * and inverse chi-square(49) LUT.  Returns the number of finite dissimilarities (0: none). */
long long dkx_nlsar_stats_build(const float* dissims, long long n, int lut_size, float* dissims_min,
                                float* dissims_max, float* quantilles, float* chi2cdf_inv);

/* despeckcl::store_nlsar_stats_collection (include/despeckcl.h:94;
 * src/compute/nlsar/training/nlsar_training.cpp:99-104): writes the entries in the reference's
 * file format, a Boost text archive of std::map<nlsar::params, nlsar::stats> (library
 * version 17), in map order. */
int dkx_nlsar_stats_store(const char* filename, const dkx_nlsar_stats* stats, int n_stats);

/* despeckcl::load_nlsar_stats_collection (include/despeckcl.h:95;
 * nlsar_training.cpp:106-113): reads a Boost text archive as written by the reference (or
 * by dkx_nlsar_stats_store).  Returns the number of entries, or a negative DKX_E* code.
 * *n_floats receives the LUT floats the entries need; when out / lut_storage are non-null
 * and max_entries / max_floats suffice, the entries are written to out in map order with
 * their LUT arrays in lut_storage (otherwise nothing is written). */
long long dkx_nlsar_stats_load(const char* filename, dkx_nlsar_stats* out, int max_entries,
                               float* lut_storage, long long max_floats, long long* n_floats);

/* Inverse CDF of the chi-square distribution (the reference uses GSL's gsl_cdf_chisq_Pinv,
 * src/compute/nlsar/stats.cpp:76). */
double dkx_chi2_inv(double p, double nu);

/* Tile shape of the single-call entry points.  NL-SAR: 0, 1, 2 = 32 x 32 output tiles with
 * 4, 2 or 1 rows per thread; -1 (default) = per pass, 2 for a pass of one patch size (the
 * default selected-key pass) and 1 for several (search windows up to 21; larger ones use the
 * 4-row variant whatever the setting).  NL-InSAR: 0 = 27 x 54 and 1 = 18 x 54 (512 threads),
 * 2 = 54 x 54 (1024 threads), -1 (default) = the autotuner's record for the parameters and
 * image size, else 1.  Results do not depend on the tile shape. */
int dkx_set_tile(int filter /* 0 NL-SAR, 1 NL-InSAR */, int tile_index);
int dkx_get_tile(int filter);

/* No counterpart in the reference, whose weights kernels write every pixel's sw^2 weights and
 * whose host smoothing (src/compute/nlinsar/smoothing.cpp:39-76) reads back those of the
 * pixels with ENL < lmin.  Which weights the NL-InSAR weights launches store for the lmin
 * smoothing: -1 automatic (the default), 0 every weight,
 * 1 those of the pixels the previous iteration smoothed and their neighbours, 2 none.  The
 * weights of a smoothed pixel that were not stored are recomputed, the same bits, so every
 * mode gives the same results; the mode changes only the memory traffic.  A batch
 * (dkx_nlinsar_batch_set) and a server (dkx_server_create) keep the mode in force when they
 * are set up. */
int dkx_set_nlinsar_weight_store(int mode);
int dkx_get_nlinsar_weight_store(void);

/* ---- autotuner -----------------------------------------------------------------------------
 * No counterpart in the reference, whose tile_size (src/compute/nlsar/tile_size.cpp:31-114)
 * picks host tile dimensions from the device memory size alone; here every image is one
 * device call and the choice that matters is the kernels' tile variant.  The entry points
 * below run the filter on the caller's image once per variant (one warm-up run, then `reps`
 * timed runs each), record the fastest (NL-SAR: among variants within 2 % of it, the one with
 * the fewest rows per thread, which moves the fewest bytes from beyond L2) for the parameter
 * set and the image size (log2 of the
 * pixel count) and return it; ms (optional, 3 entries) receives each variant's mean time, -1
 * for a variant the parameters do not allow.  While dkx_set_tile holds -1 for the filter,
 * single calls, batches (keyed by their total pixel count) and the server use a recorded
 * choice where one exists, else the built-in rule.  The outputs hold the filter's result
 * afterwards (every variant gives the same bits); NL-InSAR restores the first-iteration
 * estimates in ref/phase/coh_filt before every run. */
int dkx_nlsar_insar_device_autotune(const float* d_ampl_master, const float* d_ampl_slave,
                                    const float* d_phase, float* d_ref_filt, float* d_phase_filt,
                                    float* d_coh_filt, int height, int width, int search_window_size,
                                    const int* patch_sizes, int n_patch_sizes, const int* scale_sizes,
                                    int n_scale_sizes, const dkx_nlsar_stats* stats, int n_stats,
                                    float h_param, float c_param, int every_key, int reps, void* stream,
                                    int* tile, float* ms);
int dkx_nlinsar_device_autotune(const float* d_ampl_master, const float* d_ampl_slave,
                                const float* d_phase, float* d_ref_filt, float* d_phase_filt,
                                float* d_coh_filt, int height, int width, int search_window_size,
                                int patch_size, int niter, int lmin, int reps, void* stream, int* tile,
                                float* ms);
/* The recorded variant for a filter (0 NL-SAR: every_key selects the every-key records; 1
 * NL-InSAR: patch_sizes[0] is the patch size, scale sizes unused), -1 if none. */
int dkx_get_tuned_tile(int filter, int height, int width, int search_window_size,
                       const int* patch_sizes, int n_patch_sizes, const int* scale_sizes,
                       int n_scale_sizes, int every_key);
/* Forgets every record. */
void dkx_clear_tuned_tiles(void);

/* Stage timing.  When enabled, the main launch of every fused stage is bracketed by HIP
 * events on the caller's stream.  dkx_stage_times() synchronises with those events and
 * returns the stages of the calls since the previous dkx_stage_times() call: kinds
 * 1 = NL-SAR scale pass, 2 = NL-SAR last scale pass (statistics, weighted mean, outputs),
 * 11 = NL-InSAR weights, 12 = NL-InSAR smoothing, 13 = NL-InSAR weights of the last row /
 * column in thin blocks (batches), 14 = NL-InSAR recomputation of unstored weights; durations
 * in milliseconds.  Returns the
 * number of entries (entries beyond max_entries are dropped). */
int dkx_set_timing(int enabled);
int dkx_stage_times(int* kinds, float* ms, int max_entries);

/* Releases the device workspace cached by the library on the current device. */
void dkx_release_workspace(void);

/* ---- batched NL-SAR ----------------------------------------------------------------------
 * No counterpart in the reference, which filters one image per call (despeckcl::nlsar,
 * include/despeckcl.h:45-59) in fixed tiles of up to 2048^2 pixels, one tile after another
 * (src/compute/nlsar/nlsar.cpp:81-136, map_filter_tiles.h:37-104).  For callers that filter
 * many interferograms of any size with the same parameters: every stage runs as one launch
 * over the tiles of all images, and the whole stage sequence can be captured in a HIP graph.
 * Each image's result equals the result of dkx_nlsar_insar_device on that image alone (every
 * pass variant gives the same bits). */

/* One image of a batch: device pointers, row-major height x width float32 planes. */
typedef struct {
    const float* ampl_master;
    const float* ampl_slave;
    const float* phase;
    float* ref_filt;
    float* phase_filt;
    float* coh_filt;
    int height;
    int width;
} dkx_nlsar_image;

/* One image of a covariance batch: device pointers to 2*dim*dim float32 planes of
 * height x width, interlaced as for dkx_nlsar_covmat. */
typedef struct {
    const float* covmat;
    float* covmat_filt;
    int height;
    int width;
} dkx_nlsar_covmat_image;

typedef struct dkx_nlsar_batch_s* dkx_nlsar_batch;

int dkx_nlsar_batch_create(dkx_nlsar_batch* batch);
void dkx_nlsar_batch_destroy(dkx_nlsar_batch batch);
/* (Re)defines the images and the parameters (same meaning and checks as dkx_nlsar_insar).
 * Waits for the batch's previous run, grows its device workspace if needed and uploads the
 * image table and the weight LUTs (synchronous).  The pass variant is the one the autotuner
 * chose for this parameter set and total pixel count, or 1. */
int dkx_nlsar_batch_set(dkx_nlsar_batch batch, const dkx_nlsar_image* images, int n_images,
                        int search_window_size, const int* patch_sizes, int n_patch_sizes,
                        const int* scale_sizes, int n_scale_sizes, const dkx_nlsar_stats* stats,
                        int n_stats, float h_param, float c_param);
/* The same for covariance input of dimension 2 or 3 (dkx_nlsar_covmat's meaning and checks);
 * every image of the batch has that dimension.  A covariance batch has one pass variant per
 * dimension, so _set_tile and _autotune leave it unchanged (_autotune reports -1). */
int dkx_nlsar_batch_set_covmat(dkx_nlsar_batch batch, const dkx_nlsar_covmat_image* images,
                               int n_images, int dim, int search_window_size,
                               const int* patch_sizes, int n_patch_sizes, const int* scale_sizes,
                               int n_scale_sizes, const dkx_nlsar_stats* stats, int n_stats,
                               float h_param, float c_param);
/* Enqueues the whole batch on `stream` (a hipStream_t); capturable in a HIP graph. */
int dkx_nlsar_batch_enqueue(dkx_nlsar_batch batch, void* stream);
/* Captures the batch's stage sequence into a HIP graph on first use (after each _set), then
 * replays it on `stream` (not the null stream). */
int dkx_nlsar_batch_graph_launch(dkx_nlsar_batch batch, void* stream);
/* Pass variant: -1 = the autotuner's choice for this batch shape, else the automatic per-pass
 * choice of dkx_set_tile; 0/1/2 = 4/2/1 rows per thread.  _get_tile returns the variant in
 * use, -1 for the automatic one. */
int dkx_nlsar_batch_set_tile(dkx_nlsar_batch batch, int tile);
/* on != 0: later enqueues compute the statistics of every (patch, scale) key, which
 * dkx_nlsar_batch_ub_pixels needs; off (the default): only the key the output uses.  The
 * filtered images are the same either way. */
int dkx_nlsar_batch_set_diagnostics(dkx_nlsar_batch batch, int on);
int dkx_nlsar_batch_get_tile(dkx_nlsar_batch batch);
/* Times every pass variant on the batch (reps runs each, after one warm-up run), records the
 * fastest (within 2 %, the fewest rows per thread) for this parameter set and total pixel
 * count, and selects it.  ms (optional)
 * receives the mean time of each variant. */
int dkx_nlsar_batch_autotune(dkx_nlsar_batch batch, void* stream, int reps, int* tile,
                             float* ms);
/* Waits for the last run and reads each image's count of pixels where the reference's best
 * parameter selection is undefined (see dkx_nlsar_insar_device). */
int dkx_nlsar_batch_ub_pixels(dkx_nlsar_batch batch, unsigned long long* counts);

/* ---- batched NL-InSAR --------------------------------------------------------------------
 * No counterpart in the reference, which filters one image per call (despeckcl::nlinsar,
 * include/despeckcl.h:97-109) in 80 x 80 tiles, one tile after another with a host round trip
 * per tile (src/compute/nlinsar/nlinsar.cpp:103-139).  For callers that filter many
 * interferograms of any size with the same parameters: every stage runs as one launch over
 * all images (the weights on a persistent grid working through a queue of tiles per XCD),
 * and the whole iteration sequence can be captured in a HIP graph.  Each image's result
 * equals the result of dkx_nlinsar_device on that image alone. */

typedef dkx_nlsar_image dkx_nlinsar_image;   /* ref/phase/coh_filt: estimates in, result out */
typedef struct dkx_nlinsar_batch_s* dkx_nlinsar_batch;

int dkx_nlinsar_batch_create(dkx_nlinsar_batch* batch);
void dkx_nlinsar_batch_destroy(dkx_nlinsar_batch batch);
/* (Re)defines the images and the parameters (same meaning and checks as dkx_nlinsar).  Waits
 * for the batch's previous run, grows its device workspace if needed and uploads the image
 * tables (synchronous). */
int dkx_nlinsar_batch_set(dkx_nlinsar_batch batch, const dkx_nlinsar_image* images, int n_images,
                          int search_window_size, int patch_size, int niter, int lmin);
/* Enqueues every iteration of every image on `stream` (a hipStream_t); capturable in a HIP
 * graph. */
int dkx_nlinsar_batch_enqueue(dkx_nlinsar_batch batch, void* stream);
/* Captures the batch's stage sequence into a HIP graph on first use (after each _set), then
 * replays it on `stream` (not the null stream). */
int dkx_nlinsar_batch_graph_launch(dkx_nlinsar_batch batch, void* stream);
/* Waits for the last run and reads each image's number of pixel-iterations that went
 * through the lmin smoothing. */
int dkx_nlinsar_batch_smoothed(dkx_nlinsar_batch batch, unsigned long long* counts);

/* ---- serving ------------------------------------------------------------------------------
 * No counterpart in the reference (which has no request loop).  A server owns a fixed HBM
 * budget (one arena, sub-allocated per request: no allocation while serving), keeps the NL-SAR
 * weight LUTs resident, admits requests as memory frees up, and batches the waiting requests
 * of each filter into one launch per stage while earlier work still runs (NL-SAR and NL-InSAR
 * batches on streams of their own, overlapping).  Each request's result equals
 * the single-call result (dkx_nlsar_insar_device, or
 * dkx_nlinsar_device from zero estimates). */

typedef struct {
    int kind;                 /* 0 NL-SAR, 1 NL-InSAR (estimates start from zero) */
    const float* ampl_master; /* device pointers, row-major height x width */
    const float* ampl_slave;
    const float* phase;
    float* ref_filt;
    float* phase_filt;
    float* coh_filt;
    int height;
    int width;
} dkx_request;

typedef struct dkx_server_s* dkx_server;

/* NL-SAR parameters as dkx_nlsar_insar; NL-InSAR parameters as dkx_nlinsar. */
int dkx_server_create(unsigned long long hbm_budget_bytes, int search_window_size,
                      const int* patch_sizes, int n_patch_sizes, const int* scale_sizes,
                      int n_scale_sizes, const dkx_nlsar_stats* stats, int n_stats, float h_param,
                      float c_param, int nl_search_window_size, int nl_patch_size, int nl_niter,
                      int nl_lmin, dkx_server* server);
/* Queues a request.  arrival_time_s: its arrival on the library's steady clock (seconds),
 * or <= 0 for now; latencies are measured from it. */
int dkx_server_submit(dkx_server server, const dkx_request* request, double arrival_time_s,
                      long long* id);
/* 1 done (latency_ms from arrival to completion), 0 pending, < 0 failed. */
int dkx_server_poll(dkx_server server, long long id, double* latency_ms);
int dkx_server_wait(dkx_server server, long long id, double* latency_ms);
/* Submits requests[i] at arrival_s[i] seconds after the call (open loop), waits for all and
 * returns each latency (ms, -1 for a failed request) and the time to the last completion. */
int dkx_server_run_trace(dkx_server server, const dkx_request* requests, const double* arrival_s,
                         int n, double* latency_ms, double* makespan_s);
/* NL-SAR batches launched and the requests they held. */
int dkx_server_counters(dkx_server server, long long* batches, long long* batched_images);
/* The same for NL-InSAR. */
int dkx_server_nlinsar_counters(dkx_server server, long long* batches, long long* batched_images);
/* Finishes the queued work, then frees the server. */
void dkx_server_destroy(dkx_server server);

#ifdef __cplusplus
}
#endif
#endif
