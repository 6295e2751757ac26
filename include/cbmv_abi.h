/*
 * cbmv_abi.h -- C ABI of the MI355X-native CBMV cost-volume pipeline.
 *
 * One shared library (cbmv_amd/libcbmv_hip.so) exports everything below.
 * Every function returns an int status (0 = CBMV_OK, negative = error; the
 * message of the last error on the calling thread is cbmv_last_error()),
 * except the create functions, which return NULL on failure.
 *
 * Three groups of entry points:
 *
 *  1. Drop-ins for the reference's Boost.Python modules. They take and fill
 *     HOST buffers (numpy memory in the Python shims), allocate device memory
 *     per call and are synchronous, like the reference's calls:
 *       libmatchers     (cpp/matchers/matchers.cpp:568-582)
 *       libfeatextract  (cpp/featextract/featextract.cpp:519-540)
 *       post            (cpp/post/post.cu:905-933), on device-resident arrays
 *     plus the random-forest predict_proba that test.py:620 gets from
 *     scikit-learn.
 *
 *  2. The fused device pipeline (the hot path): u8 images already in device
 *     memory -> f32 disparity map in device memory, on a caller stream.
 *
 * Layout names: [H,W,D] means d fastest; [D,H,W] means x fastest.
 * RAND_MAX (2147483647) is the reference's "invalid cost" value and is
 * emitted exactly in every float64 output.
 */
#ifndef CBMV_ABI_H
#define CBMV_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum cbmv_status {
  CBMV_OK = 0,
  CBMV_EINVAL = -1,  /* bad argument (shape, size, window, dtype) */
  CBMV_EHIP = -2,    /* a HIP runtime call failed */
  CBMV_ENOMEM = -3,  /* device allocation failed */
  CBMV_ESTATE = -4   /* object used in the wrong state (freed, no model) */
};

/* Message of the last failure on this thread ("" if none). */
const char* cbmv_last_error(void);
/* ABI version of this header (incremented on incompatible changes). */
int cbmv_abi_version(void);

/* Device memory helpers for callers without their own HIP bindings (no
 * reference counterpart: the reference's Python side only sees numpy). */
void* cbmv_device_alloc(int64_t bytes);
int cbmv_device_free(void* p);
int cbmv_memcpy_h2d(void* dst, const void* src, int64_t bytes);  /* synchronous */
int cbmv_memcpy_d2h(void* dst, const void* src, int64_t bytes);  /* synchronous */
int cbmv_device_synchronize(void);
/* HIP events, for timing work enqueued on a stream (NULL = default stream). */
void* cbmv_event_create(void);
int cbmv_event_record(void* ev, void* stream);
int cbmv_event_elapsed_ms(void* start, void* stop, float* ms);  /* synchronises on stop */
int cbmv_event_destroy(void* ev);

/* ======================================================================
 * libmatchers -- cpp/matchers/matchers.cpp
 * ====================================================================== */

/* census(L, R, ndisp, wsize) -> f64[H,W,D]     (matchers.cpp:209-350) */
int cbmv_census(const uint8_t* left, const uint8_t* right, int64_t H, int64_t W,
                int32_t ndisp, int32_t wsize, double* out_hwd);
/* census between device images into a device volume f64[H,W,D] on
 * `stream`, allocating nothing: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under graph capture,
 * serving loops). The two images' descriptors live in a fixed 48 MiB device
 * array per device (H*W*ceil(wsize^2/64) <= 3*2^20 words per image: 1,572,864
 * pixels at wsize 11); larger sizes fail with CBMV_EINVAL. Calls that overlap
 * on one device share that array and race. */
int cbmv_census_noalloc(const uint8_t* d_left, const uint8_t* d_right, int64_t H, int64_t W,
                        int32_t ndisp, int32_t wsize, double* d_out_hwd, void* stream);
/* nccNister(L, R, ndisp, wsize) -> f64[D,H,W]  (matchers.cpp:24-205) */
int cbmv_ncc_nister(const uint8_t* left, const uint8_t* right, int64_t H, int64_t W,
                    int32_t ndisp, int32_t wsize, double* out_dhw);
/* sadsob(L, R, ndisp, wsize) -> f64[D,H,W] on f64 images (matchers.cpp:353-436) */
int cbmv_sadsob(const double* left, const double* right, int64_t H, int64_t W,
                int32_t ndisp, int32_t wsize, double* out_dhw);
/* zsad(L, R, ndisp, wsize) -> f64[D,H,W]       (matchers.cpp:440-512) */
int cbmv_zsad(const uint8_t* left, const uint8_t* right, int64_t H, int64_t W,
              int32_t ndisp, int32_t wsize, double* out_dhw);
/* sobel(img) -> f64[H,W]                        (matchers.cpp:515-555) */
int cbmv_sobel(const uint8_t* img, int64_t H, int64_t W, double* out);
/* initthreads() -> worker count                 (matchers.cpp:557-566);
 * here: the number of compute units of the current device. */
int cbmv_initthreads(void);

/* ======================================================================
 * libfeatextract -- cpp/featextract/featextract.cpp
 * N is the number of rows of a [N,D] view of an [H,W,D] volume.
 * ====================================================================== */

/* swap_axes: [D,H,W] -> [H,W,D]                 (featextract.cpp:24-53) */
int cbmv_swap_axes(const double* in_dhw, int64_t D, int64_t H, int64_t W, double* out_hwd);
/* swap_axes_back: [H,W,D] -> [D,H,W]            (featextract.cpp:55-83) */
int cbmv_swap_axes_back(const double* in_hwd, int64_t H, int64_t W, int64_t D, double* out_dhw);
/* get_cost: channel 0 of [H,W,D,2]              (featextract.cpp:85-113) */
int cbmv_get_cost(const double* prob_hwd2, int64_t H, int64_t W, int64_t D, double* out_hwd);
/* get_right_cost: out[y,x,d] = in[y,x+d,d], else in[0]   (featextract.cpp:115-152) */
int cbmv_get_right_cost(const double* in_hwd, int64_t H, int64_t W, int64_t D, double* out_hwd);
/* get_left_cost: out[y,x,d] = in[y,x-d,d], else in[0]    (featextract.cpp:452-488) */
int cbmv_get_left_cost(const double* in_hwd, int64_t H, int64_t W, int64_t D, double* out_hwd);
/* generate_d_indices(gt, maxd, thresh) -> i32[n,3]       (featextract.cpp:154-215).
 * The reference draws with unseeded rand() inside an OpenMP loop; here the
 * draws come from a counter-based generator keyed by (seed, sample index),
 * so results are reproducible. Same ranges; gt outside [0, maxd] and the
 * reference's division-by-zero case (an empty negative range) are errors. */
int cbmv_generate_d_indices(const int32_t* gt, int64_t n, int32_t maxd, int32_t thresh,
                            uint64_t seed, int32_t* out_n3);
/* get_samples(vol[N,D], samp[N,3]) -> f64[3N]   (featextract.cpp:217-252) */
int cbmv_get_samples(const double* vol, int64_t N, int64_t D, const int32_t* samp,
                     double* out_3n);
/* extract_ratio(vol, e) -> f64[N,D]             (extract_pkrn_test, featextract.cpp:303-342) */
int cbmv_extract_ratio(const double* vol, int64_t N, int64_t D, double e, double* out);
/* extract_ratio(vol, samp, e) -> f64[3N]        (extract_pkrn, featextract.cpp:254-300) */
int cbmv_extract_ratio_samples(const double* vol, int64_t N, int64_t D, const int32_t* samp,
                               double e, double* out_3n);
/* extract_likelihood(vol, sigma) -> f64[N,D]    (extract_aml_testing, featextract.cpp:402-450) */
int cbmv_extract_likelihood(const double* vol, int64_t N, int64_t D, double sigma, double* out);
/* extract_likelihood(vol, samp, sigma) -> f64[3N] (extract_aml, featextract.cpp:345-399) */
int cbmv_extract_likelihood_samples(const double* vol, int64_t N, int64_t D,
                                    const int32_t* samp, double sigma, double* out_3n);
/* generate_labels(samp[n,3]) -> i32[3n] = [1,0,0]*n (featextract.cpp:490-516) */
int cbmv_generate_labels(int64_t n, int32_t* out_3n);

/* ======================================================================
 * post -- cpp/post/post.cu. Arrays live in device memory; `shape` is the
 * logical shape that swap_axis*, CUDA_copy_double update (post.cu:612-614,
 * 663-665,702-704) and get_array returns.
 * ====================================================================== */

typedef struct cbmv_array cbmv_array;
enum cbmv_dtype { CBMV_F32 = 0, CBMV_F64 = 1 };

/* InitCUDA (post.cu:874-901): select device 0; returns 1 if a device exists. */
int cbmv_init_device(void);
/* CUDA_double_array / CUDA_float_array constructors (post.cu:63-170): copy a
 * host array of `ndims` dims (<= 4) to a new device array. */
cbmv_array* cbmv_array_create(int32_t dtype, int32_t ndims, const int64_t* shape,
                              const void* host_data);
int cbmv_array_dtype(const cbmv_array* a);
int cbmv_array_ndims(const cbmv_array* a);
int cbmv_array_shape(const cbmv_array* a, int64_t* shape_out);
/* get_array: copy the device contents to host (logical shape order). */
int cbmv_array_get(const cbmv_array* a, void* host_out);
/* set_to_zero. */
int cbmv_array_set_to_zero(cbmv_array* a);
/* free_array: release device memory; later calls on the array fail with
 * CBMV_ESTATE; calling it twice is allowed (the reference double-frees). */
int cbmv_array_free(cbmv_array* a);
/* Destructor: frees device memory if still held, then the handle. */
void cbmv_array_destroy(cbmv_array* a);

/* sgm(fl, fr, din, dout, dtmp, pi1, pi2, tau_so, alpha1, q1, q2, direction)
 * (post.cu:178-333): 4-path SGM over din [H,W,D]; dout is accumulated.
 * tmp ([W,D] in the reference) is the path state; it is not read here. */
int cbmv_post_sgm(cbmv_array* left, cbmv_array* right, cbmv_array* in_cost,
                  cbmv_array* out_cost, cbmv_array* tmp, float pi1, float pi2, float tau_so,
                  float alpha1, float q1, float q2, int32_t direction);
/* cross(fin, fout, L1, tau1)                    (post.cu:336-396) */
int cbmv_post_cross(cbmv_array* in, cbmv_array* out, int32_t L1, float tau1);
/* cbca(fleftc, frightc, din, dout, direction)   (post.cu:399-450), din [D,H,W] */
int cbmv_post_cbca(cbmv_array* leftc, cbmv_array* rightc, cbmv_array* vol,
                   cbmv_array* vol_out, int32_t direction);
/* subpixel_enchancement(fd, dcost, fout)        (post.cu:453-483), dcost [D,H,W] */
int cbmv_post_subpixel(cbmv_array* d, cbmv_array* cost, cbmv_array* out);
/* median2d(fin, fout, ksize), ksize <= 11       (post.cu:486-518) */
int cbmv_post_median2d(cbmv_array* in, cbmv_array* out, int32_t ksize);
/* mean2d(fin, fout, kernel f32[kh,kw], alpha2)  (post.cu:520-566) */
int cbmv_post_mean2d(cbmv_array* in, cbmv_array* out, const float* kernel, int32_t kh,
                     int32_t kw, float alpha2);
/* swap_axis: [D,H,W] -> [H,W,D]                 (post.cu:569-616) */
int cbmv_post_swap_axis(cbmv_array* in, cbmv_array* out);
/* swap_axis_back: [H,W,D] -> [D,H,W]            (post.cu:619-667) */
int cbmv_post_swap_axis_back(cbmv_array* in, cbmv_array* out);
/* CUDA_copy_double (copies the shape too) / CUDA_copy_float (post.cu:695-715) */
int cbmv_post_copy(cbmv_array* in, cbmv_array* out);
/* CUDA_divide_double(din, val)                  (post.cu:669-684) */
int cbmv_post_divide(cbmv_array* in, double val);
/* outlier_detection(fd0, fd1, fout, disp_max)   (post.cu:717-754) */
int cbmv_post_outlier_detection(cbmv_array* d0, cbmv_array* d1, cbmv_array* out,
                                int32_t disp_max);
/* interpolate_mismatch(fd0, foutlier, fout)     (post.cu:757-827) */
int cbmv_post_interpolate_mismatch(cbmv_array* d0, cbmv_array* outlier, cbmv_array* out);
/* interpolate_occlusion(fd0, foutlier, fout)    (post.cu:829-869) */
int cbmv_post_interpolate_occlusion(cbmv_array* d0, cbmv_array* outlier, cbmv_array* out);

/* ======================================================================
 * Random forest -- replaces sklearn RandomForestClassifier.predict_proba
 * (test.py:620; model from train.py:287-289). Trees are concatenated node
 * arrays: feature index, threshold (as sklearn stores it, float64), left and
 * right child (absolute node index, -1 at leaves), value[n_nodes, 2] = the
 * class fractions of the node (sklearn's tree_.value); roots[t] = first node
 * of tree t.
 * ====================================================================== */

typedef struct cbmv_forest cbmv_forest;
cbmv_forest* cbmv_forest_create(int32_t n_trees, const int32_t* roots, int64_t n_nodes,
                                const int32_t* feature, const double* threshold,
                                const int32_t* left, const int32_t* right,
                                const double* value, int32_t n_features);
int cbmv_forest_n_trees(const cbmv_forest* f);
/* predict_proba on host features X[N, n_features] (cast to float32 as
 * sklearn does) -> proba[N,2] float64 (column 0 = P(non-match)). */
int cbmv_forest_predict_proba(const cbmv_forest* f, const float* X, int64_t N,
                              double* proba_n2);
void cbmv_forest_destroy(cbmv_forest* f);

/* Random-forest trainer (host, std::threads) -- replaces
 * RandomForestClassifier(min_samples_leaf=700, n_estimators=50).fit
 * (train.py:287-289) where scikit-learn is not installed. CART trees with
 * Gini impurity and scikit-learn's exact best-split rules (a single tree on
 * all features without bootstrap equals DecisionTreeClassifier's), seeded:
 * the forest depends on `seed` only, not on the thread count. X[n,
 * n_features] float32 row-major, y[n] in {0, 1}; sample_weight (optional,
 * n_trees = 1 only) replaces the bootstrap. The model exports the node
 * arrays cbmv_forest_create takes (cbmv_cart_n_nodes entries each, roots
 * cbmv_cart_n_trees; value[n_nodes, 2] class fractions). */
typedef struct cbmv_cart_params {
  int32_t n_trees;          /* 50 */
  int32_t max_features;     /* candidate features per node: 4 (sqrt of 20) */
  int32_t min_samples_leaf; /* 700 */
  int32_t max_depth;        /* 0 = unlimited */
  int32_t bootstrap;        /* 1: n rows drawn with replacement per tree */
  int32_t n_threads;        /* 0: one per hardware thread */
  uint64_t seed;
  /* n_trees = 1: the splitter's random state (scikit-learn's rand_r_state,
   * 1 .. 2^31 - 1); 0 derives it from seed, as for every tree of a forest */
  uint32_t splitter_state;
} cbmv_cart_params;
typedef struct cbmv_cart_model cbmv_cart_model;
void cbmv_cart_params_default(cbmv_cart_params* p);
cbmv_cart_model* cbmv_cart_fit(const float* X, const int32_t* y, int64_t n, int32_t n_features,
                               const double* sample_weight, const cbmv_cart_params* params);
int64_t cbmv_cart_n_nodes(const cbmv_cart_model* m);
int32_t cbmv_cart_n_trees(const cbmv_cart_model* m);
int cbmv_cart_export(const cbmv_cart_model* m, int32_t* roots, int32_t* feature,
                     double* threshold, int32_t* left, int32_t* right, double* value);
void cbmv_cart_destroy(cbmv_cart_model* m);

/* ======================================================================
 * Fused device pipeline (the hot path)
 * ====================================================================== */

typedef struct cbmv_params {
  /* matcher windows (Testing defaults test.py:46; SAD-on-Sobel window 5, test.py:131) */
  int32_t census_w, ncc_w, sad_w, sob_w;
  /* post-processing (test.py:46-47) */
  float pi1, pi2, sgm_q1, sgm_q2, alpha1, tau_so;
  int32_t sgm_iters, L1;
  float tau1;
  int32_t cbca_i1, cbca_i2, median_iters, median_w;
  float blur_sigma, blur_t;
  int32_t sgm_paths;        /* 4 (reference) or 8 (adds the diagonals) */
  int32_t lr_check;         /* 0: __postprocessing_mem chain, 1: _mem_interp chain */
  int32_t test_batch;       /* rows per feature batch = floor(test_batch / W) (test.py:608) */
  int32_t classifier;       /* 0 = random forest, 1 = MLP, 2 = MLP gated to RF */
  float gate_threshold;     /* MLP gate: pixels with max_d p(match) < this use the RF */
  int32_t cbca_tile;        /* columns per workgroup of the fused aggregation: 4, 8 or 16
                               (0 = 8); a tuning choice, results do not depend on it */
} cbmv_params;

/* Fill `p` with the reference defaults. */
void cbmv_params_default(cbmv_params* p);

typedef struct cbmv_pipeline cbmv_pipeline;
/* Allocates every device buffer the pipeline needs for one H x W x D pair. */
cbmv_pipeline* cbmv_pipeline_create(int64_t H, int64_t W, int32_t ndisp, const cbmv_params* p);
/* Upload a forest (copied to the pipeline's device memory). */
int cbmv_pipeline_set_forest(cbmv_pipeline* pl, const cbmv_forest* f);
/* Hidden-unit activation of the MLP. */
enum { CBMV_ACT_GELU = 0, CBMV_ACT_RELU = 1 };
/* Upload MLP weights (float32, row-major): w1[20,128], b1[128], w2[128,64],
 * b2[64], w3[64], b3[1]; the 20 input features are standardised with
 * mean[20], inv_std[20] (and clamped to +-8). W1 is stored as bf16, W2 and w3
 * as half, and the hidden activations enter layers 2 and 3 as half values:
 * CBMV_EINVAL when a bound of the hidden pre-activations (from |W|, |b| and
 * the clamp) reaches 65504. activation: CBMV_ACT_GELU (tanh form) or
 * CBMV_ACT_RELU, as trained. */
int cbmv_pipeline_set_mlp(cbmv_pipeline* pl, const float* mean, const float* inv_std,
                          const float* w1, const float* b1, const float* w2, const float* b2,
                          const float* w3, const float* b3, int32_t activation);
/* Change the pair size of a pipeline to H x W x ndisp without reallocating:
 * each must be at most the size it was created with (buffers are sized for
 * that). For running pairs of different sizes through one set of buffers. */
int cbmv_pipeline_reshape(cbmv_pipeline* pl, int64_t H, int64_t W, int32_t ndisp);
/* Set the fused aggregation's tile width (cbmv_params.cbca_tile) of an
 * existing pipeline. */
int cbmv_pipeline_set_cbca_tile(cbmv_pipeline* pl, int32_t tile);
/* One SGM sweep per path (0) or paired sweeps (1; the default of an 8-path
 * LR-check pipeline with ndisp <= 256, whose two view chains share the
 * device). Results differ only by float32 rounding; a batch of pipelines on
 * several streams is faster with 0. Turning pairs on needs a pipeline that
 * set them up at creation (CBMV_EINVAL otherwise). */
int cbmv_pipeline_set_sgm_pairs(cbmv_pipeline* pl, int32_t on);
/* Run the whole path: d_left/d_right u8[H,W] device -> d_disp f32[H,W] device,
 * enqueued on `stream` (a hipStream_t; NULL = the default stream). Part of
 * the work runs on the pipeline's own side stream, forked from and joined
 * back to `stream` by events, so the call stays ordered on `stream` (and can
 * be captured into a graph from it); calls on one pipeline must not overlap. */
int cbmv_pipeline_run(cbmv_pipeline* pl, const uint8_t* d_left, const uint8_t* d_right,
                      float* d_disp, void* stream);
/* Run one stage of the path (for per-stage timing): 0 matchers, 1 per-pixel
 * statistics, 2 features + classifier, 3 post-processing. Stages read what
 * the earlier ones wrote into the pipeline's buffers (the matchers also
 * write the left view's minima, rows 0..3 of the left statistics, which
 * stage 1 takes as they are). */
int cbmv_pipeline_run_stage(cbmv_pipeline* pl, int32_t stage, const uint8_t* d_left,
                            const uint8_t* d_right, float* d_disp, void* stream);
/* Device pointer and size of an intermediate buffer, for tests:
 * 0 census u8[H,D,W], 1 ncc f64[H,D,W], 2 sob u16[H,D,W], 3 sad f64[H,D,W],
 * 4 left stats f64[8,H,W], 5 right stats f64[8,H,W] (rows 0..3 the minima,
 * 4..7 the sums of exponentials, of census, NCC, SAD of Sobel, zero-mean
 * SAD), 6 cost f32[H,W,D],
 * 7 left WTA f32[H,W], 8 right WTA f32[H,W], 9 outlier f32[H,W],
 * 10 best MLP match probability per pixel f32[H,W] (gated MLP only),
 * 11 gate list i32[H*W + 16]
 * (pixels sent to the forest; their count at index H*W),
 * 12 right view's cost f32[H,W,D] (LR-check chains),
 * 13 left view's final volume f32[H,W,D], the winner-take-all's input
 * (after the last aggregation with L1 = 6, or after SGM when cbca_i2 = 0),
 * 14, 15 left and right images normalised to (img - mean) / std f32[H,W]
 * (numpy's float32 mean and std, test.py:489-492). */
int cbmv_pipeline_buffer(cbmv_pipeline* pl, int32_t which, void** ptr, int64_t* nbytes);
/* Synchronises the device and returns in *count how many waits between
 * workgroups of the 8-path SGM sweeps gave up (a safety bound of 0.5 s that
 * correct operation never reaches; a nonzero count means the disparities of
 * those runs are invalid). */
int cbmv_pipeline_sgm_timeouts(cbmv_pipeline* pl, int64_t* count);
/* Features of one row of voxels, for tests: f32[D,W,20] of row y. */
int cbmv_pipeline_row_features(cbmv_pipeline* pl, int64_t y, float* d_out, void* stream);
/* Single-precision features (the MLP's inputs) of sampled voxels, for
 * training: d_samples int32[N,3] = (y, x, d) -> d_out f32[N,20]. The
 * pipeline must have run stages 0 and 1 on the pair. Like the reference's
 * training sampler (train.py:100-273, get_samples featextract.cpp:217-252)
 * but on device. */
int cbmv_pipeline_sample_features(cbmv_pipeline* pl, const int32_t* d_samples, int64_t N,
                                  float* d_out, void* stream);
void cbmv_pipeline_destroy(cbmv_pipeline* pl);

/* ======================================================================
 * Batches of pairs of different sizes (no reference counterpart: the
 * reference processes one pair per call, main.py:49-70). Pairs are spread
 * over n_streams HIP streams, each with one pipeline sized for its largest
 * pair; each pair's whole path is captured into a hipGraph once
 * (cbmv_batch_capture) and replayed by every cbmv_batch_run. The pairs'
 * u8 images and f32 disparity maps live in device buffers owned by the
 * batch (cbmv_batch_pair_buffers): write new images there and run again.
 * ====================================================================== */

typedef struct cbmv_batch cbmv_batch;
cbmv_batch* cbmv_batch_create(int32_t n, const int32_t* H, const int32_t* W, const int32_t* ndisp,
                              const cbmv_params* p, int32_t n_streams);
int cbmv_batch_set_forest(cbmv_batch* b, const cbmv_forest* f);
int cbmv_batch_set_mlp(cbmv_batch* b, const float* mean, const float* inv_std, const float* w1,
                       const float* b1, const float* w2, const float* b2, const float* w3,
                       const float* b3, int32_t activation);
/* Device buffers of pair i: u8[H,W] left and right images, f32[H,W] disparity. */
int cbmv_batch_pair_buffers(cbmv_batch* b, int32_t i, uint8_t** d_left, uint8_t** d_right,
                            float** d_disp);
/* Autotune: per (W rounded up to 256, ndisp) bucket, time the fused
 * aggregation's tile widths (4, 8, 16 columns) on the bucket's first pair's
 * post-processing (reps timed runs each) and keep the fastest for the
 * bucket; tile_out (optional, n entries) receives each pair's choice. Runs
 * on the pairs' current images; drops captured graphs. */
int cbmv_batch_tune(cbmv_batch* b, int32_t reps, int32_t* tile_out);
/* Capture every pair's path into a hipGraph (after the models are set and
 * any tuning). */
int cbmv_batch_capture(cbmv_batch* b);
/* Launch every pair's graph and wait. Optional host outputs (n floats
 * each): exec_ms = each pair's graph time, done_ms = its completion time
 * from the batch start; total_ms = the batch's time. */
int cbmv_batch_run(cbmv_batch* b, float* exec_ms, float* done_ms, float* total_ms);
/* Stream (0 .. n_streams-1) that pair i runs on, or -1. */
int32_t cbmv_batch_stream_of(const cbmv_batch* b, int32_t i);
void cbmv_batch_destroy(cbmv_batch* b);

/* ======================================================================
 * MLP training on the GPU (no reference counterpart: the reference fits a
 * scikit-learn forest, train.py:287-289). Weights are float32 in the layout
 * of cbmv_pipeline_set_mlp; the kernels run forward and backward passes on
 * bf16 MFMA with float32 accumulation and update float32 master weights with
 * Adam. Loss: binary cross-entropy of sigmoid(logit) against y in {0, 1}.
 * ====================================================================== */

typedef struct cbmv_mlp_trainer cbmv_mlp_trainer;
cbmv_mlp_trainer* cbmv_mlp_trainer_create(const float* mean, const float* inv_std,
                                          const float* w1, const float* b1, const float* w2,
                                          const float* b2, const float* w3, const float* b3,
                                          int32_t activation);
/* One Adam step on device features X f32[B,20] and labels y f32[B]; *loss
 * (optional; synchronises) receives the batch's mean loss. */
int cbmv_mlp_trainer_step(cbmv_mlp_trainer* t, const float* d_X, const float* d_y, int64_t B,
                          float lr, float* loss, void* stream);
/* p(match) f32[N] of device features X f32[N,20]. */
int cbmv_mlp_trainer_predict(cbmv_mlp_trainer* t, const float* d_X, int64_t N, float* d_p,
                             void* stream);
/* Mean loss and its gradient (float32, order w1, b1, w2, b2, w3, b3) of a
 * device batch, copied to host; the weights are not changed. */
int cbmv_mlp_trainer_gradients(cbmv_mlp_trainer* t, const float* d_X, const float* d_y,
                               int64_t B, float* grad_host, float* loss);
int cbmv_mlp_trainer_weights(cbmv_mlp_trainer* t, float* w1, float* b1, float* w2, float* b2,
                             float* w3, float* b3);
void cbmv_mlp_trainer_destroy(cbmv_mlp_trainer* t);

/* Diagnostic: C[16][16] = A[16][32] x B[32][16] through one
 * v_mfma_f32_16x16x32_bf16 (inputs rounded to bf16), host arrays. */
int cbmv_mfma_probe(const float* A, const float* B, float* C);

/* Diagnostic: the pipeline MLP's hidden activation (0 GELU, 1 ReLU) as it
 * enters layer 2: y = act(half(x)) in half precision, returned as floats;
 * n a multiple of 8, at most 2^28, host arrays. */
int cbmv_act_probe(const float* x, float* y, int64_t n, int32_t activation);

#ifdef __cplusplus
}
#endif

#endif /* CBMV_ABI_H */
