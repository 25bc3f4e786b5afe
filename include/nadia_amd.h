/*
 * nadia_amd.h -- C ABI of the MI355X phase-retrieval library (libnadia_hip.so).
 *
 * This is the boundary behind NADIA's C++ reconstruction API: the drop-in
 * classes (the headers in include/nadia: Complex_2D, Double_2D, BaseCDI, PlanarCDI,
 * FresnelCDI, FresnelCDI_WF) call only these functions, and any other host
 * language (ctypes, IDL, ...) can bind them directly.  Plain pointers and
 * sizes only.
 *
 * Conventions
 *   - Arrays cross the boundary as host pointers owned by the caller and are
 *     copied.  Complex arrays are interleaved (re, im), real arrays plain;
 *     both x-major: element (x, y) at x*ny + y, the in-memory layout of
 *     NADIA's Complex_2D / Double_2D (NADIA/include/Complex_2D.h:179-194,
 *     NADIA/include/Double_2D.h:121-143).  The element type is float for
 *     NADIA_FP32 objects and double for NADIA_FP64 objects.
 *   - Every function returns 0 on success and a negative code on error;
 *     nadia_last_error() describes the last error of the calling thread.
 *   - Device work is queued on one stream per device; calls that return host
 *     data (downloads, get_error) wait for it.  An object must not be used by
 *     two host threads at once (NADIA's own objects are not re-entrant either).
 *   - Supported lengths per axis: 2^k (8..4096) and 3*2^k (24..3072).
 */
#ifndef NADIA_AMD_H
#define NADIA_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* precision (NADIA's compile-time DOUBLE_PRECISION switch, include/types.h:5-39) */
#define NADIA_FP32 0
#define NADIA_FP64 1

/* reconstruction kinds */
#define NADIA_PLANAR 0     /* PlanarCDI      (NADIA/include/PlanarCDI.h:24-44) */
#define NADIA_FRESNEL 1    /* FresnelCDI     (NADIA/include/FresnelCDI.h:33-232) */
#define NADIA_FRESNEL_WF 2 /* FresnelCDI_WF  (NADIA/include/FresnelCDI_WF.h:31-186) */

/* named algorithms, same values as NADIA/include/BaseCDI.h:43 */
#define NADIA_ER 0
#define NADIA_BIO 1
#define NADIA_BOO 2
#define NADIA_HIO 3
#define NADIA_DM 4
#define NADIA_SF 5
#define NADIA_ASR 6
#define NADIA_HPR 7
#define NADIA_RAAR 8
#define NADIA_CUSTOM 9

/* execution paths of iterate() */
#define NADIA_PATH_FUSED 0 /* two fused FFT passes per projection (default) */
#define NADIA_PATH_PLAIN 1 /* one kernel per reference operation, reference data flow */

/* error codes */
#define NADIA_OK 0
#define NADIA_EINVAL (-1)
#define NADIA_EHIP (-2)
#define NADIA_ESIZE (-3)
#define NADIA_ESTATE (-4)

typedef struct nadia_field nadia_field; /* device-resident complex field */
typedef struct nadia_recon nadia_recon; /* one reconstruction (BaseCDI object) */
typedef struct nadia_batch nadia_batch; /* reconstructions iterated together */
typedef struct nadia_constraint nadia_constraint; /* a TransmissionConstraint */
typedef struct nadia_pd nadia_pd;                 /* a PhaseDiverseCDI */
typedef struct nadia_event nadia_event;           /* a marker in the device's queue of work */
/* host function applied to a transmission function (interleaved complex,
 * x-major, nx x ny, float or double as `precision`), in place */
typedef void (*nadia_custom_constraint_fn)(void* complex_host, int nx, int ny, int precision, void* user);

const char* nadia_last_error(void);
int nadia_version(void);
int nadia_device_count(int* count);
int nadia_set_device(int device);
/* 1 if n is a valid axis length: 1..4096 (FFTW plans any length,
 * NADIA/src/Complex_2D.c++:472-477).  Reconstructions also need even sizes
 * (Complex_2D::invert assumes them, NADIA/src/Complex_2D.c++:318-331). */
int nadia_size_supported(int n);
/* 1 if n has compiled single-launch passes (2^k 8..4096, 3*2^k 24..3072);
 * other lengths run transforms of any length and the unfused iteration.
 * No counterpart in the reference: it tells callers which sizes are fast. */
int nadia_size_native(int n);

/* ---- Complex_2D (NADIA/include/Complex_2D.h:35-515, src/Complex_2D.c++) ---- */

/* ComplexR_2D(int,int) (src/Complex_2D.c++:16-32); zero-filled */
int nadia_field_create(int nx, int ny, int precision, nadia_field** out);
/* ~ComplexR_2D (src/Complex_2D.c++:88-98) */
int nadia_field_destroy(nadia_field* f);
int nadia_field_dims(const nadia_field* f, int* nx, int* ny, int* precision);
/* host -> device / device -> host copies of the whole array */
int nadia_field_upload(nadia_field* f, const void* host);
int nadia_field_download(const nadia_field* f, void* host);
/* copy (src/Complex_2D.c++:300-313) */
int nadia_field_copy(nadia_field* dst, const nadia_field* src);
/* perform_forward_fft (dir -1) / perform_backward_fft (+1), unnormalised, in
 * place (src/Complex_2D.c++:486-503) */
int nadia_field_fft(nadia_field* f, int direction);
/* invert(bool scale) (src/Complex_2D.c++:317-353); even sizes */
int nadia_field_invert(nadia_field* f, int scale);
/* scale (src/Complex_2D.c++:169-180) */
int nadia_field_scale(nadia_field* f, double factor);
/* add(c2, scale) (src/Complex_2D.c++:183-199) */
int nadia_field_add(nadia_field* f, const nadia_field* other, double scale);
/* get_norm (src/Complex_2D.c++:259-270) */
int nadia_field_norm(const nadia_field* f, double* out);
/* get_2d(MAG, result) (src/Complex_2D.c++:158-165) into a host real array */
int nadia_field_magnitude(const nadia_field* f, void* host_real);
/* No counterpart in the reference: nadia_field_magnitude for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The magnitudes are
 * staged in a fixed device array of the library (32 MiB per GPU, held from
 * library load to process end) instead of a per-call allocation; fields
 * needing more (more than 2048 x 2048 FP64 or 4096 x 2048 FP32 pixels) are
 * refused with NADIA_ESIZE.  Calls that overlap on one GPU race on the
 * array. */
int nadia_field_magnitude_noalloc(const nadia_field* f, void* host_real);

/* ---- BaseCDI (NADIA/include/BaseCDI.h:66-522, src/BaseCDI.c++) ---- */

/* BaseCDI(Complex_2D&, n_best) (src/BaseCDI.c++:33-66) plus the subclass
 * constructor for `kind`; `estimate` is referenced (not copied) and mutated
 * by iterate(), like BaseCDI::complex. */
int nadia_recon_create(nadia_field* estimate, int kind, int n_best, nadia_recon** out);
int nadia_recon_destroy(nadia_recon* r);
/* set_support(const Double_2D&, bool soften) (src/BaseCDI.c++:95-108) */
int nadia_recon_set_support(nadia_recon* r, const void* host_real, int soften);
/* set_intensity (src/BaseCDI.c++:118-124) */
int nadia_recon_set_intensity(nadia_recon* r, const void* host_real);
/* set_beam_stop (src/BaseCDI.c++:112-116) */
int nadia_recon_set_beam_stop(nadia_recon* r, const void* host_real);
/* set_algorithm (src/BaseCDI.c++:192-237), including its early return */
int nadia_recon_set_algorithm(nadia_recon* r, int algorithm);
/* set_custom_algorithm(m1..m10) (src/BaseCDI.c++:240-255) */
int nadia_recon_set_custom_algorithm(nadia_recon* r, const double m[10]);
/* set_relaxation_parameter (include/BaseCDI.h:264-267) */
int nadia_recon_set_relaxation_parameter(nadia_recon* r, double beta);
/* the five coefficients {PSF, PFS, PS, PF, PI} and the algorithm id
 * (print_algorithm, src/BaseCDI.c++:278-295, prints these) */
int nadia_recon_get_coefficients(const nadia_recon* r, double out5[5], int* algorithm);
/* iterate() n times (src/BaseCDI.c++:297-387; FresnelCDI_WF::iterate,
 * src/FresnelCDI_WF.c++:122-150).  errors (may be NULL) receives get_error()
 * after each iteration; the estimate stays on the device. */
int nadia_recon_iterate(nadia_recon* r, int n, double* errors);
/* get_error (src/BaseCDI.c++:126-128) */
int nadia_recon_get_error(nadia_recon* r, double* out);
/* apply_shrinkwrap(gauss_width, threshold) (src/BaseCDI.c++:416-429) */
int nadia_recon_apply_shrinkwrap(nadia_recon* r, double gauss_width, double threshold);
/* get_support (include/BaseCDI.h:399-401) */
int nadia_recon_get_support(const nadia_recon* r, void* host_real);
/* initialise_estimate(seed): PlanarCDI (src/PlanarCDI.c++:39-62), FresnelCDI
 * (src/FresnelCDI.c++:152-201), FresnelCDI_WF (src/FresnelCDI_WF.c++:99-117);
 * uses the C library's srand/rand on the host, like the reference. */
int nadia_recon_initialise_estimate(nadia_recon* r, int seed);
/* public virtual operators on any field of the same size:
 * apply_support (src/BaseCDI.c++:130-134), project_intensity (:156-160),
 * scale_intensity (:163-189), propagate_to/from_detector (PlanarCDI.c++:65-73,
 * FresnelCDI.c++:238-250, FresnelCDI_WF.c++:152-175) */
int nadia_recon_apply_support(nadia_recon* r, nadia_field* c);
int nadia_recon_project_intensity(nadia_recon* r, nadia_field* c);
int nadia_recon_scale_intensity(nadia_recon* r, nadia_field* c);
int nadia_recon_propagate_to_detector(nadia_recon* r, nadia_field* c);
int nadia_recon_propagate_from_detector(nadia_recon* r, nadia_field* c);
/* get_best_result(error, index) (src/BaseCDI.c++:82-91), reset_best
 * (include/BaseCDI.h:476-479) */
int nadia_recon_get_best_result(nadia_recon* r, int index, double* error, void* host_complex);
int nadia_recon_reset_best(nadia_recon* r);
/* PlanarCDI::get_intensity_autocorrelation (src/PlanarCDI.c++:13-36) */
int nadia_recon_get_intensity_autocorrelation(nadia_recon* r, void* host_real);
/* fused (default) or plain execution of iterate(); both compute the same
 * reference algorithm */
int nadia_recon_set_path(nadia_recon* r, int path);

/* ---- FresnelCDI (NADIA/src/FresnelCDI.c++) ---- */

/* FresnelCDI(initial_guess, white_field, wavelength, focal_detector_length,
 * focal_sample_length, pixel_size, normalisation) (src/FresnelCDI.c++:20-44) */
int nadia_fresnel_setup(nadia_recon* r, const void* white_field, double wavelength,
                        double focal_detector_length, double focal_sample_length, double pixel_size,
                        double normalisation);
/* set_experimental_parameters (src/FresnelCDI.c++:57-95) */
int nadia_fresnel_set_experimental_parameters(nadia_recon* r, double wavelength,
                                              double focal_detector_length,
                                              double focal_sample_length, double pixel_size);
/* set_norm (src/FresnelCDI.c++:139-148), auto_set_norm (:129-137),
 * set_normalisation (include/FresnelCDI.h:154-156, stores only) */
int nadia_fresnel_set_norm(nadia_recon* r, double norm);
int nadia_fresnel_auto_set_norm(nadia_recon* r);
int nadia_fresnel_set_normalisation(nadia_recon* r, double norm);
/* get_illumination_at_sample (src/FresnelCDI.c++:293-296) */
int nadia_fresnel_get_illumination_at_sample(nadia_recon* r, void* host_complex);
/* get_transmission_function (src/FresnelCDI.c++:298-355) of esw (NULL = the estimate) */
int nadia_fresnel_get_transmission_function(nadia_recon* r, const nadia_field* esw, void* host_complex);
/* set_transmission_function (src/FresnelCDI.c++:253-283) into esw (NULL = the estimate) */
int nadia_fresnel_set_transmission_function(nadia_recon* r, const void* host_complex, nadia_field* esw);

/* ---- FresnelCDI_WF (NADIA/src/FresnelCDI_WF.c++) ---- */

/* FresnelCDI_WF(initial_guess, wavelength, zone_focal_length,
 * focal_detector_length, pixel_size) (src/FresnelCDI_WF.c++:17-67) */
int nadia_wf_setup(nadia_recon* r, double wavelength, double zone_focal_length,
                   double focal_detector_length, double pixel_size);
/* set_support(z_diameter, size) (src/FresnelCDI_WF.c++:177-196) */
int nadia_wf_set_support(nadia_recon* r, double z_diameter, double size);

/* ---- TransmissionConstraint (NADIA/include/TransmissionConstraint.h,
 * NADIA/src/TransmissionConstraint.c++:26-172) ----
 * create: unity magnitude and charge flipping (sign +1) enabled, like the
 * reference's constructor.  add_region = add_complex_constraint with a
 * ComplexConstraint(region, alpha1, alpha2): pixels with region > 0 (nx*ny
 * doubles) join region `index`; a pixel already claimed keeps its first
 * region.  set_fixed_c = ComplexConstraint::set_fixed_c; get_c_mean returns
 * the value the last application computed (or the fixed one).  apply =
 * apply_constraint on a field.  set_complex_constraint attaches one to a
 * reconstruction (NULL detaches): apply_support then applies it after the
 * support (to the estimate for PlanarCDI, to the transmission function for
 * FresnelCDI).  The constraint's per-region means need the whole field
 * between projection and update: ER then runs the fused passes with the
 * constraint between iterations, other algorithms the unfused path.  The
 * constraint is not owned and must outlive its use. */
int nadia_constraint_create(int nx, int ny, nadia_constraint** out);
int nadia_constraint_destroy(nadia_constraint* tc);
int nadia_constraint_set_charge_flipping(nadia_constraint* tc, int enable, int flip_sign);
int nadia_constraint_set_enforce_unity(nadia_constraint* tc, int enable);
int nadia_constraint_add_region(nadia_constraint* tc, const double* region, double alpha1, double alpha2,
                                int* index);
/* ComplexConstraint::set_alpha1 / set_alpha2 of region `index` */
int nadia_constraint_set_alphas(nadia_constraint* tc, int index, double alpha1, double alpha2);
int nadia_constraint_set_fixed_c(nadia_constraint* tc, int index, double c_mean);
int nadia_constraint_get_c_mean(nadia_constraint* tc, int index, double* c_mean);
/* delete_complex_constraint_regions */
int nadia_constraint_clear_regions(nadia_constraint* tc);
/* set_custom_constraint: fn runs on the host after the built-in steps */
int nadia_constraint_set_custom(nadia_constraint* tc, nadia_custom_constraint_fn fn, void* user);
int nadia_constraint_apply(nadia_constraint* tc, nadia_field* f);
/* BaseCDI::set_complex_constraint (NADIA/include/BaseCDI.h:471-473) */
int nadia_recon_set_complex_constraint(nadia_recon* r, nadia_constraint* tc);

/* ---- PhaseDiverseCDI (NADIA/include/PhaseDiverseCDI.h,
 * NADIA/src/PhaseDiverseCDI.c++) ----
 * Frames are PlanarCDI or FresnelCDI reconstructions (not owned, same
 * precision, same device) at positions (x, y) of one global object, which
 * grows to the frames' bounding box as they are added.  create = the
 * constructor (beta, gamma, parallel, granularity = sub-pixel scale);
 * add_position = add_new_position; set_parameters = set_iterations_per_cycle,
 * set_feedback_parameter, set_amplification_factor; set_probe_scaling;
 * initialise_estimate; iterate = one "big" iteration (frame_errors, may be
 * NULL, receives each frame's get_error()); get/set_transmission = the
 * global object (nx x ny, interleaved complex, x-major).  In parallel mode
 * the frames' iterations run as one batch. */
int nadia_pd_create(double beta, double gamma, int parallel, int granularity, nadia_pd** out);
int nadia_pd_destroy(nadia_pd* d);
int nadia_pd_add_position(nadia_pd* d, nadia_recon* frame, double x, double y, double alpha);
int nadia_pd_remove_positions(nadia_pd* d);
int nadia_pd_set_parameters(nadia_pd* d, int iterations_per_cycle, double beta, double gamma);
int nadia_pd_set_probe_scaling(nadia_pd* d, int n_probe, double alpha);
int nadia_pd_initialise_estimate(nadia_pd* d);
int nadia_pd_iterate(nadia_pd* d, double* frame_errors);
int nadia_pd_get_size(const nadia_pd* d, int* nx, int* ny);
int nadia_pd_get_transmission(nadia_pd* d, void* host);
int nadia_pd_set_transmission(nadia_pd* d, int nx, int ny, const void* host);
/* adjust_positions (NADIA/include/PhaseDiverseCDI.h:325-329,
 * NADIA/src/PhaseDiverseCDI.c++:337-719): type 0 = CROSS_CORRELATION,
 * 1 = MINIMUM_ERROR; forward = align each frame to the previous one (else to
 * the next, from the last); the search window [x_min, x_max) x [y_min,
 * y_max) and the initial step of the minimum-error search.  gave_up (may be
 * NULL, one int per frame) is set where the minimum-error search gave up and
 * restored the frame's position.  Frames must share one size.  Runs on a
 * temporary object of ones; the global object is not changed. */
int nadia_pd_adjust_positions(nadia_pd* d, int type, int forward, int x_min, int x_max, int y_min, int y_max,
                              double step_size, int* gave_up);
/* get_final_x_position / get_final_y_position (PhaseDiverseCDI.h:339-352) */
int nadia_pd_get_position(const nadia_pd* d, int n, double* x, double* y);
/* align (NADIA/include/utils.h:24-30, NADIA/src/utils.c:1151-1343): the
 * offset of img2 against img1 (real arrays, x-major, element type of
 * `precision`) maximising their weighted cross-correlation ratio; w1 / w2
 * (may be NULL) select the pixels taken (> 0).  min == max on either axis
 * searches every offset.  *found = 0 (no offset passed the overlap test)
 * leaves the offsets unchanged. */
int nadia_align(int precision, const void* img1, int nx1, int ny1, const void* img2, int nx2, int ny2, int min_x,
                int max_x, int min_y, int max_y, const void* w1, const void* w2, double overlap_fraction,
                int* offset_x, int* offset_y, int* found);

/* Device memory (bytes) a reconstruction holds besides its estimate field,
 * and an estimate for one not yet created, estimate field included (fused
 * path, beam stop set): the admission figure of a request loop.  No
 * counterpart in the reference. */
int nadia_recon_device_bytes(const nadia_recon* r, size_t* out);
int nadia_recon_bytes_estimate(int nx, int ny, int precision, int kind, int n_best, size_t* out);

/* ---- batches (no counterpart in the reference) ----
 * A batch iterates many reconstructions together: those of the same
 * precision, kind and size share every launch of the fused passes, so small
 * jobs fill the GPU and the phases of different jobs overlap.  Each job keeps
 * its own state, algorithm and error ring; its results are bit-identical to
 * iterating it alone (same kernels, same per-job reductions; a job with a
 * transmission constraint takes its own launches, as alone).  The
 * reconstructions stay owned by the caller; all must live on the batch's
 * device. */
int nadia_batch_create(nadia_batch** out);
int nadia_batch_destroy(nadia_batch* b);
int nadia_batch_add(nadia_batch* b, nadia_recon* r);
int nadia_batch_remove(nadia_batch* b, nadia_recon* r);
int nadia_batch_size(const nadia_batch* b, int* n);
/* n iterations of every job (each job's iterate()); errors (may be NULL)
 * receives n x size values, errors[k*size + j] = get_error() of job j (in the
 * order added) after iteration k. */
int nadia_batch_iterate(nadia_batch* b, int n, double* errors);
/* counts[j] iterations of job j (in the order added; 0 leaves it untouched):
 * job j takes part in the first counts[j] launches of its group, so jobs at
 * different points of their schedules share launches (continuous batching).
 * Each job's result is the same as nadia_recon_iterate(r, counts[j], NULL). */
int nadia_batch_iterate_each(nadia_batch* b, const int* counts);
/* Graph mode: with steps > 0, a nadia_batch_iterate of n >= steps iterations
 * runs as replays of one captured hipGraph holding `steps` iterations of
 * every launch group (in launch order); the graph is rebuilt only when the
 * batch's launches change.  0 (default) launches eagerly.
 * Results are bit-identical to eager launches, n-best bookkeeping included.
 * Applies when every job shares the fused launches (a job on the plain path
 * or with a transmission constraint makes the batch launch eagerly). */
int nadia_batch_set_graph(nadia_batch* b, int steps);
/* graph replays run so far (for tests and measurement) */
int nadia_batch_graph_replays(const nadia_batch* b, long long* n);

/* ---- asynchronous use (no counterpart in the reference) ----
 * A request loop keeps the next batch of work queued while it reads the
 * results of the last: events mark points in the library's queue on the
 * current device, and get_error() can be copied out without waiting. */
int nadia_event_create(nadia_event** out);
int nadia_event_destroy(nadia_event* e);
/* mark everything queued so far on the current device */
int nadia_event_record(nadia_event* e);
/* *done = 1 once the work before the last record has finished */
int nadia_event_query(nadia_event* e, int* done);
int nadia_event_synchronize(nadia_event* e);
/* page-locked host memory (the destination of asynchronous reads) */
int nadia_host_alloc(size_t bytes, void** out);
int nadia_host_free(void* p);
/* queue a copy of get_error() into *host (page-locked memory from
 * nadia_host_alloc); the value is there once an event recorded after this
 * call has completed */
int nadia_recon_get_error_async(nadia_recon* r, double* host);

/* ---- tuning and measurement ---- */

/* Tile variants of the fused passes (no counterpart in the reference).  By
 * default an autotuner times the variants on the first fused iterate() of
 * each (precision, kind, size) and keeps the fastest for the process; every
 * variant computes bit-identical results.  Setting both to -1 returns to the
 * autotuner; get returns -1 before the first fused iterate(). */
int nadia_recon_set_tile_variants(nadia_recon* r, int col_variant, int row_variant);
int nadia_recon_get_tile_variants(const nadia_recon* r, int* col_variant, int* row_variant);
/* number of column and row tile variants (valid variants are 0..count-1) */
int nadia_tile_variant_count(int* col_variants, int* row_variants);


/* Record the device time of every fused pass launched by iterate() with HIP
 * events on the library stream (on = 1); nadia_recon_pass_times returns, per
 * pass kind (0 = column pass, 1 = row pass), the summed milliseconds and the
 * number of launches since timing was switched on. */
int nadia_recon_set_timing(nadia_recon* r, int on);
int nadia_recon_pass_times(nadia_recon* r, double ms[2], long long count[2]);
/* wait for all queued device work */
int nadia_synchronize(void);

#ifdef __cplusplus
}
#endif
#endif /* NADIA_AMD_H */
