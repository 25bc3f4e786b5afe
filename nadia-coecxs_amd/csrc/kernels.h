// kernels.h -- argument blocks and launchers of the HIP kernels (host view).
#pragma once
#include <hip/hip_runtime.h>

namespace nadia {

// Row pass: transforms along y (contiguous, length ny) of every row x.
enum RowMode {
  ROW_FFT = 0,     // dst = scale * DFT_dir(pre(src)) ; pre = optional (-1)^(x+y)
  ROW_UPDATE = 1,  // inverse row DFT -> P_F x -> algorithm update of x -> forward row DFT
  ROW_WF2 = 2,     // white field: inverse row DFT -> support -> forward row DFT
  ROW_WF4 = 3,     // white field: forward row DFT -> modulus -> store x -> inverse row DFT
  ROW_TC = 4,      // transmission constraint on src (stored back), then dst = forward row DFT
  ROW_UPDTC = 5,   // ER update storing x = support(P_F x) and the constraint's region sums of x
};

// Column pass: transforms along x (stride ny, length nx) of every column y.
enum ColMode {
  COL_FFT = 0,      // buf = scale * DFT_dir(buf)
  COL_PLANAR = 1,   // forward column DFT -> far-field modulus projection -> inverse
  COL_FRESNEL = 2,  // forward column DFT -> Fresnel phase, white field, modulus -> inverse
  COL_WF1 = 3,      // white field: inverse column DFT -> x sf*coef -> inverse column DFT
  COL_WF3 = 4,      // white field: forward column DFT -> x conj(coef) -> forward column DFT
};

// Per-reconstruction data of the fused passes.  A launch covers one or more
// reconstructions of the same precision, kind and size; blockIdx.y selects
// the job (the column pass: job * nfields + field).  Lives in device memory.
struct JobDesc {
  void* x;            // estimate
  void* w;            // work field(s): nfields consecutive nx*ny fields
  const void* s;      // support
  // the support as bytes for the update pass: 0 / 1 where it is 0 / 1, 2
  // where it is neither (read s there); null for softened supports
  const unsigned char* s8;
  const void* q;      // sqrt(I) with the beam-stop sign
  const void* ill;    // Fresnel white field (detector plane)
  double* partials;   // this job's partial error sums
  double* errs;       // this job's error ring (slot = iteration index)
  double* cur_err;    // get_error()'s slot: the update pass writes each error there too (null: not)
  double* host_err;   // a page-locked host slot it also writes (null: none), read after a stream sync
  double coef[5];     // PSF, PFS, PS, PF, PI
  int er;             // error-reduction special case
  int err_field;      // field whose projection error get_error() reports
  int store_every;    // the update pass stores x after every iteration (kept best estimates read it)
  // kept best estimates (BaseCDI::update_n_best): errors by rank, rank ->
  // buffer map, the buffer the next snapshot goes to (-1: none), how many
  double* nb_err;
  int* nb_perm;
  int* nb_target;
  int nb_n;
  double phase_factor;                                   // Fresnel
  const void* fx;     // Fresnel: (cos, sin) of phase_factor (xmid - x)^2 per row x, complex T
  const void* fy;     // the same per column y
  double wf_sx, wf_sy, wf_z12, wf_z23, wf_k, wf_norm;    // white field
};

// TransmissionConstraint state on the device (NADIA/include/TransmissionConstraint.h)
struct TcDev {
  const unsigned char* map;  // region number 1..nregions per pixel, 0 outside
  int nregions;
  const double* alpha1;      // per region
  const double* alpha2;
  double* c_mean;            // per region; recomputed unless fixed
  const int* fixed;
  int charge_flip, flip_sign, unity;
  // FresnelCDI: the buffer holds the exit-surface wave; the constraint acts
  // on T = (ESW + L)/L (L = ill, the illumination at the sample) and writes
  // ESW = (T - 1) L back in the columns j < nx (set_transmission_function)
  const void* ill;
  int nx, ny;
};

struct RowArgs {
  int mode;
  const JobDesc* jobs;  // fused modes: per-job data (blockIdx.y = job)
  int iter;             // error ring slot written by this launch
  int nrows, ny;        // rows per job, row length
  const void* tw;       // twiddle table of length ny (sign -1 values)
  // ROW_FFT: one field src -> dst, or (jobs set) every job's x -> w
  const void* src;
  void* dst;
  int dir;
  int pre_sign;  // multiply by (-1)^(x+y) before the transform
  double scale;
  // ROW_UPDATE
  int two;       // the jobs carry a second field (P_F P_S x) after the first
  // split work fields: the column pass left the projection in the second
  // field and the first still holds the row transform of x, from which the
  // update rebuilds x (x itself is not read); x is stored only by the jobs
  // [xs_first, xs_first + xs_count) (xs_count < 0: all) and those with
  // store_every, the first field always
  int split, xs_first, xs_count;
  double post;   // factor applied after the inverse row transform
  int fin_n;     // number of column-pass blocks per field (error finalisation)
  // ROW_UPDATE: the workgroup that publishes a job's error also ranks it
  // against the job's kept best errors (nb_n > 0); the snapshot copy is a
  // launch of its own after the pass
  int rank;
  // ROW_TC, ROW_UPDTC
  TcDev tc;
  double* tc_part;  // ROW_UPDTC: per-workgroup region sums, [2*(block*nregions + k)]
};

struct ColArgs {
  int mode;
  const JobDesc* jobs;  // fused modes: per-job data (blockIdx.y = job * nfields + field)
  int nfields;
  int nx, ny;
  const void* tw;  // twiddle table of length nx
  // COL_FFT (one field, no job table)
  void* buf;
  int dir;
  double scale;
  // fused modes
  double sf;  // 1/sqrt(nx*ny)
  double xmid, ymid;
  // COL_WF1: the previous iteration's error (ring slot iter - 1, from fin_n
  // row-pass partial pairs) is reduced by each job's first workgroup
  int iter, fin_n;
  // COL_PLANAR / COL_FRESNEL: the projection goes to the job's second work
  // field (split work fields, see RowArgs), not back into the first
  int split;
};

struct ConvArgs {
  const void* src;  // complex (magnitude taken) or real
  int src_complex;
  void* dst;        // real
  int nx, ny;
  // (cut+1)^2 window values, rounded to the field's real type, passed by
  // value (no upload, nothing for the host to keep alive); gf: the same
  // values as floats, read by the single-precision kernels
  double g[36];
  float gf[36];
  double* partials; // one max per block
};

// 1-D transforms of any length (kern_any.inc): the sequences of a field
// along one axis, cb whole sequences per workgroup held in LDS.
constexpr int kAnyMaxStages = 16;
constexpr int kAnyLds[3] = {16 * 1024, 64 * 1024, 128 * 1024};  // LDS classes (bytes per workgroup)
// complex values a thread holds per stage (registers, at least one butterfly): 8;
// the largest prime radix of a direct plan: 13 in FP32, 7 in FP64 (larger
// ones spill registers), beyond that Bluestein
constexpr int any_max_radix(int real_bytes) { return real_bytes == 4 ? 13 : 7; }
constexpr int any_vals(int real_bytes) { return real_bytes == 4 ? 8 : 8; }
// ANY_FFT: the transforms of `buf` along one axis.  The fused modes work on
// one job's W (JobDesc w), like the single-launch passes: ANY_COL_PLANAR /
// ANY_COL_FRESNEL = forward column DFT -> modulus projection -> inverse
// (error partials per workgroup); ANY_ROW_UPDATE = inverse row DFT -> the
// algorithm's update of x (one work field) -> forward row DFT, and the
// column pass's error published by workgroup 0.
enum AnyMode { ANY_FFT = 0, ANY_COL_PLANAR = 1, ANY_COL_FRESNEL = 2, ANY_ROW_UPDATE = 3 };
struct AnyArgs {
  int mode;
  const JobDesc* jd;    // fused modes: the job (device)
  int nx, ny;           // field size (fused modes)
  double sf;            // 1/sqrt(nx*ny)
  int iter, fin_n;      // ANY_ROW_UPDATE: error ring slot, column-pass workgroups
  void* buf;            // field, transformed in place
  long long es, ss;     // element stride and sequence stride (complex elements)
  int nseq;             // sequences (rows or columns)
  int rows;             // 1: es == 1 (load along a sequence), 0: ss == 1 (across sequences)
  int n;                // transform length
  int len;              // LDS sequence length: n, or Bluestein's m >= 2n - 1 (prime factors <= 7)
  int cb, stride;       // sequences per workgroup, LDS elements between sequences
  int dir;              // -1 forward, +1 backward (unnormalised)
  int nst;              // stages of the len-point plan
  const int* stages;    // device: (radix, span, twiddle offset) per stage
  const void* tw;       // device: twiddles of the len-point plan, sign -1
  const void* chirp;    // Bluestein: exp(-i pi k^2 / n), k < n; null for a direct plan
  const void* kern;     // Bluestein: DFT_len of the conjugate chirp kernel, times 1/len
};
template <typename T> void launch_any(const AnyArgs& a, int lds_class, int threads, hipStream_t st);

// v selects the tile variant (plan.h: col_cfg / row_cfg); all variants of a
// length compute identical results
template <typename T> void launch_row(const RowArgs& a, int njobs, int v, hipStream_t st);
template <typename T> void launch_col(const ColArgs& a, int njobs, int v, hipStream_t st);
template <typename T> int col_tile(int nx, int v);
// columns per error-partial group of the column pass (ny / col_group
// partials per field, the same for every tile variant)
template <typename T> int col_group(int nx);
template <typename T> int row_blocks(int nx, int ny, int v);


// PhaseDiverseCDI frame geometry (NADIA/src/PhaseDiverseCDI.c++): local
// pixel (i_, j_) of a frame at (x, y) is global pixel
// (trunc(i_ - x - x_min) * scale, trunc(j_ - y - y_min) * scale)
struct PdFrameDev {
  const void* weight;  // lnx x lny, element type T
  int lnx, lny;
  double x, y;
};

struct PdArgs {
  void* object;        // nx x ny complex
  int nx, ny, scale, x_min, y_min;
  int parallel;
  PdFrameDev f;        // the frame (add / update)
  void* small;         // the frame's result (lnx x lny complex)
};

// frames of one update-and-set launch, passed by value (frame k = blockIdx.y)
constexpr int kPdSetMax = 16;
struct PdSetList {
  PdFrameDev f[kPdSetMax];
  void* small[kPdSetMax];      // result, lnx x lny complex
  void* est[kPdSetMax];        // the frame's estimate
  const void* ill[kPdSetMax];  // illumination at the sample (Fresnel), else null
  int n;
};

// up to kCopyMax single doubles copied device to device in one launch
constexpr int kCopyMax = 32;
struct CopyList {
  const double* src[kCopyMax];
  double* dst[kCopyMax];
  int n;
};

// element-wise and reduction kernels (kern_misc.inc)
template <typename T> void k_invert(void* buf, int nx, int ny, int scale, hipStream_t st);
template <typename T> void k_scale_intensity(void* buf, const void* q, int nx, int ny, double* partials, int nblocks, hipStream_t st);
template <typename T> void k_fresnel_add(void* buf, const void* ill, double s, long long n, hipStream_t st);
template <typename T> void k_support(void* buf, const void* s, long long n, hipStream_t st);
template <typename T> void k_combine(void* x, const void* pfs, const void* pf, const void* ps, const void* psf,
                                     const double* coef, long long n, hipStream_t st);
template <typename T> void k_mulfac_fresnel(void* buf, int nx, int ny, double factor, int dir, hipStream_t st);
template <typename T> void k_mulfac_wf(void* buf, int nx, int ny, double sx, double sy, double z12, double z23,
                                       double k, double nrm, int dir, hipStream_t st);
template <typename T> void k_scale(void* buf, double s, long long n, hipStream_t st);
template <typename T> void k_norm2(const void* buf, long long n, double* partials, int nblocks, hipStream_t st);
template <typename T> void k_magnitude(const void* buf, void* out, long long n, hipStream_t st);
template <typename T> void k_make_q(const void* intensity, const void* beam, int has_beam, void* q, long long n, hipStream_t st);
template <typename T> void k_sqrt_to_real(const void* q, void* out, long long n, hipStream_t st);
template <typename T> void k_autocorr_prep(const void* q, void* out, long long n, hipStream_t st);
template <typename T> void k_esw_from_trans(const void* trans, const void* ill, void* esw, int nx, int ny, hipStream_t st);
template <typename T> void k_trans_from_esw(const void* esw, const void* ill, void* out, long long n, hipStream_t st);
template <typename T> void k_conv(const ConvArgs& a, int cut, int* nblocks, hipStream_t st);
template <typename T> void k_real_max(const void* src, long long n, double* partials, int nblocks, hipStream_t st);
template <typename T> void k_threshold(const void* conv, void* out, int nx, int ny, const double* maxv, double thr,
                                       hipStream_t st, unsigned char* s8 = nullptr);
// dst = src / max; s8 (if set) = 0 / 1 where dst is 0 / 1, 2 elsewhere
template <typename T>
void k_normalise(const void* src, void* dst, const double* maxv, long long n, hipStream_t st,
                 unsigned char* s8 = nullptr);
template <typename T> void k_wf_disk(void* s, int nx, int ny, double centre, double dz, double radius, hipStream_t st);
// TransmissionConstraint::apply_constraint: per-region sums (partials of
// kEwBlocks workgroups), their reduction to c_mean, and the per-pixel update
template <typename T> void k_tc_stats(const void* t, const TcDev& d, long long n, double* partials, hipStream_t st);
void k_tc_final(const TcDev& d, const double* partials, int nblocks, hipStream_t st);
template <typename T> void k_tc_apply(void* t, const TcDev& d, long long n, hipStream_t st);
// PhaseDiverseCDI add_to_object / update_from_object / scale_object
template <typename T> void k_pd_add(const PdArgs& a, hipStream_t st);
template <typename T> void k_pd_update(const PdArgs& a, hipStream_t st);
template <typename T> void k_pd_update_set(const PdArgs& a, const PdSetList& l, hipStream_t st);
template <typename T> void k_pd_get_add(const PdArgs& a, const void* est, const void* ill, hipStream_t st);
template <typename T> void k_pd_get_add_all(const PdArgs& a, const PdSetList& l, hipStream_t st);
template <typename T> void k_pd_cover(const PdArgs& a, const PdFrameDev* frames, int nframes, unsigned char* covered,
                                      hipStream_t st);
template <typename T> void k_pd_scale(const PdArgs& a, const PdFrameDev* frames, int nframes,
                                      const unsigned char* covered, double factor, hipStream_t st);
// PhaseDiverseCDI::adjust_positions (NADIA/src/PhaseDiverseCDI.c++:337-579):
// the 1 - |T| images of the global sub-grid around a frame and of the frame
// (cross-correlation), the object sub-grid (minimum error), bilinear
// up-sampling (NADIA/src/utils.c:1346-1435) and the weighted
// cross-correlation of align (NADIA/src/utils.c:1151-1343)
struct XcMaps {
  const void* object;  // nx x ny complex
  const void* single;  // the frame's result, lnx x lny complex
  void* others;        // mx x my real (or complex for the sub-grid copy)
  void* tsingle;       // lnx x lny real
  int nx, ny, lnx, lny, mx, my, scale, x_min, y_min;
  double bx, by;
};
template <typename T> void k_xc_maps(const XcMaps& a, hipStream_t st);
template <typename T> void k_pd_subgrid(const XcMaps& a, hipStream_t st);
template <typename T> void k_interp(const void* src, int nx, int ny, void* dst, int bnx, int bny, hipStream_t st);
struct XcPlace {
  const void *img1, *w1, *img2, *w2;  // real images and optional weights
  int nx1, ny1, nx2, ny2;
  void *i1, *i2, *c1, *c2;            // NX x NY complex outputs (zeroed by the caller)
  int nx, ny;
  unsigned* count;                    // weight sum (zeroed by the caller)
};
template <typename T> void k_xc_place(const XcPlace& a, hipStream_t st);
template <typename T> void k_xc_mulconj(void* a, const void* b, long long n, hipStream_t st);
struct XcBest {
  double value;
  long long index;  // -1: none
};
struct XcScan {
  const void *ci, *cw;  // correlations of the images and of the weights (complex)
  int nx, ny;
  double overlap_factor, overlap_fraction;
  int min_x, max_x, min_y, max_y;
};
// best ratio (first in index order on ties) over kEwBlocks partials, then one
template <typename T> void k_xc_scan(const XcScan& a, XcBest* partials, hipStream_t st);
void k_xc_final(const XcBest* partials, int n, XcBest* out, hipStream_t st);
// *dst[k] = *src[k] for the l.n pairs, one launch
void k_copy_doubles(const CopyList& l, hipStream_t st);
// update_n_best on the device: rank the error at `err`, then copy `bytes` of
// the estimate `src` into bufs[new rank's buffer] when it placed
void k_nbest(const double* err, double* best_err, int* perm, int n, int* target, const void* src, void* const* bufs,
             long long bytes, hipStream_t st);
// the copy alone, after an update pass that ranked the error itself
void k_nbest_copy(const int* target, const void* src, void* const* bufs, long long bytes, hipStream_t st);
void k_final_sum2(const double* partials, int n, double* out_ratio, double* out_pair, hipStream_t st);
// per-job final error reduction of the white-field row pass (grid = njobs)
void k_final_jobs(const JobDesc* jobs, int njobs, int n, int iter, hipStream_t st);
void k_final_max(const double* partials, int n, double* out, hipStream_t st);
void k_final_sum1(const double* partials, int n, double* out, hipStream_t st);

constexpr int kEwBlocks = 1024;  // grid size of grid-stride reductions
}  // namespace nadia
