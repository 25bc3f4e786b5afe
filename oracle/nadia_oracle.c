/*
 * nadia_oracle.c -- CPU restatement of NADIA's iterative phase-retrieval path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library: it
 * is imported by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg, never by the product path.
 *
 * Every function follows one reference function; the citation is given as
 * NADIA/<file>:<line> (NADIA v0.5).  The arithmetic type `real` is double by
 * default (the reference built with -DDOUBLE_PRECISION, i.e. libNADIAd) and
 * float with -DORACLE_FP32 (the default libNADIA build: float arrays, double
 * temporaries and double error accumulators).
 *
 * The reference links FFTW3, which this image does not have; the 2-D DFT here
 * is this file's own mixed-radix Stockham FFT implementing FFTW's documented
 * definition (unnormalised, exponent sign -1 for FFTW_FORWARD, row-major nx x ny).
 *
 * Pinning (tests/test_oracle_*.py): brute-force DFT, numpy.fft, closed forms
 * (delta <-> constant, plane wave <-> delta, Parseval), fftshift identity,
 * unitarity / round trips of every propagator, the named-algorithm coefficient
 * table, fixed points of every algorithm, HIO's inside/outside rule, scipy's
 * 2-D convolution for the shrinkwrap window, the closed-form Fresnel and
 * white-field phases, the homogeneous-material fixed point and the alpha = 1
 * limits of the transmission constraint, PhaseDiverseCDI's geometry and its
 * one-frame and two-identical-frame limits, a brute-force weighted
 * correlation for align, the bilinear closed form of the up-sampling, and the
 * recovery of known frame offsets by adjust_positions.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#ifdef ORACLE_FP32
typedef float real;
#else
typedef double real;
#endif

typedef struct { real re, im; } cplx;

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* 1-D mixed-radix Stockham FFT (radix 4, 2, 3, 5 and generic odd factors).  */
/* ------------------------------------------------------------------------ */

#define MAXSTAGES 32

typedef struct {
  int n;
  int nst;
  int radix[MAXSTAGES];
  int span[MAXSTAGES];      /* L: product of the radices of earlier stages */
  double *tw[MAXSTAGES];    /* per stage: L*(r-1) complex twiddles, sign -1 */
} fft_plan1;

static void plan1_make(fft_plan1 *p, int n) {
  int m = n, L = 1, s = 0;
  p->n = n;
  while (m > 1) {
    int r;
    if (m % 4 == 0) r = 4;
    else if (m % 2 == 0) r = 2;
    else if (m % 3 == 0) r = 3;
    else if (m % 5 == 0) r = 5;
    else { r = 7; while (m % r) r += 2; }
    p->radix[s] = r;
    p->span[s] = L;
    p->tw[s] = (double *)malloc(sizeof(double) * 2 * (size_t)L * (r - 1) + 16);
    for (int k = 0; k < L; k++)
      for (int q = 1; q < r; q++) {
        /* exp(-2 pi i k q / (L r)), argument reduced exactly in integers */
        long t = ((long)k * q) % ((long)L * r);
        double a = -2.0 * M_PI * (double)t / (double)(L * r);
        p->tw[s][2 * (k * (r - 1) + q - 1)] = cos(a);
        p->tw[s][2 * (k * (r - 1) + q - 1) + 1] = sin(a);
      }
    L *= r;
    m /= r;
    s++;
  }
  p->nst = s;
}

static void plan1_free(fft_plan1 *p) {
  for (int s = 0; s < p->nst; s++) free(p->tw[s]);
  p->nst = 0;
}

/* Small DFT of size r on v (double), sign = -1 or +1. */
static void small_dft(double *vr, double *vi, int r, int sign) {
  if (r == 2) {
    double ar = vr[0], ai = vi[0];
    vr[0] = ar + vr[1]; vi[0] = ai + vi[1];
    vr[1] = ar - vr[1]; vi[1] = ai - vi[1];
  } else if (r == 4) {
    double t0r = vr[0] + vr[2], t0i = vi[0] + vi[2];
    double t1r = vr[0] - vr[2], t1i = vi[0] - vi[2];
    double t2r = vr[1] + vr[3], t2i = vi[1] + vi[3];
    double t3r = vr[1] - vr[3], t3i = vi[1] - vi[3];
    /* multiply t3 by sign*i */
    double ur = -sign * t3i, ui = sign * t3r;
    vr[0] = t0r + t2r; vi[0] = t0i + t2i;
    vr[2] = t0r - t2r; vi[2] = t0i - t2i;
    vr[1] = t1r + ur;  vi[1] = t1i + ui;
    vr[3] = t1r - ur;  vi[3] = t1i - ui;
  } else {
    double bufr[64], bufi[64];
    if (r > 64) { fprintf(stderr, "oracle: radix too large\n"); exit(1); }
    for (int k = 0; k < r; k++) {
      double sr = 0, si = 0;
      for (int j = 0; j < r; j++) {
        long t = ((long)j * k) % r;
        double a = sign * 2.0 * M_PI * (double)t / (double)r;
        double c = cos(a), s = sin(a);
        sr += vr[j] * c - vi[j] * s;
        si += vr[j] * s + vi[j] * c;
      }
      bufr[k] = sr; bufi[k] = si;
    }
    for (int k = 0; k < r; k++) { vr[k] = bufr[k]; vi[k] = bufi[k]; }
  }
}

/* In-place 1-D DFT of x (n points, stride 1) using scratch y (n points). */
static void fft1(const fft_plan1 *p, cplx *x, cplx *y, int sign) {
  int n = p->n;
  cplx *in = x, *out = y;
  double vr[64], vi[64];
  for (int s = 0; s < p->nst; s++) {
    int r = p->radix[s], L = p->span[s], nr = n / r;
    const double *tw = p->tw[s];
    for (int j = 0; j < nr; j++) {
      int k = j % L;
      for (int q = 0; q < r; q++) {
        double ar = in[j + q * nr].re, ai = in[j + q * nr].im;
        if (q > 0 && k > 0) {
          /* tw holds the sign -1 twiddle; sign +1 uses its conjugate */
          double c = tw[2 * (k * (r - 1) + q - 1)];
          double d = (sign < 0) ? tw[2 * (k * (r - 1) + q - 1) + 1] : -tw[2 * (k * (r - 1) + q - 1) + 1];
          double tr = ar * c - ai * d, ti = ar * d + ai * c;
          ar = tr; ai = ti;
        }
        vr[q] = ar; vi[q] = ai;
      }
      small_dft(vr, vi, r, sign);
      int base = (j / L) * L * r + k;
      for (int q = 0; q < r; q++) {
        out[base + q * L].re = (real)vr[q];
        out[base + q * L].im = (real)vi[q];
      }
    }
    cplx *t = in; in = out; out = t;
  }
  if (in != x) memcpy(x, in, sizeof(cplx) * (size_t)n);
}

/* Unnormalised in-place 2-D DFT of the nx x ny x-major array a (element
 * (i,j) at a[i*ny+j]); sign -1 = FFTW_FORWARD, +1 = FFTW_BACKWARD.
 * Replaces FFTW's fftw_plan_dft_2d/fftw_execute at NADIA/src/Complex_2D.c++:462-503. */
void or_fft2d(cplx *a, int nx, int ny, int sign) {
  fft_plan1 px, py;
  plan1_make(&py, ny);
  plan1_make(&px, nx);
  int nmax = nx > ny ? nx : ny;
  const int B = 8; /* columns gathered per block */
  /* The timing build (ORACLE_OMP) splits the row transforms and the column
   * blocks over threads, the way FFTW's threaded plans split a 2-D
   * transform (NADIA's -DMULTI_THREADED build, NADIA/src/Complex_2D.c++:465-468);
   * each transform's arithmetic is unchanged. */
#ifdef ORACLE_OMP
#pragma omp parallel
#endif
  {
    cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)nmax * B);
    cplx *scr = (cplx *)malloc(sizeof(cplx) * (size_t)nmax);
#ifdef ORACLE_OMP
#pragma omp for schedule(static)
#endif
    for (int i = 0; i < nx; i++) fft1(&py, a + (size_t)i * ny, scr, sign);
#ifdef ORACLE_OMP
#pragma omp for schedule(static)
#endif
    for (int j0 = 0; j0 < ny; j0 += B) {
      int nb = (ny - j0) < B ? (ny - j0) : B;
      for (int i = 0; i < nx; i++)
        for (int b = 0; b < nb; b++) buf[(size_t)b * nx + i] = a[(size_t)i * ny + j0 + b];
      for (int b = 0; b < nb; b++) fft1(&px, buf + (size_t)b * nx, scr, sign);
      for (int i = 0; i < nx; i++)
        for (int b = 0; b < nb; b++) a[(size_t)i * ny + j0 + b] = buf[(size_t)b * nx + i];
    }
    free(buf); free(scr);
  }
  plan1_free(&px); plan1_free(&py);
}

/* ------------------------------------------------------------------------ */
/* Complex_2D / Double_2D operations                                         */
/* ------------------------------------------------------------------------ */

/* fftshift with 1/sqrt(nx*ny) scaling: NADIA/src/Complex_2D.c++:317-353 */
void or_invert(cplx *a, int nx, int ny, int scale) {
  int mx = nx / 2, my = ny / 2;
  real sf = 1;
  if (scale) sf = (real)(1.0 / sqrt((double)(nx * ny)));
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < my; ++j) {
      int jn = j + my, in = i + mx;
      if (i >= mx) in -= 2 * mx;
      cplx t = a[(size_t)in * ny + jn];
      a[(size_t)in * ny + jn].re = a[(size_t)i * ny + j].re * sf;
      a[(size_t)in * ny + jn].im = a[(size_t)i * ny + j].im * sf;
      a[(size_t)i * ny + j].re = t.re * sf;
      a[(size_t)i * ny + j].im = t.im * sf;
    }
}

/* |c|: NADIA/include/Complex_2D.h:265-268 (computed in real) */
static inline real c_mag(cplx c) { return (real)sqrt(c.re * c.re + c.im * c.im); }

/* set_mag: NADIA/include/Complex_2D.h:205-215 */
static inline void c_set_mag(cplx *c, real value) {
  double mag = c_mag(*c);
  if (mag == 0) { c->re = value; c->im = 0; }
  else { c->re = (real)(c->re * (value / mag)); c->im = (real)(c->im * (value / mag)); }
}

/* Complex_2D::add: NADIA/src/Complex_2D.c++:183-199 */
void or_add(cplx *a, const cplx *b, size_t n, real scale) {
  for (size_t p = 0; p < n; p++) { a[p].re += scale * b[p].re; a[p].im += scale * b[p].im; }
}

/* Complex_2D::scale: NADIA/src/Complex_2D.c++:169-180 */
void or_scale(cplx *a, size_t n, real s) {
  for (size_t p = 0; p < n; p++) { a[p].re *= s; a[p].im *= s; }
}

/* Complex_2D::get_norm: NADIA/src/Complex_2D.c++:259-270 */
real or_get_norm(const cplx *a, size_t n) {
  real s = 0;
  for (size_t p = 0; p < n; p++) s += (real)(pow(a[p].im, 2) + pow(a[p].re, 2));
  return (real)sqrt(s);
}

/* Real_2D::get_max: NADIA/include/Double_2D.h:252-265 */
static real r_max(const real *a, size_t n) {
  if (n == 0) return 0;
  real m = a[0];
  for (size_t p = 0; p < n; p++) if (a[p] > m) m = a[p];
  return m;
}

/* Real_2D::get_sum: NADIA/include/Double_2D.h:210-216 */
static real r_sum(const real *a, size_t n) {
  real t = 0;
  for (size_t p = 0; p < n; p++) t += a[p];
  return t;
}

/* ------------------------------------------------------------------------ */
/* BaseCDI state                                                             */
/* ------------------------------------------------------------------------ */

enum { PSF, PFS, PS, PF, PI_ };
enum { ER, BIO, BOO, HIO, DM, SF, ASR, HPR, RAAR, CUSTOM };
enum { KIND_PLANAR = 0, KIND_FRESNEL = 1, KIND_WF = 2 };

typedef struct {
  int kind, nx, ny;
  cplx *complex;            /* the caller's estimate, mutated in place */
  real *support, *intensity_sqrt, *beam_stop;
  double beta;
  int algorithm;
  double alg[5];
  double current_error;
  cplx *tPFS, *tPF, *tPS, *tPSF;
  int n_best;
  cplx **best;
  double *best_err;
  /* FresnelCDI */
  cplx *illumination, *illum_at_sample, *coefficient, *transmission;
  double norm, wavelength, pixel, zfd, zfs;
  /* FresnelCDI_WF */
  double z12, z23;
  /* transmission constraint (not owned), NULL when none is set */
  struct otc *tc;
} ocdi;

#define NPIX(o) ((size_t)(o)->nx * (o)->ny)

static cplx *calloc_c(size_t n) { return (cplx *)calloc(n ? n : 1, sizeof(cplx)); }
static real *calloc_r(size_t n) { return (real *)calloc(n ? n : 1, sizeof(real)); }

void ocdi_set_algorithm(ocdi *o, int alg);

/* BaseCDI::BaseCDI: NADIA/src/BaseCDI.c++:33-66 */
ocdi *ocdi_create(cplx *initial_guess, int nx, int ny, int kind, int n_best) {
  ocdi *o = (ocdi *)calloc(1, sizeof(ocdi));
  o->kind = kind; o->nx = nx; o->ny = ny;
  o->complex = initial_guess;
  o->beta = 0.9;
  o->support = calloc_r(NPIX(o));
  o->intensity_sqrt = calloc_r(NPIX(o));
  o->n_best = n_best;
  if (n_best > 0) {
    o->best = (cplx **)malloc(sizeof(cplx *) * n_best);
    o->best_err = (double *)malloc(sizeof(double) * n_best);
    for (int i = 0; i < n_best; i++) { o->best[i] = calloc_c(NPIX(o)); o->best_err[i] = 1; }
  }
  o->algorithm = CUSTOM;
  ocdi_set_algorithm(o, ER);
  return o;
}

void ocdi_destroy(ocdi *o) {
  if (!o) return;
  free(o->support); free(o->intensity_sqrt); free(o->beam_stop);
  free(o->tPFS); free(o->tPF); free(o->tPS); free(o->tPSF);
  for (int i = 0; i < o->n_best; i++) free(o->best[i]);
  free(o->best); free(o->best_err);
  free(o->illumination); free(o->illum_at_sample); free(o->coefficient); free(o->transmission);
  free(o);
}

/* BaseCDI::convolve: NADIA/src/BaseCDI.c++:432-474 (unnormalised truncated
 * Gaussian, zero outside the array, window stored in real; each product
 * array.get() * gauss_dist.get() is formed in real, float * float in the
 * single-precision build, and added to a double accumulator) */
void ocdi_convolve(ocdi *o, real *array, double gauss_width, int cut) {
  int nx = o->nx, ny = o->ny;
  real *tmp = calloc_r(NPIX(o));
  real *g = calloc_r((size_t)(cut + 1) * (cut + 1));
  for (int i = 0; i <= cut; i++)
    for (int j = 0; j <= cut; j++) {
      double denom = 2.0 * gauss_width * gauss_width;
      g[i * (cut + 1) + j] = (real)exp(-1 * (i * i + j * j) / denom);
    }
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      double v = 0;
      for (int i2 = i - cut; i2 <= i + cut; i2++)
        for (int j2 = j - cut; j2 <= j + cut; j2++)
          if (i2 < nx && i2 >= 0 && j2 >= 0 && j2 < ny)
            v += (double)(real)(array[(size_t)i2 * ny + j2] * g[abs(i - i2) * (cut + 1) + abs(j - j2)]);
      tmp[(size_t)i * ny + j] = (real)v;
    }
  memcpy(array, tmp, sizeof(real) * NPIX(o));
  free(tmp); free(g);
}

/* BaseCDI::set_support: NADIA/src/BaseCDI.c++:95-108 */
void ocdi_set_support(ocdi *o, const real *s, int soften) {
  double max = r_max(s, NPIX(o));
  for (size_t p = 0; p < NPIX(o); p++) o->support[p] = (real)(s[p] / max);
  if (soften) ocdi_convolve(o, o->support, 3, 5);
}

/* BaseCDI::set_beam_stop: NADIA/src/BaseCDI.c++:112-116 */
void ocdi_set_beam_stop(ocdi *o, const real *b) {
  if (!o->beam_stop) o->beam_stop = calloc_r(NPIX(o));
  memcpy(o->beam_stop, b, sizeof(real) * NPIX(o));
}

/* BaseCDI::set_intensity: NADIA/src/BaseCDI.c++:118-124 */
void ocdi_set_intensity(ocdi *o, const real *I) {
  for (size_t p = 0; p < NPIX(o); p++) o->intensity_sqrt[p] = (real)sqrt(I[p]);
}

/* BaseCDI::get_error: NADIA/src/BaseCDI.c++:126-128 */
double ocdi_get_error(ocdi *o) { return o->current_error; }

/* BaseCDI::support_constraint: NADIA/src/BaseCDI.c++:136-154 */
static void support_constraint(ocdi *o, cplx *c) {
  for (size_t p = 0; p < NPIX(o); p++) {
    double s = o->support[p];
    if (s == 0) { c[p].re = 0; c[p].im = 0; }
    else if (s < 1) { c[p].re = (real)(c[p].re * s); c[p].im = (real)(c[p].im * s); }
  }
}

/* ---- TransmissionConstraint ------------------------------------------- */
/* NADIA/src/TransmissionConstraint.c++:26-172 with ComplexConstraint from
 * NADIA/include/TransmissionConstraint.h.  Regions are numbered 1..n in
 * region_map (the first region to claim a pixel keeps it); the custom
 * constraint callback is not modelled. */
#define OTC_MAX 32
typedef struct otc {
  int nx, ny, nregions;
  real *region_map;
  double alpha1[OTC_MAX], alpha2[OTC_MAX], c_mean[OTC_MAX];
  int fixed_c[OTC_MAX];
  int enforce_unity, charge_flip, flip_sign;
} otc;

/* TransmissionConstraint::TransmissionConstraint: unity and charge flipping
 * (sign +1) enabled */
otc *otc_create(int nx, int ny) {
  otc *t = (otc *)calloc(1, sizeof(otc));
  t->nx = nx; t->ny = ny;
  t->enforce_unity = 1; t->charge_flip = 1; t->flip_sign = 1;
  return t;
}

void otc_destroy(otc *t) { if (t) { free(t->region_map); free(t); } }

void otc_set_charge_flipping(otc *t, int enable, int flip_sign) {
  t->charge_flip = enable;
  t->flip_sign = flip_sign > 0 ? 1 : -1;
}

void otc_set_enforce_unity(otc *t, int enable) { t->enforce_unity = enable; }

/* add_complex_constraint: NADIA/src/TransmissionConstraint.c++:43-77;
 * returns the region index or -1 */
int otc_add_region(otc *t, const real *region, double alpha1, double alpha2) {
  if (t->nregions >= OTC_MAX) return -1;
  if (!t->region_map) t->region_map = calloc_r((size_t)t->nx * t->ny);
  int k = t->nregions++;
  t->alpha1[k] = alpha1; t->alpha2[k] = alpha2; t->fixed_c[k] = 0; t->c_mean[k] = 0;
  for (size_t p = 0; p < (size_t)t->nx * t->ny; p++)
    if (region[p] > 0 && !(t->region_map[p] > 0)) t->region_map[p] = (real)(k + 1);
  return k;
}

/* ComplexConstraint::set_fixed_c */
void otc_set_fixed_c(otc *t, int k, double c) { t->fixed_c[k] = 1; t->c_mean[k] = c; }
double otc_get_c_mean(otc *t, int k) { return t->c_mean[k]; }

static inline double c_phase(cplx c) { return (real)atan2(c.im, c.re); }

/* Complex_2D::set_mag / set_phase (NADIA/include/Complex_2D.h:205-230) */
static inline void c_set_mag_d(cplx *c, double value) {
  double mag = c_mag(*c);
  if (mag == 0) { c->re = (real)value; c->im = 0; }
  else { c->re = (real)(c->re * (value / mag)); c->im = (real)(c->im * (value / mag)); }
}
static inline void c_set_phase_d(cplx *c, double value) {
  double mag = c_mag(*c);
  c->re = (real)(mag * cos(value));
  c->im = (real)(mag * sin(value));
}

/* apply_constraint: NADIA/src/TransmissionConstraint.c++:81-172 */
void otc_apply(otc *t, cplx *tr) {
  size_t n = (size_t)t->nx * t->ny;
  int R = t->nregions;
  if (R > 0) {
    double lnA[OTC_MAX] = {0}, ph[OTC_MAX] = {0};
    for (size_t p = 0; p < n; p++)
      if (t->region_map[p] > 0) {
        int k = (int)t->region_map[p] - 1;
        ph[k] += fabs((double)(real)c_phase(tr[p]));
        lnA[k] += fabs(log((double)c_mag(tr[p])));
      }
    for (int k = 0; k < R; k++)
      if (!t->fixed_c[k]) t->c_mean[k] = lnA[k] / ph[k];
  }
  if (!(R > 0 || t->charge_flip || t->enforce_unity)) return;
  for (size_t p = 0; p < n; p++) {
    double phase_old = (real)c_phase(tr[p]);
    double mag_old = c_mag(tr[p]);
    if (R > 0 && t->region_map[p] > 0) {
      int k = (int)t->region_map[p] - 1;
      double c = t->c_mean[k];
      double mag_new = exp((1 - t->alpha1[k]) * log(mag_old) + t->alpha1[k] * c * phase_old);
      double phase_new = (1 - t->alpha2[k]) * phase_old + t->alpha2[k] * log(mag_old) / c;
      c_set_mag_d(&tr[p], mag_new);
      c_set_phase_d(&tr[p], phase_new);
      mag_old = mag_new;
    }
    if (t->charge_flip && t->flip_sign * tr[p].im > 0) tr[p].im = -tr[p].im;
    if (t->enforce_unity && mag_old > 1) c_set_mag_d(&tr[p], 1.0);
  }
}

/* ---- FresnelCDI transmission helpers ------------------------------------ */

/* FresnelCDI::set_transmission_function: NADIA/src/FresnelCDI.c++:253-283
 * (the loop runs j < nx, so for nx < ny the columns j >= nx keep their
 * values; for nx > ny the reference's loop runs past each row and, on the
 * last row, past the array -- undefined behaviour; here each pixel is then
 * written once, as the device kernel does) */
void ocdi_set_transmission_function(ocdi *o, const cplx *T, cplx *esw) {
  int nx = o->nx, ny = o->ny;
  const int jn = nx < ny ? nx : ny;
  if (!esw) esw = o->complex;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < jn; j++) {
      size_t p = (size_t)i * ny + j;
      double lr = o->illum_at_sample[p].re, li = o->illum_at_sample[p].im;
      double tr = T[p].re, ti = T[p].im;
      esw[p].re = (real)(lr * tr - li * ti - lr);
      esw[p].im = (real)(lr * ti + li * tr - li);
    }
}

/* FresnelCDI::get_transmission_function: NADIA/src/FresnelCDI.c++:298-355 */
void ocdi_get_transmission_function(ocdi *o, cplx *result, const cplx *esw) {
  if (!esw) esw = o->complex;
  for (size_t p = 0; p < NPIX(o); p++) {
    double lr = o->illum_at_sample[p].re, li = o->illum_at_sample[p].im;
    double den = lr * lr + li * li;
    if (den != 0) {
      double er = esw[p].re, ei = esw[p].im;
      double rn = lr * (er + lr) + li * (ei + li);
      double in = lr * (ei + li) - li * (er + lr);
      result[p].re = (real)(rn / den);
      result[p].im = (real)(in / den);
    } else { result[p].re = 1.0; result[p].im = 0.0; }
  }
  if (o->tc) otc_apply(o->tc, result);
}

/* ---- virtual methods dispatched on kind ----------------------------- */

/* FresnelCDI::multiply_factors: NADIA/src/FresnelCDI.c++:98-126, and
 * FresnelCDI_WF::multiply_factors: NADIA/src/FresnelCDI_WF.c++:70-96
 * (coefficient table nx/2 x ny/2 mirrored about (n-1)/2) */
static void multiply_factors(ocdi *o, cplx *c, int direction) {
  int nx = o->nx, ny = o->ny;
  double xm = (nx - 1) / 2.0, ym = (ny - 1) / 2.0;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      size_t p = (size_t)i * ny + j;
      double orr = c[p].re, oi = c[p].im;
      int i_ = (int)(xm - fabs(xm - i));
      int j_ = (int)(ym - fabs(ym - j));
      double cr = o->coefficient[(size_t)i_ * (ny / 2) + j_].re;
      double ci = direction * (double)o->coefficient[(size_t)i_ * (ny / 2) + j_].im;
      c[p].re = (real)(orr * cr - oi * ci);
      c[p].im = (real)(oi * cr + orr * ci);
    }
}

#define FORWARD (-1)
#define BACKWARD (+1)

/* PlanarCDI::propagate_to_detector: NADIA/src/PlanarCDI.c++:65-68
 * FresnelCDI::propagate_to_detector: NADIA/src/FresnelCDI.c++:245-250
 * FresnelCDI_WF::propagate_to_detector: NADIA/src/FresnelCDI_WF.c++:164-175 */
void ocdi_propagate_to_detector(ocdi *o, cplx *c) {
  int nx = o->nx, ny = o->ny;
  if (o->kind == KIND_PLANAR) {
    or_fft2d(c, nx, ny, -1);
    or_invert(c, nx, ny, 1);
  } else if (o->kind == KIND_FRESNEL) {
    or_invert(c, nx, ny, 1);
    or_fft2d(c, nx, ny, -1);
    multiply_factors(o, c, FORWARD);
  } else {
    or_fft2d(c, nx, ny, -1);
    multiply_factors(o, c, FORWARD);
    or_invert(c, nx, ny, 1);
    or_fft2d(c, nx, ny, -1);
  }
}

/* PlanarCDI::propagate_from_detector: NADIA/src/PlanarCDI.c++:70-73
 * FresnelCDI::propagate_from_detector: NADIA/src/FresnelCDI.c++:238-243
 * FresnelCDI_WF::propagate_from_detector: NADIA/src/FresnelCDI_WF.c++:152-162 */
void ocdi_propagate_from_detector(ocdi *o, cplx *c) {
  int nx = o->nx, ny = o->ny;
  if (o->kind == KIND_PLANAR) {
    or_invert(c, nx, ny, 1);
    or_fft2d(c, nx, ny, +1);
  } else if (o->kind == KIND_FRESNEL) {
    multiply_factors(o, c, BACKWARD);
    or_fft2d(c, nx, ny, +1);
    or_invert(c, nx, ny, 1);
  } else {
    or_fft2d(c, nx, ny, +1);
    or_invert(c, nx, ny, 1);
    multiply_factors(o, c, BACKWARD);
    or_fft2d(c, nx, ny, +1);
  }
}

/* BaseCDI::scale_intensity: NADIA/src/BaseCDI.c++:163-189 */
static void base_scale_intensity(ocdi *o, cplx *c) {
  double norm2_mag = 0, norm2_diff = 0;
  for (size_t p = 0; p < NPIX(o); p++) {
    if (o->beam_stop == NULL || o->beam_stop[p] > 0) {
      double is = o->intensity_sqrt[p];
      double mag = c_mag(c[p]);
      c_set_mag(&c[p], (real)is);
      norm2_mag += is * is;
      norm2_diff += (mag - is) * (mag - is);
    }
  }
  o->current_error = norm2_diff / norm2_mag;
}

/* BaseCDI::scale_intensity, FresnelCDI::scale_intensity
 * (NADIA/src/FresnelCDI.c++:217-236: add white field, project, subtract) */
void ocdi_scale_intensity(ocdi *o, cplx *c) {
  if (o->kind == KIND_FRESNEL) {
    or_add(c, o->illumination, NPIX(o), 1);
    base_scale_intensity(o, c);
    or_add(c, o->illumination, NPIX(o), -1.0);
  } else {
    base_scale_intensity(o, c);
  }
}

void ocdi_set_complex_constraint(ocdi *o, otc *t) { o->tc = t; }

/* BaseCDI::apply_support: NADIA/src/BaseCDI.c++:130-134;
 * FresnelCDI::apply_support: NADIA/src/FresnelCDI.c++:203-215 */
void ocdi_apply_support(ocdi *o, cplx *c) {
  support_constraint(o, c);
  if (!o->tc) return;
  if (o->kind == KIND_FRESNEL) {
    if (!o->transmission) o->transmission = calloc_c(NPIX(o));
    ocdi_get_transmission_function(o, o->transmission, c);  /* applies the constraint */
    ocdi_set_transmission_function(o, o->transmission, c);
  } else {
    otc_apply(o->tc, c);
  }
}

/* BaseCDI::project_intensity: NADIA/src/BaseCDI.c++:156-160 */
void ocdi_project_intensity(ocdi *o, cplx *c) {
  ocdi_propagate_to_detector(o, c);
  ocdi_scale_intensity(o, c);
  ocdi_propagate_from_detector(o, c);
}

/* BaseCDI::set_custom_algorithm: NADIA/src/BaseCDI.c++:240-255 plus
 * reallocate_temp_complex_memory :257-275 */
void ocdi_set_custom_algorithm(ocdi *o, double m1, double m2, double m3, double m4, double m5,
                               double m6, double m7, double m8, double m9, double m10) {
  o->alg[PSF] = m1 + m2 + m3 + m4;
  o->alg[PFS] = -m1 + m5 + m6 + m7;
  o->alg[PS] = -m3 - m6 - m8 + m10;
  o->alg[PF] = -m2 - m5 + m8 + m9;
  o->alg[PI_] = -m4 - m7 - m9 - m10;
  cplx **t[4] = {&o->tPSF, &o->tPFS, &o->tPS, &o->tPF};
  for (int n = 0; n < 4; n++) {
    if (o->alg[n] == 0 && *t[n]) { free(*t[n]); *t[n] = NULL; }
    if (o->alg[n] != 0 && !*t[n]) *t[n] = calloc_c(NPIX(o));
  }
  o->algorithm = CUSTOM;
}

/* BaseCDI::set_algorithm: NADIA/src/BaseCDI.c++:192-237 (early return when
 * the algorithm is unchanged, so beta only applies on a change) */
void ocdi_set_algorithm(ocdi *o, int alg) {
  if (o->kind == KIND_WF && o->algorithm == ER) return; /* FresnelCDI_WF.h:147-150 hides it */
  if (o->algorithm == alg) return;
  double b = o->beta;
  switch (alg) {
    case ER: ocdi_set_custom_algorithm(o, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0); break;
    case BIO: ocdi_set_custom_algorithm(o, 0, b, 0, 0, 0, 0, 0, 0, 0, 0); break;
    case BOO: ocdi_set_custom_algorithm(o, 0, b, 0, 0, 0, 0, 0, 0, 1, 0); break;
    case HIO: ocdi_set_custom_algorithm(o, 0, b, 1, 0, 0, 0, 0, 0, 0, 0); break;
    case DM: ocdi_set_custom_algorithm(o, b, 0, 1, 0, 1, 0, 0, 0, 0, 0); break;
    case SF: ocdi_set_custom_algorithm(o, 0, 1, 0, 1, 0, 0, 0, 0, 0, 0); break;
    case ASR: ocdi_set_custom_algorithm(o, 0, 1, 1, 0, 0, 0, 0, 0, 0, 0); break;
    case HPR: ocdi_set_custom_algorithm(o, 0, b, 1, 0, 0, 0, 0, 0, 0, 0); break;
    case RAAR: ocdi_set_custom_algorithm(o, 0, 1, b, 0, 0, 0, 0, 0, (1 - b), 0); break;
    default: fprintf(stdout, "Algorithm unknown\n");
  }
  o->algorithm = alg;
}

/* BaseCDI::set_relaxation_parameter: NADIA/include/BaseCDI.h:264-267 */
void ocdi_set_relaxation_parameter(ocdi *o, double beta) {
  if (o->kind == KIND_WF) return; /* FresnelCDI_WF.h:168-172 */
  o->beta = beta;
  ocdi_set_algorithm(o, o->algorithm);
}

void ocdi_get_coefficients(ocdi *o, double *out5, int *algorithm) {
  for (int i = 0; i < 5; i++) out5[i] = o->alg[i];
  if (algorithm) *algorithm = o->algorithm;
}

/* BaseCDI::update_n_best: NADIA/src/BaseCDI.c++:390-413 */
static void update_n_best(ocdi *o) {
  int place = 0;
  for (; place < o->n_best && o->current_error > o->best_err[place]; place++);
  if (o->n_best > 0 && place < o->n_best) {
    cplx *tp = o->best[o->n_best - 1];
    memcpy(tp, o->complex, sizeof(cplx) * NPIX(o));
    for (int i = o->n_best - 1; i > place; i--) {
      o->best_err[i] = o->best_err[i - 1];
      o->best[i] = o->best[i - 1];
    }
    o->best_err[place] = o->current_error;
    o->best[place] = tp;
  }
}

/* BaseCDI::get_best_result: NADIA/src/BaseCDI.c++:82-91 */
int ocdi_get_best_result(ocdi *o, int index, double *error, cplx *out) {
  if (index >= 0 && index < o->n_best) {
    *error = o->best_err[index];
    memcpy(out, o->best[index], sizeof(cplx) * NPIX(o));
    return 1;
  }
  return 0;
}

/* BaseCDI::reset_best: NADIA/include/BaseCDI.h:476-479 */
void ocdi_reset_best(ocdi *o) { for (int i = 0; i < o->n_best; i++) o->best_err[i] = 1; }

/* FresnelCDI_WF::iterate: NADIA/src/FresnelCDI_WF.c++:122-150 */
static int wf_iterate(ocdi *o) {
  ocdi_propagate_from_detector(o, o->complex);
  ocdi_apply_support(o, o->complex);
  ocdi_propagate_to_detector(o, o->complex);
  ocdi_scale_intensity(o, o->complex);
  return 1;
}

/* BaseCDI::iterate: NADIA/src/BaseCDI.c++:297-387 */
int ocdi_iterate(ocdi *o) {
  if (o->kind == KIND_WF) return wf_iterate(o);
  cplx *x = o->complex;
  size_t n = NPIX(o);
  if (o->algorithm == ER) {
    ocdi_project_intensity(o, x);
    ocdi_apply_support(o, x);
    update_n_best(o);
    return 1;
  }
  if (o->alg[PFS] != 0) {
    memcpy(o->tPFS, x, sizeof(cplx) * n);
    ocdi_apply_support(o, o->tPFS);
    ocdi_project_intensity(o, o->tPFS);
  }
  if (o->alg[PF] != 0) {
    memcpy(o->tPF, x, sizeof(cplx) * n);
    ocdi_project_intensity(o, o->tPF);
  }
  if (o->alg[PS] != 0) {
    memcpy(o->tPS, x, sizeof(cplx) * n);
    ocdi_apply_support(o, o->tPS);
  }
  if (o->alg[PSF] != 0) {
    if (o->alg[PF] != 0) {
      memcpy(o->tPSF, o->tPF, sizeof(cplx) * n);
      ocdi_apply_support(o, o->tPSF);
    } else {
      memcpy(o->tPSF, x, sizeof(cplx) * n);
      ocdi_project_intensity(o, o->tPSF);
      ocdi_apply_support(o, o->tPSF);
    }
  }
  for (size_t p = 0; p < n; p++) {
    double vr = (1 + o->alg[PI_]) * x[p].re;
    double vi = (1 + o->alg[PI_]) * x[p].im;
    if (o->alg[PFS] != 0) { vr += o->alg[PFS] * o->tPFS[p].re; vi += o->alg[PFS] * o->tPFS[p].im; }
    if (o->alg[PF] != 0) { vr += o->alg[PF] * o->tPF[p].re; vi += o->alg[PF] * o->tPF[p].im; }
    if (o->alg[PS] != 0) { vr += o->alg[PS] * o->tPS[p].re; vi += o->alg[PS] * o->tPS[p].im; }
    if (o->alg[PSF] != 0) { vr += o->alg[PSF] * o->tPSF[p].re; vi += o->alg[PSF] * o->tPSF[p].im; }
    x[p].re = (real)vr;
    x[p].im = (real)vi;
  }
  update_n_best(o);
  return 1;
}

/* BaseCDI::apply_threshold: NADIA/src/BaseCDI.c++:477-492 (loops j < nx, so
 * for nx < ny the columns j >= nx keep their values).  For nx > ny the
 * reference's loop runs past the end of each row and, on the last row, past
 * the array (undefined behaviour); here each pixel is then thresholded once,
 * as the device kernel does. */
static void apply_threshold(ocdi *o, real *a, double threshold) {
  double max = r_max(a, NPIX(o));
  const int jn = o->nx < o->ny ? o->nx : o->ny;
  for (int i = 0; i < o->nx; i++)
    for (int j = 0; j < jn; j++) {
      size_t p = (size_t)i * o->ny + j;
      if (a[p] < threshold * max) a[p] = 0.0;
      else a[p] = 1.0;
    }
}

/* BaseCDI::apply_shrinkwrap: NADIA/src/BaseCDI.c++:416-429 */
void ocdi_apply_shrinkwrap(ocdi *o, double gauss_width, double threshold) {
  real *recon = calloc_r(NPIX(o));
  for (size_t p = 0; p < NPIX(o); p++) recon[p] = c_mag(o->complex[p]);
  ocdi_convolve(o, recon, gauss_width, 4);
  apply_threshold(o, recon, threshold);
  ocdi_set_support(o, recon, 0);
  free(recon);
}

void ocdi_get_support(ocdi *o, real *out) { memcpy(out, o->support, sizeof(real) * NPIX(o)); }
void ocdi_get_intensity_sqrt(ocdi *o, real *out) { memcpy(out, o->intensity_sqrt, sizeof(real) * NPIX(o)); }

/* PlanarCDI::initialise_estimate: NADIA/src/PlanarCDI.c++:39-62 (glibc rand;
 * max sqrt(I) truncated to int) */
static void planar_initialise_estimate(ocdi *o, int seed) {
  srand(seed);
  int max_value = (int)r_max(o->intensity_sqrt, NPIX(o));
  for (size_t p = 0; p < NPIX(o); p++) {
    if (!o->support[p]) { o->complex[p].re = 0; o->complex[p].im = 0; }
    else {
      /* the reference multiplies two ints (max_value*rand()); the product
       * wraps modulo 2^32 on its target, reproduced here without UB */
      int pr = (int)((unsigned)max_value * (unsigned)rand());
      double r = o->support[p] * (pr / (double)RAND_MAX);
      int pi = (int)((unsigned)max_value * (unsigned)rand());
      double im = o->support[p] * (pi / (double)RAND_MAX);
      o->complex[p].re = (real)r;
      o->complex[p].im = (real)im;
    }
  }
}

/* FresnelCDI::initialise_estimate: NADIA/src/FresnelCDI.c++:152-201 */
static void fresnel_initialise_estimate(ocdi *o, int seed) {
  srand(seed);
  cplx *rt = calloc_c(NPIX(o));
  for (size_t p = 0; p < NPIX(o); p++) {
    rt[p].re = (real)(1 - 0.2 * rand() / ((double)RAND_MAX));
    rt[p].im = 0;
  }
  ocdi_set_transmission_function(o, rt, o->complex);
  ocdi_apply_support(o, o->complex);
  free(rt);
}

/* FresnelCDI_WF::initialise_estimate: NADIA/src/FresnelCDI_WF.c++:99-117 */
static void wf_initialise_estimate(ocdi *o, int seed) {
  srand(seed);
  for (size_t p = 0; p < NPIX(o); p++) {
    if (!o->support[p]) { o->complex[p].re = 0; o->complex[p].im = 0; }
    else { o->complex[p].re = o->intensity_sqrt[p]; o->complex[p].im = 0; }
  }
}

void ocdi_initialise_estimate(ocdi *o, int seed) {
  if (o->kind == KIND_PLANAR) planar_initialise_estimate(o, seed);
  else if (o->kind == KIND_FRESNEL) fresnel_initialise_estimate(o, seed);
  else wf_initialise_estimate(o, seed);
}

/* PlanarCDI::get_intensity_autocorrelation: NADIA/src/PlanarCDI.c++:13-36 */
void ocdi_get_intensity_autocorrelation(ocdi *o, real *autoc) {
  cplx *t = calloc_c(NPIX(o));
  for (size_t p = 0; p < NPIX(o); p++) {
    double comp = (1.0 / sqrt(2.0)) * ((double)o->intensity_sqrt[p] * o->intensity_sqrt[p]);
    t[p].re = (real)comp;
    t[p].im = (real)comp;
  }
  or_fft2d(t, o->nx, o->ny, +1);
  or_invert(t, o->nx, o->ny, 1);
  for (size_t p = 0; p < NPIX(o); p++) autoc[p] = c_mag(t[p]);
  free(t);
}

/* ---- FresnelCDI set-up ------------------------------------------------ */

/* FresnelCDI::set_experimental_parameters: NADIA/src/FresnelCDI.c++:57-95 */
void ocdi_fresnel_set_experimental_parameters(ocdi *o, double wl, double zfd, double zfs, double pixel) {
  int nx = o->nx, ny = o->ny;
  o->wavelength = wl; o->pixel = pixel; o->zfd = zfd; o->zfs = zfs;
  double xm = (nx - 1) / 2.0, ym = (ny - 1) / 2.0;
  double zsd = zfd - zfs;
  double factor = pixel * pixel * M_PI / wl * ((1 / zfd) - (1 / zsd));
  for (int i = 0; i < nx / 2.0; i++)
    for (int j = 0; j < ny / 2.0; j++) {
      double phi = factor * ((xm - i) * (xm - i) + (ym - j) * (ym - j));
      o->coefficient[(size_t)i * (ny / 2) + j].re = (real)cos(phi);
      o->coefficient[(size_t)i * (ny / 2) + j].im = (real)sin(phi);
    }
  if (!o->illum_at_sample) o->illum_at_sample = calloc_c(NPIX(o));
  memcpy(o->illum_at_sample, o->illumination, sizeof(cplx) * NPIX(o));
  ocdi_propagate_from_detector(o, o->illum_at_sample);
}

/* FresnelCDI::FresnelCDI: NADIA/src/FresnelCDI.c++:20-44 */
void ocdi_fresnel_init(ocdi *o, const cplx *white_field, double wl, double zfd, double zfs,
                       double pixel, double normalisation) {
  o->norm = normalisation;
  o->illumination = calloc_c(NPIX(o));
  o->coefficient = calloc_c((size_t)(o->nx / 2) * (o->ny / 2));
  memcpy(o->illumination, white_field, sizeof(cplx) * NPIX(o));
  or_scale(o->illumination, NPIX(o), (real)o->norm);
  ocdi_fresnel_set_experimental_parameters(o, wl, zfd, zfs, pixel);
  ocdi_set_algorithm(o, ER);
}

/* FresnelCDI::set_norm: NADIA/src/FresnelCDI.c++:139-148 */
void ocdi_fresnel_set_norm(ocdi *o, double nn) {
  double old = o->norm;
  o->norm = nn;
  or_scale(o->illumination, NPIX(o), (real)(nn / old));
  if (o->illum_at_sample) or_scale(o->illum_at_sample, NPIX(o), (real)(nn / old));
}

/* FresnelCDI::auto_set_norm: NADIA/src/FresnelCDI.c++:129-137 */
void ocdi_fresnel_auto_set_norm(ocdi *o) {
  real *t = calloc_r(NPIX(o));
  for (size_t p = 0; p < NPIX(o); p++) t[p] = c_mag(o->illumination[p]);
  double wf = r_sum(t, NPIX(o));
  double in = r_sum(o->intensity_sqrt, NPIX(o));
  free(t);
  ocdi_fresnel_set_norm(o, in / wf);
}

void ocdi_fresnel_get_illumination_at_sample(ocdi *o, cplx *out) {
  memcpy(out, o->illum_at_sample, sizeof(cplx) * NPIX(o));
}
void ocdi_fresnel_get_illumination(ocdi *o, cplx *out) {
  memcpy(out, o->illumination, sizeof(cplx) * NPIX(o));
}

/* ---- FresnelCDI_WF set-up --------------------------------------------- */

/* FresnelCDI_WF::FresnelCDI_WF: NADIA/src/FresnelCDI_WF.c++:17-67 */
void ocdi_wf_init(ocdi *o, double wl, double z12, double z23, double pixel) {
  int nx = o->nx, ny = o->ny;
  o->wavelength = wl; o->z12 = z12; o->z23 = z23; o->pixel = pixel;
  o->coefficient = calloc_c((size_t)(nx / 2) * (ny / 2));
  double xm = (nx - 1) / 2.0, ym = (ny - 1) / 2.0;
  double sx = wl * z23 / (pixel * nx), sy = wl * z23 / (pixel * ny);
  double nrm = 1 / (sqrt((double)(nx * ny)));
  for (int i = 0; i < nx / 2; i++)
    for (int j = 0; j < ny / 2; j++) {
      double rho2 = pow(sx * (xm - i), 2) + pow(sy * (ym - j), 2);
      double phi = rho2 / z12 + rho2 / z23;
      phi *= M_PI / wl;
      o->coefficient[(size_t)i * (ny / 2) + j].re = (real)(sin(phi) * nrm);
      o->coefficient[(size_t)i * (ny / 2) + j].im = (real)(-cos(phi) * nrm);
    }
}

/* FresnelCDI_WF::set_support: NADIA/src/FresnelCDI_WF.c++:177-196 (integer
 * centre (nx-1)/2, square arrays, softened and left unnormalised) */
void ocdi_wf_set_support(ocdi *o, double z_diameter, double size) {
  int nx = o->nx, ny = o->ny;
  double centre = (nx - 1) / 2;
  double dz = (o->z12 / o->z23) * o->pixel;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      double i_ = i - centre, j_ = j - centre;
      double r = dz * sqrt(i_ * i_ + j_ * j_);
      if (r < size * z_diameter / 2.0) o->support[(size_t)i * ny + j] = 1;
    }
  ocdi_convolve(o, o->support, 3, 5);
}

/* Run n iterations, recording get_error() after each (the driver loop of
 * NADIA/tools/CDI_reconstruction.c:371-439 without file output). */
void ocdi_iterate_n(ocdi *o, int n, double *errors) {
  for (int k = 0; k < n; k++) {
    ocdi_iterate(o);
    if (errors) errors[k] = o->current_error;
  }
}

int oracle_real_size(void) { return (int)sizeof(real); }

/* ------------------------------------------------------------------------ */
/* PhaseDiverseCDI: NADIA/src/PhaseDiverseCDI.c++ with                       */
/* NADIA/include/PhaseDiverseCDI.h (position refinement: see below)         */
/* ------------------------------------------------------------------------ */

#define OPD_MAX 64
typedef struct {
  double beta, gamma;
  int parallel, scale, ipc;
  int nframes;
  ocdi *frame[OPD_MAX];
  double xpos[OPD_MAX], ypos[OPD_MAX], alpha[OPD_MAX];
  cplx *result[OPD_MAX];
  real *weight[OPD_MAX];
  cplx *object;
  int nx, ny, x_min, y_min, weights_set;
} opd;

/* constructor: PhaseDiverseCDI.c++:27-42 */
opd *opd_create(double beta, double gamma, int parallel, int granularity) {
  opd *d = (opd *)calloc(1, sizeof(opd));
  d->beta = beta; d->gamma = gamma; d->parallel = parallel; d->scale = granularity; d->ipc = 1;
  return d;
}

void opd_destroy(opd *d) {
  if (!d) return;
  for (int n = 0; n < d->nframes; n++) { free(d->result[n]); free(d->weight[n]); }
  free(d->object);
  free(d);
}

void opd_set_iterations_per_cycle(opd *d, int n) { d->ipc = n; }
void opd_set_feedback_parameter(opd *d, double beta) { d->beta = beta; d->weights_set = 0; }
void opd_set_amplification_factor(opd *d, double gamma) { d->gamma = gamma; d->weights_set = 0; }
void opd_set_probe_scaling(opd *d, int n, double alpha) {
  if (n < d->nframes) d->alpha[n] = alpha;
  d->weights_set = 0;
}

/* get_global_x_pos / get_local_x_pos (PhaseDiverseCDI.h): int results of
 * double expressions (truncation toward zero) */
static int pd_gx(opd *d, int x, double off) { return (int)(x - off - d->x_min); }
static int pd_gy(opd *d, int y, double off) { return (int)(y - off - d->y_min); }
static int pd_lx(opd *d, int x, double off) { return (int)(x + off + d->x_min); }
static int pd_ly(opd *d, int y, double off) { return (int)(y + off + d->y_min); }

/* reallocate_object_memory: a fresh zero array (PhaseDiverseCDI.c++:72-82) */
static void pd_realloc(opd *d, int nx, int ny) {
  free(d->object);
  d->object = calloc_c((size_t)nx * ny);
  d->nx = nx; d->ny = ny;
}

/* add_new_position: PhaseDiverseCDI.c++:118-178 */
int opd_add_position(opd *d, ocdi *local, double x, double y, double alpha) {
  if (d->nframes >= OPD_MAX) return -1;
  d->weights_set = 0;
  int lnx = local->nx, lny = local->ny, n = d->nframes++;
  d->frame[n] = local; d->xpos[n] = x; d->ypos[n] = y; d->alpha[n] = alpha;
  d->result[n] = calloc_c((size_t)lnx * lny);
  d->weight[n] = calloc_r((size_t)lnx * lny);
  if (!d->object) {
    d->x_min = (int)(-x); d->y_min = (int)(-y);
    pd_realloc(d, lnx * d->scale, lny * d->scale);
  }
  int extra_x = 0, extra_y = 0;
  int gxmin = pd_gx(d, 0, x) * d->scale, gymin = pd_gy(d, 0, y) * d->scale;
  int gxmax = pd_gx(d, lnx, x) * d->scale, gymax = pd_gy(d, lny, y) * d->scale;
  if (gxmin < 0) { d->x_min = (int)(-x); extra_x += -gxmin; }
  if (gymin < 0) { d->y_min = (int)(-y); extra_y += -gymin; }
  if (gxmax > d->nx) extra_x += gxmax - d->nx;
  if (gymax > d->ny) extra_y += gymax - d->ny;
  if (extra_x || extra_y) pd_realloc(d, d->nx + extra_x, d->ny + extra_y);
  return n;
}

/* set_up_weights: PhaseDiverseCDI.c++:721-848 */
static void pd_weights(opd *d) {
  if (d->weights_set) return;
  d->weights_set = 1;
  for (int n = 0; n < d->nframes; n++) {
    ocdi *f = d->frame[n];
    int lnx = f->nx, lny = f->ny;
    real mx = 0;
    int fres = f->kind == KIND_FRESNEL;
    if (fres) {
      mx = c_mag(f->illum_at_sample[0]);
      for (size_t p = 0; p < (size_t)lnx * lny; p++) { real m = c_mag(f->illum_at_sample[p]); if (m > mx) mx = m; }
    }
    for (int i = 0; i < lnx; i++)
      for (int j = 0; j < lny; j++) {
        size_t p = (size_t)i * lny + j;
        double value = fres ? (double)c_mag(f->illum_at_sample[p]) / mx : 1;
        value = d->alpha[n] * pow(value, d->gamma);
        if (f->support[p] <= 0.0) value = 0;
        if (!d->parallel) value *= d->beta;
        d->weight[n][p] = (real)value;
      }
  }
  if (d->parallel) {
    real *norm = calloc_r((size_t)d->nx * d->ny);
    for (int n = 0; n < d->nframes; n++) {
      int lnx = d->frame[n]->nx, lny = d->frame[n]->ny;
      for (int i_ = 0; i_ < lnx; i_++)
        for (int j_ = 0; j_ < lny; j_++) {
          int i = pd_gx(d, i_, d->xpos[n]), j = pd_gy(d, j_, d->ypos[n]);
          if (i >= 0 && j >= 0 && i < d->nx && j < d->ny) {
            size_t g = (size_t)i * d->ny + j;
            double w = d->weight[n][(size_t)i_ * lny + j_];
            norm[g] = (real)(n == 0 ? w : w + norm[g]);
          }
        }
    }
    for (int n = 0; n < d->nframes; n++) {
      int lnx = d->frame[n]->nx, lny = d->frame[n]->ny;
      for (int i_ = 0; i_ < lnx; i_++)
        for (int j_ = 0; j_ < lny; j_++) {
          int i = pd_gx(d, i_, d->xpos[n]), j = pd_gy(d, j_, d->ypos[n]);
          if (i >= 0 && j >= 0 && i < d->nx && j < d->ny) {
            size_t p = (size_t)i_ * lny + j_;
            double nm = norm[(size_t)i * d->ny + j];
            d->weight[n][p] = (real)(nm <= 0 ? 0 : d->beta * d->weight[n][p] / nm);
          }
        }
    }
    free(norm);
  }
}

/* get_result / set_result: PhaseDiverseCDI.c++:317-335 */
static void pd_get_result(opd *d, int n) {
  ocdi *f = d->frame[n];
  if (f->kind == KIND_FRESNEL) ocdi_get_transmission_function(f, d->result[n], NULL);
  else memcpy(d->result[n], f->complex, sizeof(cplx) * NPIX(f));
}
static void pd_set_result(opd *d, int n) {
  ocdi *f = d->frame[n];
  if (f->kind == KIND_FRESNEL) ocdi_set_transmission_function(f, d->result[n], NULL);
  else memcpy(f->complex, d->result[n], sizeof(cplx) * NPIX(f));
}

/* add_to_object: PhaseDiverseCDI.c++:851-1016.  Sub-pixel points outside the
 * object are skipped (the reference does not check them). */
static void pd_add(opd *d, int n) {
  pd_get_result(d, n);
  pd_weights(d);
  double xo = d->xpos[n], yo = d->ypos[n];
  cplx *sm = d->result[n];
  int lnx = d->frame[n]->nx, lny = d->frame[n]->ny, sc = d->scale;
  for (int i_ = 1; i_ < lnx - 1; i_++)
    for (int j_ = 1; j_ < lny - 1; j_++) {
      double w = d->weight[n][(size_t)i_ * lny + j_];
      if (w == 0) continue;
      int i = pd_gx(d, i_, xo) * sc, j = pd_gy(d, j_, yo) * sc;
      if (!(i >= 0 && j >= 0 && i < d->nx && j < d->ny)) continue;
      double f00r = sm[(size_t)i_ * lny + j_].re, f00i = sm[(size_t)i_ * lny + j_].im;
      if (sc != 1) {
        double f10r = sm[(size_t)(i_ + 1) * lny + j_].re, f10i = sm[(size_t)(i_ + 1) * lny + j_].im;
        double f01r = sm[(size_t)i_ * lny + j_ + 1].re, f01i = sm[(size_t)i_ * lny + j_ + 1].im;
        double f11r = sm[(size_t)(i_ + 1) * lny + j_ + 1].re, f11i = sm[(size_t)(i_ + 1) * lny + j_ + 1].im;
        for (int di = 0; di < sc; di++)
          for (int dj = 0; dj < sc; dj++) {
            double x = (di + 0.5 * ((sc + 1) % 2)) / (double)sc, y = (dj + 0.5 * ((sc + 1) % 2)) / (double)sc;
            double x1 = 1 - x, y1 = 1 - y;
            double vr = f00r * x1 * y1 + f10r * x * y1 + f01r * x1 * y + f11r * x * y;
            double vi = f00i * x1 * y1 + f10i * x * y1 + f01i * x1 * y + f11i * x * y;
            int ni = (int)((i_ + 0.5 - xo - d->x_min) * sc + di), nj = (int)((j_ + 0.5 - yo - d->y_min) * sc + dj);
            if (ni < 0 || nj < 0 || ni >= d->nx || nj >= d->ny) continue;
            cplx *o = &d->object[(size_t)ni * d->ny + nj];
            double nr = w * vr + o->re, nim = w * vi + o->im;
            if (!d->parallel) { nr -= w * o->re; nim -= w * o->im; }
            o->re = (real)nr; o->im = (real)nim;
          }
      } else {
        cplx *o = &d->object[(size_t)i * d->ny + j];
        double nr = w * f00r + o->re, nim = w * f00i + o->im;
        if (!d->parallel) { nr -= w * o->re; nim -= w * o->im; }
        o->re = (real)nr; o->im = (real)nim;
      }
    }
}

/* update_from_object: PhaseDiverseCDI.c++:1066-1130.  Sub-pixel blocks
 * reaching past the object are not read (the reference does not check). */
static void pd_update(opd *d, int n) {
  pd_weights(d);
  double xo = d->xpos[n], yo = d->ypos[n];
  cplx *sm = d->result[n];
  int lnx = d->frame[n]->nx, lny = d->frame[n]->ny, sc = d->scale;
  for (int i_ = 0; i_ < lnx; i_++)
    for (int j_ = 0; j_ < lny; j_++) {
      if (d->weight[n][(size_t)i_ * lny + j_] == 0) continue;
      int i = pd_gx(d, i_, xo) * sc, j = pd_gy(d, j_, yo) * sc;
      if (!(i >= 0 && j >= 0 && i < d->nx && j < d->ny)) continue;
      if (sc != 1) {
        double rv = 0, iv = 0;
        for (int di = 0; di < sc; di++)
          for (int dj = 0; dj < sc; dj++)
            if (i + di < d->nx && j + dj < d->ny) {
              rv += d->object[(size_t)(i + di) * d->ny + j + dj].re;
              iv += d->object[(size_t)(i + di) * d->ny + j + dj].im;
            }
        double nm = 1 / ((double)(sc * sc));
        sm[(size_t)i_ * lny + j_].re = (real)(rv * nm);
        sm[(size_t)i_ * lny + j_].im = (real)(iv * nm);
      } else {
        sm[(size_t)i_ * lny + j_] = d->object[(size_t)i * d->ny + j];
      }
    }
  pd_set_result(d, n);
}

/* scale_object: PhaseDiverseCDI.c++:270-313 */
static void pd_scale_object(opd *d, double factor) {
  pd_weights(d);
  for (int i = 0; i < d->nx; i++)
    for (int j = 0; j < d->ny; j++) {
      int not_in = 1;
      for (int n = 0; not_in && n < d->nframes; n++) {
        int i_ = pd_lx(d, i / d->scale, d->xpos[n]), j_ = pd_ly(d, j / d->scale, d->ypos[n]);
        int lnx = d->frame[n]->nx, lny = d->frame[n]->ny;
        if (i_ >= 0 && j_ >= 0 && i_ < lnx && j_ < lny) not_in = d->weight[n][(size_t)i_ * lny + j_] == 0;
      }
      cplx *o = &d->object[(size_t)i * d->ny + j];
      if (not_in) { o->re = 1; o->im = 0; }
      else { o->re = (real)(factor * o->re); o->im = (real)(factor * o->im); }
    }
}

/* initialise_estimate: PhaseDiverseCDI.c++:182-199 */
void opd_initialise_estimate(opd *d) {
  for (size_t p = 0; p < (size_t)d->nx * d->ny; p++) { d->object[p].re = 1; d->object[p].im = 0; }
  for (int n = 0; n < d->nframes; n++) pd_add(d, n);
}

/* iterate: PhaseDiverseCDI.c++:202-267 */
void opd_iterate(opd *d) {
  for (int n = 0; n < d->nframes; n++) {
    pd_update(d, n);
    for (int k = 0; k < d->ipc; k++) ocdi_iterate(d->frame[n]);
    if (!d->parallel) pd_add(d, n);
  }
  if (d->parallel) {
    pd_scale_object(d, 1 - d->beta);
    for (int n = 0; n < d->nframes; n++) pd_add(d, n);
  }
}

void opd_get_size(opd *d, int *nx, int *ny) { *nx = d->nx; *ny = d->ny; }
void opd_get_transmission(opd *d, cplx *out) { memcpy(out, d->object, sizeof(cplx) * (size_t)d->nx * d->ny); }
/* set_transmission: PhaseDiverseCDI.c++:94-113 */
int opd_set_transmission(opd *d, const cplx *t, int nx, int ny) {
  if (!d->object) pd_realloc(d, nx, ny);
  if (nx != d->nx || ny != d->ny) return -1;
  memcpy(d->object, t, sizeof(cplx) * (size_t)nx * ny);
  return 0;
}
void opd_get_weight(opd *d, int n, real *out) {
  pd_weights(d);
  memcpy(out, d->weight[n], sizeof(real) * NPIX(d->frame[n]));
}
void opd_get_origin(opd *d, int *x_min, int *y_min) { *x_min = d->x_min; *y_min = d->y_min; }

/* ------------------------------------------------------------------------ */
/* Position refinement: PhaseDiverseCDI::adjust_positions / check_position  */
/* (NADIA/src/PhaseDiverseCDI.c++:337-719) with align and interpolate       */
/* (NADIA/src/utils.c:1151-1435, defaults NADIA/include/utils.h:24-30).      */
/* ------------------------------------------------------------------------ */

/* align: the weighted cross-correlation of two real images zero-padded to
 * twice the larger size.  FFTW's r2c / c2r pair becomes complex transforms
 * of the real arrays (same products; the imaginary parts are zero to
 * rounding and dropped).  Returns 1 when an offset was found. */
int oalign(const real *a, int nx1, int ny1, const real *b, int nx2, int ny2, int *ox, int *oy, int min_x,
           int max_x, int min_y, int max_y, const real *wa, const real *wb, double overlap_fraction) {
  int nx = 2 * (nx1 > nx2 ? nx1 : nx2), ny = 2 * (ny1 > ny2 ? ny1 : ny2);
  if (min_x == max_x || min_y == max_y) {
    min_x = (int)(1 - nx / 2.0); max_x = (int)(nx / 2.0);
    min_y = (int)(1 - ny / 2.0); max_y = (int)(ny / 2.0);
  }
  size_t np = (size_t)nx * ny;
  cplx *i1 = calloc_c(np), *i2 = calloc_c(np), *w1 = calloc_c(np), *w2 = calloc_c(np);
  double weight_sum = 0;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      if (!(i >= 0.5 * (nx - nx1) && i < 0.5 * (nx + nx1) && j >= 0.5 * (ny - ny1) && j < 0.5 * (ny + ny1))) continue;
      int i_ = (int)(i - 0.5 * (nx - nx1)), j_ = (int)(j - 0.5 * (ny - ny1));
      size_t p = (size_t)i * ny + j;
      if (!wa || wa[(size_t)i_ * ny1 + j_] > 0) { i1[p].re = a[(size_t)i_ * ny1 + j_]; w1[p].re = 1; }
      if (i_ < nx2 && j_ < ny2 && (!wb || wb[(size_t)i_ * ny2 + j_] > 0)) {
        i2[p].re = b[(size_t)i_ * ny2 + j_];
        w2[p].re = 1;
      }
      weight_sum += (nx1 * ny1 < nx2 * ny2) ? w1[p].re : w2[p].re;
    }
  or_fft2d(i1, nx, ny, -1); or_fft2d(i2, nx, ny, -1);
  or_fft2d(w1, nx, ny, -1); or_fft2d(w2, nx, ny, -1);
  for (size_t p = 0; p < np; p++) {
    /* conj(F1) * F2 */
    real ar = i1[p].re, ai = -i1[p].im, br = i2[p].re, bi = i2[p].im;
    i1[p].re = ar * br - ai * bi; i1[p].im = ar * bi + ai * br;
    ar = w1[p].re; ai = -w1[p].im; br = w2[p].re; bi = w2[p].im;
    w1[p].re = ar * br - ai * bi; w1[p].im = ar * bi + ai * br;
  }
  or_fft2d(i1, nx, ny, +1); or_fft2d(w1, nx, ny, +1);
  double overlap_factor = weight_sum * nx * ny;
  real best = 0;
  int found = 0;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      size_t p = (size_t)i * ny + j;
      real ci = i1[p].re, cw = w1[p].re;
      if (!(cw / overlap_factor > overlap_fraction)) continue;
      real ratio = ci / cw;
      if (!(ratio > best)) continue;
      int ti = -i, tj = -j;
      if (i >= nx / 2.0) ti = -i + nx;
      if (j >= ny / 2.0) tj = -j + ny;
      if (ti >= min_x && ti < max_x && tj >= min_y && tj < max_y) { best = ratio; *ox = ti; *oy = tj; found = 1; }
    }
  free(i1); free(i2); free(w1); free(w2);
  return found;
}

/* interpolate(Double_2D): bilinear up-sampling by big/original; the outer
 * ring of original pixels is not written (NADIA/src/utils.c:1346-1435). */
void ointerpolate(const real *o, int nx, int ny, real *big, int bnx, int bny) {
  int sx = bnx / nx, sy = bny / ny;
  memset(big, 0, sizeof(real) * (size_t)bnx * bny);
  for (int i = 1; i < nx - 1; i++)
    for (int j = 1; j < ny - 1; j++) {
      double f00 = o[(size_t)i * ny + j], f10 = o[(size_t)(i + 1) * ny + j];
      double f01 = o[(size_t)i * ny + j + 1], f11 = o[(size_t)(i + 1) * ny + j + 1];
      double ni = (i + 0.5) * sx, nj = (j + 0.5) * sy;
      for (int di = 0; di < sx; di++)
        for (int dj = 0; dj < sy; dj++) {
          double x = (di + 0.5 * ((sx + 1) % 2)) / (double)sx, y = (dj + 0.5 * ((sy + 1) % 2)) / (double)sy;
          double v = f00 * (1 - x) * (1 - y) + f10 * x * (1 - y) + f01 * (1 - x) * y + f11 * x * y;
          int bi = (int)(ni + di), bj = (int)(nj + dj);
          if (bi < bnx && bj < bny) big[(size_t)bi * bny + bj] = (real)v;
        }
    }
}

/* check_position: PhaseDiverseCDI.c++:582-719 (recursive 3 x 3 search with
 * the step halved when the centre wins; gives up after 10 moves). */
static int pd_check_position(opd *d, int n, double step, int min_x, int max_x, int min_y, int max_y, int tries,
                             int *gave_up) {
  double x = d->xpos[n], y = d->ypos[n];
  if (step < 1.0 / d->scale) return 1;
  if (tries > 10 || x < x + min_x || x > x + max_x || y < y + min_y || y > y + max_y) {
    if (gave_up) *gave_up = 1;
    return 0;
  }
  ocdi *single = d->frame[n];
  double best_x = x, best_y = y, best_error = 100;
  size_t no = (size_t)d->nx * d->ny, ns = NPIX(single);
  cplx *obj_copy = calloc_c(no), *single_copy = calloc_c(ns);
  memcpy(obj_copy, d->object, sizeof(cplx) * no);
  memcpy(single_copy, d->result[n], sizeof(cplx) * ns);
  d->beta = 0.5;
  d->weights_set = 0;
  for (int i = -1; i < 2; i++)
    for (int j = -1; j < 2; j++) {
      double nxp = x + i * step, nyp = y + j * step;
      d->xpos[n] = nxp; d->ypos[n] = nyp;
      pd_add(d, n);
      pd_update(d, n);
      ocdi_iterate(single);
      if (single->current_error < best_error) { best_error = single->current_error; best_x = nxp; best_y = nyp; }
      memcpy(d->object, obj_copy, sizeof(cplx) * no);
      memcpy(d->result[n], single_copy, sizeof(cplx) * ns);
      pd_set_result(d, n);
    }
  free(obj_copy); free(single_copy);
  d->xpos[n] = best_x; d->ypos[n] = best_y;
  if (best_x == x && best_y == y) step = step / 2.0;
  if (!pd_check_position(d, n, step, min_x, max_x, min_y, max_y, tries + 1, gave_up)) {
    d->xpos[n] = x; d->ypos[n] = y;
    return 0;
  }
  d->beta = 1.0;
  d->weights_set = 0;
  return 1;
}

/* adjust_positions: PhaseDiverseCDI.c++:337-579.  type 0 = cross-correlation,
 * 1 = minimum error.  gave_up[n] (optional) is set when the minimum-error
 * search gave up on frame n.  Returns -1 for an unknown type or frames of
 * different sizes. */
int opd_adjust_positions(opd *d, int type, int forward, int min_x, int max_x, int min_y, int max_y, double step,
                         int *gave_up) {
  if (type != 0 && type != 1) return -1;
  int nf = d->nframes;
  if (nf == 0 || !d->object) return -1;
  for (int k = 1; k < nf; k++)
    if (d->frame[k]->nx != d->frame[0]->nx || d->frame[k]->ny != d->frame[0]->ny) return -1;
  int parallel_c = d->parallel, nx_c = d->nx, ny_c = d->ny, xmin_c = d->x_min, ymin_c = d->y_min;
  double beta_c = d->beta;
  cplx *obj_c = d->object;
  d->parallel = 0;
  int n = 1, limit = nf;
  if (!forward) { n = nf - 2; limit = -1; }
  cplx *temp_object = calloc_c((size_t)d->nx * d->ny);
  for (size_t p = 0; p < (size_t)d->nx * d->ny; p++) { temp_object[p].re = 1; temp_object[p].im = 0; }
  d->object = temp_object;
  d->beta = 1.0;
  d->weights_set = 0;
  int first = forward ? 0 : nf - 1;
  ocdi_iterate(d->frame[first]);
  pd_get_result(d, first);
  pd_add(d, first);
  while (n != limit) {
    double bx = d->xpos[n], by = d->ypos[n];
    int lnx = d->frame[0]->nx, lny = d->frame[0]->ny, sc = d->scale;
    int mx = lnx * sc, my = lny * sc;
    ocdi_iterate(d->frame[n]);
    pd_get_result(d, n);
    if (type == 0) {
      real *tsingle = calloc_r((size_t)lnx * lny), *tothers = calloc_r((size_t)mx * my);
      for (int i = 0; i < mx; i++)
        for (int j = 0; j < my; j++) {
          int i_ = (int)(i - bx * sc - d->x_min * sc), j_ = (int)(j - by * sc - d->y_min * sc);
          if (i_ >= 0 && j_ >= 0 && i_ < d->nx && j_ < d->ny) {
            double m = c_mag(d->object[(size_t)i_ * d->ny + j_]);
            if (m < 1) tothers[(size_t)i * my + j] = (real)(1 - m);
          }
          if (i < lnx && j < lny) {
            double m = c_mag(d->result[n][(size_t)i * lny + j]);
            if (m < 1) tsingle[(size_t)i * lny + j] = (real)(1 - m);
          }
        }
      real *img = tsingle, *wgt = d->frame[n]->support, *ibig = NULL, *wbig = NULL;
      if (sc != 1) {
        ibig = calloc_r((size_t)mx * my); wbig = calloc_r((size_t)mx * my);
        ointerpolate(tsingle, lnx, lny, ibig, mx, my);
        ointerpolate(d->frame[n]->support, lnx, lny, wbig, mx, my);
        img = ibig; wgt = wbig;
      }
      int new_x = 0, new_y = 0;
      oalign(img, mx, my, tothers, mx, my, &new_x, &new_y, min_x, max_x, min_y, max_y, wgt, tothers, 0.2);
      d->xpos[n] = bx + new_x / (double)sc;
      d->ypos[n] = by + new_y / (double)sc;
      free(tsingle); free(tothers); free(ibig); free(wbig);
    } else {
      cplx *temp = calloc_c((size_t)mx * my);
      for (int i = 0; i < mx; i++)
        for (int j = 0; j < my; j++) {
          int i_ = (int)(i - bx * sc - d->x_min * sc), j_ = (int)(j - by * sc - d->y_min * sc);
          if (i_ >= 0 && j_ >= 0 && i_ < d->nx && j_ < d->ny) temp[(size_t)i * my + j] = d->object[(size_t)i_ * d->ny + j_];
        }
      d->object = temp; d->nx = mx; d->ny = my;
      d->x_min = (int)(-d->xpos[n]); d->y_min = (int)(-d->ypos[n]);
      pd_check_position(d, n, step, min_x, max_x, min_y, max_y, 0, gave_up ? &gave_up[n] : NULL);
      d->object = temp_object; d->nx = nx_c; d->ny = ny_c; d->x_min = xmin_c; d->y_min = ymin_c;
      free(temp);
    }
    d->weights_set = 0;
    pd_add(d, n);
    n += forward ? 1 : -1;
  }
  d->object = obj_c; d->nx = nx_c; d->ny = ny_c; d->x_min = xmin_c; d->y_min = ymin_c;
  d->parallel = parallel_c; d->beta = beta_c; d->weights_set = 0;
  free(temp_object);
  return 0;
}

void opd_get_position(opd *d, int n, double *x, double *y) { *x = d->xpos[n]; *y = d->ypos[n]; }
