// types.h -- constants of the NADIA drop-in API (same names and values as
// NADIA/include/types.h:43-61).  Precision is chosen at compile time with
// DOUBLE_PRECISION, as in the reference (link libNADIAd_amd instead of
// libNADIA_amd).
#ifndef NADIA_AMD_TYPES_H
#define NADIA_AMD_TYPES_H

#define FAILURE 0
#define SUCCESS 1

#define REAL 0
#define IMAG 1
#define MAG 2
#define PHASE 3
#define MAG_SQ 4

// FFTW planner flags accepted (and ignored: the device FFT needs no planning)
// by set_fftw_type, so code written for the reference still compiles.
#ifndef FFTW_MEASURE
#define FFTW_MEASURE (0U)
#endif
#ifndef FFTW_ESTIMATE
#define FFTW_ESTIMATE (1U << 6)
#endif

#ifdef DOUBLE_PRECISION
typedef double nadia_real;
#else
typedef float nadia_real;
#endif

#endif
