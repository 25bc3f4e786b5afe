// PlanarCDI.h -- far-field reconstruction of the NADIA drop-in API
// (interface of NADIA/include/PlanarCDI.h:24-44).
#ifndef NADIA_AMD_PLANARCDI_H
#define NADIA_AMD_PLANARCDI_H

#include "BaseCDI.h"

class PlanarCDI : public BaseCDI {
 public:
  PlanarCDI(Complex_2D& complex, unsigned int n_best = 0) : BaseCDI(complex, n_best, NADIA_PLANAR) {}
  virtual void get_intensity_autocorrelation(Double_2D& autoc);
  virtual void initialise_estimate(int seed = 0);
  virtual void propagate_to_detector(Complex_2D& c);
  virtual void propagate_from_detector(Complex_2D& c);
};

#endif
