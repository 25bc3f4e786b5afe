// FresnelCDI.h -- Fresnel reconstruction of the NADIA drop-in API
// (interface of NADIA/include/FresnelCDI.h:33-232).  The white field, its
// image at the sample and the quadratic phase live on the device; the phase
// is evaluated in the kernels from its closed form instead of a table.
#ifndef NADIA_AMD_FCDI_H
#define NADIA_AMD_FCDI_H

#include "BaseCDI.h"

class FresnelCDI : public BaseCDI {
 protected:
  Complex_2D illumination_at_sample;  // host copy returned by get_illumination_at_sample
  double wavelength;
  double pixel_length;
  double focal_detector_length;
  double focal_sample_length;
  double norm;

 public:
  FresnelCDI(Complex_2D& initial_guess, Complex_2D& white_field, double beam_wavelength,
             double focal_detector_length, double focal_sample_length, double pixel_size,
             double normalisation = 1.0, int n_best = 0);
  virtual ~FresnelCDI() {}
  virtual void initialise_estimate(int seed = 0);
  virtual void auto_set_norm();
  virtual void get_transmission_function(Complex_2D& result, Complex_2D* esw = 0);
  virtual void apply_support(Complex_2D& complex);
  virtual void set_transmission_function(Complex_2D& transmission, Complex_2D* esw = 0);
  virtual void scale_intensity(Complex_2D& c);
  virtual void propagate_from_detector(Complex_2D& c);
  virtual void propagate_to_detector(Complex_2D& c);
  void set_normalisation(double normalisation);
  void set_experimental_parameters(double beam_wavelength, double focal_detector_length,
                                   double focal_sample_length, double pixel_size);
  void set_norm(double new_normalisation);
  const Complex_2D& get_illumination_at_sample();
};

#endif
