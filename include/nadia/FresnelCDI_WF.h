// FresnelCDI_WF.h -- white-field (illumination) recovery of the NADIA drop-in
// API (interface of NADIA/include/FresnelCDI_WF.h:31-186).  The algorithm is
// fixed: the setters below hide BaseCDI's and only warn, as in the reference.
#ifndef NADIA_AMD_FCDI_WF_H
#define NADIA_AMD_FCDI_WF_H

#include <iostream>

#include "BaseCDI.h"

class FresnelCDI_WF : public BaseCDI {
 protected:
  double wavelength;
  double zone_to_focal_length;
  double focal_to_detector_length;
  double pixel_length;

 public:
  FresnelCDI_WF(Complex_2D& initial_guess, double beam_wavelength, double zone_focal_length,
                double focal_detector_length, double pixel_size, int n_best = 0);
  virtual ~FresnelCDI_WF() {}
  int iterate();
  virtual void propagate_from_detector(Complex_2D& c);
  virtual void propagate_to_detector(Complex_2D& c);
  virtual void initialise_estimate(int seed = 0);
  virtual void set_support(double z_diameter, double size = 1.01);
  void set_algorithm(int) {
    std::cout << "WARNING: Algorithm can not be set when performing Fresnel "
              << "CDI white field recovery" << std::endl;
  }
  void print_algorithm() {
    std::cout << "Using the default Fresnel CDI "
              << "white field recovery algorithm" << std::endl;
  }
  void set_relaxation_parameter(double) {
    std::cout << "WARNING: The relaxation parameter can not "
              << "set when performing Fresnel white-field "
              << "reconstruction" << std::endl;
  }
  void set_custom_algorithm(double, double, double, double, double, double, double, double, double, double) {
    std::cout << "WARNING: Algorithm can not be set when performing Fresnel "
              << "CDI white field recovery" << std::endl;
  }
  // c * (sin phi -/+ i cos phi)/sqrt(N) (NADIA/src/FresnelCDI_WF.c++:70-96)
  void multiply_factors(Complex_2D& c, int direction);
};

#endif
