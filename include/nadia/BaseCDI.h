// BaseCDI.h -- reconstruction base class of the NADIA drop-in API.
//
// Public interface of NADIA's BaseCDI (NADIA/include/BaseCDI.h:66-522): the
// same methods, enums and defaults.  Each object owns a device reconstruction
// (nadia_recon, include/nadia_amd.h) bound to the caller's Complex_2D, which
// it references (not copies) and updates in place, as the reference does
// (NADIA/include/BaseCDI.h:93).  iterate() queues the work on the device and
// returns; get_error() and the element accessors of the estimate wait for it.
#ifndef NADIA_AMD_BASECDI_H
#define NADIA_AMD_BASECDI_H

#include <map>
#include <string>
#include <vector>

#include "Complex_2D.h"
#include "Double_2D.h"
#include "types.h"

#define NALGORITHMS 9
#define NTERMS 5
#define NVECTORS 10

enum { PSF, PFS, PS, PF, PI };
enum { ER, BIO, BOO, HIO, DM, SF, ASR, HPR, RAAR, CUSTOM };

class TransmissionConstraint;

class BaseCDI {
 protected:
  Complex_2D& complex;
  int nx;
  int ny;
  double beta;
  Double_2D support;         // host view of the device support (get_support)
  Double_2D intensity_sqrt;  // host copy of sqrt(I)
  int n_best;
  std::vector<Complex_2D*> best_array;  // host copies handed out by get_best_result
  TransmissionConstraint* transmission_constraint;
  nadia_recon* recon;

  static std::map<std::string, int>* algNameMap;
  static std::map<std::string, int>* set_up_algorithm_name_map();

  // constructor used by the subclasses to pick the propagator
  BaseCDI(Complex_2D& complex, unsigned int n_best, int kind);

  void apply_threshold(Double_2D& array, double threshold);
  void convolve(Double_2D& array, double gauss_width, int pixel_cut_off = 4);
  void support_constraint(Complex_2D& c);
  nadia_field* dev(Complex_2D& c);  // c's device field, current

 public:
  BaseCDI(Complex_2D& complex, unsigned int n_best = 0);
  virtual ~BaseCDI();

  virtual int iterate();
  virtual void initialise_estimate(int seed = 0) = 0;

  Complex_2D* get_best_result(double& error, int index = 0);
  void set_support(const Double_2D& object_support, bool soften = false);
  void set_intensity(const Double_2D& detector_intensity);
  void set_beam_stop(const Double_2D& beam_stop_region);
  void set_relaxation_parameter(double relaxation_parameter);

  int get_size_x() { return nx; }
  int get_size_y() { return ny; }

  const Complex_2D& get_exit_surface_wave() { return complex; }
  void set_exit_surface_wave(Complex_2D& esw) { complex.copy(esw); }

  void set_algorithm(int alg);
  void set_custom_algorithm(double m1, double m2, double m3, double m4, double m5, double m6, double m7,
                            double m8, double m9, double m10);
  void print_algorithm();
  double get_error();

  static int getAlgFromName(std::string algorithm_name) {
    std::map<std::string, int>::iterator alg = algNameMap->find(algorithm_name);
    if (alg == algNameMap->end()) return -1;
    return alg->second;
  }

  virtual void apply_shrinkwrap(double gauss_width = 1.5, double threshold = 0.1);
  const Double_2D& get_support();

  virtual void apply_support(Complex_2D& c);
  virtual void project_intensity(Complex_2D& c);
  virtual void scale_intensity(Complex_2D& c);
  virtual void propagate_to_detector(Complex_2D& c) = 0;
  virtual void propagate_from_detector(Complex_2D& c) = 0;

  void set_fftw_type(int type);
  virtual void set_complex_constraint(TransmissionConstraint& trans_constraint);
  void reset_best();

  // ---- extensions of this build (no counterpart in the reference) ----
  // Run n iterations in one call; errors (may be null) receives get_error()
  // after each.  Equivalent to n calls of iterate() and get_error().
  int iterate(int n, double* errors);
  // 0 = fused passes (default), 1 = one kernel per reference operation
  void set_execution_path(int path);
  nadia_recon* device_handle() { return recon; }
  // For callers that drive the device reconstruction directly
  // (PhaseDiverseCDI): bring the estimate's device copy up to date, and
  // record that device work changed it.
  nadia_recon* device_sync() {
    dev(complex);
    return recon;
  }
  void device_changed() { complex.device_modified(); }
};

#endif
