// PhaseDiverseCDI.h -- phase-diverse (multi-frame) reconstruction of the NADIA
// drop-in API.
//
// Interface of NADIA's PhaseDiverseCDI (NADIA/include/PhaseDiverseCDI.h,
// NADIA/src/PhaseDiverseCDI.c++): frames (PlanarCDI or FresnelCDI objects)
// at transverse positions of one global transmission function, merged with
// weights alpha * (|illumination| / max)^gamma on each frame's support, in
// series or parallel mode, optionally on a sub-pixel grid (granularity).
// The global object, the frames' results and weights live on the device
// (nadia_pd, include/nadia_amd.h); in parallel mode the frames' iterations
// run as one batch.  adjust_positions (cross-correlation or minimum-error
// refinement of the frame offsets) runs on the device and prints the
// reference's per-frame messages.
#ifndef NADIA_AMD_PHASED_H
#define NADIA_AMD_PHASED_H

#include <vector>

#include "BaseCDI.h"

class PhaseDiverseCDI {
 protected:
  std::vector<BaseCDI*> singleCDI;
  std::vector<double> x_position;
  std::vector<double> y_position;
  double beta;
  double gamma;
  std::vector<double> alpha;
  Complex_2D* object;  // host copy handed out by get_transmission
  int iterations_per_cycle;
  int scale;
  bool parallel;
  nadia_pd* h;

  void push_parameters();
  void frames_changed();

 public:
  enum { CROSS_CORRELATION, MINIMUM_ERROR };

  PhaseDiverseCDI(double beta = 1.0, double gamma = 1.0, bool parallel = false, int granularity = 1);
  ~PhaseDiverseCDI();

  void add_new_position(BaseCDI* local, double x = 0, double y = 0, double alpha = 1);
  void remove_positions();
  void initialise_estimate();
  void iterate();
  void set_iterations_per_cycle(int iterations);
  void set_feedback_parameter(double beta);
  void set_amplification_factor(double gamma);
  void set_probe_scaling(int n_probe, double alpha);
  Complex_2D* get_transmission();
  void set_transmission(Complex_2D& new_transmission);
  void adjust_positions(int type = CROSS_CORRELATION, bool forwards = true, int x_min = -50, int x_max = +50,
                        int y_min = -50, int y_max = +50, double step_size = 4);
  double get_final_x_position(int n_probe) { return x_position.at(n_probe); }
  double get_final_y_position(int n_probe) { return y_position.at(n_probe); }
};

#endif
