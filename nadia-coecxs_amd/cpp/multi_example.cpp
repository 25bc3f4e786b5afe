// multi_example.cpp -- two PlanarCDI frames (HIO) merged by PhaseDiverseCDI
// in parallel mode, frame 0 with a TransmissionConstraint (one complex
// region), written against the drop-in headers.  Inputs in <dir> (.dbin /
// .cplx, n x n): intensity_k.dbin, support_k.dbin, x0_k.cplx for k = 0, 1
// and region.dbin; frame 1 sits at (x1, y1).  After <iters> big iterations
// the global transmission goes to <dir>/object.cplx and frame 0's estimate to
// <dir>/frame0.cplx.  Drop-in C++ API only.
//
// usage: multi_example <dir> <n> <x1> <y1> <iters>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "PhaseDiverseCDI.h"
#include "PlanarCDI.h"
#include "TransmissionConstraint.h"
#include "io.h"

int main(int argc, char* argv[]) {
  if (argc != 6) {
    std::cout << "usage: multi_example <dir> <n> <x1> <y1> <iters>" << std::endl;
    return 1;
  }
  const std::string dir = argv[1];
  const int n = std::atoi(argv[2]);
  const double x1 = std::atof(argv[3]), y1 = std::atof(argv[4]);
  const int iters = std::atoi(argv[5]);

  Double_2D intensity[2], support[2];
  Complex_2D* est[2];
  PlanarCDI* frame[2];
  for (int k = 0; k < 2; k++) {
    const std::string s = std::to_string(k);
    if (!read_dbin(dir + "/intensity_" + s + ".dbin", n, n, intensity[k]) ||
        !read_dbin(dir + "/support_" + s + ".dbin", n, n, support[k]))
      return 1;
    est[k] = new Complex_2D(n, n);
    if (!read_cplx(dir + "/x0_" + s + ".cplx", *est[k])) return 1;
    frame[k] = new PlanarCDI(*est[k]);
    frame[k]->set_support(support[k]);
    frame[k]->set_intensity(intensity[k]);
    frame[k]->set_algorithm(HIO);
  }

  Double_2D region;
  if (!read_dbin(dir + "/region.dbin", n, n, region)) return 1;
  TransmissionConstraint tc;
  ComplexConstraint cc(region, 0.5, 0.5);
  tc.add_complex_constraint(cc);
  tc.set_enforce_unity(false);
  frame[0]->set_complex_constraint(tc);

  PhaseDiverseCDI pd(0.8, 1.0, true);
  pd.add_new_position(frame[0], 0, 0);
  pd.add_new_position(frame[1], x1, y1);
  pd.initialise_estimate();
  for (int i = 0; i < iters; i++) pd.iterate();
  std::printf("c_mean %.17g\n", cc.get_c_mean());

  Complex_2D* object = pd.get_transmission();
  if (!write_cplx(dir + "/object.cplx", *object) || !write_cplx(dir + "/frame0.cplx", *est[0])) return 1;
  std::printf("object %d %d\n", object->get_size_x(), object->get_size_y());
  for (int k = 0; k < 2; k++) {
    delete frame[k];
    delete est[k];
  }
  return 0;
}
