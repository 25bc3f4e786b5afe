// align_example.cpp -- PhaseDiverseCDI position refinement through the
// drop-in headers: PlanarCDI frames (ER) read from <dir> (x0_k.cplx,
// intensity_k.dbin, support_k.dbin for k = 0 .. nframes-1, n x n) at the
// starting positions given on the command line, merged in series (beta 1),
// then adjust_positions(type, forwards).  Prints "position k x y" for every
// frame (%.17g).
//
// usage: align_example <dir> <n> <type 0|1> <forwards 0|1> <x0> <y0> [<x1> <y1> ...]
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "PhaseDiverseCDI.h"
#include "PlanarCDI.h"
#include "io.h"

int main(int argc, char* argv[]) {
  if (argc < 7 || (argc - 5) % 2 != 0) {
    std::cout << "usage: align_example <dir> <n> <type 0|1> <forwards 0|1> <x0> <y0> [<x1> <y1> ...]" << std::endl;
    return 1;
  }
  const std::string dir = argv[1];
  const int n = std::atoi(argv[2]), type = std::atoi(argv[3]);
  const bool forwards = std::atoi(argv[4]) != 0;
  const int nf = (argc - 5) / 2;
  std::vector<Double_2D> intensity(nf), support(nf);
  std::vector<Complex_2D*> est(nf);
  std::vector<PlanarCDI*> frame(nf);
  PhaseDiverseCDI pd(1.0, 1.0, false);
  for (int k = 0; k < nf; k++) {
    const std::string s = std::to_string(k);
    if (!read_dbin(dir + "/intensity_" + s + ".dbin", n, n, intensity[k]) ||
        !read_dbin(dir + "/support_" + s + ".dbin", n, n, support[k]))
      return 1;
    est[k] = new Complex_2D(n, n);
    if (!read_cplx(dir + "/x0_" + s + ".cplx", *est[k])) return 1;
    frame[k] = new PlanarCDI(*est[k]);
    frame[k]->set_support(support[k]);
    frame[k]->set_intensity(intensity[k]);
    frame[k]->set_algorithm(ER);
    pd.add_new_position(frame[k], std::atof(argv[5 + 2 * k]), std::atof(argv[6 + 2 * k]));
  }
  pd.initialise_estimate();
  pd.adjust_positions(type, forwards);
  for (int k = 0; k < nf; k++)
    std::printf("position %d %.17g %.17g\n", k, pd.get_final_x_position(k), pd.get_final_y_position(k));
  for (int k = 0; k < nf; k++) {
    delete frame[k];
    delete est[k];
  }
  return 0;
}
