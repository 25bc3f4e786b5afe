// planar_example.cpp -- the reconstruction loop of NADIA's
// examples/PlanarCDI_example.c written against the drop-in headers: ER, then
// HIO, then ER, with shrinkwrap every `sw` iterations and the error printed
// after every iteration.  Inputs are raw x-major files in <dir>:
//   intensity.raw, support.raw (nx*ny doubles), x0.raw (nx*ny complex doubles)
// and the final estimate is written to <dir>/out.raw (complex doubles).
//
// usage: planar_example <dir> <nx> <ny> <n_er1> <n_hio> <n_er2> <sw>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "PlanarCDI.h"

static std::vector<double> read_raw(const std::string& path, size_t n) {
  std::vector<double> v(n);
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f || std::fread(v.data(), sizeof(double), n, f) != n) {
    std::cout << "failed to read " << path << std::endl;
    std::exit(1);
  }
  std::fclose(f);
  return v;
}

int main(int argc, char* argv[]) {
  if (argc != 8) {
    std::cout << "usage: planar_example <dir> <nx> <ny> <n_er1> <n_hio> <n_er2> <sw>" << std::endl;
    return 1;
  }
  const std::string dir = argv[1];
  const int nx = std::atoi(argv[2]), ny = std::atoi(argv[3]);
  const int er1 = std::atoi(argv[4]), hio = std::atoi(argv[5]), er2 = std::atoi(argv[6]);
  const int sw = std::atoi(argv[7]);
  const size_t n = (size_t)nx * ny;
  std::vector<double> I = read_raw(dir + "/intensity.raw", n);
  std::vector<double> S = read_raw(dir + "/support.raw", n);
  std::vector<double> X = read_raw(dir + "/x0.raw", 2 * n);

  Double_2D data(nx, ny), support(nx, ny);
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      data.set(i, j, I[(size_t)i * ny + j]);
      support.set(i, j, S[(size_t)i * ny + j]);
    }

  // the estimate that the reconstruction updates in place
  Complex_2D object_estimate(nx, ny);
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      object_estimate.set_real(i, j, X[2 * ((size_t)i * ny + j)]);
      object_estimate.set_imag(i, j, X[2 * ((size_t)i * ny + j) + 1]);
    }

  PlanarCDI planar(object_estimate, 4);
  planar.set_support(support);
  planar.set_intensity(data);
  planar.set_algorithm(ER);
  planar.print_algorithm();

  int it = 0;
  const int phases[3] = {er1, hio, er2};
  const int algs[3] = {ER, PlanarCDI::getAlgFromName("HIO"), ER};
  for (int ph = 0; ph < 3; ph++) {
    planar.set_algorithm(algs[ph]);
    for (int k = 0; k < phases[ph]; k++, it++) {
      planar.iterate();
      std::printf("iteration %d error %.17g\n", it, planar.get_error());
      if (sw > 0 && it % sw == sw - 1) planar.apply_shrinkwrap(2.0, 0.1);
    }
  }

  // reading elements brings the estimate back from the device
  std::vector<double> out(2 * n);
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      out[2 * ((size_t)i * ny + j)] = object_estimate.get_real(i, j);
      out[2 * ((size_t)i * ny + j) + 1] = object_estimate.get_imag(i, j);
    }
  FILE* f = std::fopen((dir + "/out.raw").c_str(), "wb");
  if (!f) return 1;
  std::fwrite(out.data(), sizeof(double), out.size(), f);
  std::fclose(f);
  double err = 0;
  planar.get_best_result(err, 0);
  std::printf("best error %.17g\n", err);
  return 0;
}
