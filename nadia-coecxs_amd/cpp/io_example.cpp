// io_example.cpp -- reads <dir>/in.dbin and <dir>/in.cplx (nx x ny) with the
// drop-in read_dbin / read_cplx, prints a few elements by (x, y), doubles
// them and writes <dir>/out.dbin and <dir>/out.cplx with write_dbin /
// write_cplx.  Host-side only (no device work).
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "Complex_2D.h"
#include "Double_2D.h"
#include "io.h"

int main(int argc, char** argv) {
  if (argc < 4) {
    std::cout << "usage: io_example dir nx ny" << std::endl;
    return 2;
  }
  const std::string dir = argv[1];
  const int nx = atoi(argv[2]), ny = atoi(argv[3]);
  Double_2D d;
  if (!read_dbin(dir + "/in.dbin", nx, ny, d)) return 1;
  Complex_2D c(nx, ny);
  if (!read_cplx(dir + "/in.cplx", c)) return 1;
  for (int k = 0; k < 3; k++) {
    const int x = (k * 5 + 1) % nx, y = (k * 3 + 2) % ny;
    printf("element %d %d %.17g %.17g %.17g\n", x, y, d.get(x, y), (double)c.get_real(x, y), (double)c.get_imag(x, y));
  }
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      d.set(i, j, 2 * d.get(i, j));
      c.set_real(i, j, 2 * c.get_real(i, j));
      c.set_imag(i, j, 2 * c.get_imag(i, j));
    }
  if (!write_dbin(dir + "/out.dbin", d) || !write_cplx(dir + "/out.cplx", c)) return 1;
  // a missing file reports FAILURE
  Double_2D e;
  if (read_dbin(dir + "/missing.dbin", nx, ny, e)) return 1;
  return 0;
}
