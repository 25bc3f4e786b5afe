// Element accessors and whole-array operations of the drop-in Complex_2D
// (include/nadia/Complex_2D.h), printed for the tests to check against
// NADIA's semantics (NADIA/include/Complex_2D.h:167-337,
// NADIA/src/Complex_2D.c++:169-199, 259-270).
//
//   complex_probe            host accessors only (no device call)
//   complex_probe device     also scale / add / get_norm / get_2d on the
//                            device copy, read back through the accessors
#include <cstdio>
#include <cstring>

#include "nadia/Complex_2D.h"

int main(int argc, char** argv) {
  const bool device = argc > 1 && std::strcmp(argv[1], "device") == 0;
  Complex_2D c(8, 24);  // a transform size: device fields need one
  // set_mag on a zero element: (value, 0)
  c.set_mag(0, 0, 2.5);
  // set_mag keeps the phase
  c.set_real(1, 2, 3.0);
  c.set_imag(1, 2, -4.0);
  c.set_mag(1, 2, 10.0);
  // set_phase keeps the magnitude
  c.set_real(2, 3, -1.0);
  c.set_imag(2, 3, 1.0);
  c.set_phase(2, 3, 0.75);
  // set_value by type; get_phase in (-pi, pi]
  c.set_value(3, 5, REAL, -2.0);
  c.set_value(3, 5, IMAG, 0.0);
  c.set_value(3, 4, MAG, 1.5);
  c.set_value(3, 4, PHASE, -2.0);
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 6; j++)
      std::printf("E %d %d %.17g %.17g %.17g %.17g %.17g\n", i, j, (double)c.get_real(i, j), (double)c.get_imag(i, j),
                  (double)c.get_mag(i, j), (double)c.get_phase(i, j), (double)c.get_value(i, j, MAG_SQ));
  if (device) {
    Complex_2D d(8, 24);
    for (int i = 0; i < 8; i++)
      for (int j = 0; j < 24; j++) {
        d.set_real(i, j, 0.5 * i - 0.25 * j);
        d.set_imag(i, j, 1.0 + 0.125 * i * j);
      }
    c.scale(-1.5);      // device
    c.add(d, 0.75);     // device
    Double_2D m(8, 24);
    c.get_2d(MAG, m);   // pulls the device copy
    std::printf("N %.17g\n", (double)c.get_norm());
    for (int i = 0; i < 8; i++)
      for (int j = 0; j < 24; j++)
        std::printf("D %d %d %.17g %.17g %.17g\n", i, j, (double)c.get_real(i, j), (double)c.get_imag(i, j),
                    (double)m.get(i, j));
  }
  return 0;
}
