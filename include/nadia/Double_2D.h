// Double_2D.h -- real 2-D field of the NADIA drop-in API.
//
// Same interface and storage as NADIA's Real_2D<T> (NADIA/include/Double_2D.h:
// 32-356): nx x ny values of T, x-major (element (x, y) at x*ny + y),
// zero-initialised; Double_2D is Real_2D<float>, or Real_2D<double> with
// DOUBLE_PRECISION.  It stays a host container: reconstruction objects copy
// it to the device when it is passed to set_support / set_intensity /
// set_beam_stop.
#ifndef NADIA_AMD_DOUBLE_2D_H
#define NADIA_AMD_DOUBLE_2D_H

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iostream>

template <class T>
class Real_2D {
 protected:
  T* array;
  int nx;
  int ny;

 public:
  template <class TT> friend class ComplexR_2D;

  Real_2D() : array(0), nx(0), ny(0) {}
  Real_2D(int x_size, int y_size) : array(0), nx(0), ny(0) { allocate_memory(x_size, y_size); }
  Real_2D(const Real_2D& o) : array(0), nx(0), ny(0) {
    allocate_memory(o.nx, o.ny);
    std::memcpy(array, o.array, sizeof(T) * (size_t)nx * ny);
  }
  ~Real_2D() { delete[] array; }

  Real_2D& operator=(const Real_2D& rhs) {
    if (this == &rhs) return *this;
    delete[] array;
    array = 0;
    allocate_memory(rhs.nx, rhs.ny);
    std::memcpy(array, rhs.array, sizeof(T) * (size_t)nx * ny);
    return *this;
  }

  void allocate_memory(int x_size, int y_size) {
    delete[] array;
    nx = x_size;
    ny = y_size;
    array = new T[(size_t)nx * ny]();
  }

  void copy(const Real_2D<T>& other) { std::memcpy(array, other.array, sizeof(T) * (size_t)nx * ny); }

  inline void set(int x, int y, T value) { array[(size_t)x * ny + y] = value; }

  // out-of-range reads return 0 (NADIA/include/Double_2D.h:135-143)
  inline T get(int x, int y) const {
    if (x < 0 || x >= nx || y < 0 || y >= ny) return 0;
    return array[(size_t)x * ny + y];
  }

  inline int get_size_x() const { return nx; }
  inline int get_size_y() const { return ny; }

  // raw x-major storage (for I/O and the C ABI)
  T* data() { return array; }
  const T* data() const { return array; }

  void mirror() {
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny / 2; ++j) {
        T a = get(i, ny - j - 1);
        set(i, ny - j - 1, get(i, j));
        set(i, j, a);
      }
  }

  T get_sum() const {
    T total = 0;
    for (size_t p = 0; p < (size_t)nx * ny; ++p) total += array[p];
    return total;
  }

  T get_abs_sum() const {
    T total = 0;
    for (size_t p = 0; p < (size_t)nx * ny; ++p) total += std::fabs(array[p]);
    return total;
  }

  T get_ave() const { return get_sum() / (get_size_x() * get_size_y()); }

  T get_max() const {
    if (nx == 0 || ny == 0) return 0;
    T m = array[0];
    for (size_t p = 0; p < (size_t)nx * ny; ++p)
      if (array[p] > m) m = array[p];
    return m;
  }

  T get_min() const {
    if (nx == 0 || ny == 0) return 0;
    T m = array[0];
    for (size_t p = 0; p < (size_t)nx * ny; ++p)
      if (array[p] < m) m = array[p];
    return m;
  }

  void add(const Real_2D<T>& other, double norm = 1.0) {
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++) array[(size_t)i * ny + j] += norm * other.get(i, j);
  }

  void scale(double s) {
    for (size_t p = 0; p < (size_t)nx * ny; ++p) array[p] *= s;
  }

  void square() {
    for (size_t p = 0; p < (size_t)nx * ny; ++p) array[p] *= array[p];
  }

  void sq_root() {
    for (size_t p = 0; p < (size_t)nx * ny; ++p) array[p] = std::sqrt(array[p]);
  }
};

#ifndef DOUBLE_PRECISION
typedef Real_2D<float> Double_2D;
#else
typedef Real_2D<double> Double_2D;
#endif

#endif
