// Complex_2D.h -- complex 2-D field of the NADIA drop-in API.
//
// Interface of NADIA's ComplexR_2D<T> (NADIA/include/Complex_2D.h:35-515,
// NADIA/src/Complex_2D.c++); Complex_2D is ComplexR_2D<float>, or
// ComplexR_2D<double> with DOUBLE_PRECISION.  The data live in two places:
// a host array (interleaved re/im, x-major: element (x, y) at x*ny + y, the
// reference layout) used by the element accessors, and a device field used by
// the transforms and by the reconstruction objects.  Each side is copied to
// the other only when it is stale, so a reconstruction loop that never reads
// elements never leaves the device.
#ifndef NADIA_AMD_COMPLEX_2D_H
#define NADIA_AMD_COMPLEX_2D_H

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "../nadia_amd.h"
#include "Double_2D.h"
#include "types.h"

#ifndef NUM_THREADS
#define NUM_THREADS 1
#endif

// Fatal C-ABI failure: print and exit, the reference's error convention
// (e.g. NADIA/src/Complex_2D.c++:186-191).
inline void nadia_cxx_check(int rc, const char* what) {
  if (rc != 0) {
    std::cout << "NADIA (HIP) error in " << what << ": " << nadia_last_error() << std::endl;
    std::exit(1);
  }
}

template <class T>
class ComplexR_2D {
  int nx, ny;
  int fftw_type;
  T* array;                  // host copy, 2*nx*ny values
  mutable nadia_field* dev;  // device copy (created on first device use)
  mutable int valid;         // bit 0: host copy current, bit 1: device copy current

  enum { HOST = 1, DEVICE = 2 };
  static int precision() { return sizeof(T) == sizeof(double) ? NADIA_FP64 : NADIA_FP32; }

  void alloc(int x, int y) {
    nx = x;
    ny = y;
    array = new T[2 * (size_t)nx * ny]();
    dev = 0;
    valid = HOST;
    fftw_type = FFTW_MEASURE;
  }

  void pull() const {
    if (dev) nadia_cxx_check(nadia_field_download(dev, array), "Complex_2D download");
    valid |= HOST;
  }
  inline void host_read() const {
    if (!(valid & HOST)) pull();
  }
  inline void host_write() {
    if (!(valid & HOST)) pull();
    valid = HOST;
  }

 public:
  ComplexR_2D(int x_size, int y_size) { alloc(x_size, y_size); }

  ComplexR_2D(const ComplexR_2D& o) {
    alloc(o.nx, o.ny);
    o.host_read();
    std::memcpy(array, o.array, sizeof(T) * 2 * (size_t)nx * ny);
  }

  ComplexR_2D(const Double_2D& o) {
    alloc(o.get_size_x(), o.get_size_y());
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++) {
        array[2 * ((size_t)i * ny + j)] = o.get(i, j);
        array[2 * ((size_t)i * ny + j) + 1] = 0;
      }
  }

  ~ComplexR_2D() {
    if (dev) nadia_field_destroy(dev);
    delete[] array;
  }

  ComplexR_2D& operator=(const ComplexR_2D& rhs) {
    if (this == &rhs) return *this;
    if (dev) nadia_field_destroy(dev);
    delete[] array;
    alloc(rhs.nx, rhs.ny);
    rhs.host_read();
    std::memcpy(array, rhs.array, sizeof(T) * 2 * (size_t)nx * ny);
    return *this;
  }

  ComplexR_2D& operator=(const Double_2D& rhs) {
    if (dev) nadia_field_destroy(dev);
    delete[] array;
    alloc(rhs.get_size_x(), rhs.get_size_y());
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++) array[2 * ((size_t)i * ny + j)] = rhs.get(i, j);
    return *this;
  }

  // ---- element access (NADIA/include/Complex_2D.h:167-305) ----
  void set_value(int x, int y, int type, T value) {
    if (x < 0 || x >= nx || y < 0 || y >= ny) {
      std::cout << "can not set value out of array bounds" << std::endl;
      std::exit(1);
    }
    switch (type) {
      case REAL: set_real(x, y, value); break;
      case IMAG: set_imag(x, y, value); break;
      case MAG: set_mag(x, y, value); break;
      case PHASE: set_phase(x, y, value); break;
      default:
        std::cout << "Value type in Complex_2D::set_value is unknown: " << type
                  << ". Must be REAL, IMAG, MAG or PHASE" << std::endl;
        std::exit(1);
    }
  }

  inline void set_real(int x, int y, T value) {
    host_write();
    array[2 * ((size_t)x * ny + y)] = value;
  }
  inline void set_imag(int x, int y, T value) {
    host_write();
    array[2 * ((size_t)x * ny + y) + 1] = value;
  }
  // keeps the phase; a zero value becomes (value, 0)
  inline void set_mag(int x, int y, T value) {
    double mag = get_mag(x, y);
    host_write();
    T* c = array + 2 * ((size_t)x * ny + y);
    if (mag == 0) {
      c[0] = value;
      c[1] = 0;
    } else {
      c[0] *= value / mag;
      c[1] *= value / mag;
    }
  }
  inline void set_phase(int x, int y, T value) {
    double mag = get_mag(x, y);
    host_write();
    array[2 * ((size_t)x * ny + y)] = mag * std::cos(value);
    array[2 * ((size_t)x * ny + y) + 1] = mag * std::sin(value);
  }
  inline T get_real(int x, int y) const {
    host_read();
    return array[2 * ((size_t)x * ny + y)];
  }
  inline T get_imag(int x, int y) const {
    host_read();
    return array[2 * ((size_t)x * ny + y) + 1];
  }
  inline T get_mag(int x, int y) const {
    host_read();
    const T* c = array + 2 * ((size_t)x * ny + y);
    return std::sqrt(c[0] * c[0] + c[1] * c[1]);
  }
  inline T get_phase(int x, int y) const {
    T phase = std::atan2(get_imag(x, y), get_real(x, y));
    if (phase > M_PI) return phase - 2 * M_PI;
    return phase;
  }

  T get_value(int x, int y, int type) const {
    switch (type) {
      case MAG: return get_mag(x, y);
      case REAL: return get_real(x, y);
      case IMAG: return get_imag(x, y);
      case PHASE: return get_phase(x, y);
      case MAG_SQ: return std::pow(get_mag(x, y), 2);
      default:
        std::cout << "value type in Complex_2D::get_value is unknown" << std::endl;
        std::exit(1);
    }
  }

  int get_size_x() const { return nx; }
  int get_size_y() const { return ny; }

  void get_2d(int type, Double_2D& result) const {
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++) result.set(i, j, get_value(i, j, type));
  }

  // ---- whole-array operations ----
  void scale(T factor) {
    nadia_cxx_check(nadia_field_scale(device_field(), factor), "Complex_2D::scale");
    device_modified();
  }

  void add(ComplexR_2D& c2, T s = 1) {
    if (nx != c2.nx || ny != c2.ny) {
      std::cout << "in Complex_2D::add, the dimensions of the input Complex_2D do not match the "
                   "dimensions of this Complex_2D object"
                << std::endl;
      std::exit(1);
    }
    nadia_cxx_check(nadia_field_add(device_field(), c2.device_field(), s), "Complex_2D::add");
    device_modified();
  }

  void multiply(ComplexR_2D& c2, T s = 1) {
    if (nx != c2.nx || ny != c2.ny) {
      std::cout << "in Complex_2D::multiply, the dimensions of the input Complex_2D do not match "
                   "the dimensions of this Complex_2D object"
                << std::endl;
      std::exit(1);
    }
    host_write();
    c2.host_read();
    for (size_t p = 0; p < (size_t)nx * ny; ++p) {
      T ar = array[2 * p], ai = array[2 * p + 1], br = c2.array[2 * p], bi = c2.array[2 * p + 1];
      T re = br * ar - bi * ai;
      T im = br * ai + bi * ar;
      array[2 * p] = s * re;
      array[2 * p + 1] = s * im;
    }
  }

  void multiply(Double_2D& d2, T s = 1) {
    if (nx != d2.get_size_x() || ny != d2.get_size_y()) {
      std::cout << "in Complex_2D::multiply, the dimensions of the input Double_2D do not match "
                   "the dimensions of this Complex_2D object"
                << std::endl;
      std::exit(1);
    }
    host_write();
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++) {
        T* c = array + 2 * ((size_t)i * ny + j);
        T re = d2.get(i, j) * c[0], im = d2.get(i, j) * c[1];
        c[0] = s * re;
        c[1] = s * im;
      }
  }

  T get_norm() const {
    double v = 0;
    nadia_cxx_check(nadia_field_norm(const_cast<ComplexR_2D*>(this)->device_field(), &v), "Complex_2D::get_norm");
    return (T)v;
  }

  void conjugate() {
    host_write();
    for (size_t p = 0; p < (size_t)nx * ny; ++p) array[2 * p + 1] = -array[2 * p + 1];
  }

  ComplexR_2D* clone() const {
    ComplexR_2D* n = new ComplexR_2D(nx, ny);
    n->copy(*this);
    return n;
  }

  void copy(const ComplexR_2D& c) {
    if (c.nx != nx || c.ny != ny) {
      std::cout << "Trying to copy an array with different dimensions... exiting" << std::endl;
      std::exit(1);
    }
    if (&c == this) return;
    if ((c.valid & DEVICE) && !(c.valid & HOST)) {
      nadia_cxx_check(nadia_field_copy(device_field_no_upload(), c.dev), "Complex_2D::copy");
      valid = DEVICE;
    } else {
      c.host_read();
      host_write();
      std::memcpy(array, c.array, sizeof(T) * 2 * (size_t)nx * ny);
    }
  }

  // fftshift with optional 1/sqrt(nx*ny) (NADIA/src/Complex_2D.c++:317-353)
  void invert(bool scale = false) {
    if (nx % 2 == 1 || ny % 2 == 1)
      std::cout << "WARNING: The array dimensions are odd but we have assumed they are even when "
                   "inverting an array after FFT. This will probably cause you issues..."
                << std::endl;
    nadia_cxx_check(nadia_field_invert(device_field(), scale ? 1 : 0), "Complex_2D::invert");
    device_modified();
  }

  void perform_forward_fft() {
    nadia_cxx_check(nadia_field_fft(device_field(), -1), "Complex_2D::perform_forward_fft");
    device_modified();
  }

  void perform_backward_fft() {
    nadia_cxx_check(nadia_field_fft(device_field(), +1), "Complex_2D::perform_backward_fft");
    device_modified();
  }

  void mirror() {
    host_write();
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny / 2; ++j) {
        T* a = array + 2 * ((size_t)i * ny + (ny - j - 1));
        T* b = array + 2 * ((size_t)i * ny + j);
        T t0 = a[0], t1 = a[1];
        a[0] = b[0];
        a[1] = b[1];
        b[0] = t0;
        b[1] = t1;
      }
  }

  // Real-to-complex / complex-to-real transforms in FFTW's packed layout
  // (nx x (ny/2+1) leading values), as NADIA/src/Complex_2D.c++:506-523.
  void perform_forward_fft_real(Double_2D& input) {
    ComplexR_2D full(input);
    full.perform_forward_fft();
    full.host_read();
    host_write();
    const int h = ny / 2 + 1;
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < h; j++) {
        array[2 * ((size_t)i * h + j)] = full.array[2 * ((size_t)i * ny + j)];
        array[2 * ((size_t)i * h + j) + 1] = full.array[2 * ((size_t)i * ny + j) + 1];
      }
  }

  void perform_backward_fft_real(Double_2D& result) {
    host_read();
    const int h = ny / 2 + 1;
    ComplexR_2D full(nx, ny);
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++) {
        T re, im;
        if (j < h) {
          re = array[2 * ((size_t)i * h + j)];
          im = array[2 * ((size_t)i * h + j) + 1];
        } else {  // Hermitian symmetry of a real signal's transform
          const int ii = (nx - i) % nx, jj = ny - j;
          re = array[2 * ((size_t)ii * h + jj)];
          im = -array[2 * ((size_t)ii * h + jj) + 1];
        }
        full.array[2 * ((size_t)i * ny + j)] = re;
        full.array[2 * ((size_t)i * ny + j) + 1] = im;
      }
    full.perform_backward_fft();
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++) result.set(i, j, full.get_real(i, j));
  }

  void set_fftw_type(int type) { fftw_type = type; }

  ComplexR_2D get_padded(int x_add, int y_add) {
    ComplexR_2D padded(nx + 2 * x_add, ny + 2 * y_add);
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++) {
        padded.set_real(i + x_add, j + y_add, get_real(i, j));
        padded.set_imag(i + x_add, j + y_add, get_imag(i, j));
      }
    return padded;
  }

  ComplexR_2D get_unpadded(int x_add, int y_add) {
    ComplexR_2D unpadded(nx - 2 * x_add, ny - 2 * y_add);
    for (int i = 0; i < nx - 2 * x_add; i++)
      for (int j = 0; j < ny - 2 * y_add; j++) {
        unpadded.set_real(i, j, get_real(i + x_add, j + y_add));
        unpadded.set_imag(i, j, get_imag(i + x_add, j + y_add));
      }
    return unpadded;
  }

  // raw x-major interleaved host data (brought up to date first)
  const T* host_data() const {
    host_read();
    return array;
  }
  T* host_data_for_write() {
    host_write();
    return array;
  }

  // ---- device side (used by the reconstruction classes) ----
  // The device field, brought up to date with the host copy.
  nadia_field* device_field() {
    device_field_no_upload();
    if (!(valid & DEVICE)) {
      nadia_cxx_check(nadia_field_upload(dev, array), "Complex_2D upload");
      valid |= DEVICE;
    }
    return dev;
  }
  // Record that device work changed the field (the host copy is stale).
  void device_modified() { valid = DEVICE; }

 private:
  nadia_field* device_field_no_upload() {
    if (!dev) nadia_cxx_check(nadia_field_create(nx, ny, precision(), &dev), "Complex_2D device allocation");
    return dev;
  }
};

#ifndef DOUBLE_PRECISION
typedef ComplexR_2D<float> Complex_2D;
#else
typedef ComplexR_2D<double> Complex_2D;
#endif

#endif
