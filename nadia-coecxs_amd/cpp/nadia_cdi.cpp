// nadia_cdi.cpp -- bodies of the drop-in reconstruction classes
// (include/nadia/BaseCDI.h, PlanarCDI.h, FresnelCDI.h, FresnelCDI_WF.h).
// Every numerical operation goes through the C ABI (include/nadia_amd.h);
// this file keeps the reference's object model, defaults and messages.
// Built twice: libNADIA_amd.so (float) and, with -DDOUBLE_PRECISION,
// libNADIAd_amd.so (double), like the reference's libNADIA / libNADIAd.
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <cstdio>
#include <string>
#include <vector>

#include "BaseCDI.h"
#include "FresnelCDI.h"
#include "FresnelCDI_WF.h"
#include "PlanarCDI.h"
#include "TransmissionConstraint.h"
#include "io.h"
#include "PhaseDiverseCDI.h"

using std::cout;
using std::endl;

// ---------------------------------------------------------------- BaseCDI --

std::map<std::string, int>* BaseCDI::set_up_algorithm_name_map() {
  std::map<std::string, int>* m = new std::map<std::string, int>;
  const char* names[] = {"ER", "BIO", "BOO", "HIO", "DM", "SF", "ASR", "HPR", "RAAR"};
  for (int i = 0; i < NALGORITHMS; i++) (*m)[names[i]] = i;
  return m;
}

std::map<std::string, int>* BaseCDI::algNameMap = BaseCDI::set_up_algorithm_name_map();

BaseCDI::BaseCDI(Complex_2D& initial_guess, unsigned int nbest, int kind)
    : complex(initial_guess),
      nx(initial_guess.get_size_x()),
      ny(initial_guess.get_size_y()),
      beta(0.9),
      support(nx, ny),
      intensity_sqrt(nx, ny),
      n_best((int)nbest),
      best_array(nbest, (Complex_2D*)0),
      transmission_constraint(0),
      recon(0) {
  nadia_cxx_check(nadia_recon_create(initial_guess.device_field(), kind, n_best, &recon), "BaseCDI constructor");
}

BaseCDI::BaseCDI(Complex_2D& initial_guess, unsigned int nbest) : BaseCDI(initial_guess, nbest, NADIA_PLANAR) {}

BaseCDI::~BaseCDI() {
  for (size_t i = 0; i < best_array.size(); i++) delete best_array[i];
  if (recon) nadia_recon_destroy(recon);
}

nadia_field* BaseCDI::dev(Complex_2D& c) {
  if (c.get_size_x() != nx || c.get_size_y() != ny) {
    cout << "the dimensions of the Complex_2D do not match the reconstruction" << endl;
    std::exit(1);
  }
  return c.device_field();
}

int BaseCDI::iterate() { return iterate(1, 0); }

int BaseCDI::iterate(int n, double* errors) {
  dev(complex);
  nadia_cxx_check(nadia_recon_iterate(recon, n, errors), "iterate");
  complex.device_modified();
  return SUCCESS;
}

void BaseCDI::set_execution_path(int path) { nadia_cxx_check(nadia_recon_set_path(recon, path), "set_execution_path"); }

Complex_2D* BaseCDI::get_best_result(double& error, int index) {
  if (index < 0 || index >= n_best) {
    cout << "index out of bounds in BaseCDI::get_best_result" << endl;
    return 0;
  }
  if (!best_array[index]) best_array[index] = new Complex_2D(nx, ny);
  Complex_2D* b = best_array[index];
  nadia_cxx_check(nadia_recon_get_best_result(recon, index, &error, b->host_data_for_write()), "get_best_result");
  return b;
}

void BaseCDI::set_support(const Double_2D& object_support, bool soften) {
  nadia_cxx_check(nadia_recon_set_support(recon, object_support.data(), soften ? 1 : 0), "set_support");
}

void BaseCDI::set_intensity(const Double_2D& detector_intensity) {
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) intensity_sqrt.set(i, j, std::sqrt(detector_intensity.get(i, j)));
  nadia_cxx_check(nadia_recon_set_intensity(recon, detector_intensity.data()), "set_intensity");
}

void BaseCDI::set_beam_stop(const Double_2D& beam_stop_region) {
  nadia_cxx_check(nadia_recon_set_beam_stop(recon, beam_stop_region.data()), "set_beam_stop");
}

void BaseCDI::set_relaxation_parameter(double relaxation_parameter) {
  beta = relaxation_parameter;
  nadia_cxx_check(nadia_recon_set_relaxation_parameter(recon, relaxation_parameter), "set_relaxation_parameter");
}

void BaseCDI::set_algorithm(int alg) { nadia_cxx_check(nadia_recon_set_algorithm(recon, alg), "set_algorithm"); }

void BaseCDI::set_custom_algorithm(double m1, double m2, double m3, double m4, double m5, double m6, double m7,
                                   double m8, double m9, double m10) {
  const double m[10] = {m1, m2, m3, m4, m5, m6, m7, m8, m9, m10};
  nadia_cxx_check(nadia_recon_set_custom_algorithm(recon, m), "set_custom_algorithm");
}

// same text as NADIA/src/BaseCDI.c++:278-295
void BaseCDI::print_algorithm() {
  double c[5];
  int alg = 0;
  nadia_cxx_check(nadia_recon_get_coefficients(recon, c, &alg), "print_algorithm");
  if (alg == ER)
    cout << "Currently using error-reduction: "
         << "x(k+1) = Ps Pf x(k)" << endl;
  else
    cout << "Currently using the algorithm: "
         << "x(k+1) = x(k) + (" << c[PSF] << "*PsPf + " << c[PFS] << "*PfPs + " << c[PS] << "*Ps + " << c[PF]
         << "*Pf + " << c[PI] << "*I"
         << ")x(k)" << endl;
  cout << "Ps - support constraint, Pf - modulus constraint" << endl;
}

double BaseCDI::get_error() {
  double e = 0;
  nadia_cxx_check(nadia_recon_get_error(recon, &e), "get_error");
  return e;
}

void BaseCDI::apply_shrinkwrap(double gauss_width, double threshold) {
  dev(complex);
  nadia_cxx_check(nadia_recon_apply_shrinkwrap(recon, gauss_width, threshold), "apply_shrinkwrap");
}

const Double_2D& BaseCDI::get_support() {
  nadia_cxx_check(nadia_recon_get_support(recon, support.data()), "get_support");
  return support;
}

void BaseCDI::apply_support(Complex_2D& c) {
  nadia_cxx_check(nadia_recon_apply_support(recon, dev(c)), "apply_support");
  c.device_modified();
}

void BaseCDI::support_constraint(Complex_2D& c) { BaseCDI::apply_support(c); }

void BaseCDI::project_intensity(Complex_2D& c) {
  nadia_cxx_check(nadia_recon_project_intensity(recon, dev(c)), "project_intensity");
  c.device_modified();
}

void BaseCDI::scale_intensity(Complex_2D& c) {
  nadia_cxx_check(nadia_recon_scale_intensity(recon, dev(c)), "scale_intensity");
  c.device_modified();
}

void BaseCDI::set_fftw_type(int type) { complex.set_fftw_type(type); }

// BaseCDI::set_complex_constraint (NADIA/include/BaseCDI.h:471-473): the
// constraint is held by pointer and applied on the device by apply_support
void BaseCDI::set_complex_constraint(TransmissionConstraint& trans_constraint) {
  transmission_constraint = &trans_constraint;
  nadia_cxx_check(nadia_recon_set_complex_constraint(recon, trans_constraint.device_handle(nx, ny)),
                  "set_complex_constraint");
}

// ---- TransmissionConstraint (NADIA/src/TransmissionConstraint.c++) ----------

TransmissionConstraint::TransmissionConstraint()
    : do_enforce_unity(true), do_charge_flip(true), flip_sign(1), custom_constraint(0), h(0), nx(0), ny(0) {}

TransmissionConstraint::~TransmissionConstraint() {
  if (h) nadia_constraint_destroy(h);
}

void TransmissionConstraint::push_settings() {
  if (!h) return;
  nadia_cxx_check(nadia_constraint_set_enforce_unity(h, do_enforce_unity ? 1 : 0), "set_enforce_unity");
  nadia_cxx_check(nadia_constraint_set_charge_flipping(h, do_charge_flip ? 1 : 0, flip_sign), "set_charge_flipping");
  nadia_cxx_check(nadia_constraint_set_custom(h, custom_constraint ? custom_trampoline : 0, this),
                  "set_custom_constraint");
}

nadia_constraint* TransmissionConstraint::device_handle(int x, int y) {
  if (!h) {
    nadia_cxx_check(nadia_constraint_create(x, y, &h), "TransmissionConstraint");
    nx = x;
    ny = y;
    push_settings();
  } else if (x != nx || y != ny) {
    cout << "Dimensions of the transmission function do not match the complex constraint regions" << endl;
    exit(1);
  }
  return h;
}

void TransmissionConstraint::delete_complex_constraint_regions() {
  while (!complex_constraint_list.empty()) {
    delete complex_constraint_list.back();
    complex_constraint_list.pop_back();
  }
  if (h) nadia_cxx_check(nadia_constraint_clear_regions(h), "delete_complex_constraint_regions");
}

// add_complex_constraint (NADIA/src/TransmissionConstraint.c++:43-77)
void TransmissionConstraint::add_complex_constraint(ComplexConstraint& c) {
  Double_2D* r = c.get_region();
  const int rx = r->get_size_x(), ry = r->get_size_y();
  if (h && (rx != nx || ry != ny)) {
    cout << "Dimensions of the complex constraint region do not match the others given previously. "
         << "Ignoring this complex constraint." << endl;
    return;
  }
  device_handle(rx, ry);
  std::vector<double> reg((size_t)rx * ry);
  for (int i = 0; i < rx; i++)
    for (int j = 0; j < ry; j++) reg[(size_t)i * ry + j] = r->get(i, j);
  int k = -1;
  nadia_cxx_check(nadia_constraint_add_region(h, reg.data(), c.alpha1, c.alpha2, &k), "add_complex_constraint");
  if (c.fixed_c) nadia_cxx_check(nadia_constraint_set_fixed_c(h, k, c.c_mean), "set_fixed_c");
  c.owner = this;
  c.index = k;
  complex_constraint_list.push_back(&c);
}

void TransmissionConstraint::set_charge_flipping(bool enable, int sign) {
  do_charge_flip = enable;
  flip_sign = sign > 0 ? 1 : -1;
  push_settings();
}

void TransmissionConstraint::set_enforce_unity(bool enable) {
  do_enforce_unity = enable;
  push_settings();
}

void TransmissionConstraint::set_custom_constraint(void (*fn)(Complex_2D& transmission)) {
  custom_constraint = fn;
  push_settings();
}

// the custom constraint sees the transmission function as a Complex_2D
void TransmissionConstraint::custom_trampoline(void* data, int x, int y, int precision, void* user) {
  TransmissionConstraint* self = static_cast<TransmissionConstraint*>(user);
  if (!self->custom_constraint) return;
  Complex_2D t(x, y);
  nadia_real* a = t.host_data_for_write();
  const size_t n = 2 * (size_t)x * y;
  if (precision == NADIA_FP64) {
    const double* d = static_cast<const double*>(data);
    for (size_t i = 0; i < n; i++) a[i] = (nadia_real)d[i];
  } else {
    const float* d = static_cast<const float*>(data);
    for (size_t i = 0; i < n; i++) a[i] = (nadia_real)d[i];
  }
  self->custom_constraint(t);
  const nadia_real* b = t.host_data();
  if (precision == NADIA_FP64) {
    double* d = static_cast<double*>(data);
    for (size_t i = 0; i < n; i++) d[i] = (double)b[i];
  } else {
    float* d = static_cast<float*>(data);
    for (size_t i = 0; i < n; i++) d[i] = (float)b[i];
  }
}

// apply_constraint (NADIA/src/TransmissionConstraint.c++:81-172)
void TransmissionConstraint::apply_constraint(Complex_2D& transmission) {
  nadia_constraint* d = device_handle(transmission.get_size_x(), transmission.get_size_y());
  nadia_cxx_check(nadia_constraint_apply(d, transmission.device_field()), "apply_constraint");
  transmission.device_modified();
}

void ComplexConstraint::set_fixed_c(double c) {
  fixed_c = true;
  c_mean = c;
  if (owner) nadia_cxx_check(nadia_constraint_set_fixed_c(owner->h, index, c), "set_fixed_c");
}

double ComplexConstraint::get_c_mean() {
  if (owner && !fixed_c) nadia_cxx_check(nadia_constraint_get_c_mean(owner->h, index, &c_mean), "get_c_mean");
  return c_mean;
}

void ComplexConstraint::set_alpha1(double a) {
  alpha1 = a;
  if (owner) nadia_cxx_check(nadia_constraint_set_alphas(owner->h, index, alpha1, alpha2), "set_alpha1");
}

void ComplexConstraint::set_alpha2(double a) {
  alpha2 = a;
  if (owner) nadia_cxx_check(nadia_constraint_set_alphas(owner->h, index, alpha1, alpha2), "set_alpha2");
}

void BaseCDI::reset_best() { nadia_cxx_check(nadia_recon_reset_best(recon), "reset_best"); }

// Host versions of the reference's protected helpers, for derived classes
// (NADIA/src/BaseCDI.c++:432-492).
void BaseCDI::convolve(Double_2D& array, double gauss_width, int cut) {
  Double_2D tmp(nx, ny);
  Double_2D g(cut + 1, cut + 1);
  for (int i = 0; i <= cut; i++)
    for (int j = 0; j <= cut; j++) g.set(i, j, std::exp(-1 * (i * i + j * j) / (2.0 * gauss_width * gauss_width)));
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      double v = 0;
      for (int i2 = i - cut; i2 <= i + cut; i2++)
        for (int j2 = j - cut; j2 <= j + cut; j2++)
          if (i2 < nx && i2 >= 0 && j2 >= 0 && j2 < ny) v += array.get(i2, j2) * g.get(std::abs(i - i2), std::abs(j - j2));
      tmp.set(i, j, v);
    }
  array.copy(tmp);
}

// BaseCDI::apply_threshold (NADIA/src/BaseCDI.c++:477-492) loops j < nx: for
// nx < ny the columns j >= nx keep their values; for nx > ny the reference's
// loop writes past each row and, on the last row, past the array (undefined
// behaviour), so here each pixel is then thresholded once, as on the device.
void BaseCDI::apply_threshold(Double_2D& array, double threshold) {
  double max = array.get_max();
  const int jn = nx < ny ? nx : ny;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < jn; j++) array.set(i, j, array.get(i, j) < threshold * max ? 0.0 : 1.0);
}

// -------------------------------------------------------------- PlanarCDI --

void PlanarCDI::initialise_estimate(int seed) {
  nadia_cxx_check(nadia_recon_initialise_estimate(recon, seed), "initialise_estimate");
  complex.device_modified();
}

void PlanarCDI::propagate_to_detector(Complex_2D& c) {
  nadia_cxx_check(nadia_recon_propagate_to_detector(recon, dev(c)), "propagate_to_detector");
  c.device_modified();
}

void PlanarCDI::propagate_from_detector(Complex_2D& c) {
  nadia_cxx_check(nadia_recon_propagate_from_detector(recon, dev(c)), "propagate_from_detector");
  c.device_modified();
}

void PlanarCDI::get_intensity_autocorrelation(Double_2D& autoc) {
  nadia_cxx_check(nadia_recon_get_intensity_autocorrelation(recon, autoc.data()), "get_intensity_autocorrelation");
}

// ------------------------------------------------------------- FresnelCDI --

FresnelCDI::FresnelCDI(Complex_2D& initial_guess, Complex_2D& white_field, double beam_wavelength,
                       double focal_detector_length, double focal_sample_length, double pixel_size,
                       double normalisation, int nbest)
    : BaseCDI(initial_guess, nbest, NADIA_FRESNEL),
      illumination_at_sample(initial_guess.get_size_x(), initial_guess.get_size_y()),
      wavelength(beam_wavelength),
      pixel_length(pixel_size),
      focal_detector_length(focal_detector_length),
      focal_sample_length(focal_sample_length),
      norm(normalisation) {
  nadia_cxx_check(nadia_fresnel_setup(recon, white_field.host_data(), beam_wavelength, focal_detector_length,
                                      focal_sample_length, pixel_size, normalisation),
                  "FresnelCDI constructor");
}

void FresnelCDI::set_experimental_parameters(double wl, double fd, double fs, double px) {
  wavelength = wl;
  focal_detector_length = fd;
  focal_sample_length = fs;
  pixel_length = px;
  nadia_cxx_check(nadia_fresnel_set_experimental_parameters(recon, wl, fd, fs, px), "set_experimental_parameters");
}

void FresnelCDI::initialise_estimate(int seed) {
  nadia_cxx_check(nadia_recon_initialise_estimate(recon, seed), "initialise_estimate");
  complex.device_modified();
}

void FresnelCDI::auto_set_norm() { nadia_cxx_check(nadia_fresnel_auto_set_norm(recon), "auto_set_norm"); }

void FresnelCDI::set_norm(double n) {
  norm = n;
  nadia_cxx_check(nadia_fresnel_set_norm(recon, n), "set_norm");
}

void FresnelCDI::set_normalisation(double n) {
  norm = n;
  nadia_cxx_check(nadia_fresnel_set_normalisation(recon, n), "set_normalisation");
}

void FresnelCDI::get_transmission_function(Complex_2D& result, Complex_2D* esw) {
  nadia_field* e = esw ? dev(*esw) : dev(complex);
  nadia_cxx_check(nadia_fresnel_get_transmission_function(recon, e, result.host_data_for_write()),
                  "get_transmission_function");
}

void FresnelCDI::set_transmission_function(Complex_2D& transmission, Complex_2D* esw) {
  Complex_2D& e = esw ? *esw : complex;
  nadia_cxx_check(nadia_fresnel_set_transmission_function(recon, transmission.host_data(), dev(e)),
                  "set_transmission_function");
  e.device_modified();
}

void FresnelCDI::apply_support(Complex_2D& c) { BaseCDI::apply_support(c); }

void FresnelCDI::scale_intensity(Complex_2D& c) { BaseCDI::scale_intensity(c); }

void FresnelCDI::propagate_from_detector(Complex_2D& c) {
  nadia_cxx_check(nadia_recon_propagate_from_detector(recon, dev(c)), "propagate_from_detector");
  c.device_modified();
}

void FresnelCDI::propagate_to_detector(Complex_2D& c) {
  nadia_cxx_check(nadia_recon_propagate_to_detector(recon, dev(c)), "propagate_to_detector");
  c.device_modified();
}

const Complex_2D& FresnelCDI::get_illumination_at_sample() {
  nadia_cxx_check(nadia_fresnel_get_illumination_at_sample(recon, illumination_at_sample.host_data_for_write()),
                  "get_illumination_at_sample");
  return illumination_at_sample;
}

// ---------------------------------------------------------- FresnelCDI_WF --

FresnelCDI_WF::FresnelCDI_WF(Complex_2D& initial_guess, double beam_wavelength, double zone_focal_length,
                             double focal_detector_length, double pixel_size, int nbest)
    : BaseCDI(initial_guess, nbest, NADIA_FRESNEL_WF),
      wavelength(beam_wavelength),
      zone_to_focal_length(zone_focal_length),
      focal_to_detector_length(focal_detector_length),
      pixel_length(pixel_size) {
  nadia_cxx_check(nadia_wf_setup(recon, beam_wavelength, zone_focal_length, focal_detector_length, pixel_size),
                  "FresnelCDI_WF constructor");
}

int FresnelCDI_WF::iterate() { return BaseCDI::iterate(1, 0); }

void FresnelCDI_WF::propagate_from_detector(Complex_2D& c) {
  nadia_cxx_check(nadia_recon_propagate_from_detector(recon, dev(c)), "propagate_from_detector");
  c.device_modified();
}

void FresnelCDI_WF::propagate_to_detector(Complex_2D& c) {
  nadia_cxx_check(nadia_recon_propagate_to_detector(recon, dev(c)), "propagate_to_detector");
  c.device_modified();
}

void FresnelCDI_WF::initialise_estimate(int seed) {
  nadia_cxx_check(nadia_recon_initialise_estimate(recon, seed), "initialise_estimate");
  complex.device_modified();
}

void FresnelCDI_WF::set_support(double z_diameter, double size) {
  nadia_cxx_check(nadia_wf_set_support(recon, z_diameter, size), "set_support");
}

void FresnelCDI_WF::multiply_factors(Complex_2D& c, int direction) {
  const double xm = (nx - 1) / 2.0, ym = (ny - 1) / 2.0;
  const double sx = wavelength * focal_to_detector_length / (pixel_length * nx);
  const double sy = wavelength * focal_to_detector_length / (pixel_length * ny);
  const double nrm = 1 / (std::sqrt((double)(nx * ny)));
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      const double ax = sx * (xm - i), ay = sy * (ym - j);
      const double rho2 = ax * ax + ay * ay;
      double phi = rho2 / zone_to_focal_length + rho2 / focal_to_detector_length;
      phi *= M_PI / wavelength;
      const nadia_real cr = std::sin(phi) * nrm, ci = direction * (nadia_real)(-std::cos(phi) * nrm);
      const double re = c.get_real(i, j), im = c.get_imag(i, j);
      c.set_real(i, j, re * cr - im * ci);
      c.set_imag(i, j, im * cr + re * ci);
    }
}

// ---- .dbin / .cplx files (NADIA/src/io_dbin.c:24-100, NADIA/src/io_cplx.c:25-101)
// Headerless doubles, y-major: element (x, y) at y*nx + x.

int read_dbin(std::string file_name, int nx, int ny, Double_2D& data) {
  FILE* file = fopen(file_name.c_str(), "rb");
  if (!file) {
    cout << "Could not open the file " << file_name << "." << endl;
    return FAILURE;
  }
  std::vector<double> buffer((size_t)nx * ny);
  const size_t got = fread(buffer.data(), sizeof(double), buffer.size(), file);
  fclose(file);
  if (got != buffer.size()) {
    cout << "Could not correctly read the file" << file_name << ". Perhaps the dimensions specifies are wrong"
         << " or the format type is not double?" << endl;
    return FAILURE;
  }
  if (data.get_size_x() == 0) data.allocate_memory(nx, ny);
  if (data.get_size_x() != nx || data.get_size_y() != ny) {
    cout << "The Double_2D object supplied has the wrong dimensions" << endl;
    return FAILURE;
  }
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) data.set(i, j, buffer[(size_t)j * nx + i]);
  return SUCCESS;
}

int write_dbin(std::string file_name, const Double_2D& data) {
  FILE* file = fopen(file_name.c_str(), "w+b");
  if (!file) {
    cout << "Could not open the file " << file_name << endl;
    return FAILURE;
  }
  const int nx = data.get_size_x(), ny = data.get_size_y();
  std::vector<double> buffer((size_t)nx * ny);
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) buffer[(size_t)j * nx + i] = data.get(i, j);
  fwrite(buffer.data(), sizeof(double), buffer.size(), file);
  fclose(file);
  return SUCCESS;
}

int read_cplx(std::string file_name, Complex_2D& complex) {
  const int nx = complex.get_size_x(), ny = complex.get_size_y();
  FILE* file = fopen(file_name.c_str(), "rb");
  if (!file) {
    cout << "Could not open the file " << file_name << endl;
    return FAILURE;
  }
  std::vector<double> buffer(2 * (size_t)nx * ny);
  const size_t got = fread(buffer.data(), sizeof(double), buffer.size(), file);
  fclose(file);
  if (got != buffer.size()) {
    cout << "Could not correctly read the file " << file_name << ". Perhaps the dimensions specifies are wrong?"
         << endl;
    return FAILURE;
  }
  nadia_real* a = complex.host_data_for_write();
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      const size_t p = (size_t)j * nx + i, q = (size_t)i * ny + j;
      a[2 * q] = (nadia_real)buffer[2 * p];
      a[2 * q + 1] = (nadia_real)buffer[2 * p + 1];
    }
  return SUCCESS;
}

int write_cplx(std::string file_name, const Complex_2D& complex) {
  const int nx = complex.get_size_x(), ny = complex.get_size_y();
  FILE* file = fopen(file_name.c_str(), "w+b");
  if (!file) {
    cout << "Could not open the file " << file_name << endl;
    return FAILURE;
  }
  std::vector<double> buffer(2 * (size_t)nx * ny);
  const nadia_real* a = complex.host_data();
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      const size_t p = (size_t)j * nx + i, q = (size_t)i * ny + j;
      buffer[2 * p] = a[2 * q];
      buffer[2 * p + 1] = a[2 * q + 1];
    }
  fwrite(buffer.data(), sizeof(double), buffer.size(), file);
  fclose(file);
  return SUCCESS;
}

// ---- read_image / write_image (NADIA/include/io.h:213-284) -------------------
// The format comes from a substring of the file name, tested in the
// reference's order; only .dbin is handled here.

namespace {

bool name_has(const std::string& name, const char* ext) { return name.find(ext) != std::string::npos; }

// the image formats whose readers and writers are not part of this library
const char* other_format(const std::string& name) {
  if (name_has(name, ".tiff") || name_has(name, ".tif")) return "TIFF";
  if (name_has(name, ".ppm")) return "PPM";
  if (name_has(name, ".txt")) return "text";
  if (name_has(name, ".hdf")) return "HDF";
  return 0;
}

}  // namespace

void read_image(std::string file_name, Double_2D& data, int nx, int ny, const char*) {
  int status = FAILURE;
  if (const char* f = other_format(file_name))
    cout << f << " files are not read by this library (use .dbin)" << endl;
  if (name_has(file_name, ".dbin")) {
    if (nx == 0 || ny == 0)
      cout << "Please pass the dimensions of the dbin file you wish to read to the read_image method." << endl;
    else
      status = read_dbin(file_name, nx, ny, data);
  }
  if (status == FAILURE) {
    cout << "Failed to read the file: " << file_name << ". Exiting now.." << endl;
    exit(0);
  }
}

void write_image(std::string file_name, Double_2D& data, bool, double, double) {
  int status = FAILURE;
  if (const char* f = other_format(file_name))
    cout << f << " files are not written by this library (use .dbin)" << endl;
  if (name_has(file_name, ".dbin")) status = write_dbin(file_name, data);
  if (status == FAILURE) {
    cout << "Failed to write to the file: " << file_name << ". Exiting now.." << endl;
    exit(0);
  }
}

// ---- PhaseDiverseCDI (NADIA/src/PhaseDiverseCDI.c++) ------------------------

PhaseDiverseCDI::PhaseDiverseCDI(double beta, double gamma, bool parallel, int granularity)
    : beta(beta), gamma(gamma), object(0), iterations_per_cycle(1), scale(granularity), parallel(parallel), h(0) {
  nadia_cxx_check(nadia_pd_create(beta, gamma, parallel ? 1 : 0, granularity, &h), "PhaseDiverseCDI");
}

PhaseDiverseCDI::~PhaseDiverseCDI() {
  if (object) delete object;
  nadia_pd_destroy(h);
}

void PhaseDiverseCDI::push_parameters() {
  nadia_cxx_check(nadia_pd_set_parameters(h, iterations_per_cycle, beta, gamma), "PhaseDiverseCDI parameters");
}

// the frames' estimates were changed on the device
void PhaseDiverseCDI::frames_changed() {
  for (size_t i = 0; i < singleCDI.size(); i++) singleCDI[i]->device_changed();
}

void PhaseDiverseCDI::add_new_position(BaseCDI* local, double x, double y, double a) {
  nadia_cxx_check(nadia_pd_add_position(h, local->device_sync(), x, y, a), "add_new_position");
  singleCDI.push_back(local);
  x_position.push_back(x);
  y_position.push_back(y);
  alpha.push_back(a);
  cout << "Added position " << singleCDI.size() - 1 << endl;
}

void PhaseDiverseCDI::remove_positions() {
  nadia_cxx_check(nadia_pd_remove_positions(h), "remove_positions");
  singleCDI.clear();
  x_position.clear();
  y_position.clear();
  alpha.clear();
}

void PhaseDiverseCDI::initialise_estimate() {
  for (size_t i = 0; i < singleCDI.size(); i++) singleCDI[i]->device_sync();
  nadia_cxx_check(nadia_pd_initialise_estimate(h), "initialise_estimate");
}

void PhaseDiverseCDI::iterate() {
  static int total_iterations = 0;
  cout << "Iteration " << total_iterations << endl;
  for (size_t i = 0; i < singleCDI.size(); i++) singleCDI[i]->device_sync();
  std::vector<double> errors(singleCDI.size());
  nadia_cxx_check(nadia_pd_iterate(h, errors.data()), "PhaseDiverseCDI::iterate");
  frames_changed();
  for (size_t i = 0; i < errors.size(); i++) cout << "Error for frame " << i << " is " << errors[i] << endl;
  total_iterations++;
}

void PhaseDiverseCDI::set_iterations_per_cycle(int iterations) {
  iterations_per_cycle = iterations;
  push_parameters();
}

void PhaseDiverseCDI::set_feedback_parameter(double b) {
  beta = b;
  push_parameters();
}

void PhaseDiverseCDI::set_amplification_factor(double g) {
  gamma = g;
  push_parameters();
}

void PhaseDiverseCDI::set_probe_scaling(int n_probe, double a) {
  nadia_cxx_check(nadia_pd_set_probe_scaling(h, n_probe, a), "set_probe_scaling");
  if (n_probe >= 0 && n_probe < (int)alpha.size()) alpha[n_probe] = a;
}

Complex_2D* PhaseDiverseCDI::get_transmission() {
  int nx = 0, ny = 0;
  nadia_cxx_check(nadia_pd_get_size(h, &nx, &ny), "get_transmission");
  if (nx == 0 || ny == 0) return 0;
  if (!object || object->get_size_x() != nx || object->get_size_y() != ny) {
    if (object) delete object;
    object = new Complex_2D(nx, ny);
  }
  nadia_cxx_check(nadia_pd_get_transmission(h, object->host_data_for_write()), "get_transmission");
  return object;
}

void PhaseDiverseCDI::set_transmission(Complex_2D& t) {
  int rc = nadia_pd_set_transmission(h, t.get_size_x(), t.get_size_y(), t.host_data());
  if (rc == NADIA_ESIZE)
    cout << "Can not set the transmission function in PhaseDiverseCDI because the complex array given does not "
         << "have the same dimensions as the current transmission function" << endl;
  else
    nadia_cxx_check(rc, "set_transmission");
}

// adjust_positions (NADIA/src/PhaseDiverseCDI.c++:337-579)
void PhaseDiverseCDI::adjust_positions(int type, bool forwards, int x_min, int x_max, int y_min, int y_max,
                                       double step_size) {
  if (type != CROSS_CORRELATION && type != MINIMUM_ERROR) {
    std::cerr << "The alignment type given to PhaseDiverseCDI::adjust_positions is not a known type. Please "
              << "consider using PhaseDiverseCDI::CROSS_CORRELATION or  PhaseDiverseCDI::MINIMUM_ERROR." << endl;
    return;
  }
  const int nf = (int)singleCDI.size();
  if (nf == 0) return;
  for (int i = 0; i < nf; i++) singleCDI[i]->device_sync();
  std::vector<int> gave_up(nf, 0);
  nadia_cxx_check(nadia_pd_adjust_positions(h, type, forwards ? 1 : 0, x_min, x_max, y_min, y_max, step_size,
                                            gave_up.data()),
                  "adjust_positions");
  frames_changed();
  // the reference reports each aligned frame as it goes: all but the first
  // (forwards) or the last (backwards), in alignment order
  for (int k = 1; k < nf; k++) {
    const int n = forwards ? k : nf - 1 - k;
    const double before_x = x_position[n], before_y = y_position[n];
    nadia_cxx_check(nadia_pd_get_position(h, n, &x_position[n], &y_position[n]), "adjust_positions");
    if (gave_up[n])
      cout << "Giving up on probe " << n << ". Could not find it's position. Returning to the original. " << endl;
    cout << "moving prob " << n << " from (" << before_x << "," << before_y << ") -> (" << x_position[n] << ","
         << y_position[n] << ")." << endl;
  }
}
