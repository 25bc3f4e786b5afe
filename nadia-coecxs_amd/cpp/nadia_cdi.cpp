// nadia_cdi.cpp -- bodies of the drop-in reconstruction classes
// (include/nadia/BaseCDI.h, PlanarCDI.h, FresnelCDI.h, FresnelCDI_WF.h).
// Every numerical operation goes through the C ABI (include/nadia_amd.h);
// this file keeps the reference's object model, defaults and messages.
// Built twice: libNADIA_amd.so (float) and, with -DDOUBLE_PRECISION,
// libNADIAd_amd.so (double), like the reference's libNADIA / libNADIAd.
#include <cmath>
#include <cstdlib>
#include <iostream>

#include "BaseCDI.h"
#include "FresnelCDI.h"
#include "FresnelCDI_WF.h"
#include "PlanarCDI.h"

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

void BaseCDI::set_complex_constraint(TransmissionConstraint& trans_constraint) {
  transmission_constraint = &trans_constraint;
  cout << "WARNING: complex (transmission) constraints are not applied by the device "
       << "reconstruction; the constraint is ignored" << endl;
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

void BaseCDI::apply_threshold(Double_2D& array, double threshold) {
  double max = array.get_max();
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < nx; j++) array.set(i, j, array.get(i, j) < threshold * max ? 0.0 : 1.0);
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
