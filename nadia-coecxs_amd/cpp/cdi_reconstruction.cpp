// cdi_reconstruction.cpp -- the configuration-file reconstruction tool of the
// NADIA drop-in API (NADIA/tools/CDI_reconstruction.c:94-466).
//
// usage: cdi_reconstruction <config file> <reco_type> <seed>
//
// reco_type: "planar" (default), "fresnel_wf" (white-field reconstruction,
// 3-plane propagation) or "fresnel" (object reconstruction with the white
// field reconstructed before).  The partially coherent types of the reference
// ("partial", "partchar", "poly") are not part of this library.  The
// configuration keys, console messages, iteration schedule (algorithm list
// with iteration counts, shrinkwrap every shrinkwrap_iterations, magnitude
// image every output_iterations) and output files (<prefix>.cplx, or
// white_field_reco_file_name for fresnel_wf; transmission_function.cplx for
// fresnel) follow the reference tool.  Data, support and images are .dbin
// files (see read_image).  Deliberate differences: output_iterations = 0
// writes no intermediate images (the reference divides by zero), and the
// diffraction image is named <prefix>_diffraction_<i>.<type> (the reference
// appends that name to the object image's name).
#include <cstdlib>
#include <iostream>
#include <list>
#include <sstream>
#include <string>

#include "BaseCDI.h"
#include "Complex_2D.h"
#include "Config.h"
#include "Double_2D.h"
#include "FresnelCDI.h"
#include "FresnelCDI_WF.h"
#include "PlanarCDI.h"
#include "io.h"

using std::cout;
using std::endl;
using std::string;

namespace {

enum { OUTPUT_MINIMAL = 0, OUTPUT_ITER = 1, OUTPUT_ERROR = 2 };

const string kPlanar = "planar", kFresnel = "fresnel", kFresnelWF = "fresnel_wf";

void print_usage() {
  cout << "Usage: " << endl
       << endl
       << "cdi_reconstruction <config filename> <reco_type> <seed>" << endl
       << endl
       << "where <reco_type> may be: " << kPlanar << ", " << kFresnel << " or " << kFresnelWF << endl
       << "<seed> should be an integer" << endl
       << endl
       << "If <reco_type> and <seed> do not need to be specified."
       << "if they are not, <reco_type> will default to " << kPlanar << " and <seed> to 0." << endl;
}

string image_name(const string& prefix, const string& what, int i, const string& type) {
  std::ostringstream s;
  s << prefix << what << "_" << i << "." << type;
  return s.str();
}

}  // namespace

int main(int argc, char* argv[]) {
  if (argc == 1) {
    cout << endl << "No config file given, ";
    print_usage();
    return 0;
  }
  if (argc > 4) {
    cout << endl << "Wrong number of arguments given. ";
    print_usage();
    return 0;
  }
  const string config_file = argv[1];
  string reco_type = kPlanar;
  int seed = 0;
  if (argc <= 2)
    cout << "No reconstruction type given... using planar" << endl;
  else
    reco_type = argv[2];
  if (argc <= 3)
    cout << "No seed value given... using the default (0)" << endl;
  else
    seed = std::atoi(argv[3]);

  Config c(config_file);
  if (c.getStatus() == FAILURE) {
    cout << "Could not open the file " << config_file << ". Exiting" << endl;
    return 0;
  }

  // required keys
  const int output_iterations = (int)c.getDouble("output_iterations");
  std::list<string>* algorithms = c.getStringList("algorithm");
  std::list<int>* iterations = c.getIntList("iterations");
  const int nx = c.getInt("pixels_x");
  const int ny = c.getInt("pixels_y");
  if (c.getStatus() == FAILURE) {
    cout << "Problem reading the configuration file. Exiting" << endl;
    return 0;
  }
  if (algorithms->size() != iterations->size()) {
    cout << "The number of algorithms and the number of iterations do not match. Exiting" << endl;
    return 0;
  }

  // optional keys (a missing one reads as 0 / "")
  const int output_level = c.getInt("info_level");
  const int output_diffraction_estimate = c.getInt("output_diffraction_estimate");
  const int log_diffraction = c.getInt("use_log_scale_for_diffraction");
  const int log_object = c.getInt("use_log_scale_for_object");
  int shrinkwrap_iterations = c.getInt("shrinkwrap_iterations");
  const double shrinkwrap_gauss_width = c.getDouble("shrinkwrap_gauss_width");
  const double shrinkwrap_threshold = c.getDouble("shrinkwrap_threshold");
  const string output_file_type = c.getString("output_file_type");

  Complex_2D object_estimate(nx, ny);
  const string starting_point = c.getString("starting_point_file_name");
  if (!starting_point.empty() && !read_cplx(starting_point, object_estimate)) {
    cout << "Can not process the file " << starting_point << ".. exiting" << endl;
    return 1;
  }

  string data_file_name = c.getString("data_file_name");
  string support_file_name = c.getString("support_file_name");
  const string prefix = c.getString("output_file_name_prefix");
  string output_file_name = prefix + ".cplx";

  BaseCDI* proj = 0;
  Complex_2D* white_field = 0;
  if (reco_type == kPlanar) {
    proj = new PlanarCDI(object_estimate);
  } else if (reco_type == kFresnel) {
    const double wavelength = c.getDouble("beam_wavelength");
    const double z_fd = c.getDouble("focal_detector_length");
    const double z_fs = c.getDouble("focal_sample_length");
    const double pixel = c.getDouble("pixel_size");
    const double norm = c.getDouble("normalisation");
    white_field = new Complex_2D(nx, ny);
    if (!read_cplx(c.getString("white_field_reco_file_name"), *white_field)) {
      cout << "failed to read white-field data .. exiting" << endl;
      return 1;
    }
    cout << white_field->get_value(0, 0, MAG_SQ) << endl;
    proj = new FresnelCDI(object_estimate, *white_field, wavelength, z_fd, z_fs, pixel, norm);
  } else if (reco_type == kFresnelWF) {
    const double wavelength = c.getDouble("beam_wavelength");
    const double z_zone = c.getDouble("zone_focal_length");
    const double z_fd = c.getDouble("focal_detector_length");
    const double pixel = c.getDouble("pixel_size");
    proj = new FresnelCDI_WF(object_estimate, wavelength, z_zone, z_fd, pixel);
    output_file_name = c.getString("white_field_reco_file_name");
    data_file_name = c.getString("white_field_data_file_name");
    support_file_name = c.getString("white_field_support_file_name");
    // one fixed algorithm; the name only keeps the schedule's shape
    algorithms->assign(1, "ER");
    iterations->assign(1, c.getInt("wf_iterations"));
    shrinkwrap_iterations = 0;
  } else {
    cout << "the reconstruction type specified (" << reco_type << ") is unrecognised. Please use: " << kPlanar
         << ", " << kFresnel << " or " << kFresnelWF
         << " (the partially coherent types are not part of this library)" << endl;
    return 1;
  }

  Double_2D data;
  read_image(data_file_name, data, nx, ny);  // exits when the file cannot be read
  if (nx != data.get_size_x() || ny != data.get_size_y()) {
    cout << "Dimensions of the data to not match those given ... exiting" << endl;
    return 1;
  }
  Double_2D support;
  read_image(support_file_name, support, nx, ny);
  if (nx != support.get_size_x() || ny != support.get_size_y()) {
    cout << "Dimensions of the support to not match ... exiting" << endl;
    return 1;
  }
  proj->set_support(support);
  proj->set_intensity(data);
  if (starting_point.empty()) proj->initialise_estimate(seed);

  Double_2D result(nx, ny);
  int i = 0, cumulative = 0;
  std::list<int>::const_iterator n_it = iterations->begin();
  for (std::list<string>::const_iterator a_it = algorithms->begin(); a_it != algorithms->end(); ++a_it, ++n_it) {
    if (output_level != OUTPUT_MINIMAL) cout << "Switching to the " << *a_it << " algorithm" << endl;
    const int alg = PlanarCDI::getAlgFromName(*a_it);
    if (alg == -1) {
      cout << "Could not find reconstruction algorithm with the name " << *a_it << ". Exiting" << endl;
      return 0;
    }
    proj->set_algorithm(alg);
    cumulative += *n_it;
    for (; i < cumulative; i++) {
      if (output_level != OUTPUT_MINIMAL) cout << "Iteration " << i << endl;
      proj->iterate();
      // get_error() belongs to the estimate the iteration started from
      if (output_level == OUTPUT_ERROR && i > 0)
        cout << "Error for iteration " << (i - 1) << " is " << proj->get_error() << endl;
      if (output_iterations > 0 && i % output_iterations == 0) {
        object_estimate.get_2d(MAG, result);
        write_image(image_name(prefix, "", i, output_file_type), result, log_object);
        if (output_diffraction_estimate) {
          Complex_2D* temp = object_estimate.clone();
          proj->propagate_to_detector(*temp);
          temp->get_2d(MAG_SQ, result);
          write_image(image_name(prefix, "_diffraction", i, output_file_type), result, log_diffraction);
          delete temp;
        }
      }
      if (shrinkwrap_iterations != 0 && i % shrinkwrap_iterations == shrinkwrap_iterations - 1) {
        proj->apply_shrinkwrap(shrinkwrap_gauss_width, shrinkwrap_threshold);
        cout << "Applying shrink-wrap at iteration " << i << endl;
      }
    }
  }

  write_cplx(output_file_name, object_estimate);
  if (reco_type == kFresnel) {
    static_cast<FresnelCDI*>(proj)->get_transmission_function(object_estimate);
    write_cplx("transmission_function.cplx", object_estimate);
  }

  delete algorithms;
  delete iterations;
  delete proj;
  delete white_field;
  return 0;
}
