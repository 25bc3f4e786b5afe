// TransmissionConstraint.h -- complex (transmission) constraints of the NADIA
// drop-in API.
//
// Interface of NADIA's ComplexConstraint and TransmissionConstraint
// (NADIA/include/TransmissionConstraint.h, NADIA/src/TransmissionConstraint.c++).
// The constraint itself runs on the device (nadia_constraint,
// include/nadia_amd.h): a TransmissionConstraint creates its device object
// when the array size is first known (first region added, first application
// or attachment to a reconstruction) and keeps the settings made before in
// host members until then.  A ComplexConstraint added to it reads its region
// mean c from the device.
#ifndef NADIA_AMD_TRANS_H
#define NADIA_AMD_TRANS_H

#include <cstdlib>
#include <iostream>
#include <vector>

#include "Complex_2D.h"
#include "Double_2D.h"

class TransmissionConstraint;

class ComplexConstraint {
 protected:
  double alpha1;
  double alpha2;
  Double_2D* region;
  double c_mean;
  bool fixed_c;
  TransmissionConstraint* owner;  // set by add_complex_constraint
  int index;

 public:
  ComplexConstraint(Double_2D& region, double alpha1, double alpha2)
      : alpha1(alpha1), alpha2(alpha2), region(&region), c_mean(0), fixed_c(false), owner(0), index(-1) {}
  virtual ~ComplexConstraint() {}

  void set_fixed_c(double c_mean);
  void set_c_mean(double c_mean) {
    if (!fixed_c) this->c_mean = c_mean;
  }
  // the value of the last application (read from the device once added)
  double get_c_mean();
  void set_alpha1(double alpha1);
  void set_alpha2(double alpha2);
  double get_new_mag(double old_mag, double old_phase) {
    return exp((1 - alpha1) * log(old_mag) + alpha1 * c_mean * old_phase);
  }
  double get_new_phase(double old_mag, double old_phase) {
    return (1 - alpha2) * old_phase + alpha2 * log(old_mag) / c_mean;
  }
  Double_2D* get_region() { return region; }

  friend class TransmissionConstraint;
};

class TransmissionConstraint {
 protected:
  std::vector<ComplexConstraint*> complex_constraint_list;
  bool do_enforce_unity;
  bool do_charge_flip;
  int flip_sign;
  void (*custom_constraint)(Complex_2D&);
  nadia_constraint* h;
  int nx, ny;

  static void custom_trampoline(void* data, int nx, int ny, int precision, void* user);
  void push_settings();

 public:
  TransmissionConstraint();
  virtual ~TransmissionConstraint();

  void delete_complex_constraint_regions();
  virtual void add_complex_constraint(ComplexConstraint& new_constraint);
  void set_charge_flipping(bool enable, int flip_sign = 1);
  void set_enforce_unity(bool enable);
  void set_custom_constraint(void (*custom_constraint)(Complex_2D& transmission));
  virtual void apply_constraint(Complex_2D& transmission);

  // the device object for arrays of nx x ny (created on first use)
  nadia_constraint* device_handle(int nx, int ny);

  friend class ComplexConstraint;
};

#endif
