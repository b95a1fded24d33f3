// Algorithm parameters (the reference's include/cc_alg_configuration.h:1-38,
// src/cc_alg_configuration.cpp:1-46).
#pragma once
#include <iostream>
#include <string>

class CCAlgConfiguration {
  std::string _disk_dir = ".";
  double _sketches_factor = 1;  // multiplies the number of samples (columns) per sketch
  double _batch_factor = 1;     // update batch size relative to one sketch's bytes

 public:
  CCAlgConfiguration() {}
  CCAlgConfiguration& disk_dir(std::string disk_dir);
  CCAlgConfiguration& sketches_factor(double factor);
  CCAlgConfiguration& batch_factor(double factor);

  std::string get_disk_dir() const { return _disk_dir; }
  double get_sketches_factor() const { return _sketches_factor; }
  double get_batch_factor() const { return _batch_factor; }

  friend std::ostream& operator<<(std::ostream& out, const CCAlgConfiguration& conf);
  CCAlgConfiguration(const CCAlgConfiguration&) = default;
  CCAlgConfiguration(CCAlgConfiguration&&) = default;
  CCAlgConfiguration& operator=(const CCAlgConfiguration&) = default;
};
