// Driver parameters (the reference's include/driver_configuration.h:1-48 and the
// GutteringConfiguration of its guttering dependency). Every GutterSystem value selects the
// in-memory per-vertex gutters of this library (no on-disk gutter tree).
#pragma once
#include <cstddef>
#include <iostream>
#include <string>

enum GutterSystem { GUTTERTREE, STANDALONE, CACHETREE };

class GutteringConfiguration {
  size_t _gutter_bytes = uninit_param;
  size_t _queue_factor = 8;      // work queue holds queue_factor * workers elements
  size_t _wq_batch_per_elm = 1;  // vertex batches handed to a worker at once
  size_t _buffer_exp = 20;
  size_t _page_factor = 1;

 public:
  static constexpr size_t uninit_param = (size_t)-1;
  GutteringConfiguration& gutter_bytes(size_t b) { _gutter_bytes = b; return *this; }
  GutteringConfiguration& queue_factor(size_t f) { _queue_factor = f ? f : 1; return *this; }
  GutteringConfiguration& wq_batch_per_elm(size_t b) { _wq_batch_per_elm = b ? b : 1; return *this; }
  GutteringConfiguration& buffer_exp(size_t e) { _buffer_exp = e; return *this; }
  GutteringConfiguration& page_factor(size_t f) { _page_factor = f; return *this; }
  size_t get_gutter_bytes() const { return _gutter_bytes; }
  size_t get_queue_factor() const { return _queue_factor; }
  size_t get_wq_batch_per_elm() const { return _wq_batch_per_elm; }
  size_t get_buffer_exp() const { return _buffer_exp; }
  size_t get_page_factor() const { return _page_factor; }
};

class DriverConfiguration {
  GutterSystem _gutter_sys = STANDALONE;
  std::string _disk_dir = ".";
  size_t _num_worker_threads = 1;
  GutteringConfiguration _gutter_conf;

 public:
  DriverConfiguration() {}
  DriverConfiguration& gutter_sys(GutterSystem gutter_sys);
  DriverConfiguration& disk_dir(std::string disk_dir);
  DriverConfiguration& worker_threads(size_t num_workers);
  GutteringConfiguration& gutter_conf() { return _gutter_conf; }

  GutterSystem get_gutter_sys() const { return _gutter_sys; }
  std::string get_disk_dir() const { return _disk_dir; }
  size_t get_worker_threads() const { return _num_worker_threads; }

  friend std::ostream& operator<<(std::ostream& out, const DriverConfiguration& conf);
  DriverConfiguration& operator=(const DriverConfiguration&) = delete;
  DriverConfiguration(const DriverConfiguration&) = default;
  DriverConfiguration(DriverConfiguration&&) = default;
};
