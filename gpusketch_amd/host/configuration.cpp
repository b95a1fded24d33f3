// Configuration setters and printers (src/cc_alg_configuration.cpp:1-46,
// src/driver_configuration.cpp:1-40 of the reference): out-of-range values fall back to the
// default with a message, as there.
#include "gpusketch/cc_alg_configuration.h"
#include "gpusketch/driver_configuration.h"

CCAlgConfiguration& CCAlgConfiguration::disk_dir(std::string d) {
  _disk_dir = std::move(d);
  return *this;
}

CCAlgConfiguration& CCAlgConfiguration::sketches_factor(double f) {
  if (!(f > 0)) {
    std::cout << "sketches_factor=" << f << " is out of bounds. (0, infty)Defaulting to 1."
              << std::endl;
    f = 1;
  }
  _sketches_factor = f;
  return *this;
}

CCAlgConfiguration& CCAlgConfiguration::batch_factor(double f) {
  if (!(f > 0)) {
    std::cout << "batch factor=" << f << " is out of bounds. (0, infty)Defaulting to 1."
              << std::endl;
    f = 1;
  }
  _batch_factor = f;
  return *this;
}

std::ostream& operator<<(std::ostream& out, const CCAlgConfiguration& conf) {
  out << "Connected Components Algorithm Configuration:" << std::endl;
#ifdef L0_SAMPLING
  out << " Sketching algorithm   = CubeSketch" << std::endl;
#else
  out << " Sketching algorithm   = CameoSketch" << std::endl;
#endif
  out << " Using Eager DSU       = True" << std::endl;
  out << " Num sketches factor   = " << conf._sketches_factor << std::endl;
  out << " Batch size factor     = " << conf._batch_factor << std::endl;
  out << " On disk data location = " << conf._disk_dir;
  return out;
}

DriverConfiguration& DriverConfiguration::gutter_sys(GutterSystem g) {
  _gutter_sys = g;
  return *this;
}

DriverConfiguration& DriverConfiguration::disk_dir(std::string d) {
  _disk_dir = std::move(d);
  return *this;
}

DriverConfiguration& DriverConfiguration::worker_threads(size_t n) {
  if (n < 1) {
    std::cout << "num_worker_threads=" << n << " is out of bounds. [1, infty)Defaulting to 1."
              << std::endl;
    n = 1;
  }
  _num_worker_threads = n;
  return *this;
}

std::ostream& operator<<(std::ostream& out, const DriverConfiguration& conf) {
  out << "GraphSketchDriver Configuration:" << std::endl;
  out << " Guttering system      = in-memory vertex gutters" << std::endl;
  out << " Worker thread count   = " << conf._num_worker_threads << std::endl;
  out << " On disk data location = " << conf._disk_dir;
  return out;
}
