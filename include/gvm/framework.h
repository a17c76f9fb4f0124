// Process-wide state of the imaging objective.  The reference passes these
// through ~40 extern globals (src/mfs.cu:4-38, src/functions.cu:37-71,
// src/frprmn.cu:46); here they live in one struct that the synthesizer-side
// code fills before composing the objective.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "gvm_hip.h"

namespace gvm {

// Raised by the C++ layer when a C-ABI call fails (the reference prints and
// exits instead).
class Error : public std::runtime_error {
 public:
  Error(const std::string& what, int status) : std::runtime_error(what), status(status) {}
  int status;
};

// Throws gvm::Error if st is not GVM_OK.
void check(gvm_status st, const char* what);

struct Globals {
  long M = 0, N = 0;
  int image_count = 2;
  std::vector<float> penalizators;    // -Z lambdas, indexed by penalizatorIndex
  int flag_opt = 0;                   // optimisation schedule flag
  float eta = -1.0f;                  // -e
  std::vector<float> initial_values;  // [0] is MINPIX
  bool nopositivity = false;          // -x
  bool verbose = false;
  gvm_ctx* ctx = nullptr;             // current device context
};

Globals& globals();

}  // namespace gvm
