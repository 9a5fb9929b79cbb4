// TEST INFRASTRUCTURE — the CPU oracle (FP64 OpenMP restatement of the reference's CPU backend). Loaded only by
// tests/, the smoke check and bench.py's CPU-baseline leg, as the checker; the product never uses it.
#pragma once

#include <memory>

#include "../cpp-spo-dvr_amd/host/backend.hpp"

namespace oracle {

std::unique_ptr<spo::PropagationBackend> make_oracle_backend(const spo::BackendInput& input);

}  // namespace oracle
