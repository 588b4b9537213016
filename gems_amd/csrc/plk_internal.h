// Internal declarations shared by the libplk translation units.
#pragma once

#include <cstdint>
#include <string>

#define PLK_MAX_STATES 64
#define PLK_MAX_CATS 16

namespace plk {

// Record a thread-local error message and return the code (plk_last_error reads it).
int fail(int code, const std::string &msg);
void clear_error();

// tree/phylotree.h:74-80
constexpr int kScalingThresholdExp = 256;
constexpr double kScalingThreshold = 8.636168555094444625386e-78;  // 2^-256
constexpr double kLogScalingThreshold = -177.4456782233459932741;  // log(2^-256)
constexpr double kLn2 = 0.69314718055994530941723212145818;

// NCBI standard genetic code over codons AAA..TTT (A,C,G,T order), '*' = stop.
extern const char kStandardGeneticCode[65];

}  // namespace plk
