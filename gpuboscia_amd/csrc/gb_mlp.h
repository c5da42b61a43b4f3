// Learned-branching MLP: sizes and parameter offsets (floats).
#pragma once

#include "../../include/gpuboscia_amd.h"

namespace gb {

constexpr int kMlpFeatures = 8;
constexpr int kMlpHidden = 64;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpFeatures * kMlpHidden;
constexpr int kOffW2 = kOffB1 + kMlpHidden;
constexpr int kOffB2 = kOffW2 + kMlpHidden * kMlpHidden;
constexpr int kOffW3 = kOffB2 + kMlpHidden;
constexpr int kOffB3 = kOffW3 + kMlpHidden * kMlpHidden;
constexpr int kOffW4 = kOffB3 + kMlpHidden;
constexpr int kOffB4 = kOffW4 + kMlpHidden;
constexpr int kMlpParams = kOffB4 + 1;
static_assert(kMlpParams == GB_MLP_PARAMS, "parameter count mismatch");

}  // namespace gb
