// Cone kinds; the numeric values are part of the C ABI (include/mumblebee_hip.h).
#pragma once

namespace mb {
constexpr int CONE_POS = 0, CONE_FREE = 1, CONE_SOC = 2, CONE_PSD = 3, CONE_EXP = 4, CONE_POW = 5;
}
