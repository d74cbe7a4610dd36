// e8q_host_math.h -- host-side tables of the E8Q codec and the rotator:
// everything E8qIndex computes once per index on the CPU before handing the
// device its tables.  Reference: e8q_codec.cpp:17-44 (Acklam Phi^-1),
// 427-524 (scale schedule, level tables, class means), 933-946 (refine
// tables), rotator.cpp:13-28 (sign bytes), e8q_internal.hpp:77-105 (lambda).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace e8qlib {
namespace hostmath {

constexpr int kMaxQuantizeBits = 13;
constexpr double kSigmaC = 1.1180339887498949;  // sqrt(5/4)

double FixedScale(int bits);
std::vector<double> BranchLevels(int bits, int branch);
void ScanClassMeans(int bits, double w_cls[2]);
std::vector<float> RefineTables(int bits, size_t padded_dim);
// gate_mid[2*d + t]: class mean of top-bit class t priced at dim d of a
// block; dim 7 is zero at b == 1 (that scan bit is the branch).
void GateMid(int bits, double out[16]);
// FHT-Kac-4 sign bytes: 4 rounds x padded/8 bytes from mt19937_64(seed).
std::vector<uint8_t> RotatorSigns(size_t dim, uint64_t seed);
size_t PaddedDim(size_t dim);
// lambda = max(Phi^-1(1 - min(alpha/k_eff, 0.49)), 1).
float OrthoQuantile(double alpha, size_t k_eff);
// E8Q_FLUSH_EVERY or 32.
int FlushEvery();

}  // namespace hostmath
}  // namespace e8qlib
