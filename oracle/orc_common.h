// orc_common.h -- pieces of the CPU oracle shared by the E8Q restatement
// (e8q_oracle.cpp) and the ME8Q restatement (me8q_oracle.cpp).  TEST
// INFRASTRUCTURE ONLY: nothing in the product includes or links it.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <random>
#include <vector>

namespace orc {

constexpr int kMaxBits = 13;

// AVX horizontal sum order (simd_helpers.hpp:12-18).
inline float Hsum8(const float v[8]) {
    const float s0 = v[0] + v[4], s1 = v[1] + v[5], s2 = v[2] + v[6], s3 = v[3] + v[7];
    return (s0 + s2) + (s1 + s3);
}

// FHT-Kac-4 rotator (rotator.cpp:13-80, kernels_generic.cpp:49-78)
struct Rotator {
    size_t dim = 0, padded = 0, trunc = 0;
    float fac = 0.0F;
    std::vector<uint8_t> signs;

    Rotator() = default;
    Rotator(size_t d, uint64_t seed) : dim(d) {
        padded = (d + 63) / 64 * 64;
        size_t lg = 0;
        while ((size_t{1} << (lg + 1)) <= d) ++lg;
        trunc = size_t{1} << lg;
        fac = 1.0F / std::sqrt(static_cast<float>(trunc));
        signs.resize(4 * padded / 8);
        std::mt19937_64 gen(seed);
        std::uniform_int_distribution<int> dist(0, 255);
        for (auto& s : signs) s = static_cast<uint8_t>(dist(gen));
    }

    static void Fht(float* x, size_t len) {
        for (size_t h = 1; h < len; h <<= 1)
            for (size_t i = 0; i < len; i += 2 * h)
                for (size_t j = i; j < i + h; ++j) {
                    const float a = x[j], b = x[j + h];
                    x[j] = a + b;
                    x[j + h] = a - b;
                }
    }

    void Rotate(const float* src, float* dst) const {
        for (size_t i = 0; i < padded; ++i) dst[i] = i < dim ? src[i] : 0.0F;
        const size_t per_round = padded / 8;
        for (int r = 0; r < 4; ++r) {
            const uint8_t* sg = signs.data() + r * per_round;
            for (size_t i = 0; i < padded; ++i)
                if ((sg[i >> 3] >> (i & 7)) & 1) dst[i] = -dst[i];
            float* win = dst + ((trunc == padded || r % 2 == 0) ? 0 : padded - trunc);
            Fht(win, trunc);
            for (size_t i = 0; i < trunc; ++i) win[i] *= fac;
            if (trunc != padded) {
                const size_t half = padded / 2;
                for (size_t i = 0; i < half; ++i) {
                    const float a = dst[i], b = dst[i + half];
                    dst[i] = a + b;
                    dst[i + half] = a - b;
                }
            }
        }
        if (trunc != padded)
            for (size_t i = 0; i < padded; ++i) dst[i] *= 0.25F;
    }
};

double FixedScale(int bits);
std::vector<double> BranchLevels(int bits, int branch);
// Whole-vector encode at a fixed scale (EncodeVectorAtScaleInto): `padded`
// coordinates, scaled by s * sqrt(padded).
void EncodeVector(const float* unit, int padded, int bits, double s, uint16_t* digits, uint8_t* branch);
float OrthoQuantile(double alpha, size_t k_eff);
float QueryNorm2(const float* q, size_t P);
float RouteDot(const float* q, const float* c, size_t P);
void CoarseAccumulate(const uint8_t* rowa, const uint8_t* lut, size_t blocks, int32_t out[32]);

}  // namespace orc
