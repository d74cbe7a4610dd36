// e8q_host_math.cpp -- see e8q_host_math.h.  Compiled with
// -ffp-contract=off: the fused multiply-adds the reference's GCC build
// forms (Horner steps of Phi^-1, the branch offset, Newton's Phi - p) are
// written out as std::fma so the tables are bit-identical to it.
#include "e8q_host_math.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <random>

namespace e8qlib {
namespace hostmath {
namespace {

double Horner(const double* c, int n, double x) {
    double r = c[0];
    for (int i = 1; i < n; ++i) r = std::fma(r, x, c[i]);
    return r;
}

// Acklam's rational approximation of Phi^-1 (|rel err| < 1.15e-9).
double InvNormalCdf(double p) {
    static const double a[6] = {-3.969683028665376e+01, 2.209460984245205e+02, -2.759285104469687e+02,
                                1.383577518672690e+02,  -3.066479806614716e+01, 2.506628277459239e+00};
    static const double b[6] = {-5.447609879822406e+01, 1.615858368580409e+02, -1.556989798598866e+02,
                                6.680131188771972e+01,  -1.328068155288572e+01, 1.0};
    static const double c[6] = {-7.784894002430293e-03, -3.223964580411365e-01, -2.400758277161838e+00,
                                -2.549732539343734e+00, 4.374664141464968e+00,  2.938163982698783e+00};
    static const double d[5] = {7.784695709041462e-03, 3.224671290700398e-01, 2.445134137142996e+00,
                                3.754408661907416e+00, 1.0};
    const double lo = 0.02425;
    if (p < lo) {
        const double q = std::sqrt(-2 * std::log(std::max(p, 1e-300)));
        return Horner(c, 6, q) / Horner(d, 5, q);
    }
    if (p <= 1 - lo) {
        const double q = p - 0.5;
        const double r = q * q;
        return Horner(a, 6, r) * q / Horner(b, 6, r);
    }
    const double q = std::sqrt(-2 * std::log(std::max(1 - p, 1e-300)));
    return -(Horner(c, 6, q) / Horner(d, 5, q));
}

double PhiInv(double p) {
    if (p <= 0.0) return -8.0;
    if (p >= 1.0) return 8.0;
    double x = std::sqrt(-2.0 * std::log(std::min(p, 1.0 - p)));
    if (p < 0.5) x = -x;
    for (int it = 0; it < 12; ++it) {
        const double phi = std::exp(x * (x * -0.5)) * 0.3989422804014327;
        x -= std::fma(std::erf(x * 0.7071067811865476) + 1.0, 0.5, -p) / phi;
    }
    return x;
}

}  // namespace

double FixedScale(int bits) {
    static const std::array<double, kMaxQuantizeBits + 1> sched = [] {
        std::array<double, kMaxQuantizeBits + 1> s = {0.0,   1.118034, 0.895, 0.825, 0.785, 0.755, 0.735,
                                                      0.735, 0.735,    0.725, 0.715, 0.705, 0.710, 0.705};
        if (const char* env = std::getenv("E8Q_SCALE_SCHEDULE")) {
            const char* p = env;
            for (int i = 1; i <= kMaxQuantizeBits && *p != '\0'; ++i) {
                char* end = nullptr;
                const double v = std::strtod(p, &end);
                if (end == p) break;
                s[static_cast<size_t>(i)] = v;
                p = end;
                while (*p == ',' || *p == ' ') ++p;
            }
        }
        return s;
    }();
    return sched[static_cast<size_t>(std::clamp(bits, 1, kMaxQuantizeBits))];
}

std::vector<double> BranchLevels(int bits, int branch) {
    const int q = 1 << bits;
    const double eta = std::fma(0.5, static_cast<double>(branch), 0.25);
    const int t0 = static_cast<int>(std::ceil(std::fma(-0.5, static_cast<double>(q), -eta)));
    std::vector<double> out(static_cast<size_t>(q));
    for (int i = 0; i < q; ++i) {
        const long t = t0 + i;
        const double p = 0.5 + (static_cast<double>(t) + eta) / static_cast<double>(q);
        out[static_cast<size_t>(((t % q) + q) % q)] =
            kSigmaC * InvNormalCdf(std::min(1.0 - 1e-12, std::max(1e-12, p)));
    }
    return out;
}

void ScanClassMeans(int bits, double w_cls[2]) {
    const auto l0 = BranchLevels(bits, 0), l1 = BranchLevels(bits, 1);
    const int cv = 1 << (bits - 1);
    for (int t = 0; t < 2; ++t) {
        double s0 = 0.0, s1 = 0.0;
        for (int lo = 0; lo < cv; ++lo) {
            s0 += l0[static_cast<size_t>((t << (bits - 1)) | lo)];
            s1 += l1[static_cast<size_t>((t << (bits - 1)) | lo)];
        }
        w_cls[t] = 0.5 * (s0 / cv + s1 / cv);
    }
}

std::vector<float> RefineTables(int bits, size_t padded_dim) {
    const size_t q = size_t{1} << bits, stride = std::max<size_t>(256, q);
    std::vector<float> tab(2 * stride, 0.0F);
    const double rd = std::sqrt(static_cast<double>(padded_dim));
    for (int br = 0; br < 2; ++br) {
        const auto lev = BranchLevels(bits, br);
        for (size_t d = 0; d < q; ++d) tab[br * stride + d] = static_cast<float>(lev[d] / rd);
    }
    return tab;
}

void GateMid(int bits, double out[16]) {
    double w[2];
    ScanClassMeans(bits, w);
    for (int d = 0; d < 8; ++d) {
        out[2 * d] = (bits == 1 && d == 7) ? 0.0 : w[0];
        out[2 * d + 1] = (bits == 1 && d == 7) ? 0.0 : w[1];
    }
}

size_t PaddedDim(size_t dim) { return (dim + 63) / 64 * 64; }

std::vector<uint8_t> RotatorSigns(size_t dim, uint64_t seed) {
    std::vector<uint8_t> s(4 * PaddedDim(dim) / 8);
    std::mt19937_64 gen(seed);
    std::uniform_int_distribution<int> dist(0, 255);
    for (auto& v : s) v = static_cast<uint8_t>(dist(gen));
    return s;
}

float OrthoQuantile(double alpha, size_t k_eff) {
    const double tail = std::min(alpha / static_cast<double>(std::max<size_t>(k_eff, 1)), 0.49);
    return static_cast<float>(std::max(PhiInv(1.0 - tail), 1.0));
}

int FlushEvery() {
    const char* e = std::getenv("E8Q_FLUSH_EVERY");
    const long v = e != nullptr ? std::strtol(e, nullptr, 10) : 0;
    return v > 0 ? static_cast<int>(v) : 32;
}

}  // namespace hostmath
}  // namespace e8qlib
