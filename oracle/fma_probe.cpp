// fma_probe.cpp -- which product GCC fuses in `a*b + c*d + ...` under the
// reference's own flags (-O3 -march=x86-64-v3, C++17; contraction on).  The
// reference's fixtures cannot tell the orders apart where a rounding
// difference is absorbed by a later u8 quantization (the ME8Q class-mean
// LUT, me8q_fast.cpp:651-658), so the order the oracle restates is pinned
// here instead: the expression shapes are compiled exactly as the reference
// writes them and compared with both explicit fused forms on random inputs.
// TEST INFRASTRUCTURE ONLY (tests/test_oracle_me8q.py builds and runs it).
#include <cmath>
#include <cstdio>
#include <random>

__attribute__((noinline)) float Lut4(const float* qb, float a4, float a5, float a6, float a7) {
    return qb[4] * a4 + qb[5] * a5 + qb[6] * a6 + qb[7] * a7;
}
__attribute__((noinline)) double Rx2(double coef, double n2e, double wx, double n2w, double dew) {
    return 4.0 * coef * coef * n2e + wx * wx * n2w - 4.0 * coef * wx * dew;
}

int main() {
    std::mt19937 g(1);
    std::uniform_real_distribution<float> u(-2.0F, 2.0F);
    const int N = 200000;
    int lut_first = 0, lut_second = 0, rx_first = 0, rx_second = 0;
    for (int it = 0; it < N; ++it) {
        float qb[8];
        for (float& x : qb) x = u(g);
        const float a4 = u(g), a5 = u(g), a6 = u(g), a7 = u(g);
        const float r = Lut4(qb, a4, a5, a6, a7);
        const float p5 = qb[5] * a5, p4 = qb[4] * a4;
        lut_first += r == std::fma(qb[7], a7, std::fma(qb[6], a6, std::fma(qb[4], a4, p5)));
        lut_second += r == std::fma(qb[7], a7, std::fma(qb[6], a6, std::fma(qb[5], a5, p4)));
        const double c = u(g), n2e = u(g), wx = u(g), n2w = u(g), dew = u(g);
        const double rr = Rx2(c, n2e, wx, n2w, dew);
        const double c4 = 4.0 * c, ww = wx * wx, t1 = (4.0 * c * c) * n2e;
        const double t2 = ww * n2w;
        rx_first += rr == std::fma(-(c4 * wx), dew, std::fma(c4 * c, n2e, t2));
        rx_second += rr == std::fma(-(c4 * wx), dew, std::fma(ww, n2w, t1));
    }
    std::printf("%d %d %d %d %d\n", N, lut_first, lut_second, rx_first, rx_second);
    return 0;
}
