// Host check of lue_amd::pow_positive (the block Newton's FP64 power), full and short, against long-double
// powl. Prints the worst relative error of each over random and edge-case arguments; exit status 0.
#include "kinematic_wave.hpp"

#include <cmath>
#include <cstdio>
#include <random>

int main()
{
    std::mt19937_64 rng(20261203);
    std::uniform_real_distribution<double> log_x(-103.0, 88.0);  // FP32 normal and subnormal range
    std::uniform_real_distribution<double> exponent(-1.0, 1.5);
    double worst = 0.0;
    double worst_short = 0.0;
    auto check = [&](double x, double y) {
        long double const want = powl(static_cast<long double>(x), static_cast<long double>(y));
        double const got = lue_amd::pow_positive(x, y);
        double const err = static_cast<double>(fabsl((static_cast<long double>(got) - want) / want));
        if (!(err <= worst))
            worst = err;
        double const got_short = lue_amd::pow_positive<true, true>(x, y);
        double const err_short = static_cast<double>(fabsl((static_cast<long double>(got_short) - want) / want));
        if (!(err_short <= worst_short))
            worst_short = err_short;
    };
    for (int i = 0; i < 2000000; ++i)
        check(std::exp(log_x(rng)), exponent(rng));
    double const edges[] = {1e-45, 1.17549435e-38, 0.5, 0.7071067811865475, 0.7071067811865476, 1.0, 1.0 + 0x1p-52,
                            2.0, 3.0, 1e4, 3.4e38};
    for (double x : edges)
        for (double y : {-1.0, -0.4, 0.0, 0.6, 1.0, 1.5})
            check(x, y);
    std::printf("worst_relative_error_short %.6e\n", worst_short);
    std::printf("worst_relative_error %.6e\n", worst);
    return 0;
}
