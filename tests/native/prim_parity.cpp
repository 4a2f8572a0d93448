// Host build of the kernels' __host__ __device__ primitives (ugx_dev.hpp)
// compared bit for bit with the oracle's (orc_prim.h).  Prints the number of
// mismatching results.
#include <cstdio>
#include <cstring>
#include <random>

#include "ugx_dev.hpp"
extern "C" {
#include "orc_prim.h"
}

static bool same(double a, double b) { return std::memcmp(&a, &b, 8) == 0; }

int main()
{
    std::mt19937_64 g(12345);
    std::uniform_real_distribution<double> u01(0.0, 1.0), uexp(-700.0, 700.0), ulog(-300.0, 300.0);
    long bad = 0, n = 0;
    for (int i = 0; i < 400000; ++i) {
        double x = (i & 1) ? u01(g) : std::pow(10.0, ulog(g));
        if (x <= 0) x = 1e-300;
        bad += !same(ugx::dlog(x), orc_log(x)); ++n;
        double y = uexp(g);
        bad += !same(ugx::dexp(y), orc_exp(y)); ++n;
        double s1, c1, s2, c2;
        double w = u01(g);
        ugx::dsincos2pi(w, s1, c1);
        orc_sincos2pi(w, &s2, &c2);
        bad += !same(s1, s2) + !same(c1, c2); n += 2;
        double p = u01(g) * 10.0 + 1e-9, q = u01(g) * 6.0 - 3.0;
        bad += !same(ugx::dpow(p, q), orc_pow(p, q)); ++n;
    }
    // streams
    for (uint32_t id = 0; id < 2000; ++id) {
        ugx::Rng a;
        orc_rng b;
        ugx::rng_init(a, 0x9E3779B9u, 0x12345u, 3u, id & 7u, id, 17u);
        orc_rng_init(&b, 0x000123459E3779B9ull, 3u, id & 7u, id, 17u);
        for (int k = 0; k < 50; ++k) {
            bad += !same(ugx::uniform(a), orc_uniform(&b)); ++n;
            bad += ugx::uindex(a, 1000u + id) != orc_index(&b, 1000u + id); ++n;
            double g0, g1, h0, h1;
            ugx::gauss2(a, g0, g1);
            orc_gauss2(&b, &h0, &h1);
            bad += !same(g0, h0) + !same(g1, h1); n += 2;
        }
        double e1 = ugx::equip_erot(a, 300.0, 3), e2 = 0.0;
        (void)e2;
    }
    std::printf("checked %ld mismatches %ld\n", n, bad);
    return 0;
}
