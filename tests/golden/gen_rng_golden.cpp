// gen_rng_golden.cpp -- prints libstdc++ <random> draws as JSON: the streams the reference's
// initialisation consumes (std::mt19937_64 + uniform_real_distribution<double>(0,1) seeded 0
// and 1, std::default_random_engine + normal_distribution<double>(5, 0.5)). These vectors pin the
// oracle's C restatement of the same engines and distributions.
#include <cstdio>
#include <random>

int main() {
    std::printf("{\n");
    for (int seed = 0; seed < 2; seed++) {
        std::mt19937_64 g(seed);
        std::printf("  \"mt64_raw_seed%d\": [", seed);
        for (int i = 0; i < 16; i++) std::printf("%s\"%llu\"", i ? ", " : "", (unsigned long long)g());
        std::printf("],\n");
        std::mt19937_64 g2(seed);
        std::uniform_real_distribution<double> dis(0.0, 1.0);
        std::printf("  \"uniform_seed%d\": [", seed);
        for (int i = 0; i < 2000; i++) std::printf("%s\"%a\"", i ? ", " : "", dis(g2));
        std::printf("],\n");
    }
    std::default_random_engine e;
    std::printf("  \"minstd_raw\": [");
    for (int i = 0; i < 16; i++) std::printf("%s%u", i ? ", " : "", (unsigned)e());
    std::printf("],\n");
    std::default_random_engine e2;
    std::normal_distribution<double> nd(5.0, 0.5);
    std::printf("  \"normal_5_0.5\": [");
    for (int i = 0; i < 2001; i++) std::printf("%s\"%a\"", i ? ", " : "", nd(e2));
    std::printf("]\n}\n");
    return 0;
}
