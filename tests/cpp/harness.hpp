// harness.hpp -- minimal self-registering test harness (TEST_CASE / EXPECT_*).
#pragma once
#include <cmath>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

namespace th {
struct Case {
    const char* name;
    std::function<void(int&)> fn;
};
inline std::vector<Case>& registry() {
    static std::vector<Case> r;
    return r;
}
struct Reg {
    Reg(const char* n, std::function<void(int&)> f) { registry().push_back({n, std::move(f)}); }
};
}  // namespace th

#define TEST_CASE(name)                                    \
    static void name(int& th_fail);                        \
    static th::Reg th_reg_##name(#name, name);             \
    static void name(int& th_fail)

#define EXPECT_TRUE(c)                                                              \
    do {                                                                            \
        if (!(c)) {                                                                 \
            std::printf("    %s:%d EXPECT_TRUE(%s)\n", __FILE__, __LINE__, #c);     \
            ++th_fail;                                                              \
        }                                                                           \
    } while (0)
#define EXPECT_EQ(a, b) EXPECT_TRUE((a) == (b))
#define EXPECT_NEAR(a, b, tol)                                                                       \
    do {                                                                                             \
        const double th_a = (a), th_b = (b);                                                         \
        if (!(std::fabs(th_a - th_b) <= (tol))) {                                                    \
            std::printf("    %s:%d EXPECT_NEAR(%s=%.17g, %s=%.17g, %g)\n", __FILE__, __LINE__, #a, th_a, \
                        #b, th_b, (double)(tol));                                                    \
            ++th_fail;                                                                               \
        }                                                                                            \
    } while (0)

inline int th_run_all(const char* filter) {
    int failed = 0, run = 0;
    for (auto& c : th::registry()) {
        if (filter && std::string(c.name).find(filter) == std::string::npos) continue;
        int f = 0;
        c.fn(f);
        ++run;
        std::printf("%s %s\n", f ? "FAIL" : "PASS", c.name);
        std::fflush(stdout);
        failed += f ? 1 : 0;
    }
    std::printf("%d/%d passed\n", run - failed, run);
    return failed ? 1 : 0;
}
