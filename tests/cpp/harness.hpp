// harness.hpp -- minimal self-registering test harness (TEST_CASE / EXPECT_*).
// A case is a scenario (run once per backend pass), GPU-only or host-only.
#pragma once
#include <cmath>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

namespace th {
enum Kind { SCENARIO, GPU_ONLY, HOST_ONLY };
struct Case {
    const char* name;
    std::function<void(int&)> fn;
    Kind kind;
};
inline std::vector<Case>& registry() {
    static std::vector<Case> r;
    return r;
}
struct Reg {
    Reg(const char* n, std::function<void(int&)> f, Kind k) { registry().push_back({n, std::move(f), k}); }
};
}  // namespace th

#define TH_CASE(name, kind)                                \
    static void name(int& th_fail);                        \
    static th::Reg th_reg_##name(#name, name, kind);       \
    static void name(int& th_fail)
#define TEST_CASE(name) TH_CASE(name, th::SCENARIO)
#define TEST_CASE_GPU(name) TH_CASE(name, th::GPU_ONLY)
#define TEST_CASE_HOST(name) TH_CASE(name, th::HOST_ONLY)

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

// Runs the cases whose kind is in `kinds` (and whose name contains
// `filter`), each reported as "PASS/FAIL [tag] name"; returns the failures.
inline int th_run(const char* filter, std::vector<th::Kind> kinds, const char* tag, int& run) {
    int failed = 0;
    for (auto& c : th::registry()) {
        bool want = false;
        for (th::Kind k : kinds) want = want || c.kind == k;
        if (!want) continue;
        if (filter && std::string(c.name).find(filter) == std::string::npos) continue;
        int f = 0;
        c.fn(f);
        ++run;
        std::printf("%s [%s] %s\n", f ? "FAIL" : "PASS", tag, c.name);
        std::fflush(stdout);
        failed += f ? 1 : 0;
    }
    return failed;
}
