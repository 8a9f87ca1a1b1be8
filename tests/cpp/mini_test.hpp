// mini_test.hpp — a minimal test harness (Catch2 is not available offline): TEST_CASE
// registers a function, CHECK* record failures, main() runs everything and returns the
// number of failed checks.
#pragma once
#include <cmath>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

namespace mt {
struct Case { const char* name; std::function<void()> fn; };
inline std::vector<Case>& registry() { static std::vector<Case> r; return r; }
inline int& failures() { static int f = 0; return f; }
struct Reg { Reg(const char* n, std::function<void()> f) { registry().push_back({n, std::move(f)}); } };
inline void fail(const char* file, int line, const std::string& what) {
    ++failures();
    std::fprintf(stderr, "  FAILED %s:%d: %s\n", file, line, what.c_str());
}
inline int run_all(const char* filter = nullptr) {
    int ran = 0;
    for (auto& c : registry()) {
        if (filter && std::string(c.name).find(filter) == std::string::npos) continue;
        const int before = failures();
        std::printf("[ RUN  ] %s\n", c.name);
        std::fflush(stdout);
        try {
            c.fn();
        } catch (const std::exception& e) {
            fail(__FILE__, __LINE__, std::string("unexpected exception: ") + e.what());
        }
        std::printf("[ %s ] %s\n", failures() == before ? " OK " : "FAIL", c.name);
        ++ran;
    }
    std::printf("%d test cases, %d failed checks\n", ran, failures());
    return failures();
}
}  // namespace mt

#define MT_CAT2(a, b) a##b
#define MT_CAT(a, b) MT_CAT2(a, b)
#define TEST_CASE(name) \
    static void MT_CAT(mt_fn_, __LINE__)(); \
    static mt::Reg MT_CAT(mt_reg_, __LINE__)(name, MT_CAT(mt_fn_, __LINE__)); \
    static void MT_CAT(mt_fn_, __LINE__)()
#define CHECK(cond) do { if (!(cond)) mt::fail(__FILE__, __LINE__, #cond); } while (0)
#define CHECK_NEAR(a, b, tol) do { const double mt_a = (a), mt_b = (b); \
    if (!(std::abs(mt_a - mt_b) <= (tol))) mt::fail(__FILE__, __LINE__, std::string(#a " ~ " #b ": ") + \
        std::to_string(mt_a) + " vs " + std::to_string(mt_b)); } while (0)
#define CHECK_THROWS_AS(expr, T) do { bool mt_ok = false; try { (void)(expr); } catch (const T&) { mt_ok = true; } \
    catch (...) {} if (!mt_ok) mt::fail(__FILE__, __LINE__, "expected " #T " from " #expr); } while (0)
#define CHECK_NOTHROW(expr) do { try { (void)(expr); } catch (const std::exception& e) { \
    mt::fail(__FILE__, __LINE__, std::string(#expr " threw ") + e.what()); } } while (0)
