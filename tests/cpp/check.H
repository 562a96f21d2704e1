// Minimal assertion helpers for the C++ API tests (no gtest in this image).
#pragma once
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <vector>

static int g_failures = 0;
static std::vector<std::pair<std::string, std::function<void()>>>& registry()
{
    static std::vector<std::pair<std::string, std::function<void()>>> r;
    return r;
}
struct Register {
    Register(const char* n, std::function<void()> f) { registry().emplace_back(n, f); }
};
#define TEST(name)                                   \
    static void name();                              \
    static Register reg_##name(#name, name);         \
    static void name()
#define CHECK(cond)                                                                  \
    do {                                                                             \
        if (!(cond)) {                                                               \
            std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);          \
            g_failures++;                                                            \
        }                                                                            \
    } while (0)
#define CHECK_EQ(a, b) CHECK((a) == (b))
#define CHECK_NEAR(a, b, tol) CHECK(std::abs((double)(a) - (double)(b)) <= (tol))
static int run_all()
{
    for (auto& t : registry()) {
        std::printf("[ RUN ] %s\n", t.first.c_str());
        std::fflush(stdout);
        const int before = g_failures;
        try {
            t.second();
        } catch (const std::exception& e) {
            std::printf("  EXCEPTION %s\n", e.what());
            g_failures++;
        }
        std::printf("[ %s ] %s\n", g_failures == before ? "OK" : "FAIL", t.first.c_str());
        std::fflush(stdout);
    }
    std::printf("%d failure(s)\n", g_failures);
    return g_failures == 0 ? 0 : 1;
}
