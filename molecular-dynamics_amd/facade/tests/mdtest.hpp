#pragma once

// A minimal test harness (no GoogleTest offline): TEST registers a function; EXPECT_* records a failure and
// continues; ASSERT_* records a failure and leaves the test. main() runs all tests, or those whose name contains
// argv[1], and returns the number of failed tests.
#include <cmath>
#include <cstdio>
#include <exception>
#include <functional>
#include <string>
#include <vector>

namespace mdtest {

struct Case {
    std::string name;
    std::function<void()> fn;
};

inline std::vector<Case>& registry() {
    static std::vector<Case> cases;
    return cases;
}

struct Registrar {
    Registrar(const char* name, std::function<void()> fn) {
        registry().push_back({name, std::move(fn)});
    }
};

inline int& failures() {
    static int f = 0;
    return f;
}

struct AssertAbort {};

inline void report(const char* file, int line, const std::string& what) {
    ++failures();
    std::printf("    FAIL %s:%d: %s\n", file, line, what.c_str());
}

inline std::string num(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%.17g", v);
    return b;
}

}  // namespace mdtest

#define MDTEST_CAT(a, b) a##_##b
#define TEST(suite, name)                                                                                      \
    static void MDTEST_CAT(suite, name)();                                                                     \
    static mdtest::Registrar MDTEST_CAT(reg_##suite, name)(#suite "." #name, MDTEST_CAT(suite, name));         \
    static void MDTEST_CAT(suite, name)()

#define MDTEST_CHECK(cond, what, fatal)                                                                        \
    do {                                                                                                       \
        if (!(cond)) {                                                                                         \
            mdtest::report(__FILE__, __LINE__, what);                                                          \
            if (fatal) throw mdtest::AssertAbort {};                                                           \
        }                                                                                                      \
    } while (0)

#define EXPECT_TRUE(c) MDTEST_CHECK((c), "expected true: " #c, false)
#define EXPECT_FALSE(c) MDTEST_CHECK(!(c), "expected false: " #c, false)
#define ASSERT_TRUE(c) MDTEST_CHECK((c), "expected true: " #c, true)
#define EXPECT_EQ(a, b) MDTEST_CHECK((a) == (b), "expected " #a " == " #b, false)
#define ASSERT_EQ(a, b) MDTEST_CHECK((a) == (b), "expected " #a " == " #b, true)
#define ASSERT_NE(a, b) MDTEST_CHECK((a) != (b), "expected " #a " != " #b, true)
#define EXPECT_GT(a, b) MDTEST_CHECK((a) > (b), "expected " #a " > " #b, false)
#define EXPECT_GE(a, b) MDTEST_CHECK((a) >= (b), "expected " #a " >= " #b, false)
#define ASSERT_GE(a, b) MDTEST_CHECK((a) >= (b), "expected " #a " >= " #b, true)
#define EXPECT_LT(a, b) MDTEST_CHECK((a) < (b), "expected " #a " < " #b " (" + mdtest::num(a) + ")", false)
#define EXPECT_LE(a, b) MDTEST_CHECK((a) <= (b), "expected " #a " <= " #b " (" + mdtest::num(a) + ")", false)
#define EXPECT_NEAR(a, b, tol)                                                                                 \
    MDTEST_CHECK(std::abs((a) - (b)) <= (tol),                                                                 \
                 "expected |" #a " - " #b "| <= " #tol ": " + mdtest::num(a) + " vs " + mdtest::num(b), false)
#define EXPECT_DOUBLE_EQ(a, b)                                                                                 \
    MDTEST_CHECK(std::abs((a) - (b)) <= 4 * 2.220446049250313e-16 * std::abs(b) || (a) == (b),                \
                 "expected " #a " ~= " #b ": " + mdtest::num(a) + " vs " + mdtest::num(b), false)
#define EXPECT_THROW(stmt, type)                                                                               \
    do {                                                                                                       \
        bool thrown_ = false;                                                                                  \
        try {                                                                                                  \
            stmt;                                                                                              \
        } catch (const type&) {                                                                                \
            thrown_ = true;                                                                                    \
        }                                                                                                      \
        MDTEST_CHECK(thrown_, "expected " #stmt " to throw " #type, false);                                    \
    } while (0)
