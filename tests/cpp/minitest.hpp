// minitest.hpp - a small TEST/EXPECT harness in the style of googletest (not available in this image).
// Each test binary runs every registered test (or those whose "Suite.Name" contains argv filter),
// prints one line per test and exits non-zero if any expectation failed.
#pragma once
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

namespace minitest {
struct Case {
  const char *suite, *name;
  std::function<void()> fn;
};
inline std::vector<Case> &registry() {
  static std::vector<Case> r;
  return r;
}
inline int &failures() {
  static int f = 0;
  return f;
}
struct Reg {
  Reg(const char *s, const char *n, std::function<void()> f) { registry().push_back({s, n, std::move(f)}); }
};
inline void fail(const char *file, int line, const std::string &msg) {
  std::printf("    FAILED %s:%d: %s\n", file, line, msg.c_str());
  std::fflush(stdout);
  failures()++;
}
}  // namespace minitest

#define TEST(S, N)                                                            \
  static void S##_##N##_body();                                               \
  static minitest::Reg S##_##N##_reg(#S, #N, S##_##N##_body);                 \
  static void S##_##N##_body()

#define EXPECT_TRUE(c) \
  do { if (!(c)) minitest::fail(__FILE__, __LINE__, "expected true: " #c); } while (0)
#define EXPECT_FALSE(c) EXPECT_TRUE(!(c))
#define EXPECT_EQ(a, b)                                                                                        \
  do {                                                                                                         \
    if (!((a) == (b))) minitest::fail(__FILE__, __LINE__, std::string("expected equal: ") + #a + " == " + #b); \
  } while (0)
#define EXPECT_NEAR(a, b, tol)                                                                           \
  do {                                                                                                   \
    const double a_ = (a), b_ = (b);                                                                     \
    if (!(std::abs(a_ - b_) <= (tol))) {                                                                 \
      char buf_[256];                                                                                    \
      std::snprintf(buf_, sizeof buf_, "|%s - %s| = %.3e > %.3e (%.17g vs %.17g)", #a, #b,               \
                    std::abs(a_ - b_), (double)(tol), a_, b_);                                           \
      minitest::fail(__FILE__, __LINE__, buf_);                                                          \
    }                                                                                                    \
  } while (0)
#define EXPECT_DOUBLE_EQ(a, b) EXPECT_NEAR(a, b, 4 * 2.220446049250313e-16 * std::max(std::abs((double)(a)), std::abs((double)(b))))
#define EXPECT_THROW(stmt)                                                            \
  do {                                                                                \
    bool thrown_ = false;                                                             \
    try { stmt; } catch (...) { thrown_ = true; }                                     \
    if (!thrown_) minitest::fail(__FILE__, __LINE__, "expected an exception: " #stmt); \
  } while (0)

inline int minitest_main(int argc, char **argv) {
  const char *filter = nullptr;
  for (int i = 1; i < argc; i++)
    if (!std::strncmp(argv[i], "--filter=", 9)) filter = argv[i] + 9;
  int run = 0;
  for (auto &c : minitest::registry()) {
    const std::string full = std::string(c.suite) + "." + c.name;
    if (filter && full.find(filter) == std::string::npos) continue;
    std::printf("[ RUN  ] %s\n", full.c_str());
    std::fflush(stdout);
    const int before = minitest::failures();
    try {
      c.fn();
    } catch (const std::exception &e) {
      minitest::fail(__FILE__, __LINE__, std::string("exception: ") + e.what());
    }
    std::printf("[ %s ] %s\n", minitest::failures() == before ? " OK " : "FAIL", full.c_str());
    std::fflush(stdout);
    run++;
  }
  std::printf("%d tests, %d failed expectations\n", run, minitest::failures());
  return minitest::failures() ? 1 : 0;
}
