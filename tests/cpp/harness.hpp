// Minimal test harness for the C++ drop-in API (Catch2 is not available).
// TEST_CASE bodies need a GPU; CPU_TEST_CASE bodies do not.  Run a binary
// with --cpu-only to run the CPU cases alone.
#pragma once

#include <cstdio>
#include <cstring>
#include <exception>
#include <functional>
#include <string>
#include <vector>

namespace harness {
  struct Case {
    const char* name;
    std::function<void()> fn;
    bool gpu;
  };
  inline std::vector<Case>& registry() {
    static std::vector<Case> r;
    return r;
  }
  struct Reg {
    Reg(const char* n, std::function<void()> f, bool gpu) { registry().push_back({n, std::move(f), gpu}); }
  };
  struct Failure {
    std::string msg;
  };
  inline long& checks() {
    static long c = 0;
    return c;
  }
  inline std::string where(const char* file, int line) { return std::string(file) + ":" + std::to_string(line); }

  inline int run(int argc, char** argv) {
    bool cpu_only = argc > 1 && std::strcmp(argv[1], "--cpu-only") == 0;
    int failed = 0, ran = 0;
    for (auto& c : registry()) {
      if (cpu_only && c.gpu) continue;
      ++ran;
      std::printf("[ RUN  ] %s\n", c.name);
      std::fflush(stdout);
      try {
        c.fn();
        std::printf("[  OK  ] %s\n", c.name);
      } catch (const Failure& f) {
        ++failed;
        std::printf("[ FAIL ] %s: %s\n", c.name, f.msg.c_str());
      } catch (const std::exception& e) {
        ++failed;
        std::printf("[ FAIL ] %s: unexpected exception: %s\n", c.name, e.what());
      }
      std::fflush(stdout);
    }
    std::printf("%d cases, %d failed, %ld checks\n", ran, failed, checks());
    return failed ? 1 : 0;
  }
}

#define HX_CAT2(a, b) a##b
#define HX_CAT(a, b) HX_CAT2(a, b)
#define HX_CASE(name, gpu)                                                             \
  static void HX_CAT(hx_case_, __LINE__)();                                            \
  static harness::Reg HX_CAT(hx_reg_, __LINE__)(name, HX_CAT(hx_case_, __LINE__), gpu); \
  static void HX_CAT(hx_case_, __LINE__)()
#define TEST_CASE(name) HX_CASE(name, true)
#define CPU_TEST_CASE(name) HX_CASE(name, false)

#define REQUIRE(cond)                                                                              \
  do {                                                                                             \
    ++harness::checks();                                                                           \
    if (!(cond)) throw harness::Failure{harness::where(__FILE__, __LINE__) + ": REQUIRE(" #cond ")"}; \
  } while (0)

#define REQUIRE_NOTHROW(expr)                                                                       \
  do {                                                                                              \
    ++harness::checks();                                                                            \
    try {                                                                                           \
      expr;                                                                                         \
    } catch (const std::exception& hx_e) {                                                          \
      throw harness::Failure{harness::where(__FILE__, __LINE__) + ": threw " + hx_e.what()};        \
    }                                                                                               \
  } while (0)

#define REQUIRE_THROWS_AS(expr, T)                                                                 \
  do {                                                                                             \
    ++harness::checks();                                                                           \
    bool hx_thrown = false;                                                                        \
    try {                                                                                          \
      expr;                                                                                        \
    } catch (const T&) {                                                                           \
      hx_thrown = true;                                                                            \
    }                                                                                              \
    if (!hx_thrown) throw harness::Failure{harness::where(__FILE__, __LINE__) + ": no " #T " from " #expr}; \
  } while (0)

#define REQUIRE_THROWS(expr) REQUIRE_THROWS_AS(expr, std::exception)

#define HARNESS_MAIN \
  int main(int argc, char** argv) { return harness::run(argc, argv); }
