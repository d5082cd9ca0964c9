// common.h -- types and constants of the GpuSolver C++ interface.
// Mirrors the reference's include/common.h:11-40 (u8/u32/u64/i64, INF_DIST,
// NO_PARENT, HostTimer) so that callers written against the reference compile
// unchanged.
#pragma once
#include <chrono>
#include <cstdint>
#include <string>
#include <vector>

using u8 = uint8_t;
using u32 = uint32_t;
using u64 = uint64_t;
using i64 = int64_t;

constexpr u32 INF_DIST = 0xFFFFFFFFu;   // unreachable distance
constexpr u32 NO_PARENT = 0xFFFFFFFFu;  // no predecessor

class HostTimer {
public:
    void start() { t0_ = std::chrono::steady_clock::now(); }
    double elapsedMs() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
    }

private:
    std::chrono::steady_clock::time_point t0_ = std::chrono::steady_clock::now();
};
