// Common_Triangulation.h -- records and vector types of the C++ drop-in
// interface over libmcleap_hip.so (Common_Triangulation.h:42-57 of the
// reference: Vertex{pos_index, he}, HalfEdge{next, v, t, op}, Triangle{he};
// the twin of half-edge i is i^1, t = -1 marks the outer face).
//
// Positions are glm::vec2 when the caller includes glm first (the
// reference's type), else a layout-identical {float x, y}.
#pragma once

#include <cstdio>
#include <stdexcept>
#include <string>

#include "../mcleap_hip.h"

#if defined(GLM_VERSION)
namespace mcleap {
using vec2 = glm::vec2;
using ivec2 = glm::ivec2;
}  // namespace mcleap
#else
// callable from the caller's device functors when the translation unit is
// compiled by hipcc
#if defined(__HIP__)
#define MCLEAP_HD __host__ __device__
#else
#define MCLEAP_HD
#endif
namespace mcleap {
struct vec2 {
    float x, y;
    MCLEAP_HD vec2() : x(0.f), y(0.f) {}
    MCLEAP_HD explicit vec2(float s) : x(s), y(s) {}
    MCLEAP_HD vec2(float a, float b) : x(a), y(b) {}
    MCLEAP_HD vec2 operator+(const vec2& o) const { return vec2(x + o.x, y + o.y); }
    MCLEAP_HD vec2 operator-(const vec2& o) const { return vec2(x - o.x, y - o.y); }
    MCLEAP_HD vec2& operator+=(const vec2& o) { x += o.x; y += o.y; return *this; }
};
struct ivec2 {
    int x, y;
    MCLEAP_HD ivec2() : x(0), y(0) {}
    MCLEAP_HD explicit ivec2(int s) : x(s), y(s) {}
    MCLEAP_HD ivec2(int a, int b) : x(a), y(b) {}
};
}  // namespace mcleap
#endif

static_assert(sizeof(mcleap::vec2) == 8, "vec2 must be two packed floats");

struct Vertex {
    int pos_index;
    int he;
};
struct HalfEdge {
    int next;
    int v;
    int t;
    int op;
};
struct Triangle {
    int he;
};

namespace mcleap {
// library status -> the reference's bool results; messages go to stderr
inline bool ok(int rc, const char* what) {
    if (rc == MCL_OK) return true;
    std::fprintf(stderr, "mcleap: %s failed (status %d)\n", what, rc);
    return false;
}
// construction failures cannot be reported through a bool
inline void must(int rc, const char* what) {
    if (rc != MCL_OK) throw std::runtime_error(std::string("mcleap: ") + what + " failed, status " + std::to_string(rc));
}
}  // namespace mcleap
