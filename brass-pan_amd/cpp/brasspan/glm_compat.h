// glm_compat.h -- the few glm types brass-pan's Scene/ParticleSolver API uses
// (glm::vec3, glm::ivec3), for builds without glm.  When <glm/glm.hpp> is
// included first, the real glm types are used instead.
#pragma once
#if !defined(GLM_VERSION)
namespace glm {
struct vec3 {
    float x = 0.f, y = 0.f, z = 0.f;
    vec3() = default;
    explicit vec3(float s) : x(s), y(s), z(s) {}
    vec3(float a, float b, float c) : x(a), y(b), z(c) {}
};
struct ivec3 {
    int x = 0, y = 0, z = 0;
    ivec3() = default;
    explicit ivec3(int s) : x(s), y(s), z(s) {}
    ivec3(int a, int b, int c) : x(a), y(b), z(c) {}
};
inline vec3 normalize(const vec3 &v) {
    float l = __builtin_sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    return vec3(v.x / l, v.y / l, v.z / l);
}
}  // namespace glm
#endif
