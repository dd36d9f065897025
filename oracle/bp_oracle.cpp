// bp_oracle.cpp -- CPU oracle of the brass-pan particle-solver substep.
//
// TEST INFRASTRUCTURE ONLY (tests/, __graft_entry__.smoke(), bench.py
// cpu_baseline).  Never linked into the product library.
//
// Each function cites the reference file:line it restates.  Operation order
// follows the reference source (left-to-right float3 operators of
// ext/helper_math.h); build with -ffp-contract=off so no FMA contraction
// changes the FP32 bits.
#include "bp_oracle.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <numeric>
#include <type_traits>
#include <vector>

// Per-particle loops whose iteration i writes only slot i's outputs (every
// neighbour pass in gather form, and the density and contact passes in both
// forms) run over OpenMP threads; results do not depend on the thread count.
#define OR_PARALLEL_FOR _Pragma("omp parallel for schedule(dynamic, 256)")

namespace {

// ---------------------------------------------------------------- dual numbers
// Forward-mode derivative type for the FP64 tangent (or_rollout_loss): value
// v and kDirs directional derivatives.  Comparisons and branches use the
// value, so a Dual rollout takes exactly the branches of the double rollout
// and its derivative is the derivative of that rollout with every branch,
// clamp and neighbour set held where it is.
constexpr int kDirs = 8;
struct Dual {
    double v = 0.0;
    double d[kDirs] = {};
    Dual() = default;
    Dual(double x) : v(x) {}
    explicit operator double() const { return v; }
    explicit operator float() const { return (float)v; }
};
inline Dual operator+(const Dual &a, const Dual &b) {
    Dual r(a.v + b.v);
    for (int k = 0; k < kDirs; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
inline Dual operator-(const Dual &a, const Dual &b) {
    Dual r(a.v - b.v);
    for (int k = 0; k < kDirs; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
inline Dual operator-(const Dual &a) {
    Dual r(-a.v);
    for (int k = 0; k < kDirs; ++k) r.d[k] = -a.d[k];
    return r;
}
inline Dual operator*(const Dual &a, const Dual &b) {
    Dual r(a.v * b.v);
    for (int k = 0; k < kDirs; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
inline Dual operator/(const Dual &a, const Dual &b) {
    Dual r(a.v / b.v);
    for (int k = 0; k < kDirs; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
    return r;
}
inline Dual &operator+=(Dual &a, const Dual &b) { return a = a + b; }
inline Dual &operator-=(Dual &a, const Dual &b) { return a = a - b; }
inline bool operator<(const Dual &a, const Dual &b) { return a.v < b.v; }
inline bool operator<=(const Dual &a, const Dual &b) { return a.v <= b.v; }
inline bool operator>(const Dual &a, const Dual &b) { return a.v > b.v; }
inline bool operator>=(const Dual &a, const Dual &b) { return a.v >= b.v; }
inline bool operator==(const Dual &a, const Dual &b) { return a.v == b.v; }
inline bool operator!=(const Dual &a, const Dual &b) { return a.v != b.v; }
inline Dual chain(const Dual &a, double val, double dfda) {
    Dual r(val);
    for (int k = 0; k < kDirs; ++k) r.d[k] = (a.d[k] == 0.0) ? 0.0 : dfda * a.d[k];
    return r;
}
// sqrt'(0) is taken as 0 (lengths of zero vectors)
inline Dual sqrt(const Dual &a) {
    const double r = std::sqrt(a.v);
    return chain(a, r, r > 0.0 ? 0.5 / r : 0.0);
}
inline Dual pow(const Dual &a, const Dual &b) {
    const double val = std::pow(a.v, b.v);
    Dual r(val);
    for (int k = 0; k < kDirs; ++k) {
        double t = (a.d[k] == 0.0) ? 0.0 : b.v * std::pow(a.v, b.v - 1.0) * a.d[k];
        if (b.d[k] != 0.0) t += val * std::log(a.v) * b.d[k];
        r.d[k] = t;
    }
    return r;
}
inline Dual fabs(const Dual &a) { return a.v < 0.0 ? -a : a; }
inline Dual sin(const Dual &a) { return chain(a, std::sin(a.v), std::cos(a.v)); }
inline Dual cos(const Dual &a) { return chain(a, std::cos(a.v), -std::sin(a.v)); }
inline Dual acos(const Dual &a) { return chain(a, std::acos(a.v), -1.0 / std::sqrt(1.0 - a.v * a.v)); }
// fminf / fmaxf: the operand selected by value (NaN loses)
inline Dual fmin(const Dual &a, const Dual &b) { return (std::isnan(a.v) || b.v < a.v) ? b : a; }
inline Dual fmax(const Dual &a, const Dual &b) { return (std::isnan(a.v) || b.v > a.v) ? b : a; }

// math on the real type R (float, double or Dual)
// value without derivative (identity for float / double)
template <class R> inline R nograd(R a) { return a; }
template <> inline Dual nograd<Dual>(Dual a) { return Dual(a.v); }

template <class R> inline R sqrt_(R a) { using std::sqrt; return sqrt(a); }
template <class R> inline R pow_(R a, R b) { using std::pow; return pow(a, b); }
template <class R> inline R fabs_(R a) { using std::fabs; return fabs(a); }
template <class R> inline R sin_(R a) { using std::sin; return sin(a); }
template <class R> inline R cos_(R a) { using std::cos; return cos(a); }
template <class R> inline R acos_(R a) { using std::acos; return acos(a); }

template <class R> struct V3 {
    R x, y, z;
};
template <class R> inline V3<R> mk(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline V3<R> &operator+=(V3<R> &a, V3<R> b) { a = a + b; return a; }
template <class R> inline V3<R> &operator-=(V3<R> &a, V3<R> b) { a = a - b; return a; }
// helper_math.h dot/length2/length/cross semantics (:1253,1308,1296,1437)
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline R length2(V3<R> a) { return dot(a, a); }
template <class R> inline R length(V3<R> a) { return sqrt_(dot(a, a)); }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// device normalize = v * rsqrt(dot(v,v)) (helper_math.h:1326-1330); host uses 1/sqrt.
template <class R> inline V3<R> normalize(V3<R> v) { R inv = R(1) / sqrt_(dot(v, v)); return v * inv; }

template <class R> struct Q4 {
    R x, y, z, w;
};

// fminf / fmaxf semantics for the reference's float min/max (Q11).
template <class R> inline R fmin_(R a, R b) { using std::fmin; return fmin(a, b); }
template <class R> inline R fmax_(R a, R b) { using std::fmax; return fmax(a, b); }

// ---------------------------------------------------------------- grid hash
// calcGridPos (grid.cuh:9-13): int3((p + 50000) / cellSize), FP32, truncation.
struct Grid {
    float cs[3];
    int dims[3];
    uint32_t mask[3];
    int num_cells() const { return dims[0] * dims[1] * dims[2]; }
    inline void pos(const float p[3], int g[3]) const {
        for (int a = 0; a < 3; ++a) g[a] = (int)((p[a] + 50000.0f) / cs[a]);
    }
    // calcGridAddress (grid.cuh:15-22), x-major, extended to per-axis dims
    inline int addr(int x, int y, int z) const {
        uint32_t cx = (uint32_t)x & mask[0], cy = (uint32_t)y & mask[1], cz = (uint32_t)z & mask[2];
        return (int)(cx * (uint32_t)(dims[1] * dims[2]) + cy * (uint32_t)dims[2] + cz);
    }
    inline void decompose(int a, int c[3]) const {
        c[2] = a % dims[2];
        c[1] = (a / dims[2]) % dims[1];
        c[0] = a / (dims[1] * dims[2]);
    }
};

template <class R> inline void to_f(const V3<R> &v, float p[3]) {
    p[0] = (float)v.x; p[1] = (float)v.y; p[2] = (float)v.z;
}

// --------------------------------------------------------------- SPH kernels
// SetKernelRadius (fluid.cuh:11-27), computed in the real type.
template <class R> struct KConst {
    R c1, c2, c3, c4, h, h2, halfh;
    void set(R hh) {
        const R pi = R(3.141592f);
        c1 = R(315) / R(64) / pi / pow_(hh, R(9));
        c2 = R(-45) / pi / pow_(hh, R(6));
        c3 = R(32) / pi / pow_(hh, R(9));
        c4 = pow_(hh, R(6)) / R(64);
        h2 = hh * hh;
        halfh = hh * R(0.5f);
        h = hh;
    }
    // poly6Kernel (fluid.cuh:31-42), range check off (USE_KERNEL_DISTANCE_CHECK undefined)
    inline R poly6(R r2) const { R t = h2 - r2; return c1 * t * t * t; }
    // gradientSpikyKernel (fluid.cuh:45-56)
    inline V3<R> grad(V3<R> v, R r2) const {
        R r = sqrt_(r2);
        R t = h - r;
        R s = c2 * t * t;
        return (s * v) / r;
    }
    // akinciSplineC (fluid.cuh:59-73)
    inline R akinci(R r) const {
        R t = (h - r) * r;
        R t3 = t * t * t;
        if (r >= halfh) return c3 * t3;
        return R(2) * c3 * t3 - c4;
    }
};

// ------------------------------------------------------------- quaternions
// cudamatrix.cuh:112-172
template <class R> inline Q4<R> qmul(Q4<R> a, Q4<R> b) {
    Q4<R> r;
    r.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
    r.y = a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x;
    r.z = a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w;
    r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    return r;
}
template <class R> inline Q4<R> angle_axis(V3<R> d, R angle) {
    R s = sin_(angle * R(0.5f)), c = cos_(angle * R(0.5f));
    return Q4<R>{d.x * s, d.y * s, d.z * s, c};
}
template <class R> inline Q4<R> qnormalize(Q4<R> q) {
    R inv = R(1) / sqrt_(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    return Q4<R>{q.x * inv, q.y * inv, q.z * inv, q.w * inv};
}
template <class R> struct M3 {
    V3<R> c[3];
};
template <class R> inline M3<R> quat_to_mat(Q4<R> q) {
    R x2 = q.x * q.x, y2 = q.y * q.y, z2 = q.z * q.z;
    M3<R> m;
    m.c[0].x = R(1) - R(2) * y2 - R(2) * z2;
    m.c[0].y = R(2) * q.x * q.y + R(2) * q.w * q.z;
    m.c[0].z = R(2) * q.x * q.z - R(2) * q.w * q.y;
    m.c[1].x = R(2) * q.x * q.y - R(2) * q.w * q.z;
    m.c[1].y = R(1) - R(2) * x2 - R(2) * z2;
    m.c[1].z = R(2) * q.y * q.z + R(2) * q.w * q.x;
    m.c[2].x = R(2) * q.x * q.z + R(2) * q.w * q.y;
    m.c[2].y = R(2) * q.y * q.z - R(2) * q.w * q.x;
    m.c[2].z = R(1) - R(2) * x2 - R(2) * y2;
    return m;
}
// matrix3 * float3 (cudamatrix.cuh:87-95)
template <class R> inline V3<R> mat_vec(const M3<R> &m, V3<R> a) {
    V3<R> o = m.c[0] * a.x;
    o += m.c[1] * a.y;
    o += m.c[2] * a.z;
    return o;
}
// Rotation extraction loop of shapeMatchingAlphaOne (shapematching.cuh:70-88)
template <class R> Q4<R> extract_rotation(const V3<R> A[3], Q4<R> q, int iters) {
    for (int it = 0; it < iters; ++it) {
        M3<R> Rm = quat_to_mat(q);
        V3<R> num = cross(Rm.c[0], A[0]) + cross(Rm.c[1], A[1]) + cross(Rm.c[2], A[2]);
        R den = R(1) / fabs_(dot(Rm.c[0], A[0]) + dot(Rm.c[1], A[1]) + dot(Rm.c[2], A[2])) + R(1e-9f);
        V3<R> om = num * den;
        R w2 = length2(om);
        if (w2 <= R(1e-9f)) break;
        R w = sqrt_(w2);
        q = qmul(angle_axis(om / w, w), q);
        q = qnormalize(q);
    }
    return q;
}

// ------------------------------------------------------- stable key sort
// cub::DeviceRadixSort::SortPairs over all 32 bits of non-negative int keys
// (particlesolver.cuh:721-738): stable ascending.  A counting sort over the
// cell range gives the identical permutation.
void stable_sort_keys(const std::vector<int> &keys, int num_cells, std::vector<int> &sorted_keys,
                      std::vector<int> &perm) {
    const int n = (int)keys.size();
    std::vector<int> count((size_t)num_cells + 1, 0);
    for (int i = 0; i < n; ++i) count[(size_t)keys[i] + 1]++;
    for (int c = 0; c < num_cells; ++c) count[(size_t)c + 1] += count[(size_t)c];
    sorted_keys.assign(n, 0);
    perm.assign(n, 0);
    for (int i = 0; i < n; ++i) {
        int dst = count[(size_t)keys[i]]++;
        sorted_keys[dst] = keys[i];
        perm[dst] = i;
    }
}

// findStartEndId (grid.cuh:70-99)
void find_ranges(const std::vector<int> &sk, std::vector<int> &start, std::vector<int> &end) {
    const int n = (int)sk.size();
    for (int i = 0; i < n; ++i) {
        int u = sk[i];
        if (i > 0) {
            int v = sk[i - 1];
            if (u != v) { start[u] = i; end[v] = i; }
        } else {
            start[u] = i;
        }
        if (i == n - 1) end[u] = n;
    }
}
// resetStartEndId (grid.cuh:101-132)
void reset_ranges(const std::vector<int> &sk, std::vector<int> &start, std::vector<int> &end) {
    const int n = (int)sk.size();
    for (int i = 0; i < n; ++i) {
        int u = sk[i];
        if (i > 0) {
            int v = sk[i - 1];
            if (u != v) { start[u] = -1; end[v] = -1; }
        } else {
            start[u] = -1;
        }
        if (i == n - 1) end[u] = -1;
    }
}

struct SolverBase {
    virtual ~SolverBase() {}
    virtual int n() const = 0;
    virtual int step(int substeps, float dt, int picked, const float *pp, const float *pv) = 0;
    virtual void set_params(const bp_params &p) = 0;
    virtual void get_state(double *pos, double *vel, int *map) = 0;
    virtual void set_state(const double *pos, const double *vel, const int *map) = 0;
    virtual void get_rigid(double *q, double *cm) = 0;
    virtual void set_rigid(const double *q) = 0;
    virtual void fetch(int which, double *out) = 0;
    virtual void fetch_int(int which, int *out) = 0;
};

template <class R> struct Solver : SolverBase {
    bp_params P;
    Grid grid;
    KConst<R> K;
    int N = 0;
    int mode = 0;
    R radius = 0, rest_density = 1;
    // per-slot state (sorted order, ParticleSolver dev* members :1078-1091)
    std::vector<V3<R>> xs, x, v;
    std::vector<R> mass;
    std::vector<int> phase, orig, group;
    std::vector<int> to_new;
    // scratch
    std::vector<V3<R>> dx, omega, normal, vtmp;
    std::vector<R> lambda, density, invw;
    std::vector<int> keys, sorted_keys, perm, cstart, cend;
    // scene (original ids)
    std::vector<V3<R>> plane_o, plane_n;
    std::vector<int> dpairs;
    std::vector<R> dparams;
    std::vector<int> bends, faces, rranges, immov;
    std::vector<V3<R>> rrest;
    std::vector<Q4<R>> rrot;
    std::vector<V3<R>> rcm;
    bool has_solid = false, has_fluid = false;
    // differentiable parameters (wind vector, drag, lift, XSPH c), from P
    R th_wind[3] = {R(0), R(0), R(0)}, th_cd = R(0), th_cl = R(0), th_xsph = R(0);
    void theta_from_params() {
        for (int a = 0; a < 3; ++a) th_wind[a] = R(P.wind[a]);
        th_cd = R(P.drag_coeff);
        th_cl = R(P.lift_coeff);
        th_xsph = R(P.xsph_c);
    }

    int n() const override { return N; }

    void init(const bp_scene_desc &s, const bp_params &p, int m) {
        P = p;
        theta_from_params();
        mode = m;
        N = s.num_particles;
        for (int a = 0; a < 3; ++a) {
            grid.cs[a] = s.cell_size[a];
            grid.dims[a] = s.grid_dims[a];
            grid.mask[a] = (uint32_t)s.grid_dims[a] - 1u;
        }
        K.set(R(s.kernel_radius));
        radius = R(s.particle_radius);
        rest_density = R(p.fluid_rest_density);
        xs.assign(N, {});
        x.resize(N);
        v.assign(N, V3<R>{0, 0, 0});
        mass.resize(N);
        phase.resize(N);
        orig.resize(N);
        group.assign(N, 0);
        to_new.resize(N);
        for (int i = 0; i < N; ++i) {
            x[i] = mk(R(s.positions[3 * i]), R(s.positions[3 * i + 1]), R(s.positions[3 * i + 2]));
            if (s.velocities)
                v[i] = mk(R(s.velocities[3 * i]), R(s.velocities[3 * i + 1]), R(s.velocities[3 * i + 2]));
            mass[i] = R(s.masses[i]);
            phase[i] = s.phases[i];
            if (s.group_ids) group[i] = s.group_ids[i];
            orig[i] = i;       // initOrder_int (util.cuh:312-328)
            to_new[i] = i;
            if (phase[i] > 0) has_solid = true; else has_fluid = true;
        }
        dx.assign(N, {}); omega.assign(N, {}); normal.assign(N, {}); vtmp.assign(N, {});
        lambda.assign(N, R(0)); density.assign(N, R(0)); invw.assign(N, R(0));
        cstart.assign(grid.num_cells(), -1);
        cend.assign(grid.num_cells(), -1);
        for (int i = 0; i < s.num_planes; ++i) {
            plane_o.push_back(mk(R(s.plane_origins[3 * i]), R(s.plane_origins[3 * i + 1]), R(s.plane_origins[3 * i + 2])));
            plane_n.push_back(mk(R(s.plane_normals[3 * i]), R(s.plane_normals[3 * i + 1]), R(s.plane_normals[3 * i + 2])));
        }
        dpairs.assign(s.distance_pairs, s.distance_pairs + 2 * s.num_distance);
        for (int i = 0; i < 2 * s.num_distance; ++i) dparams.push_back(R(s.distance_params[i]));
        bends.assign(s.bendings, s.bendings + 4 * s.num_bendings);
        faces.assign(s.faces, s.faces + 3 * s.num_faces);
        rranges.assign(s.rigid_ranges, s.rigid_ranges + 2 * s.num_rigid);
        for (int i = 0; i < s.rigid_rest_len; ++i)
            rrest.push_back(mk(R(s.rigid_rest_positions[3 * i]), R(s.rigid_rest_positions[3 * i + 1]),
                               R(s.rigid_rest_positions[3 * i + 2])));
        rrot.assign(s.num_rigid, Q4<R>{0, 0, 0, 1});
        rcm.assign(s.num_rigid, V3<R>{0, 0, 0});
        immov.assign(s.immovables, s.immovables + s.num_immovables);
    }

    void set_params(const bp_params &p) override {
        P = p;
        theta_from_params();
        rest_density = R(p.fluid_rest_density);
    }

    inline int cur_cell_coords(const V3<R> &p, int g[3]) const {
        float f[3];
        to_f(p, f);
        grid.pos(f, g);
        return grid.addr(g[0], g[1], g[2]);
    }

    // ------------------------------------------------------------ grid
    // updateGrid (particlesolver.cuh:712-765) + inverseMapping (util.cuh:348-359)
    void update_grid() {
        keys.resize(N);
        for (int i = 0; i < N; ++i) {  // updateGridId (grid.cuh:30-47)
            int g[3];
            keys[i] = cur_cell_coords(xs[i], g);
        }
        stable_sort_keys(keys, grid.num_cells(), sorted_keys, perm);
        find_ranges(sorted_keys, cstart, cend);
        // reorderParticlesData (grid.cuh:134-164) + the swaps (:758-764)
        std::vector<V3<R>> nxs(N), nx(N), nv(N);
        std::vector<R> nm(N);
        std::vector<int> np(N), no(N), ng(N);
        for (int i = 0; i < N; ++i) {
            int o = perm[i];
            nxs[i] = xs[o]; nx[i] = x[o]; nv[i] = v[o]; nm[i] = mass[o];
            np[i] = phase[o]; no[i] = orig[o]; ng[i] = group[o];
        }
        xs.swap(nxs); x.swap(nx); v.swap(nv); mass.swap(nm);
        phase.swap(np); orig.swap(no); group.swap(ng);
        for (int i = 0; i < N; ++i) to_new[orig[i]] = i;
    }

    // ------------------------------------------------------------ neighbours
    template <class F> inline void for_candidates(const V3<R> &center, F &&f) const {
        int g[3];
        float c[3];
        to_f(center, c);
        grid.pos(c, g);
        if (mode == 0) {
            // loop order z, y, x as in every fluid kernel (fluid.cuh:108-110)
            for (int z = g[2] - 1; z <= g[2] + 1; ++z)
                for (int y = g[1] - 1; y <= g[1] + 1; ++y)
                    for (int xx = g[0] - 1; xx <= g[0] + 1; ++xx) {
                        int a = grid.addr(xx, y, z);
                        int s = cstart[a], e = cend[a];
                        for (int j = s; j < e; ++j) f(j);
                    }
        } else {
            // gather mode: the GPU kernels' order -- y, x, then the z run
            for (int y = g[1] - 1; y <= g[1] + 1; ++y)
                for (int xx = g[0] - 1; xx <= g[0] + 1; ++xx)
                    for (int z = g[2] - 1; z <= g[2] + 1; ++z) {
                        int a = grid.addr(xx, y, z);
                        int s = cstart[a], e = cend[a];
                        for (int j = s; j < e; ++j) f(j);
                    }
        }
    }
    // true iff wrapped cell `cell` lies in the 27-cell block around grid position g
    inline bool in_block(const int g[3], int cell) const {
        int c[3];
        grid.decompose(cell, c);
        for (int a = 0; a < 3; ++a) {
            bool ok = false;
            for (int d = -1; d <= 1; ++d)
                if ((int)(((uint32_t)(g[a] + d)) & grid.mask[a]) == c[a]) ok = true;
            if (!ok) return false;
        }
        return true;
    }

    // ---------------------------------------------------------- passes
    // computeInvScaledMasses (particlesolver.cuh:84-98)
    void inv_scaled_masses() {
        const R e = R(2.7182818284f);
        const R k = R(P.mass_scaling_k);
        for (int i = 0; i < N; ++i) {
            R scale = pow_(e, -k * x[i].y);
            invw[i] = R(1) / (scale * mass[i]);
        }
    }
    // planeStabilize (particleplane.cuh:3-19)
    void plane_stabilize(V3<R> o, V3<R> nrm) {
        OR_PARALLEL_FOR
        for (int i = 0; i < N; ++i) {
            V3<R> o2p = o - x[i];
            R d = dot(o2p, nrm) + radius;
            if (d <= R(0)) continue;
            x[i] += d * nrm;
            xs[i] += d * nrm;
        }
    }
    // particlePlaneCollisionConstraint (particleplane.cuh:21-58)
    void plane_collide(V3<R> o, V3<R> nrm) {
        const R fs = R(P.friction_static), fd = R(P.friction_dynamic);
        OR_PARALLEL_FOR
        for (int i = 0; i < N; ++i) {
            V3<R> o2p = o - xs[i];
            R d = dot(o2p, nrm) + radius;
            if (d <= R(0)) continue;
            V3<R> pos = x[i];
            V3<R> diff = xs[i] - pos;
            R dn = dot(diff, nrm);
            V3<R> dtan = diff - dn * nrm;
            R dtl = length(dtan);
            V3<R> resolved = d * nrm + xs[i];
            V3<R> dX = resolved - pos;
            V3<R> tdx = dX - dot(dX, nrm) * nrm;
            if (dtl < fs * dn) {
                xs[i] = resolved - tdx;
            } else if (dtl == R(0) && !(fd * -dn > R(0))) {
                // no tangential motion: fminf(NaN, 1) selects 1 but tdx is
                // zero, and any tangential motion gives the factor fd (-dn) /
                // dtl <= 0; the value is kept, the derivative is the limit's
                V3<R> t = tdx * fmin_(fd * -dn / dtl, R(1));
                xs[i] = resolved - V3<R>{nograd(t.x), nograd(t.y), nograd(t.z)};
            } else {
                xs[i] = resolved - tdx * fmin_(fd * -dn / dtl, R(1));
            }
        }
    }
    // particleParticleCollisionConstraint (particleparticle.cuh:5-101) + Δx apply
    void particle_particle() {
        const R fs = R(P.friction_static), fd = R(P.friction_dynamic);
        const R r = radius;
        for (int i = 0; i < N; ++i) dx[i] = V3<R>{0, 0, 0};
        OR_PARALLEL_FOR
        for (int i = 0; i < N; ++i) {
            const V3<R> xi = x[i], xip = xs[i];
            V3<R> sumD{0, 0, 0}, sumF{0, 0, 0};
            const R wi = invw[i];
            const int pi = phase[i];
            int cnt = 0;
            for_candidates(xip, [&](int j) {
                int pj = phase[j];
                if (i != j && pi != pj && (pi > 0 || pj > 0)) {
                    V3<R> xjp = xs[j];
                    V3<R> diff = xip - xjp;
                    R d2 = length2(diff);
                    if (d2 < r * r * R(4) && d2 > R(1e-5f)) {
                        R dist = sqrt_(d2);
                        R wj = invw[j];
                        R w1 = wi / (wi + wj);
                        R w2 = wj / (wi + wj);
                        V3<R> proj = diff * (R(2) * r / dist - R(1));
                        V3<R> dxi = w1 * proj;
                        V3<R> xistar = dxi + xip;
                        sumD += dxi;
                        V3<R> dxj = -w2 * proj;
                        R l2 = length2(dxi);
                        if (l2 > r * r * R(0.001f) * R(0.001f)) {
                            cnt += 1;
                            V3<R> xj = x[j];
                            V3<R> xjstar = dxj + xjp;
                            V3<R> term1 = (xistar - xi) - (xjstar - xj);
                            V3<R> n = diff / dist;
                            V3<R> tan = term1 - dot(term1, n) * n;
                            R tl2 = length2(tan);
                            if (tl2 <= (fs * fs) * l2)
                                sumF -= w1 * tan;
                            else
                                sumF -= w1 * tan * fmin_(fd * sqrt_(l2 / tl2), R(1));
                        }
                    }
                }
            });
            dx[i] += (cnt == 0) ? sumD : sumD + sumF / R(cnt);
        }
        for (int i = 0; i < N; ++i) xs[i] += dx[i];  // accDevArr_float3 (util.cuh:261-269)
    }
    // fluidLambda (fluid.cuh:78-159)
    void fluid_lambda() {
        const R sds = R(P.solid_density_scale);
        OR_PARALLEL_FOR
        for (int i = 0; i < N; ++i) {
            if (phase[i] > 0) {  // solids carry lambda 0 (the reference leaves stale values)
                lambda[i] = R(0);
                density[i] = R(0);
                continue;
            }
            const V3<R> pi = xs[i];
            R dens = 0, sumg2 = 0;
            V3<R> gi{0, 0, 0};
            for_candidates(pi, [&](int j) {
                if (i <= j) {  // Q1: j >= i only (fluid.cuh:117)
                    V3<R> pj = xs[j];
                    V3<R> diff = pi - pj;
                    R d2 = length2(pi - pj);
                    if (d2 <= K.h2) {
                        R mj = mass[j];
                        R ld = mj * K.poly6(d2);
                        dens += (phase[j] < 0) ? ld : sds * ld;
                        if (P.lambda_mode == 1 && d2 > R(0)) {
                            V3<R> g = (-mj * K.grad(diff, d2)) / rest_density;
                            sumg2 += dot(g, g);
                            gi -= g;
                        }
                    }
                }
            });
            if (P.lambda_mode == 1) sumg2 += dot(gi, gi);
            R C = dens / rest_density - R(1);
            if (P.akinci) C = fmax_(C, R(0));
            R lam = (P.lambda_mode == 1) ? -C / (sumg2 + R(P.lambda_relaxation))
                                         : -C / R(P.lambda_denominator);
            lambda[i] = lam;
            density[i] = dens;
        }
    }
    // fluidPosition pair term (fluid.cuh:199-221) for the pair (a,b) seen from a
    inline V3<R> pair_delta(const V3<R> &diff, R d2, R lam_sum, R wdq) const {
        R sl = lam_sum;
        if (!P.akinci) sl += -R(P.scorr_k) * pow_(K.poly6(d2) / wdq, R(P.scorr_n));
        V3<R> delta = sl * K.grad(diff, d2);
        const R clamp = R(P.delta_clamp);
        R len = length(delta);
        return (len < clamp) ? delta : delta * fmin_(len, clamp) / len;
    }
    // fluidPosition (fluid.cuh:161-228) + Δx zero/apply (particlesolver.cuh:919-932)
    long long missed_pairs = 0;
    void fluid_position() {
        const R inv_rho = R(1) / rest_density;
        const R wdq = K.poly6(pow_(R(P.scorr_dq) * K.h, R(2)));
        for (int i = 0; i < N; ++i) dx[i] = V3<R>{0, 0, 0};
        if (mode == 0) {
            for (int i = 0; i < N; ++i) {
                if (phase[i] > 0) continue;
                const V3<R> pi = xs[i];
                V3<R> sum{0, 0, 0};
                const R mi = mass[i], li = lambda[i];
                for_candidates(pi, [&](int j) {
                    if (i < j) {
                        V3<R> pj = xs[j];
                        V3<R> diff = pi - pj;
                        R d2 = length2(diff);
                        if (d2 <= K.h2 && d2 > R(0)) {
                            R mj = mass[j];
                            V3<R> delta = pair_delta(diff, d2, li + lambda[j], wdq);
                            sum += mj * delta;
                            dx[j] += -mi * delta * inv_rho;
                        }
                    }
                });
                dx[i] += sum * inv_rho;
            }
        } else {
            // gather: each k sums m_j δ(k,j) over the pairs the reference forms
            // that its own search finds; a pair its owner k finds but the
            // partner j's search misses is delivered to j in 2^-40 fixed point
            // (order-independent integer sums), as the GPU kernel does.
            std::vector<long long> corr((size_t)3 * N, 0);
            std::vector<V3<R>> sums(N);
            OR_PARALLEL_FOR
            for (int k = 0; k < N; ++k) {
                const V3<R> pk = xs[k];
                const int bk = sorted_keys[k];
                const bool k_fluid = phase[k] <= 0;
                V3<R> sum{0, 0, 0};
                const R lk = lambda[k];
                for_candidates(pk, [&](int j) {
                    if (j == k) return;
                    if (j > k && !k_fluid) return;      // pair owned by solid k: none
                    if (j < k && phase[j] > 0) return;  // pair owned by solid j: none
                    V3<R> pj = xs[j];
                    V3<R> diff = pk - pj;
                    R d2 = length2(diff);
                    if (!(d2 <= K.h2 && d2 > R(0))) return;
                    int gj[3];
                    cur_cell_coords(pj, gj);
                    const bool partner_sees_k = in_block(gj, bk);
                    if (j < k && !partner_sees_k) return;  // owner j's search does not reach k
                    V3<R> delta = pair_delta(diff, d2, lk + lambda[j], wdq);
                    sum += mass[j] * delta;
                    if (j > k && !partner_sees_k) {
                        V3<R> give = -mass[k] * delta;
                        const long long gx = std::llrint((double)give.x * 1099511627776.0);
                        const long long gy = std::llrint((double)give.y * 1099511627776.0);
                        const long long gz = std::llrint((double)give.z * 1099511627776.0);
#pragma omp atomic
                        corr[3 * (size_t)j + 0] += gx;
#pragma omp atomic
                        corr[3 * (size_t)j + 1] += gy;
#pragma omp atomic
                        corr[3 * (size_t)j + 2] += gz;
                    }
                });
                sums[k] = sum;
            }
            const R inv = R(1.0 / 1099511627776.0);
            for (int k = 0; k < N; ++k) {
                V3<R> sum = sums[k];
                const long long *c = &corr[3 * (size_t)k];
                if (c[0] | c[1] | c[2]) {
                    sum.x += R(c[0]) * inv;
                    sum.y += R(c[1]) * inv;
                    sum.z += R(c[2]) * inv;
                }
                dx[k] = sum * inv_rho;
            }
        }
        for (int i = 0; i < N; ++i) xs[i] += dx[i];
    }
    // distanceConstraints (distance.cuh:3-41) + Δx apply
    void distance() {
        for (int i = 0; i < N; ++i) dx[i] = V3<R>{0, 0, 0};
        const int nd = (int)dpairs.size() / 2;
        for (int c = 0; c < nd; ++c) {
            int id1 = to_new[dpairs[2 * c]], id2 = to_new[dpairs[2 * c + 1]];
            R L = dparams[2 * c], stiff = dparams[2 * c + 1];
            V3<R> x12 = xs[id1] - xs[id2];
            R d = length(x12);
            if (d <= R(1e-5f)) continue;
            V3<R> n = x12 / d;
            R w1 = invw[id1], w2 = invw[id2];
            R sw = w1 + w2;
            V3<R> k = n * ((d - L) / sw) * stiff;
            dx[id1] += -w1 * k;
            dx[id2] += w2 * k;
        }
        for (int i = 0; i < N; ++i) xs[i] += dx[i];
    }
    // bendingConstraints (bending.cuh:3-76) + Δx apply.  The hinge is
    // evaluated in the type W (double for the FP32 instance): near a flat
    // hinge (d -> -1) sqrt(1 - d^2)(acos d - pi) and sum w|Q|^2 both vanish
    // and in FP32 their ratio is dominated by rounding (FP32 against FP64
    // per-substep differences of 2e-3 on a rippled flag), so the reference's
    // FP32 arithmetic does not define the result there.  The GPU kernel
    // computes the same hinge in FP64; the corrections are then rounded to R.
    void bending() {
        using W = std::conditional_t<std::is_same<R, float>::value, double, R>;
        auto wide = [](V3<R> a) { return V3<W>{W(a.x), W(a.y), W(a.z)}; };
        for (int i = 0; i < N; ++i) dx[i] = V3<R>{0, 0, 0};
        const int nb = (int)bends.size() / 4;
        const W kb = W(P.bend_stiffness);
        for (int c = 0; c < nb; ++c) {
            int id1 = to_new[bends[4 * c]], id2 = to_new[bends[4 * c + 1]];
            int id3 = to_new[bends[4 * c + 2]], id4 = to_new[bends[4 * c + 3]];
            V3<W> P1 = wide(xs[id1]);
            V3<W> P2 = wide(xs[id2]) - P1, P3 = wide(xs[id3]) - P1, P4 = wide(xs[id4]) - P1;
            V3<W> c23 = cross(P2, P3), c24 = cross(P2, P4);
            W l23 = length(c23);
            if (l23 <= W(1e-5f)) continue;
            W l24 = length(c24);
            if (l24 <= W(1e-5f)) continue;
            V3<W> N1 = c23 * (W(1) / l23);
            V3<W> N2 = c24 * (W(1) / l24);
            W d = dot(N1, N2);
            d = fmin_(fmax_(d, W(-1)), W(1));
            V3<W> c2n1 = cross(P2, N1), c2n2 = cross(P2, N2);
            V3<W> c3n1 = cross(P3, N1), c3n2 = cross(P3, N2);
            V3<W> c4n1 = cross(P4, N1), c4n2 = cross(P4, N2);
            V3<W> Q3 = (c2n2 - c2n1 * d) / l23;
            V3<W> Q4v = (c2n1 - c2n2 * d) / l24;
            V3<W> Q2 = (-(c3n2 - c3n1 * d)) / l23 - (c4n1 - c4n2 * d) / l24;
            V3<W> Q1 = -Q2 - Q3 - Q4v;
            W w1 = W(invw[id1]), w2 = W(invw[id2]), w3 = W(invw[id3]), w4 = W(invw[id4]);
            W den = w1 * length2(Q1) + w2 * length2(Q2) + w3 * length2(Q3) + w4 * length2(Q4v);
            W num = sqrt_(W(1) - d * d) * (acos_(d) - W(3.141592f));
            if (den <= W(1e-5f)) continue;
            W sc = num / den;
            auto narrow = [](V3<W> a) { return V3<R>{R(a.x), R(a.y), R(a.z)}; };
            dx[id1] += narrow(-w1 * Q1 * sc * kb);
            dx[id2] += narrow(-w2 * Q2 * sc * kb);
            dx[id3] += narrow(-w3 * Q3 * sc * kb);
            dx[id4] += narrow(-w4 * Q4v * sc * kb);
        }
        for (int i = 0; i < N; ++i) xs[i] += dx[i];
    }
    // shapeMatchingAlphaOne (shapematching.cuh:8-99), one body at a time
    void shape_matching() {
        const int nr = (int)rranges.size() / 2;
        for (int b = 0; b < nr; ++b) {
            int s = rranges[2 * b], e = rranges[2 * b + 1];
            int np = e - s;
            V3<R> sum{0, 0, 0};
            for (int t = 0; t < np; ++t) sum += xs[to_new[s + t]];
            V3<R> cm = sum / R(np);
            rcm[b] = cm;
            V3<R> A[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            for (int t = 0; t < np; ++t) {
                V3<R> q = rrest[s + t];
                V3<R> p = xs[to_new[s + t]] - cm;
                A[0] += p * q.x;
                A[1] += p * q.y;
                A[2] += p * q.z;
            }
            Q4<R> q = extract_rotation(A, rrot[b], P.shape_matching_iterations);
            M3<R> Rm = quat_to_mat(q);
            rrot[b] = q;
            for (int t = 0; t < np; ++t) xs[to_new[s + t]] = mat_vec(Rm, rrest[s + t]) + cm;
        }
    }
    // applyWindForce (wind.cuh:3-50) + Δx apply (particlesolver.cuh:838-858)
    void wind(R dt) {
        for (int i = 0; i < N; ++i) dx[i] = V3<R>{0, 0, 0};
        const int nf = (int)faces.size() / 3;
        const V3<R> W = mk(th_wind[0], th_wind[1], th_wind[2]);
        const R cd = th_cd, cl = th_cl;
        for (int f = 0; f < nf; ++f) {
            int id1 = to_new[faces[3 * f]], id2 = to_new[faces[3 * f + 1]], id3 = to_new[faces[3 * f + 2]];
            V3<R> p1 = xs[id1], p2 = xs[id2], p3 = xs[id3];
            V3<R> a = p2 - p1, b = p3 - p1;
            V3<R> vel = (v[id1] + v[id2] + v[id3]) / R(3);
            vel = vel - W;  // relative air velocity; W = 0 is the reference
            R lv = length(vel);
            if (lv <= R(1e-5f)) continue;
            V3<R> vhat = vel / lv;
            V3<R> fn = cross(a, b);
            R lfn = length(fn);
            if (lfn <= R(1e-5f)) continue;
            fn = fn / lfn;
            fn = (dot(fn, vel) > R(0)) ? fn : -fn;
            R area = R(0.5f) * lfn;
            V3<R> drag = cd * (lv * lv) * area * dot(fn, vhat) * (-vhat);
            V3<R> uhat = cross(cross(fn, vhat), vhat);
            V3<R> lift = cl * (lv * lv) * area * dot(fn, vhat) * uhat;
            V3<R> force = drag + lift;
            V3<R> move = force * dt * dt;
            dx[id1] += move / mass[id1];
            dx[id2] += move / mass[id2];
            dx[id3] += move / mass[id3];
        }
        for (int i = 0; i < N; ++i) xs[i] += dx[i];
    }
    // fluidOmega (fluid.cuh:231-279)
    void fluid_omega() {
        OR_PARALLEL_FOR
        for (int i = 0; i < N; ++i) {
            const int pi = phase[i];
            if (pi >= 0) continue;
            const V3<R> p = xs[i], vi = v[i];
            V3<R> om{0, 0, 0};
            for_candidates(p, [&](int j) {
                if (i != j && pi == phase[j]) {
                    V3<R> diff = p - xs[j];
                    R d2 = length2(diff);
                    if (d2 <= K.h2 && d2 > R(0)) om += cross(v[j] - vi, K.grad(diff, d2));
                }
            });
            omega[i] = om;
        }
    }
    // fluidVorticity (fluid.cuh:281-336), in place on v
    void fluid_vorticity(R dt) {
        const R eps = R(P.vorticity_eps);
        OR_PARALLEL_FOR
        for (int i = 0; i < N; ++i) {
            const int pi = phase[i];
            if (pi >= 0) continue;
            const V3<R> p = xs[i];
            V3<R> eta{0, 0, 0};
            for_candidates(p, [&](int j) {
                if (i != j && pi == phase[j]) {
                    V3<R> diff = p - xs[j];
                    R d2 = length2(diff);
                    if (d2 <= K.h2 && d2 > R(0)) eta += length(omega[j]) * K.grad(diff, d2);
                }
            });
            if (length2(eta) > R(1e-3f)) {
                V3<R> nrm = normalize(eta);
                v[i] += eps * cross(nrm, omega[i]) * dt;
            }
        }
    }
    // fluidNormal (fluid.cuh:387-434)
    void fluid_normal() {
        OR_PARALLEL_FOR
        for (int i = 0; i < N; ++i) {
            const int pi = phase[i];
            if (pi >= 0) continue;
            const V3<R> p = xs[i];
            V3<R> nn{0, 0, 0};
            for_candidates(p, [&](int j) {
                if (i != j && pi == phase[j]) {
                    V3<R> diff = p - xs[j];
                    R d2 = length2(diff);
                    if (d2 <= K.h2 && d2 > R(0)) nn += R(1) / density[j] * K.grad(diff, d2);
                }
            });
            normal[i] = K.h * nn;
        }
    }
    // fluidAkinciTension (fluid.cuh:436-489) -> vtmp, then swap (particlesolver.cuh:1048).
    // Particles the kernel skips (non-fluids) keep their velocity; the
    // reference leaves whatever the swap buffer held in those slots.
    void fluid_akinci(R dt) {
        const R g = R(P.tension_gamma);
        OR_PARALLEL_FOR
        for (int i = 0; i < N; ++i) {
            const int pi = phase[i];
            vtmp[i] = v[i];
            if (pi >= 0) continue;
            const V3<R> p = xs[i];
            V3<R> f{0, 0, 0};
            for_candidates(p, [&](int j) {
                if (i != j && pi == phase[j]) {
                    V3<R> pj = xs[j];
                    V3<R> diff = p - pj;
                    R d2 = length2(diff);
                    if (d2 <= K.h2 && d2 > R(0)) {
                        R dist = sqrt_(d2);
                        V3<R> fcoh = -g * K.akinci(dist) * (p - pj) / dist;
                        V3<R> fcur = -g * (normal[i] - normal[j]);
                        R kij = R(2) * rest_density / (density[i] + density[j]);
                        f += kij * (fcoh + fcur);
                    }
                }
            });
            vtmp[i] = v[i] + f * dt;
        }
        v.swap(vtmp);
    }
    // fluidXSph (fluid.cuh:338-385) -> vtmp, then swap (particlesolver.cuh:1060)
    void fluid_xsph() {
        const R c = th_xsph;
        OR_PARALLEL_FOR
        for (int i = 0; i < N; ++i) {
            const int pi = phase[i];
            vtmp[i] = v[i];
            if (pi >= 0) continue;
            const V3<R> p = xs[i];
            const V3<R> vi = v[i];
            V3<R> vn{0, 0, 0};
            for_candidates(p, [&](int j) {
                if (i != j && pi == phase[j]) {
                    V3<R> diff = p - xs[j];
                    R d2 = length2(diff);
                    if (d2 <= K.h2) vn += (v[j] - vi) * K.poly6(d2);
                }
            });
            vtmp[i] = vi + c * vn;
        }
        v.swap(vtmp);
    }

    // ------------------------------------------------------------ update
    // ParticleSolver::update (particlesolver.cuh:772-1072)
    int step(int substeps, float dtf, int picked, const float *pp, const float *pv) override {
        if (N <= 0) return 0;
        const float subdt_f = dtf / (float)substeps;
        const R subdt = R(subdt_f);
        const R inv_dt = R(1.0f / subdt_f);
        const bool pick = picked >= 0 && picked < N;
        V3<R> ppos{0, 0, 0}, pvel{0, 0, 0};
        if (pick) {
            ppos = mk(R(pp[0]), R(pp[1]), R(pp[2]));
            if (pv) pvel = mk(R(pv[0]), R(pv[1]), R(pv[2]));
        }
        const V3<R> grav = mk(R(P.gravity[0]), R(P.gravity[1]), R(P.gravity[2]));
        for (int s = 0; s < substeps; ++s) {
            if (pick) {  // :790-794
                int slot = to_new[picked];
                x[slot] = ppos;
                v[slot] = V3<R>{0, 0, 0};
            }
            // predictPositions + applyGravity (:796-804; particlesolver.cuh:49-69)
            for (int i = 0; i < N; ++i) xs[i] = x[i] + v[i] * subdt;
            const V3<R> gstep = grav * subdt * subdt;
            for (int i = 0; i < N; ++i) xs[i] += gstep;
            update_grid();        // :807
            inv_scaled_masses();  // :812-816
            for (int it = 0; it < P.stabilization_passes; ++it)  // :822-833
                for (size_t pl = 0; pl < plane_o.size(); ++pl) plane_stabilize(plane_o[pl], plane_n[pl]);
            if (P.wind_enabled && !faces.empty()) wind(subdt);  // :838-858
            for (int it = 0; it < P.iterations; ++it) {         // :862
                if (pick) xs[to_new[picked]] = ppos, v[to_new[picked]] = V3<R>{0, 0, 0};
                for (size_t pl = 0; pl < plane_o.size(); ++pl) plane_collide(plane_o[pl], plane_n[pl]);
                if (has_solid) particle_particle();
                if (has_fluid) {
                    fluid_lambda();
                    fluid_position();
                }
                if (!dpairs.empty()) distance();
                if (!bends.empty()) bending();
                if (!rranges.empty()) shape_matching();
                for (size_t m = 0; m < immov.size(); ++m) {  // immovableConstraints (immovable.cuh:3-14)
                    int id = to_new[immov[m]];
                    xs[id] = x[id];
                }
            }
            // updateVelocity + updatePositions (:996-1002)
            for (int i = 0; i < N; ++i) v[i] = (xs[i] - x[i]) * inv_dt;
            const R thr = R(P.sleep_epsilon);
            for (int i = 0; i < N; ++i) {
                R d2 = length2(xs[i] - x[i]);
                x[i] = (d2 >= thr * thr || phase[i] < 0) ? xs[i] : x[i];
            }
            if (has_fluid) {
                if (P.vorticity_eps != 0.0f) {
                    fluid_omega();
                    fluid_vorticity(subdt);
                }
                if (P.akinci) {
                    fluid_normal();
                    fluid_akinci(subdt);
                }
                if (P.xsph_c != 0.0f) fluid_xsph();
            }
            reset_ranges(sorted_keys, cstart, cend);  // :1062
        }
        if (pick) {  // :1066-1071
            int slot = to_new[picked];
            v[slot] = pvel;
        }
        return 0;
    }

    void get_state(double *pos, double *vel, int *map) override {
        for (int i = 0; i < N; ++i) {
            if (pos) { pos[3 * i] = (double)x[i].x; pos[3 * i + 1] = (double)x[i].y; pos[3 * i + 2] = (double)x[i].z; }
            if (vel) { vel[3 * i] = (double)v[i].x; vel[3 * i + 1] = (double)v[i].y; vel[3 * i + 2] = (double)v[i].z; }
            if (map) map[i] = orig[i];
        }
    }
    // Sets the slot-order state; masses / phases / group ids follow the map.
    void set_state(const double *pos, const double *vel, const int *map) override {
        std::vector<R> m0(N);
        std::vector<int> p0(N), g0(N);
        for (int i = 0; i < N; ++i) { m0[orig[i]] = mass[i]; p0[orig[i]] = phase[i]; g0[orig[i]] = group[i]; }
        for (int i = 0; i < N; ++i) {
            if (pos) x[i] = mk(R(pos[3 * i]), R(pos[3 * i + 1]), R(pos[3 * i + 2]));
            if (vel) v[i] = mk(R(vel[3 * i]), R(vel[3 * i + 1]), R(vel[3 * i + 2]));
            if (map) orig[i] = map[i];
        }
        for (int i = 0; i < N; ++i) {
            mass[i] = m0[orig[i]]; phase[i] = p0[orig[i]]; group[i] = g0[orig[i]];
            to_new[orig[i]] = i;
        }
    }
    void get_rigid(double *q, double *cm) override {
        for (size_t b = 0; b < rrot.size(); ++b) {
            if (q) { q[4 * b] = (double)rrot[b].x; q[4 * b + 1] = (double)rrot[b].y; q[4 * b + 2] = (double)rrot[b].z; q[4 * b + 3] = (double)rrot[b].w; }
            if (cm) { cm[3 * b] = (double)rcm[b].x; cm[3 * b + 1] = (double)rcm[b].y; cm[3 * b + 2] = (double)rcm[b].z; }
        }
    }
    void set_rigid(const double *q) override {
        for (size_t b = 0; b < rrot.size(); ++b)
            rrot[b] = Q4<R>{R(q[4 * b]), R(q[4 * b + 1]), R(q[4 * b + 2]), R(q[4 * b + 3])};
    }
    void fetch(int which, double *out) override {
        for (int i = 0; i < N; ++i) {
            switch (which) {
            case 0: out[i] = (double)density[i]; break;
            case 1: out[i] = (double)lambda[i]; break;
            case 4: out[3 * i] = (double)omega[i].x; out[3 * i + 1] = (double)omega[i].y; out[3 * i + 2] = (double)omega[i].z; break;
            case 5: out[3 * i] = (double)normal[i].x; out[3 * i + 1] = (double)normal[i].y; out[3 * i + 2] = (double)normal[i].z; break;
            default: break;
            }
        }
    }
    void fetch_int(int which, int *out) override {
        for (int i = 0; i < N; ++i) out[i] = (which == 2) ? sorted_keys[i] : perm[i];
    }
};

// Loss of a rollout, L = sum_i |x_i(T) - target_i|^2 / norm, from the scene's
// initial state in reference order (scatter mode).  theta (6 doubles, may be
// null) replaces wind xyz, drag, lift, XSPH c; v0 (3N doubles, original
// order, may be null) replaces the initial velocities.  With dirs != null
// (FP64 tangent): dL along direction k for k < 6 the k-th parameter, and for
// k = 6, 7 the initial velocities along vdirs[k - 6] (3N each), into dloss.
template <class R>
double rollout_loss(const bp_scene_desc *scene, const bp_params *params, int substeps, float subdt,
                    const double *target, double norm, const double *theta, const double *v0, const double *vdirs,
                    int ndirs_v, double *dloss) {
    Solver<R> s;
    s.init(*scene, *params, 0);
    if (theta) {
        for (int a = 0; a < 3; ++a) s.th_wind[a] = R(theta[a]);
        s.th_cd = R(theta[3]);
        s.th_cl = R(theta[4]);
        s.th_xsph = R(theta[5]);
    }
    const int N = s.N;
    if (v0)
        for (int i = 0; i < N; ++i) s.v[i] = mk(R(v0[3 * i]), R(v0[3 * i + 1]), R(v0[3 * i + 2]));
    if constexpr (std::is_same<R, Dual>::value) {
        for (int a = 0; a < 3; ++a) s.th_wind[a].d[a] = 1.0;
        s.th_cd.d[3] = 1.0;
        s.th_cl.d[4] = 1.0;
        s.th_xsph.d[5] = 1.0;
        for (int k = 0; k < ndirs_v && k < kDirs - 6; ++k)
            for (int i = 0; i < N; ++i) {  // slot i holds original id i at creation
                s.v[i].x.d[6 + k] = vdirs[(size_t)k * 3 * N + 3 * i];
                s.v[i].y.d[6 + k] = vdirs[(size_t)k * 3 * N + 3 * i + 1];
                s.v[i].z.d[6 + k] = vdirs[(size_t)k * 3 * N + 3 * i + 2];
            }
    }
    for (int k = 0; k < substeps; ++k) s.step(1, subdt, -1, nullptr, nullptr);
    R L = R(0);
    for (int i = 0; i < N; ++i) {
        const int o = s.orig[i];
        const V3<R> d = s.x[i] - mk(R(target[3 * o]), R(target[3 * o + 1]), R(target[3 * o + 2]));
        L = L + dot(d, d);
    }
    L = L / R(norm);
    if constexpr (std::is_same<R, Dual>::value) {
        if (dloss)
            for (int k = 0; k < 6 + ndirs_v && k < kDirs; ++k) dloss[k] = L.d[k];
        return L.v;
    } else {
        return (double)L;
    }
}

}  // namespace

struct or_solver {
    std::unique_ptr<SolverBase> impl;
};

extern "C" {

or_solver *or_create(const bp_scene_desc *scene, const bp_params *params, int fp64, int mode) {
    if (!scene || !params) return nullptr;
    or_solver *s = new or_solver;
    if (fp64) {
        auto *p = new Solver<double>();
        p->init(*scene, *params, mode);
        s->impl.reset(p);
    } else {
        auto *p = new Solver<float>();
        p->init(*scene, *params, mode);
        s->impl.reset(p);
    }
    return s;
}
void or_destroy(or_solver *s) { delete s; }
int or_num_particles(or_solver *s) { return s->impl->n(); }
int or_step(or_solver *s, int substeps, float dt, int picked, const float *pp, const float *pv) {
    return s->impl->step(substeps, dt, picked, pp, pv);
}
void or_set_params(or_solver *s, const bp_params *p) { s->impl->set_params(*p); }
void or_get_state(or_solver *s, double *pos, double *vel, int *map) { s->impl->get_state(pos, vel, map); }
void or_set_state(or_solver *s, const double *pos, const double *vel, const int *map) {
    s->impl->set_state(pos, vel, map);
}
void or_get_rigid(or_solver *s, double *q, double *cm) { s->impl->get_rigid(q, cm); }
void or_set_rigid(or_solver *s, const double *q) { s->impl->set_rigid(q); }
void or_debug_fetch(or_solver *s, int which, double *out) { s->impl->fetch(which, out); }
void or_debug_fetch_int(or_solver *s, int which, int *out) { s->impl->fetch_int(which, out); }

double or_rollout_loss(const bp_scene_desc *scene, const bp_params *params, int substeps, float subdt,
                       const double *target, double norm, const double *theta, const double *v0) {
    return rollout_loss<double>(scene, params, substeps, subdt, target, norm, theta, v0, nullptr, 0, nullptr);
}

double or_rollout_tangent(const bp_scene_desc *scene, const bp_params *params, int substeps, float subdt,
                          const double *target, double norm, const double *vdirs, int ndirs_v, double *dloss) {
    if (ndirs_v > kDirs - 6) ndirs_v = kDirs - 6;
    return rollout_loss<Dual>(scene, params, substeps, subdt, target, norm, nullptr, nullptr, vdirs, ndirs_v, dloss);
}

int or_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}

void or_neighbor_grid(const float *xs, int n, const float cs[3], const int dims[3], int *keys, int *perm,
                      int *cell_start, int *cell_end) {
    Grid g;
    for (int a = 0; a < 3; ++a) { g.cs[a] = cs[a]; g.dims[a] = dims[a]; g.mask[a] = (uint32_t)dims[a] - 1u; }
    std::vector<int> k(n), sk, pm;
    for (int i = 0; i < n; ++i) {
        int gp[3];
        g.pos(xs + 3 * i, gp);
        k[i] = g.addr(gp[0], gp[1], gp[2]);
    }
    stable_sort_keys(k, g.num_cells(), sk, pm);
    std::vector<int> st(g.num_cells(), -1), en(g.num_cells(), -1);
    find_ranges(sk, st, en);
    if (keys) std::copy(sk.begin(), sk.end(), keys);
    if (perm) std::copy(pm.begin(), pm.end(), perm);
    if (cell_start) std::copy(st.begin(), st.end(), cell_start);
    if (cell_end) std::copy(en.begin(), en.end(), cell_end);
}

int or_cell_address(const float p[3], const float cs[3], const int dims[3]) {
    Grid g;
    for (int a = 0; a < 3; ++a) { g.cs[a] = cs[a]; g.dims[a] = dims[a]; g.mask[a] = (uint32_t)dims[a] - 1u; }
    int gp[3];
    g.pos(p, gp);
    return g.addr(gp[0], gp[1], gp[2]);
}

double or_poly6(double r2, double h, int fp64) {
    if (fp64) { KConst<double> k; k.set(h); return k.poly6(r2); }
    KConst<float> k; k.set((float)h); return (double)k.poly6((float)r2);
}
void or_grad_spiky(const double v[3], double r2, double h, int fp64, double out[3]) {
    if (fp64) {
        KConst<double> k; k.set(h);
        V3<double> g = k.grad(V3<double>{v[0], v[1], v[2]}, r2);
        out[0] = g.x; out[1] = g.y; out[2] = g.z;
    } else {
        KConst<float> k; k.set((float)h);
        V3<float> g = k.grad(V3<float>{(float)v[0], (float)v[1], (float)v[2]}, (float)r2);
        out[0] = g.x; out[1] = g.y; out[2] = g.z;
    }
}
double or_akinci_c(double r, double h, int fp64) {
    if (fp64) { KConst<double> k; k.set(h); return k.akinci(r); }
    KConst<float> k; k.set((float)h); return (double)k.akinci((float)r);
}
void or_extract_rotation(const double A[9], double q[4], int iterations, int fp64) {
    if (fp64) {
        V3<double> a[3] = {{A[0], A[1], A[2]}, {A[3], A[4], A[5]}, {A[6], A[7], A[8]}};
        Q4<double> r = extract_rotation(a, Q4<double>{q[0], q[1], q[2], q[3]}, iterations);
        q[0] = r.x; q[1] = r.y; q[2] = r.z; q[3] = r.w;
    } else {
        V3<float> a[3] = {{(float)A[0], (float)A[1], (float)A[2]}, {(float)A[3], (float)A[4], (float)A[5]},
                          {(float)A[6], (float)A[7], (float)A[8]}};
        Q4<float> r = extract_rotation(a, Q4<float>{(float)q[0], (float)q[1], (float)q[2], (float)q[3]}, iterations);
        q[0] = r.x; q[1] = r.y; q[2] = r.z; q[3] = r.w;
    }
}

}  // extern "C"
