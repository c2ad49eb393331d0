// Seeded analytic integrands that stand in for the reference's OptiX radiance
// programs (SDK/cuda/whitted.cu:164-293 `__raygen__pinhole_mdas_dof/_mb`,
// SDK/cuda/ao.cu:112-152).  They honour the same buffer contract
// (SDK/cuda/whitted.h:76-81): sample i is read as `dim` consecutive floats at
// coords[dim*i], mapped to NDC as d = 2*(x/scale.x, y/scale.y) - 1, the other
// dimensions are time / lens / hemisphere, and one float3 is written to
// values[i].
//
// The header is compiled twice: by hipcc into the library's kernels and by g++
// into the CPU restatement used by the tests, which needs the same inputs.
// Every operation is a single IEEE-754 operation (+ - * / sqrt, min/max,
// compares); the two trigonometric functions are evaluated by a fixed Taylor
// polynomial in double precision.  With FP contraction disabled on both
// compilers the float results are bit-identical on CPU and GPU, so a whole
// MDAS run (which feeds these values back into the error estimate) can be
// compared bit for bit.
//
// The `real` template parameter also gives an FP64 evaluation of the same
// scene, used for high-sample-count ground-truth images.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define MDAS_HD __host__ __device__ inline
#else
#define MDAS_HD inline
#endif

namespace mscene {

enum Kind {
    SCENE_CONSTANT = 0,   // constant colour; closed-form test integrand
    SCENE_DISKS_MB = 1,   // 16 moving opaque disks over black, (x, y, t)
    SCENE_SPHERES_DOF = 2,  // thin lens + Lambertian spheres + plane, (x, y, u, v)
    SCENE_SPHERES_MBDOF = 3,  // moving spheres + thin lens, (x, y, t, u, v)
    SCENE_AO = 4,         // thin lens + ambient occlusion, (x, y, u, v, h1, h2)
    SCENE_RAMP = 5,       // value = (x/sX, y/sY, t); closed-form test integrand
};

static const int MAX_OBJ = 256;
static const int OBJ_STRIDE = 12;  // floats per object record

// Flat parameter block.  Layout is part of the C ABI (mdas_scene_params()).
struct Params {
    int32_t kind;
    int32_t n_obj;
    int32_t dim;
    int32_t pad0;
    float scale_x, scale_y;   // sample-space extent in x and y
    float eye[3];             // camera position; the camera looks along +z
    float tan_x, tan_y;       // |U| and |V| of the camera frame (|W| = 1)
    float lens_radius;        // 0 = pinhole
    float focal_distance;
    float light_dir[3];       // unit vector towards the directional light
    float light_color[3];
    float ambient;
    float plane_y;            // ground plane height
    float plane_albedo[3];
    float ao_radius;
    float constant[3];        // SCENE_CONSTANT colour
    float pad1[5];
    // disks:   cx cy r vx vy R G B
    // spheres: cx cy cz r vx vy vz R G B
    float obj[MAX_OBJ * OBJ_STRIDE];
};

// ---------------------------------------------------------------- helpers --

// cos and sin of |phi| <= pi/4 by Taylor series in double precision (the
// series truncation error is below 1e-17 there), rounded once to `real`.
template <typename real>
MDAS_HD void cos_sin_quarter(real phi, real* c, real* s) {
    double x = (double)phi;
    double x2 = x * x;
    // sin x = x (1 - x2/3! (1 - x2/(4*5) (1 - x2/(6*7) (...))))
    double sp = 1.0 - x2 / 342.0;           // 18*19
    sp = 1.0 - x2 / 272.0 * sp;             // 16*17
    sp = 1.0 - x2 / 210.0 * sp;             // 14*15
    sp = 1.0 - x2 / 156.0 * sp;             // 12*13
    sp = 1.0 - x2 / 110.0 * sp;             // 10*11
    sp = 1.0 - x2 / 72.0 * sp;              // 8*9
    sp = 1.0 - x2 / 42.0 * sp;              // 6*7
    sp = 1.0 - x2 / 20.0 * sp;              // 4*5
    sp = 1.0 - x2 / 6.0 * sp;               // 2*3
    double sn = x * sp;
    double cp = 1.0 - x2 / 306.0;           // 17*18
    cp = 1.0 - x2 / 240.0 * cp;             // 15*16
    cp = 1.0 - x2 / 182.0 * cp;             // 13*14
    cp = 1.0 - x2 / 132.0 * cp;             // 11*12
    cp = 1.0 - x2 / 90.0 * cp;              // 9*10
    cp = 1.0 - x2 / 56.0 * cp;              // 7*8
    cp = 1.0 - x2 / 30.0 * cp;              // 5*6
    cp = 1.0 - x2 / 12.0 * cp;              // 3*4
    cp = 1.0 - x2 / 2.0 * cp;               // 1*2
    *c = (real)cp;
    *s = (real)sn;
}

template <typename real>
MDAS_HD real r_sqrt(real x);
template <>
MDAS_HD float r_sqrt<float>(float x) { return __builtin_sqrtf(x); }
template <>
MDAS_HD double r_sqrt<double>(double x) { return __builtin_sqrt(x); }

template <typename real>
MDAS_HD real r_abs(real x) { return x < (real)0 ? -x : x; }

// Shirley-Chiu concentric map of [0,1)^2 onto the unit disk.
template <typename real>
MDAS_HD void concentric_disk(real a, real b, real* px, real* py) {
    real ox = (real)2 * a - (real)1;
    real oy = (real)2 * b - (real)1;
    if (ox == (real)0 && oy == (real)0) { *px = (real)0; *py = (real)0; return; }
    const real pi4 = (real)0.78539816339744830962;
    real r, c, s;
    if (r_abs(ox) > r_abs(oy)) {
        r = ox;
        cos_sin_quarter(pi4 * (oy / ox), &c, &s);
    } else {
        r = oy;
        // theta = pi/2 - phi: cos(theta) = sin(phi), sin(theta) = cos(phi)
        cos_sin_quarter(pi4 * (ox / oy), &s, &c);
    }
    *px = r * c;
    *py = r * s;
}

template <typename real>
struct V3 { real x, y, z; };

template <typename real>
MDAS_HD V3<real> v3(real x, real y, real z) { V3<real> r; r.x = x; r.y = y; r.z = z; return r; }
template <typename real>
MDAS_HD real dot3(V3<real> a, V3<real> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename real>
MDAS_HD V3<real> normalize3(V3<real> a) {
    real inv = (real)1 / r_sqrt(dot3(a, a));
    return v3(a.x * inv, a.y * inv, a.z * inv);
}

// Ray / sphere: smallest t in (tmin, tmax); returns tmax when missed.
template <typename real>
MDAS_HD real hit_sphere(V3<real> o, V3<real> d, real cx, real cy, real cz, real r,
                        real tmin, real tmax) {
    V3<real> oc = v3(o.x - cx, o.y - cy, o.z - cz);
    real b = dot3(oc, d);
    real c = dot3(oc, oc) - r * r;
    real disc = b * b - c;
    if (disc < (real)0) return tmax;
    real sq = r_sqrt(disc);
    real t = -b - sq;
    if (t <= tmin) t = -b + sq;
    if (t <= tmin || t >= tmax) return tmax;
    return t;
}

// Camera ray for NDC (dx, dy) and lens sample (lu, lv); camera looks along +z.
template <typename real>
MDAS_HD void camera_ray(const Params& P, real dx, real dy, real lu, real lv, bool lens,
                        V3<real>* o, V3<real>* d) {
    if (!lens || P.lens_radius <= 0.0f) {
        *o = v3((real)P.eye[0], (real)P.eye[1], (real)P.eye[2]);
        *d = normalize3(v3(dx * (real)P.tan_x, dy * (real)P.tan_y, (real)1));
        return;
    }
    real lx, ly;
    concentric_disk(lu, lv, &lx, &ly);
    lx = (real)P.lens_radius * lx;
    ly = (real)P.lens_radius * ly;
    V3<real> dc = normalize3(v3(dx * (real)P.tan_x, dy * (real)P.tan_y, (real)1));
    real ft = (real)P.focal_distance / dc.z;
    V3<real> dir = v3(ft * dc.x - lx, ft * dc.y - ly, ft * dc.z);
    *o = v3((real)P.eye[0] + lx, (real)P.eye[1] + ly, (real)P.eye[2]);
    *d = normalize3(dir);
}

// ------------------------------------------------ grouped sphere tests --
// The spheres of a scene in groups of up to GRP_SIZE (median splits of their
// centres), each group with a bounding sphere (centre Gc, radius R) that
// holds its members over the whole shutter t in [0, 1].  A ray skips a group
// only where no member's own test (hit_sphere, in the same arithmetic) can
// report a hit that counts, so the closest hit over the other members, ties
// to the lowest index, is the one a loop over all spheres finds, bit for bit.
//
// The bound.  With X = o - centre, the member test finds a root when
// (X.d)^2 - |X|^2 + r^2 >= 0 in rounded arithmetic.  Writing |d|^2 = 1 + e,
// |X|^2 - (X.d)^2 = |X_perp|^2 - e X_par^2, and the rounding of that
// expression stays below 2e-6 |X|^2 (about 13 unit roundoffs), so a member
// can report a root only if |X_perp| <= r + sqrt(G) with
// G = (2e-6 + |e|) (|o - Gc| + R)^2; its roots then lie within r + sqrt(G)
// of its centre's projection, which is within R - r of the group's.  The group
// test widens R by S = sqrt(G) (with G taken as (2e-6 + |e|) 2 (l2 + R^2),
// l2 = |o - Gc|^2), a further |e| R + kCullSlack, and discounts its own
// rounding and the direction's length from the squared line distance.
static const int MAX_GRP = 64;
static const int GRP_SIZE = 8;
constexpr float kCullSlack = 0.01f;

struct Accel {
    int32_t n_grp;               // 0: no groups, every sphere is tested
    int32_t pad[3];
    float grp[MAX_GRP * 4];      // bounding sphere of each group: centre, radius
    int32_t range[MAX_GRP * 2];  // first member position and member count
    int32_t member[MAX_OBJ];     // sphere indices, group by group
};

// A scene as the integrand kernels read it: the parameters, then the groups.
struct SceneBlock {
    Params p;
    Accel accel;
};

// Whether a member of group g may report a hit in (tmin, tmax) for the ray
// (o, d); e = |d.d - 1|.
template <typename real>
MDAS_HD bool group_may_hit(const float* g, V3<real> o, V3<real> d, real e, real tmin,
                           real tmax) {
    const real gx = (real)g[0] - o.x, gy = (real)g[1] - o.y, gz = (real)g[2] - o.z;
    const real R = (real)g[3];
    const real l2 = gx * gx + gy * gy + gz * gz;
    const real p = gx * d.x + gy * d.y + gz * d.z;  // centre's distance along the ray
    const real k = (real)2e-6 + e;
    const real S = r_sqrt(k * (real)2 * (l2 + R * R)) + e * R + (real)kCullSlack;
    const real W = R + S;
    return l2 - p * p - k * l2 <= W * W && p + W >= tmin && p - W <= tmax;
}

// Closest hit among spheres (centres moved to time t) and the ground plane.
// Returns the hit distance (tmax on a miss); *id = sphere index, -1 plane, -2 miss.
// With groups (A != null, A->n_grp > 0) the spheres of the groups the ray can
// meet are tested group by group; the result is the same.
template <typename real>
MDAS_HD real trace(const Params& P, const float* obj, V3<real> o, V3<real> d, real t,
                   real tmin, real tmax, int* id, const Accel* A = nullptr) {
    real best = tmax;
    *id = -2;
    if (A != nullptr && A->n_grp > 0) {
        const real e = r_abs(dot3(d, d) - (real)1);
        for (int g = 0; g < A->n_grp; ++g) {
            if (!group_may_hit(A->grp + 4 * g, o, d, e, tmin, best)) continue;
            const int j0 = A->range[2 * g], j1 = j0 + A->range[2 * g + 1];
            for (int j = j0; j < j1; ++j) {
                const int k = A->member[j];
                const float* s = obj + k * OBJ_STRIDE;
                real cx = (real)s[0] + (real)s[4] * t;
                real cy = (real)s[1] + (real)s[5] * t;
                real cz = (real)s[2] + (real)s[6] * t;
                real th = hit_sphere(o, d, cx, cy, cz, (real)s[3], tmin, tmax);
                if (th < best || (th == best && th < tmax && k < *id)) { best = th; *id = k; }
            }
        }
    } else {
        for (int k = 0; k < P.n_obj; ++k) {
            const float* s = obj + k * OBJ_STRIDE;
            real cx = (real)s[0] + (real)s[4] * t;
            real cy = (real)s[1] + (real)s[5] * t;
            real cz = (real)s[2] + (real)s[6] * t;
            real th = hit_sphere(o, d, cx, cy, cz, (real)s[3], tmin, best);
            if (th < best) { best = th; *id = k; }
        }
    }
    if (d.y != (real)0) {
        real tp = ((real)P.plane_y - o.y) / d.y;
        if (tp > tmin && tp < best) { best = tp; *id = -1; }
    }
    return best;
}

// Any hit in (tmin, tmax) among spheres and the plane (with groups: among
// the spheres of the groups the ray can meet, the same answer).
template <typename real>
MDAS_HD bool occluded(const Params& P, const float* obj, V3<real> o, V3<real> d, real t,
                      real tmin, real tmax, bool with_plane, const Accel* A = nullptr) {
    if (A != nullptr && A->n_grp > 0) {
        const real e = r_abs(dot3(d, d) - (real)1);
        for (int g = 0; g < A->n_grp; ++g) {
            if (!group_may_hit(A->grp + 4 * g, o, d, e, tmin, tmax)) continue;
            const int j0 = A->range[2 * g], j1 = j0 + A->range[2 * g + 1];
            for (int j = j0; j < j1; ++j) {
                const float* s = obj + A->member[j] * OBJ_STRIDE;
                real cx = (real)s[0] + (real)s[4] * t;
                real cy = (real)s[1] + (real)s[5] * t;
                real cz = (real)s[2] + (real)s[6] * t;
                if (hit_sphere(o, d, cx, cy, cz, (real)s[3], tmin, tmax) < tmax) return true;
            }
        }
    } else {
        for (int k = 0; k < P.n_obj; ++k) {
            const float* s = obj + k * OBJ_STRIDE;
            real cx = (real)s[0] + (real)s[4] * t;
            real cy = (real)s[1] + (real)s[5] * t;
            real cz = (real)s[2] + (real)s[6] * t;
            if (hit_sphere(o, d, cx, cy, cz, (real)s[3], tmin, tmax) < tmax) return true;
        }
    }
    if (with_plane && d.y != (real)0) {
        real tp = ((real)P.plane_y - o.y) / d.y;
        if (tp > tmin && tp < tmax) return true;
    }
    return false;
}

// Lambertian shading of a directional-light scene (DOF and MB+DOF scenes).
template <typename real>
MDAS_HD V3<real> shade_spheres(const Params& P, const float* obj, V3<real> o, V3<real> d,
                               real t, const Accel* A) {
    int id;
    const real big = (real)1e30;
    real th = trace(P, obj, o, d, t, (real)1e-3, big, &id, A);
    if (id == -2) {
        real a = (real)0.5 * (d.y + (real)1);
        return v3((real)1 - a * (real)0.5, (real)1 - a * (real)0.3, (real)1);
    }
    V3<real> p = v3(o.x + th * d.x, o.y + th * d.y, o.z + th * d.z);
    V3<real> n;
    real ar, ag, ab;
    if (id >= 0) {
        const float* s = obj + id * OBJ_STRIDE;
        real cx = (real)s[0] + (real)s[4] * t;
        real cy = (real)s[1] + (real)s[5] * t;
        real cz = (real)s[2] + (real)s[6] * t;
        real inv = (real)1 / (real)s[3];
        n = v3((p.x - cx) * inv, (p.y - cy) * inv, (p.z - cz) * inv);
        ar = (real)s[7]; ag = (real)s[8]; ab = (real)s[9];
    } else {
        n = v3((real)0, (real)1, (real)0);
        ar = (real)P.plane_albedo[0]; ag = (real)P.plane_albedo[1]; ab = (real)P.plane_albedo[2];
    }
    V3<real> l = v3((real)P.light_dir[0], (real)P.light_dir[1], (real)P.light_dir[2]);
    real ndl = dot3(n, l);
    real direct = (real)0;
    if (ndl > (real)0) {
        real eps = (real)1e-3;
        V3<real> so = v3(p.x + eps * n.x, p.y + eps * n.y, p.z + eps * n.z);
        if (!occluded(P, obj, so, l, t, (real)0, big, false, A)) direct = ndl;
    }
    real amb = (real)P.ambient;
    return v3(ar * (amb + direct * (real)P.light_color[0]),
              ag * (amb + direct * (real)P.light_color[1]),
              ab * (amb + direct * (real)P.light_color[2]));
}

// Evaluate one sample.  `x` holds P.dim coordinates; A: the scene's sphere
// groups, or null to test every sphere.
template <typename real>
MDAS_HD V3<real> eval(const Params& P, const float* obj, const real* x,
                      const Accel* A = nullptr) {
    real dx = (real)2 * (x[0] / (real)P.scale_x) - (real)1;
    real dy = (real)2 * (x[1] / (real)P.scale_y) - (real)1;
    switch (P.kind) {
    case SCENE_CONSTANT:
        return v3((real)P.constant[0], (real)P.constant[1], (real)P.constant[2]);
    case SCENE_RAMP:
        return v3(x[0] / (real)P.scale_x, x[1] / (real)P.scale_y,
                  P.dim > 2 ? x[2] : (real)0);
    case SCENE_DISKS_MB: {
        // screen space [0,1)^2, disks moving linearly over the shutter t in [0,1);
        // the highest index on top.
        real u = x[0] / (real)P.scale_x;
        real v = x[1] / (real)P.scale_y;
        real t = x[2];
        for (int k = P.n_obj - 1; k >= 0; --k) {
            const float* s = obj + k * OBJ_STRIDE;
            real cx = (real)s[0] + (real)s[3] * t;
            real cy = (real)s[1] + (real)s[4] * t;
            real ex = u - cx, ey = v - cy;
            real r = (real)s[2];
            if (ex * ex + ey * ey <= r * r) return v3((real)s[5], (real)s[6], (real)s[7]);
        }
        return v3((real)0, (real)0, (real)0);
    }
    case SCENE_SPHERES_DOF: {
        V3<real> o, d;
        camera_ray(P, dx, dy, x[2], x[3], true, &o, &d);
        return shade_spheres(P, obj, o, d, (real)0, A);
    }
    case SCENE_SPHERES_MBDOF: {
        V3<real> o, d;
        camera_ray(P, dx, dy, x[3], x[4], true, &o, &d);
        return shade_spheres(P, obj, o, d, x[2], A);
    }
    case SCENE_AO: {
        V3<real> o, d;
        camera_ray(P, dx, dy, x[2], x[3], true, &o, &d);
        int id;
        const real big = (real)1e30;
        real th = trace(P, obj, o, d, (real)0, (real)1e-3, big, &id, A);
        if (id == -2) return v3((real)1, (real)1, (real)1);
        V3<real> p = v3(o.x + th * d.x, o.y + th * d.y, o.z + th * d.z);
        V3<real> n;
        if (id >= 0) {
            const float* s = obj + id * OBJ_STRIDE;
            real inv = (real)1 / (real)s[3];
            n = v3((p.x - (real)s[0]) * inv, (p.y - (real)s[1]) * inv, (p.z - (real)s[2]) * inv);
        } else {
            n = v3((real)0, (real)1, (real)0);
        }
        // orthonormal basis around n (Duff et al. 2017, branchless ONB)
        real sign = n.z >= (real)0 ? (real)1 : (real)-1;
        real a = (real)-1 / (sign + n.z);
        real b = n.x * n.y * a;
        V3<real> t1 = v3((real)1 + sign * n.x * n.x * a, sign * b, -sign * n.x);
        V3<real> t2 = v3(b, sign + n.y * n.y * a, -n.y);
        // cosine-weighted direction from the concentric disk
        real hx, hy;
        concentric_disk(x[4], x[5], &hx, &hy);
        real hz2 = (real)1 - hx * hx - hy * hy;
        real hz = hz2 > (real)0 ? r_sqrt(hz2) : (real)0;
        V3<real> w = v3(hx * t1.x + hy * t2.x + hz * n.x,
                        hx * t1.y + hy * t2.y + hz * n.y,
                        hx * t1.z + hy * t2.z + hz * n.z);
        real eps = (real)1e-3;
        V3<real> so = v3(p.x + eps * n.x, p.y + eps * n.y, p.z + eps * n.z);
        bool occ = occluded(P, obj, so, w, (real)0, (real)0, (real)P.ao_radius, true, A);
        real vis = occ ? (real)0 : (real)1;
        return v3(vis, vis, vis);
    }
    default:
        return v3((real)0, (real)0, (real)0);
    }
}

// ------------------------------------------------------ scene generation --
// Host-side generator of the seeded scenes (a 4-round TEA hash of the seed
// starts a 32-bit LCG; draws are float(state & 0xFFFFFF) / 2^24).

struct SceneRng {
    uint32_t s;
    explicit SceneRng(uint32_t seed) {
        uint32_t v0 = seed, v1 = 0x5ce7e5u, sum = 0;
        for (int n = 0; n < 4; ++n) {
            sum += 0x9e3779b9u;
            v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + sum) ^ ((v1 >> 5) + 0xc8013ea4u);
            v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + sum) ^ ((v0 >> 5) + 0x7e95761eu);
        }
        s = v0;
    }
    float u() { s = 1664525u * s + 1013904223u; return (float)(s & 0x00FFFFFFu) / 16777216.0f; }
    float u(float lo, float hi) { return lo + (hi - lo) * u(); }
};

inline void make_scene(Params* P, int kind, uint32_t seed, int dim, float scale_x, float scale_y) {
    char* raw = (char*)P;
    for (unsigned i = 0; i < sizeof(Params); ++i) raw[i] = 0;
    P->kind = kind;
    P->dim = dim;
    P->scale_x = scale_x;
    P->scale_y = scale_y;
    // camera: 45 degree vertical field of view, looking along +z
    float aspect = scale_x / scale_y;
    P->tan_y = 0.41421356237f;  // tan(22.5 deg)
    P->tan_x = P->tan_y * aspect;
    P->lens_radius = 0.05f;
    P->focal_distance = 5.0f;
    // light direction normalize(-0.3, 1, -0.5)
    P->light_dir[0] = -0.259160519f; P->light_dir[1] = 0.863868415f; P->light_dir[2] = -0.431934208f;
    P->light_color[0] = 1.0f; P->light_color[1] = 0.95f; P->light_color[2] = 0.9f;
    P->ambient = 0.1f;
    P->plane_y = -2.0f;
    P->plane_albedo[0] = 0.5f; P->plane_albedo[1] = 0.5f; P->plane_albedo[2] = 0.5f;
    P->ao_radius = 1.0f;
    P->constant[0] = 0.25f; P->constant[1] = 0.5f; P->constant[2] = 0.75f;
    SceneRng rng(seed);
    switch (kind) {
    case SCENE_DISKS_MB:
        P->n_obj = 16;
        for (int k = 0; k < 16; ++k) {
            float* s = P->obj + k * OBJ_STRIDE;
            s[0] = rng.u(); s[1] = rng.u();
            s[2] = rng.u(0.02f, 0.1f);
            s[3] = rng.u(-0.1f, 0.1f); s[4] = rng.u(-0.1f, 0.1f);
            s[5] = rng.u(); s[6] = rng.u(); s[7] = rng.u();
        }
        break;
    case SCENE_SPHERES_DOF:
    case SCENE_SPHERES_MBDOF:
        P->n_obj = 64;
        for (int k = 0; k < 64; ++k) {
            float* s = P->obj + k * OBJ_STRIDE;
            s[0] = rng.u(-3.0f, 3.0f); s[1] = rng.u(-2.0f, 2.0f); s[2] = rng.u(3.0f, 12.0f);
            s[3] = rng.u(0.2f, 0.8f);
            float vx = rng.u(-0.3f, 0.3f), vy = rng.u(-0.3f, 0.3f), vz = rng.u(-0.3f, 0.3f);
            if (kind == SCENE_SPHERES_MBDOF) { s[4] = vx; s[5] = vy; s[6] = vz; }
            s[7] = rng.u(); s[8] = rng.u(); s[9] = rng.u();
        }
        break;
    case SCENE_AO:
        // the lens of the depth-of-field scenes (radius 0.05, focus distance
        // 5); the camera steps back to z = -12 so that the [-5, 5]^3 box of
        // spheres lies in front of it
        P->n_obj = 256;
        P->eye[2] = -12.0f;
        P->plane_y = -5.5f;
        for (int k = 0; k < 256; ++k) {
            float* s = P->obj + k * OBJ_STRIDE;
            s[0] = rng.u(-5.0f, 5.0f); s[1] = rng.u(-5.0f, 5.0f); s[2] = rng.u(-5.0f, 5.0f);
            s[3] = rng.u(0.1f, 0.5f);
            s[7] = 1.0f; s[8] = 1.0f; s[9] = 1.0f;
        }
        break;
    default:
        break;
    }
}

// Groups of a sphere scene (none for the others): each sphere as its sweep
// over t in [0, 1] (centre c + v/2, radius r + |v|/2), split at the median
// along the longest extent of the sweep centres until at most GRP_SIZE
// remain, groups ordered by the depth (z) of their centre so that a camera
// ray, which looks along +z, meets the near ones first.  Bounds in double,
// the radius rounded up.
inline void build_accel(const Params& P, Accel* A) {
    char* raw = (char*)A;
    for (unsigned i = 0; i < sizeof(Accel); ++i) raw[i] = 0;
    if (P.kind != SCENE_SPHERES_DOF && P.kind != SCENE_SPHERES_MBDOF && P.kind != SCENE_AO) return;
    const int n = P.n_obj;
    if (n <= 0 || n > MAX_OBJ) return;
    double mid[MAX_OBJ][3], ext[MAX_OBJ];
    int idx[MAX_OBJ];
    for (int k = 0; k < n; ++k) {
        const float* s = P.obj + k * OBJ_STRIDE;
        double vv = 0.0;
        for (int a = 0; a < 3; ++a) {
            mid[k][a] = (double)s[a] + 0.5 * (double)s[4 + a];
            vv += (double)s[4 + a] * (double)s[4 + a];
        }
        ext[k] = (double)s[3] + 0.5 * __builtin_sqrt(vv);
        idx[k] = k;
    }
    int sb[64], se[64], sp = 0, ng = 0;
    int gb[MAX_GRP], ge[MAX_GRP];
    sb[sp] = 0;
    se[sp++] = n;
    while (sp > 0) {
        const int b = sb[--sp], e = se[sp];
        if (e - b <= GRP_SIZE) {
            gb[ng] = b;
            ge[ng++] = e;
            continue;
        }
        int axis = 0;
        double best = -1.0;
        for (int a = 0; a < 3; ++a) {
            double lo = mid[idx[b]][a], hi = lo;
            for (int j = b + 1; j < e; ++j) {
                lo = mid[idx[j]][a] < lo ? mid[idx[j]][a] : lo;
                hi = mid[idx[j]][a] > hi ? mid[idx[j]][a] : hi;
            }
            if (hi - lo > best) { best = hi - lo; axis = a; }
        }
        for (int j = b + 1; j < e; ++j) {  // insertion sort by (coordinate, index)
            const int v = idx[j];
            int q = j - 1;
            while (q >= b && (mid[idx[q]][axis] > mid[v][axis] ||
                              (mid[idx[q]][axis] == mid[v][axis] && idx[q] > v))) {
                idx[q + 1] = idx[q];
                --q;
            }
            idx[q + 1] = v;
        }
        const int h = (b + e) / 2;
        sb[sp] = h;
        se[sp++] = e;
        sb[sp] = b;
        se[sp++] = h;
    }
    float gc[MAX_GRP][4];
    for (int g = 0; g < ng; ++g) {
        double lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = mid[idx[gb[g]]][a] - ext[idx[gb[g]]];
            hi[a] = mid[idx[gb[g]]][a] + ext[idx[gb[g]]];
            for (int j = gb[g] + 1; j < ge[g]; ++j) {
                const double l = mid[idx[j]][a] - ext[idx[j]], u = mid[idx[j]][a] + ext[idx[j]];
                lo[a] = l < lo[a] ? l : lo[a];
                hi[a] = u > hi[a] ? u : hi[a];
            }
            gc[g][a] = (float)(0.5 * (lo[a] + hi[a]));
        }
        double R = 0.0;
        for (int j = gb[g]; j < ge[g]; ++j) {
            double d2 = 0.0;
            for (int a = 0; a < 3; ++a) {
                const double t = mid[idx[j]][a] - (double)gc[g][a];
                d2 += t * t;
            }
            const double r = __builtin_sqrt(d2) + ext[idx[j]];
            R = r > R ? r : R;
        }
        gc[g][3] = (float)(R * (1.0 + 1e-6) + 1e-6);
    }
    int order[MAX_GRP];
    for (int g = 0; g < ng; ++g) {  // by centre depth, ties in split order
        int q = g - 1;
        while (q >= 0 && gc[order[q]][2] > gc[g][2]) {
            order[q + 1] = order[q];
            --q;
        }
        order[q + 1] = g;
    }
    int pos = 0;
    for (int r = 0; r < ng; ++r) {
        const int g = order[r];
        for (int a = 0; a < 4; ++a) A->grp[4 * r + a] = gc[g][a];
        A->range[2 * r] = pos;
        A->range[2 * r + 1] = ge[g] - gb[g];
        for (int j = gb[g]; j < ge[g]; ++j) A->member[pos++] = idx[j];
    }
    A->n_grp = ng;
}

}  // namespace mscene
