// Depth image -> the solver's measurement maps: CPU restatement (the ORACLE).
//
// TEST INFRASTRUCTURE ONLY (as sh_oracle.hpp): the checker for the HIP
// measurement-map kernels.  Templated on the arithmetic type T of the geometry
// (centroids, covariances, eigenvectors); the depth conversion and the clip /
// reliability thresholds are float, as in the reference, so that both
// precisions take the reference's decisions on the same data.
//
// Path (examples/mono_star/measure/MeasureProcessorOffline.cu:81 ->
// SurfelMapInitializer::InitFromRGBDImage, SurfelMapInitializer.cu:213-301):
//   filterUnreliableDepth        raw 16-bit mm -> filtered depth [m]
//   computeRawVertexFromDepth    filtered depth -> raw vertex map
//   filterAndScaleVertex         nearest downsample + near/far clip
//   createNormalRadiusMap        7x7 PCA normal + surfel radius
#pragma once
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <vector>

#include "sh_oracle.hpp"

namespace sho
{
struct DepthCamera
{
    int width, height;         // raw depth image
    float fx, fy, cx, cy;      // raw intrinsics
    float clip_near, clip_far; // metres
    float scale;               // downsample scale
    float radius_scale;        // surfel_radius_scale
};

// is_zero_vertex (libs/star/math/vector_ops.hpp:122-126): float components
// compared with the double literal 1e-3.
inline bool zero_vertex(float x, float y, float z)
{
    return std::fabs(double(x)) < 1e-3 && std::fabs(double(y)) < 1e-3 && std::fabs(double(z)) < 1e-3;
}
template <class T>
inline bool zero_vertex_t(T x, T y, T z)
{
    return std::fabs(double(x)) < 1e-3 && std::fabs(double(y)) < 1e-3 && std::fabs(double(z)) < 1e-3;
}

// ------------------------------------------------ smallest eigenvector of a PSD 3x3
// eigen33 (libs/star/math/eigen33.hpp), the PCL closed-form solver.  The
// thresholds are the float constants of the reference in both precisions.
template <class T>
struct Eig33
{
    static constexpr T eps() { return T(FLT_EPSILON); }

    // roots of x^2 - b x + c (eigen33.hpp:6-17); roots.x = 0
    static void roots2(T b, T c, V3<T> &r)
    {
        r.x = T(0);
        T d = b * b - T(4) * c;
        if (d < T(0))
            d = T(0);
        const T sd = std::sqrt(d);
        r.z = T(0.5) * (b + sd);
        r.y = T(0.5) * (b - sd);
    }

    // roots of x^3 - c2 x^2 + c1 x - c0 in ascending order (eigen33.hpp:58-98)
    static void roots3(T c0, T c1, T c2, V3<T> &r)
    {
        if (std::fabs(c0) < eps())
        {
            roots2(c2, c1, r);
            return;
        }
        const T inv3 = T(1) / T(3);
        const T sqrt3 = std::sqrt(T(3));
        const T c2_3 = c2 * inv3;
        T a_3 = (c1 - c2 * c2_3) * inv3;
        if (a_3 > T(0))
            a_3 = T(0);
        const T half_b = T(0.5) * (c0 + c2_3 * (T(2) * c2_3 * c2_3 - c1));
        T q = half_b * half_b + a_3 * a_3 * a_3;
        if (q > T(0))
            q = T(0);
        const T rho = std::sqrt(-a_3);
        const T theta = std::atan2(std::sqrt(-q), half_b) * inv3;
        const T ct = std::cos(theta), st = std::sin(theta);
        r.x = c2_3 + T(2) * rho * ct;
        r.y = c2_3 - rho * (ct + sqrt3 * st);
        r.z = c2_3 - rho * (ct - sqrt3 * st);
        if (r.x >= r.y)
            std::swap(r.x, r.y);
        if (r.y >= r.z)
        {
            std::swap(r.y, r.z);
            if (r.x >= r.y)
                std::swap(r.x, r.y);
        }
        if (r.x <= T(0))
            roots2(c2, c1, r);
    }

    static bool much_smaller(T x, T y) { return x * x <= eps() * eps() * y * y; }

    // a unit vector orthogonal to src (eigen33.hpp:19-56)
    static V3<T> unit_orthogonal(V3<T> s)
    {
        if (!much_smaller(s.x, s.z) || !much_smaller(s.y, s.z))
        {
            const T inv = T(1) / std::sqrt(s.x * s.x + s.y * s.y);
            return {-s.y * inv, s.x * inv, T(0)};
        }
        const T inv = T(1) / std::sqrt(s.z * s.z + s.y * s.y);
        return {T(0), -s.z * inv, s.y * inv};
    }

    static V3<T> normalized(V3<T> v)
    {
        const T inv = T(1) / std::sqrt(dot(v, v));
        return {v.x * inv, v.y * inv, v.z * inv};
    }

    // eigenvector of lambda: the longest of the three row cross products of
    // (A - lambda I), normalised; `chosen` is its squared length, `third` the
    // third product's (the reference's min/max bookkeeping reads it).
    static V3<T> null_vector(const T m[6], T lambda, T &chosen, T &third)
    {
        const V3<T> r0{m[0] - lambda, m[1], m[2]}, r1{m[1], m[3] - lambda, m[4]}, r2{m[2], m[4], m[5] - lambda};
        const V3<T> v0 = cross(r0, r1), v1 = cross(r0, r2), v2 = cross(r1, r2);
        const T l1 = dot(v0, v0), l2 = dot(v1, v1), l3 = dot(v2, v2);
        third = l3;
        V3<T> v;
        if (l1 >= l2 && l1 >= l3)
        {
            chosen = l1;
            v = v0;
        }
        else if (l2 >= l1 && l2 >= l3)
        {
            chosen = l2;
            v = v1;
        }
        else
        {
            chosen = l3;
            v = v2;
        }
        const T inv = T(1) / std::sqrt(chosen);
        return {v.x * inv, v.y * inv, v.z * inv};
    }

    // evecs[0], the eigenvector of the smallest eigenvalue (eigen33.hpp:100-306)
    static V3<T> smallest(const T cov[6])
    {
        T m[6];
        T scale = T(0);
        for (int i = 0; i < 6; ++i)
            scale = std::max(scale, std::fabs(cov[i]));
        if (scale <= T(FLT_MIN))
            scale = T(1);
        for (int i = 0; i < 6; ++i)
            m[i] = cov[i] / scale;
        // m = [m00 m01 m02; . m11 m12; . . m22] packed as 0 1 2 / 3 4 / 5
        const T c0 = m[0] * m[3] * m[5] + T(2) * m[1] * m[2] * m[4] - m[0] * m[4] * m[4] - m[3] * m[2] * m[2] -
                     m[5] * m[1] * m[1];
        const T c1 = m[0] * m[3] - m[1] * m[1] + m[0] * m[5] - m[2] * m[2] + m[3] * m[5] - m[4] * m[4];
        const T c2 = m[0] + m[3] + m[5];
        V3<T> ev;
        roots3(c0, c1, c2, ev);
        V3<T> e[3];
        T chosen, third;
        if (ev.z - ev.x <= eps())
            return {T(1), T(0), T(0)};
        if (ev.y - ev.x <= eps())
        {
            // the two smallest coincide
            e[2] = null_vector(m, ev.z, chosen, third);
            e[1] = unit_orthogonal(e[2]);
            return cross(e[1], e[2]);
        }
        if (ev.z - ev.y <= eps())
            return null_vector(m, ev.x, chosen, third); // the two largest coincide
        T mmax[3];
        unsigned min_el = 2, max_el = 2;
        e[2] = null_vector(m, ev.z, mmax[2], third);
        e[1] = null_vector(m, ev.y, mmax[1], third);
        min_el = mmax[1] <= mmax[min_el] ? 1u : min_el;
        max_el = mmax[1] > mmax[max_el] ? 1u : max_el;
        e[0] = null_vector(m, ev.x, mmax[0], third);
        // the reference compares the third product's length here in every branch
        min_el = third <= mmax[min_el] ? 0u : min_el;
        max_el = third > mmax[max_el] ? 0u : max_el;
        const unsigned mid_el = 3u - min_el - max_el;
        e[min_el] = normalized(cross(e[(min_el + 1) % 3], e[(min_el + 2) % 3]));
        e[mid_el] = normalized(cross(e[(mid_el + 1) % 3], e[(mid_el + 2) % 3]));
        return e[0];
    }
};

// -------------------------------------------------------------- 7x7 normal
// The window normal of createNormalRadiusMapKernel (normal_map.cu:32-117) and
// of filterUnreliableDepthKernel (surfel_reliable_filter.cu:64-121): centroid
// of the valid points of the 7x7 window, covariance over the 6x6 window
// [x-3, x+2] x [y-3, y+2] (the loops' "<" bound), smallest eigenvector,
// oriented against the centre vertex.  `vert(x, y, V3&)` returns false for
// a zero vertex.  Returns false when the centre is zero or at most 24
// points are valid.
template <class T, class F>
bool window_normal(int x, int y, V3<T> center, F &&vert, V3<T> &normal)
{
    V3<T> c{T(0), T(0), T(0)};
    int counter = 0;
    for (int cy = y - 3; cy <= y + 3; ++cy)
        for (int cx = x - 3; cx <= x + 3; ++cx)
        {
            V3<T> p;
            if (vert(cx, cy, p))
            {
                c = c + p;
                ++counter;
            }
        }
    if (counter <= 24)
        return false;
    const T inv = T(1) / T(counter);
    c = {c.x * inv, c.y * inv, c.z * inv};
    T cov[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
    for (int cy = y - 3; cy < y + 3; ++cy)
        for (int cx = x - 3; cx < x + 3; ++cx)
        {
            V3<T> p;
            if (vert(cx, cy, p))
            {
                const V3<T> d{p.x - c.x, p.y - c.y, p.z - c.z};
                cov[0] += d.x * d.x;
                cov[1] += d.x * d.y;
                cov[2] += d.x * d.z;
                cov[3] += d.y * d.y;
                cov[4] += d.y * d.z;
                cov[5] += d.z * d.z;
            }
        }
    V3<T> n = Eig33<T>::smallest(cov);
    if (dot(n, center) >= T(0))
        n = {-n.x, -n.y, -n.z};
    normal = n;
    return true;
}

// compute_radius (normal_map.cu:9-22): [mm]
template <class T>
inline T surfel_radius(T depth, T nz, float radius_scale)
{
    const T radius = depth * T(1000 * 1.414f / 570.0f) * T(radius_scale);
    nz = std::fabs(nz);
    T rn = T(2) * radius;
    if (nz > T(0.5))
        rn = radius / nz;
    return rn;
}

// ---------------------------------------------------- filterUnreliableDepth
// (surfel_reliable_filter.cu:37-191).  Raw depth [mm] at texel (x, y), zero
// outside the image (border addressing).  out: 0 = unreliable, 10 =
// background (beyond clip_far), else depth [m].  camera_dot (optional): the
// reliability statistic, NaN where it was not evaluated.
template <class T>
void filter_depth(const uint16_t *raw, const DepthCamera &cam, float *out, float *camera_dot)
{
    const int W = cam.width, H = cam.height;
    const float inv_fx = 1.f / cam.fx, inv_fy = 1.f / cam.fy;
    auto texel = [&](int x, int y) -> uint32_t {
        return (x < 0 || y < 0 || x >= W || y >= H) ? 0u : raw[size_t(y) * W + x];
    };
    // get_vertex_confid (:14-27)
    auto vertex = [&](int x, int y) -> V3<T> {
        const T z = T(float(texel(x, y)) / 1000.f);
        return {(T(x) - T(cam.cx)) * T(inv_fx) * z, (T(y) - T(cam.cy)) * T(inv_fy) * z, z};
    };
    auto valid_vertex = [&](int x, int y, V3<T> &p) {
        p = vertex(x, y);
        return !zero_vertex_t(p.x, p.y, p.z);
    };
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
        {
            const uint32_t rd = texel(x, y);
            const float dm = float(rd) / 1000.f;
            bool reliable = true, background = false;
            float cdot = NAN;
            if (std::fabs(dm) < cam.clip_near)
                reliable = false;
            if (std::fabs(dm) > cam.clip_far)
                background = true;
            else
            {
                V3<T> nc{T(0), T(0), T(0)};
                const V3<T> vc = vertex(x, y);
                if (!zero_vertex_t(vc.x, vc.y, vc.z))
                {
                    V3<T> n;
                    if (window_normal<T>(x, y, vc, valid_vertex, n))
                        nc = n;
                }
                const T cn = std::sqrt(dot(vc, vc));
                const T ci = T(1e-6f) + cn;
                const V3<T> cd{vc.x / ci, vc.y / ci, vc.z / ci};
                const T nn = std::sqrt(dot(nc, nc));
                const T ni = T(1e-6f) + nn;
                const V3<T> nd{nc.x / ni, nc.y / ni, nc.z / ni};
                const T d = dot(cd, nd);
                cdot = float(d);
                if (d > T(-0.1f))
                    reliable = false;
            }
            float o;
            if (!reliable)
                o = 0.f;
            else if (background)
                o = 10.f;
            else
                o = float(rd) / 1000.f;
            out[size_t(y) * W + x] = o;
            if (camera_dot)
                camera_dot[size_t(y) * W + x] = cdot;
        }
}

// Output size of filterAndScaleVertex (SurfelMapInitializer.cu:319-320)
inline void scaled_size(const DepthCamera &cam, int &w, int &h)
{
    w = int(unsigned(float(cam.width) * cam.scale));
    h = int(unsigned(float(cam.height) * cam.scale));
}

// ------------------------------------------- the downsampled vertex map
// computeRawVertexFromDepth (ComputeScaledVertexFromDepthKernel with
// scale_inv = 1, SurfelMapInitializer.cu:83-110, 303-317) and
// filterAndScaleVertex (:112-137, 319-343): the vertex of raw texel
// (int(x / scale), int(y / scale)), zero (w = 0 too) outside [clip_near,
// clip_far].  vertex4: scaled w * h * 4.
template <class T>
void scaled_vertex_map(const float *filtered, const DepthCamera &cam, T *vertex4)
{
    int w, h;
    scaled_size(cam, w, h);
    const float sinv = 1.f / cam.scale;
    const int W = cam.width, H = cam.height;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
        {
            const int rx = int(float(x) * sinv), ry = int(float(y) * sinv);
            const float d = (rx < W && ry < H) ? filtered[size_t(ry) * W + rx] : 0.f;
            T *o = vertex4 + (size_t(y) * w + x) * 4;
            if (std::fabs(d) > cam.clip_far || std::fabs(d) < cam.clip_near)
            {
                o[0] = o[1] = o[2] = o[3] = T(0);
                continue;
            }
            // the raw vertex (its own clip at 1e-6 and 10 m never fires inside [near, far])
            if (std::fabs(d) < 1e-6f || std::fabs(d) > 10.f)
            {
                o[0] = o[1] = o[2] = T(0);
                o[3] = T(1);
                continue;
            }
            o[0] = T((float(rx) - cam.cx) / cam.fx * d);
            o[1] = T((float(ry) - cam.cy) / cam.fy * d);
            o[2] = T(d);
            o[3] = T(1);
        }
}

// --------------------------------------------------- createNormalRadiusMap
// (normal_map.cu:32-136) on a w x h float4 vertex map (zero outside).
template <class T>
void normal_radius_map(const T *vertex4, int w, int h, float radius_scale, T *normal4)
{
    auto at = [&](int x, int y, V3<T> &p) {
        if (x < 0 || y < 0 || x >= w || y >= h)
        {
            p = {T(0), T(0), T(0)};
            return false;
        }
        const T *v = vertex4 + (size_t(y) * w + x) * 4;
        p = {v[0], v[1], v[2]};
        return !zero_vertex_t(p.x, p.y, p.z);
    };
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
        {
            T *o = normal4 + (size_t(y) * w + x) * 4;
            o[0] = o[1] = o[2] = o[3] = T(0);
            V3<T> vc;
            if (!at(x, y, vc))
                continue;
            V3<T> n;
            if (!window_normal<T>(x, y, vc, at, n))
                continue;
            o[0] = n.x;
            o[1] = n.y;
            o[2] = n.z;
            o[3] = surfel_radius<T>(vc.z, n.z, radius_scale);
        }
}

// createVertexConfidMap (vertex_map.cu:13-33): vertex of every raw texel, w = 1
template <class T>
void vertex_confid_map(const uint16_t *raw, int W, int H, float fx, float fy, float cx, float cy, T *vertex4)
{
    const float ifx = float(1.0 / double(fx)), ify = float(1.0 / double(fy)); // inverse(), common_utils.h:51-58
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
        {
            const float z = float(raw[size_t(y) * W + x]) / 1000.f;
            T *o = vertex4 + (size_t(y) * W + x) * 4;
            o[0] = T((float(x) - cx) * ifx * z);
            o[1] = T((float(y) - cy) * ify * z);
            o[2] = T(z);
            o[3] = T(1);
        }
}
} // namespace sho
