// Seeded synthetic tracking scene for the warp-field solver benchmarks/tests.
//
// Workload generator, not part of the solve: it produces the inputs a StarHub
// frame hands the solver (reference solver maps, measured vertex/normal maps,
// optical flow, surfels and nodes) from a noisy sphere deformed by a smooth
// per-node SE(3) field.  Host C++ with OpenMP; deterministic for a given seed
// and thread count independent.
//
//  * surfels: uniform on the unit sphere, radius scaled by 1 + amp * n(p) with
//    n a 3-octave value noise in [-1, 1]; normals from the surface gradient;
//  * nodes: farthest-point subsampling of the surfels (first candidate = surfel 0);
//  * sigma = 2 x mean nearest-node spacing, node_radius_square = sigma^2;
//  * ground truth: rotation about each node centre by a smooth random field
//    (|w| <= rot_max) plus a smooth translation (|t| <= trans_max);
//  * target: surfels warped by the dual-quaternion blend of their K nearest
//    nodes, rendered as oriented disks (ray/plane depth) into a z-buffer;
//    depth noise N(0, depth_noise) is added along the ray; normals are the
//    warped surfel normals (analytic), in the camera frame;
//  * reference maps: the unwarped surfels rendered the same way, storing the
//    surfel's world vertex / normal and its index;
//  * flow (optional): projected ground-truth displacement of the surfel seen
//    at each reference pixel + N(0, flow_noise), a fraction masked to (0, 0).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace
{
// ------------------------------------------------------------------- RNG
struct SplitMix
{
    uint64_t s;
    explicit SplitMix(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal()
    {
        double u1 = uniform();
        const double u2 = uniform();
        if (u1 < 1e-300)
            u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

inline uint64_t hash3(int64_t x, int64_t y, int64_t z, uint64_t seed)
{
    uint64_t h = seed ^ 0x2545f4914f6cdd1dull;
    h ^= uint64_t(x) * 0x9e3779b185ebca87ull;
    h = (h ^ (h >> 29)) * 0xbf58476d1ce4e5b9ull;
    h ^= uint64_t(y) * 0xc2b2ae3d27d4eb4full;
    h = (h ^ (h >> 32)) * 0x94d049bb133111ebull;
    h ^= uint64_t(z) * 0x165667b19e3779f9ull;
    h = (h ^ (h >> 29)) * 0xbf58476d1ce4e5b9ull;
    return h ^ (h >> 32);
}

struct D3
{
    double x, y, z;
};
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator*(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline D3 normalize(D3 a)
{
    const double n = std::sqrt(dot(a, a));
    return (1.0 / n) * a;
}

// value noise in [-1, 1]
double value_noise(D3 p, uint64_t seed)
{
    const double fx = std::floor(p.x), fy = std::floor(p.y), fz = std::floor(p.z);
    const int64_t ix = int64_t(fx), iy = int64_t(fy), iz = int64_t(fz);
    auto sm = [](double t) { return t * t * (3.0 - 2.0 * t); };
    const double tx = sm(p.x - fx), ty = sm(p.y - fy), tz = sm(p.z - fz);
    double c[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 2; ++d)
                c[a][b][d] = double(hash3(ix + a, iy + b, iz + d, seed) >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    const double x00 = c[0][0][0] + tx * (c[1][0][0] - c[0][0][0]);
    const double x10 = c[0][1][0] + tx * (c[1][1][0] - c[0][1][0]);
    const double x01 = c[0][0][1] + tx * (c[1][0][1] - c[0][0][1]);
    const double x11 = c[0][1][1] + tx * (c[1][1][1] - c[0][1][1]);
    const double y0 = x00 + ty * (x10 - x00);
    const double y1 = x01 + ty * (x11 - x01);
    return y0 + tz * (y1 - y0);
}

double fbm(D3 p, int octaves, uint64_t seed)
{
    double s = 0.0, amp = 1.0, norm = 0.0, f = 2.0;
    for (int o = 0; o < octaves; ++o)
    {
        s += amp * value_noise(f * p, seed + 7919ull * uint64_t(o));
        norm += amp;
        amp *= 0.5;
        f *= 2.0;
    }
    return s / norm;
}

// ------------------------------------------------------------ DQ helpers
struct Qd
{
    double w, x, y, z;
};
inline Qd qmul(Qd a, Qd b)
{
    return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x, a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}
struct Rt
{
    double R[3][3];
    D3 t;
};
inline D3 apply(const Rt &T, D3 p)
{
    return {T.R[0][0] * p.x + T.R[0][1] * p.y + T.R[0][2] * p.z + T.t.x,
            T.R[1][0] * p.x + T.R[1][1] * p.y + T.R[1][2] * p.z + T.t.y,
            T.R[2][0] * p.x + T.R[2][1] * p.y + T.R[2][2] * p.z + T.t.z};
}
inline D3 rotate(const Rt &T, D3 p)
{
    return {T.R[0][0] * p.x + T.R[0][1] * p.y + T.R[0][2] * p.z, T.R[1][0] * p.x + T.R[1][1] * p.y + T.R[1][2] * p.z,
            T.R[2][0] * p.x + T.R[2][1] * p.y + T.R[2][2] * p.z};
}
Rt axis_angle(D3 w)
{
    Rt T{};
    const double a = std::sqrt(dot(w, w));
    if (a < 1e-15)
    {
        T.R[0][0] = T.R[1][1] = T.R[2][2] = 1.0;
        return T;
    }
    const D3 k = (1.0 / a) * w;
    const double c = std::cos(a), s = std::sin(a), t = 1.0 - c;
    T.R[0][0] = t * k.x * k.x + c;
    T.R[0][1] = t * k.x * k.y - k.z * s;
    T.R[0][2] = t * k.x * k.z + k.y * s;
    T.R[1][0] = t * k.x * k.y + k.z * s;
    T.R[1][1] = t * k.y * k.y + c;
    T.R[1][2] = t * k.y * k.z - k.x * s;
    T.R[2][0] = t * k.x * k.z - k.y * s;
    T.R[2][1] = t * k.y * k.z + k.x * s;
    T.R[2][2] = t * k.z * k.z + c;
    return T;
}
Qd quat_from_rot(const double R[3][3])
{
    const double tr = R[0][0] + R[1][1] + R[2][2];
    Qd q;
    if (tr > 0)
    {
        const double s = std::sqrt(tr + 1.0) * 2;
        q = {0.25 * s, (R[2][1] - R[1][2]) / s, (R[0][2] - R[2][0]) / s, (R[1][0] - R[0][1]) / s};
    }
    else if (R[0][0] > R[1][1] && R[0][0] > R[2][2])
    {
        const double s = std::sqrt(1.0 + R[0][0] - R[1][1] - R[2][2]) * 2;
        q = {(R[2][1] - R[1][2]) / s, 0.25 * s, (R[0][1] + R[1][0]) / s, (R[0][2] + R[2][0]) / s};
    }
    else if (R[1][1] > R[2][2])
    {
        const double s = std::sqrt(1.0 + R[1][1] - R[0][0] - R[2][2]) * 2;
        q = {(R[0][2] - R[2][0]) / s, (R[0][1] + R[1][0]) / s, 0.25 * s, (R[1][2] + R[2][1]) / s};
    }
    else
    {
        const double s = std::sqrt(1.0 + R[2][2] - R[0][0] - R[1][1]) * 2;
        q = {(R[1][0] - R[0][1]) / s, (R[0][2] + R[2][0]) / s, (R[1][2] + R[2][1]) / s, 0.25 * s};
    }
    return q;
}
struct DQd
{
    Qd q0, q1;
};
DQd dq_from_rt(const Rt &T)
{
    const Qd r = quat_from_rot(T.R);
    const Qd tv{0, 0.5 * T.t.x, 0.5 * T.t.y, 0.5 * T.t.z};
    return {r, qmul(tv, r)};
}
Rt rt_from_dq(DQd d)
{
    const double n = std::sqrt(d.q0.w * d.q0.w + d.q0.x * d.q0.x + d.q0.y * d.q0.y + d.q0.z * d.q0.z);
    const double inv = 1.0 / n;
    Qd a{d.q0.w * inv, d.q0.x * inv, d.q0.y * inv, d.q0.z * inv};
    Qd b{d.q1.w * inv, d.q1.x * inv, d.q1.y * inv, d.q1.z * inv};
    const double ab = a.w * b.w + a.x * b.x + a.y * b.y + a.z * b.z;
    b = {b.w - ab * a.w, b.x - ab * a.x, b.y - ab * a.y, b.z - ab * a.z};
    Rt T;
    T.R[0][0] = 1 - 2 * a.y * a.y - 2 * a.z * a.z;
    T.R[0][1] = 2 * a.x * a.y - 2 * a.z * a.w;
    T.R[0][2] = 2 * a.x * a.z + 2 * a.y * a.w;
    T.R[1][0] = 2 * a.x * a.y + 2 * a.z * a.w;
    T.R[1][1] = 1 - 2 * a.x * a.x - 2 * a.z * a.z;
    T.R[1][2] = 2 * a.y * a.z - 2 * a.x * a.w;
    T.R[2][0] = 2 * a.x * a.z - 2 * a.y * a.w;
    T.R[2][1] = 2 * a.y * a.z + 2 * a.x * a.w;
    T.R[2][2] = 1 - 2 * a.x * a.x - 2 * a.y * a.y;
    const Qd tq = qmul(Qd{2 * b.w, 2 * b.x, 2 * b.y, 2 * b.z}, Qd{a.w, -a.x, -a.y, -a.z});
    T.t = {tq.x, tq.y, tq.z};
    return T;
}

// ------------------------------------------------------------ node grid KNN
struct Grid
{
    double cell;
    int nx, ny, nz;
    D3 lo;
    std::vector<int> start, items;
    void build(const std::vector<D3> &pts, double cell_size)
    {
        cell = cell_size;
        D3 mn{1e30, 1e30, 1e30}, mx{-1e30, -1e30, -1e30};
        for (const D3 &p : pts)
        {
            mn = {std::min(mn.x, p.x), std::min(mn.y, p.y), std::min(mn.z, p.z)};
            mx = {std::max(mx.x, p.x), std::max(mx.y, p.y), std::max(mx.z, p.z)};
        }
        lo = mn - D3{cell, cell, cell};
        nx = int((mx.x - lo.x) / cell) + 2;
        ny = int((mx.y - lo.y) / cell) + 2;
        nz = int((mx.z - lo.z) / cell) + 2;
        std::vector<int> cnt(size_t(nx) * ny * nz + 1, 0);
        for (const D3 &p : pts)
            ++cnt[cell_of(p) + 1];
        for (size_t i = 1; i < cnt.size(); ++i)
            cnt[i] += cnt[i - 1];
        start = cnt;
        items.resize(pts.size());
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (size_t i = 0; i < pts.size(); ++i)
            items[fill[cell_of(pts[i])]++] = int(i);
    }
    int cx(double v, double l, int n) const { return std::min(n - 1, std::max(0, int((v - l) / cell))); }
    size_t cell_of(D3 p) const
    {
        return (size_t(cx(p.z, lo.z, nz)) * ny + cx(p.y, lo.y, ny)) * nx + cx(p.x, lo.x, nx);
    }
    // exact K nearest (ties by index), ascending
    void knn(const std::vector<D3> &pts, D3 q, int K, int *out_idx, double *out_d2) const
    {
        std::vector<std::pair<double, int>> best;
        const int ix = cx(q.x, lo.x, nx), iy = cx(q.y, lo.y, ny), iz = cx(q.z, lo.z, nz);
        for (int ring = 0;; ++ring)
        {
            best.clear();
            for (int z = std::max(0, iz - ring); z <= std::min(nz - 1, iz + ring); ++z)
                for (int y = std::max(0, iy - ring); y <= std::min(ny - 1, iy + ring); ++y)
                    for (int x = std::max(0, ix - ring); x <= std::min(nx - 1, ix + ring); ++x)
                    {
                        const size_t c = (size_t(z) * ny + y) * nx + x;
                        for (int it = start[c]; it < start[c + 1]; ++it)
                        {
                            const int i = items[it];
                            const D3 d = pts[i] - q;
                            best.push_back({dot(d, d), i});
                        }
                    }
            const bool all = (ix - ring <= 0 && iy - ring <= 0 && iz - ring <= 0 && ix + ring >= nx - 1 &&
                              iy + ring >= ny - 1 && iz + ring >= nz - 1);
            if (int(best.size()) >= K)
            {
                std::partial_sort(best.begin(), best.begin() + K, best.end());
                // everything within `ring * cell` of q has been seen
                const double safe = double(ring) * cell;
                if (best[K - 1].first <= safe * safe || all)
                    break;
            }
            else if (all)
                break;
        }
        for (int k = 0; k < K; ++k)
        {
            out_idx[k] = best[k].second;
            out_d2[k] = best[k].first;
        }
    }
};
} // namespace

extern "C"
{
    struct shs_config
    {
        int num_surfels, num_nodes, width, height;
        float fx, fy;
        float noise_amp;
        int noise_octaves;
        float rot_max_deg, trans_max;
        float depth_noise;
        int with_flow;
        float flow_noise_px, flow_mask_frac;
        float silhouette_frac; // sphere image radius / image height
        int K;
        unsigned long long seed;
    };

    struct shs_scene
    {
        float *surfel_pos;      // S*4
        float *surfel_normal;   // S*4
        float *nodes;           // N*4
        float *gt_dq;           // N*8  (q0 wxyz, q1 wxyz)
        float *cam2world;       // 16, row-major
        float *node_radius_square; // 1
        float *ref_vertex;      // W*H*4, world
        float *ref_normal;      // W*H*4, world
        uint32_t *index_map;    // W*H
        float *measure_vertex;  // W*H*4, camera
        float *measure_normal;  // W*H*4, camera
        float *flow;            // W*H*2 or null
    };

    int shs_generate(const shs_config *cfg, shs_scene *out)
    {
        const int S = cfg->num_surfels, N = cfg->num_nodes, W = cfg->width, H = cfg->height, K = cfg->K;
        if (S <= 0 || N <= 0 || N > 65535 || W <= 0 || H <= 0 || K < 1 || K > 16 || N < K || S < N)
            return -1;
        SplitMix rng(cfg->seed * 0x9e3779b97f4a7c15ull + 12345ull);
        const uint64_t noise_seed = rng.next();

        // ---- surfels
        std::vector<D3> pos(S), nrm(S);
        std::vector<double> uz(S), uphi(S);
        for (int i = 0; i < S; ++i)
        {
            uz[i] = 2.0 * rng.uniform() - 1.0;
            uphi[i] = 6.283185307179586 * rng.uniform();
        }
        auto surf = [&](D3 u) {
            const double rho = 1.0 + double(cfg->noise_amp) * fbm(u, cfg->noise_octaves, noise_seed);
            return rho * u;
        };
#pragma omp parallel for schedule(static)
        for (int i = 0; i < S; ++i)
        {
            const double z = uz[i], rxy = std::sqrt(std::max(0.0, 1.0 - z * z));
            const D3 u{rxy * std::cos(uphi[i]), rxy * std::sin(uphi[i]), z};
            pos[i] = surf(u);
            const D3 a = std::fabs(u.x) < 0.9 ? D3{1, 0, 0} : D3{0, 1, 0};
            const D3 t1 = normalize(cross(u, a)), t2 = cross(u, t1);
            const double h = 1e-4;
            const D3 d1 = surf(normalize(u + h * t1)) - surf(normalize(u - h * t1));
            const D3 d2 = surf(normalize(u + h * t2)) - surf(normalize(u - h * t2));
            D3 n = normalize(cross(d1, d2));
            if (dot(n, u) < 0)
                n = -1.0 * n;
            nrm[i] = n;
        }

        // ---- nodes: farthest-point subsampling over the first M surfels
        const int M = std::min(S, std::max(32 * N, 4096));
        std::vector<int> node_ids(N);
        {
            std::vector<double> mind(M, 1e300);
            int cur = 0;
            for (int n = 0; n < N; ++n)
            {
                node_ids[n] = cur;
                const D3 c = pos[cur];
                int best = -1;
                double bestd = -1.0;
#pragma omp parallel
                {
                    int lb = -1;
                    double ld = -1.0;
#pragma omp for schedule(static)
                    for (int i = 0; i < M; ++i)
                    {
                        const D3 d = pos[i] - c;
                        const double dd = dot(d, d);
                        if (dd < mind[i])
                            mind[i] = dd;
                        if (mind[i] > ld)
                        {
                            ld = mind[i];
                            lb = i;
                        }
                    }
#pragma omp critical
                    {
                        if (ld > bestd || (ld == bestd && lb < best))
                        {
                            bestd = ld;
                            best = lb;
                        }
                    }
                }
                cur = best;
            }
        }
        std::vector<D3> nodes(N);
        for (int n = 0; n < N; ++n)
            nodes[n] = pos[node_ids[n]];
        // mean nearest-node spacing
        double spacing = 0.0;
        {
            std::vector<double> nn(N, 1e300);
#pragma omp parallel for schedule(static)
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < N; ++j)
                    if (i != j)
                    {
                        const D3 d = nodes[i] - nodes[j];
                        nn[i] = std::min(nn[i], dot(d, d));
                    }
            for (int i = 0; i < N; ++i)
                spacing += std::sqrt(nn[i]);
            spacing /= N;
        }
        const double sigma = 2.0 * spacing;
        const float r2 = float(sigma * sigma);

        // ---- ground-truth per-node SE(3): smooth random fields
        D3 kw[3], kt[3];
        double pw[3], pt[3];
        for (int c = 0; c < 3; ++c)
        {
            kw[c] = {2.0 * rng.normal(), 2.0 * rng.normal(), 2.0 * rng.normal()};
            kt[c] = {2.0 * rng.normal(), 2.0 * rng.normal(), 2.0 * rng.normal()};
            pw[c] = 6.283185307179586 * rng.uniform();
            pt[c] = 6.283185307179586 * rng.uniform();
        }
        const double rot_max = double(cfg->rot_max_deg) * 3.141592653589793 / 180.0;
        const double inv_sqrt3 = 1.0 / std::sqrt(3.0);
        std::vector<Rt> gt(N);
        std::vector<DQd> gtdq(N);
        for (int n = 0; n < N; ++n)
        {
            const D3 g = nodes[n];
            const D3 w{rot_max * inv_sqrt3 * std::sin(dot(kw[0], g) + pw[0]),
                       rot_max * inv_sqrt3 * std::sin(dot(kw[1], g) + pw[1]),
                       rot_max * inv_sqrt3 * std::sin(dot(kw[2], g) + pw[2])};
            const D3 tr{double(cfg->trans_max) * inv_sqrt3 * std::sin(dot(kt[0], g) + pt[0]),
                        double(cfg->trans_max) * inv_sqrt3 * std::sin(dot(kt[1], g) + pt[1]),
                        double(cfg->trans_max) * inv_sqrt3 * std::sin(dot(kt[2], g) + pt[2])};
            Rt T = axis_angle(w);
            // rotation about the node centre: x' = R (x - g) + g + t
            T.t = g - rotate(T, g) + tr;
            gt[n] = T;
            gtdq[n] = dq_from_rt(T);
        }

        // ---- warp surfels with the DQ blend of their K nearest nodes
        Grid grid;
        grid.build(nodes, std::max(spacing, 1e-3));
        std::vector<D3> wpos(S), wnrm(S);
#pragma omp parallel for schedule(dynamic, 256)
        for (int i = 0; i < S; ++i)
        {
            int idx[16];
            double d2[16];
            grid.knn(nodes, pos[i], K, idx, d2);
            DQd acc{{0, 0, 0, 0}, {0, 0, 0, 0}};
            double wsum = 0;
            double wk[16];
            for (int k = 0; k < K; ++k)
            {
                wk[k] = std::exp(-d2[k] / (2.0 * sigma * sigma));
                wsum += wk[k];
            }
            for (int k = 0; k < K; ++k)
            {
                const double w = wk[k] / wsum;
                DQd d = gtdq[idx[k]];
                // keep quaternions in one hemisphere for blending
                if (d.q0.w < 0)
                {
                    d.q0 = {-d.q0.w, -d.q0.x, -d.q0.y, -d.q0.z};
                    d.q1 = {-d.q1.w, -d.q1.x, -d.q1.y, -d.q1.z};
                }
                acc.q0 = {acc.q0.w + w * d.q0.w, acc.q0.x + w * d.q0.x, acc.q0.y + w * d.q0.y, acc.q0.z + w * d.q0.z};
                acc.q1 = {acc.q1.w + w * d.q1.w, acc.q1.x + w * d.q1.x, acc.q1.y + w * d.q1.y, acc.q1.z + w * d.q1.z};
            }
            const Rt T = rt_from_dq(acc);
            wpos[i] = apply(T, pos[i]);
            wnrm[i] = normalize(rotate(T, nrm[i]));
        }

        // ---- camera looking at the origin
        const double f = double(cfg->fy);
        const double cxp = 0.5 * W, cyp = 0.5 * H;
        const double tan_t = double(cfg->silhouette_frac) * H / f;
        const double dist = 1.0 / std::sin(std::atan(tan_t));
        const D3 dir = normalize(D3{0.25, -0.15, -1.0});
        const D3 cam_pos = dist * dir;
        const D3 zc = normalize(-1.0 * cam_pos);
        const D3 xc = normalize(cross(D3{0, 1, 0}, zc));
        const D3 yc = cross(zc, xc);
        Rt c2w;
        c2w.R[0][0] = xc.x;
        c2w.R[1][0] = xc.y;
        c2w.R[2][0] = xc.z;
        c2w.R[0][1] = yc.x;
        c2w.R[1][1] = yc.y;
        c2w.R[2][1] = yc.z;
        c2w.R[0][2] = zc.x;
        c2w.R[1][2] = zc.y;
        c2w.R[2][2] = zc.z;
        c2w.t = cam_pos;
        Rt w2c;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                w2c.R[a][b] = c2w.R[b][a];
        w2c.t = -1.0 * rotate(w2c, cam_pos);

        // ---- oriented-disk z-buffer renderer
        const double surfel_radius = 0.75 * std::sqrt(4.0 * 3.141592653589793 / S);
        auto render = [&](const std::vector<D3> &P, const std::vector<D3> &Nn, std::vector<double> &zbuf,
                          std::vector<int> &ibuf) {
            zbuf.assign(size_t(W) * H, 1e30);
            ibuf.assign(size_t(W) * H, -1);
            for (int i = 0; i < S; ++i)
            {
                const D3 pc = apply(w2c, P[i]);
                const D3 nc = rotate(w2c, Nn[i]);
                if (pc.z <= 1e-6 || dot(nc, pc) >= 0.0)
                    continue;
                const double u = double(cfg->fx) * pc.x / pc.z + cxp;
                const double v = double(cfg->fy) * pc.y / pc.z + cyp;
                const double rad = std::max(0.75, f * surfel_radius / pc.z);
                const int x0 = int(std::floor(u - rad)), x1 = int(std::ceil(u + rad));
                const int y0 = int(std::floor(v - rad)), y1 = int(std::ceil(v + rad));
                const double ndp = dot(nc, pc);
                for (int y = std::max(0, y0); y <= std::min(H - 1, y1); ++y)
                    for (int x = std::max(0, x0); x <= std::min(W - 1, x1); ++x)
                    {
                        const D3 ray{(x - cxp) / double(cfg->fx), (y - cyp) / double(cfg->fy), 1.0};
                        const double nr = dot(nc, ray);
                        if (nr >= -1e-9)
                            continue;
                        const double z = ndp / nr;
                        const D3 hit = z * ray;
                        const D3 off = hit - pc;
                        if (dot(off, off) > surfel_radius * surfel_radius * 1.0001 &&
                            (std::fabs(x - u) > 0.75 || std::fabs(y - v) > 0.75))
                            continue;
                        const size_t pix = size_t(y) * W + x;
                        if (z < zbuf[pix])
                        {
                            zbuf[pix] = z;
                            ibuf[pix] = i;
                        }
                    }
            }
        };
        std::vector<double> zref, ztgt;
        std::vector<int> iref, itgt;
        render(pos, nrm, zref, iref);
        render(wpos, wnrm, ztgt, itgt);

        // ---- outputs
        for (int i = 0; i < S; ++i)
        {
            float *p = out->surfel_pos + 4 * size_t(i);
            p[0] = float(pos[i].x);
            p[1] = float(pos[i].y);
            p[2] = float(pos[i].z);
            p[3] = 1.f;
            float *q = out->surfel_normal + 4 * size_t(i);
            q[0] = float(nrm[i].x);
            q[1] = float(nrm[i].y);
            q[2] = float(nrm[i].z);
            q[3] = float(surfel_radius);
        }
        for (int n = 0; n < N; ++n)
        {
            float *p = out->nodes + 4 * size_t(n);
            p[0] = float(nodes[n].x);
            p[1] = float(nodes[n].y);
            p[2] = float(nodes[n].z);
            p[3] = 1.f;
            float *d = out->gt_dq + 8 * size_t(n);
            const DQd &g = gtdq[n];
            const double v[8] = {g.q0.w, g.q0.x, g.q0.y, g.q0.z, g.q1.w, g.q1.x, g.q1.y, g.q1.z};
            for (int c = 0; c < 8; ++c)
                d[c] = float(v[c]);
        }
        for (int a = 0; a < 3; ++a)
        {
            for (int b = 0; b < 3; ++b)
                out->cam2world[4 * a + b] = float(c2w.R[a][b]);
        }
        out->cam2world[3] = float(c2w.t.x);
        out->cam2world[7] = float(c2w.t.y);
        out->cam2world[11] = float(c2w.t.z);
        out->cam2world[12] = out->cam2world[13] = out->cam2world[14] = 0.f;
        out->cam2world[15] = 1.f;
        *out->node_radius_square = r2;

        SplitMix noise(cfg->seed * 0xd1b54a32d192ed03ull + 99ull);
        for (size_t pix = 0; pix < size_t(W) * H; ++pix)
        {
            const int i = iref[pix];
            float *rv = out->ref_vertex + 4 * pix;
            float *rn = out->ref_normal + 4 * pix;
            if (i >= 0)
            {
                rv[0] = float(pos[i].x);
                rv[1] = float(pos[i].y);
                rv[2] = float(pos[i].z);
                rv[3] = 1.f;
                rn[0] = float(nrm[i].x);
                rn[1] = float(nrm[i].y);
                rn[2] = float(nrm[i].z);
                rn[3] = float(surfel_radius);
                out->index_map[pix] = uint32_t(i);
            }
            else
            {
                std::memset(rv, 0, 16);
                std::memset(rn, 0, 16);
                out->index_map[pix] = 0xffffffffu;
            }
            const int j = itgt[pix];
            float *mv = out->measure_vertex + 4 * pix;
            float *mn = out->measure_normal + 4 * pix;
            const double dn = noise.normal();
            if (j >= 0)
            {
                const int x = int(pix % W), y = int(pix / W);
                const D3 ray{(x - cxp) / double(cfg->fx), (y - cyp) / double(cfg->fy), 1.0};
                const double z = ztgt[pix] + double(cfg->depth_noise) * dn;
                const D3 v = z * ray;
                mv[0] = float(v.x);
                mv[1] = float(v.y);
                mv[2] = float(v.z);
                mv[3] = 1.f;
                const D3 nc = rotate(w2c, wnrm[j]);
                mn[0] = float(nc.x);
                mn[1] = float(nc.y);
                mn[2] = float(nc.z);
                mn[3] = float(surfel_radius);
            }
            else
            {
                std::memset(mv, 0, 16);
                std::memset(mn, 0, 16);
            }
        }
        if (cfg->with_flow && out->flow)
        {
            SplitMix fr(cfg->seed * 0x8cb92ba72f3d8dd7ull + 7ull);
            for (size_t pix = 0; pix < size_t(W) * H; ++pix)
            {
                const double n1 = fr.normal(), n2 = fr.normal(), m = fr.uniform();
                float *fl = out->flow + 2 * pix;
                fl[0] = fl[1] = 0.f;
                const int i = iref[pix];
                if (i < 0 || m < double(cfg->flow_mask_frac))
                    continue;
                const D3 pc = apply(w2c, wpos[i]);
                const double u = double(cfg->fx) * pc.x / pc.z + cxp;
                const double v = double(cfg->fy) * pc.y / pc.z + cyp;
                const int x = int(pix % W), y = int(pix / W);
                fl[0] = float(u - x + double(cfg->flow_noise_px) * n1);
                fl[1] = float(v - y + double(cfg->flow_noise_px) * n2);
            }
        }
        else if (out->flow)
        {
            std::memset(out->flow, 0, sizeof(float) * 2 * size_t(W) * H);
        }
        return 0;
    }
}
