// C-ABI over the CPU restatement (sh_oracle.hpp) so that the Python tests and
// the CPU-baseline leg of the benchmark can drive it through ctypes.
// TEST INFRASTRUCTURE ONLY: never linked into the product library.
#include "sh_measure.hpp"
#include "sh_oracle.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <string>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C"
{
    struct sho_inputs
    {
        int width, height;
        float fx, fy, cx, cy;
        float cam2world[16];
        const float *measure_vertex, *measure_normal, *ref_vertex, *ref_normal;
        const uint32_t *index_map;
        const float *flow;
        int num_surfels, K;
        const uint16_t *surfel_knn;
        const float *surfel_spatial, *surfel_connect;
        int num_nodes;
        const float *node_coords;
        int num_edges;
        const uint16_t *edges;
        const float *node_connect;
        const float *node_motion_pred;
        float node_radius_square;
    };

    struct sho_params
    {
        int gn_iters, pcg_iters;
        float lambda_depth, lambda_flow, lambda_reg, lambda_node_translation, damping;
        const uint8_t *rot_cut; // NULL, or gn_iters * num_nodes replayed rotation-cutoff decisions
    };
}

namespace
{
struct Handle
{
    sho::Result res;
    double solve_ms = 0.0;
    int threads = 1;
};

sho::Inputs to_inputs(const sho_inputs *c)
{
    sho::Inputs in;
    in.width = c->width;
    in.height = c->height;
    in.fx = c->fx;
    in.fy = c->fy;
    in.cx = c->cx;
    in.cy = c->cy;
    std::memcpy(in.cam2world, c->cam2world, sizeof(in.cam2world));
    in.measure_vertex = c->measure_vertex;
    in.measure_normal = c->measure_normal;
    in.ref_vertex = c->ref_vertex;
    in.ref_normal = c->ref_normal;
    in.index_map = c->index_map;
    in.flow = c->flow;
    in.num_surfels = c->num_surfels;
    in.K = c->K;
    in.surfel_knn = c->surfel_knn;
    in.surfel_spatial = c->surfel_spatial;
    in.surfel_connect = c->surfel_connect;
    in.num_nodes = c->num_nodes;
    in.node_coords = c->node_coords;
    in.num_edges = c->num_edges;
    in.edges = c->edges;
    in.node_connect = c->node_connect;
    in.node_motion_pred = c->node_motion_pred;
    in.node_radius_square = c->node_radius_square;
    return in;
}

sho::Params to_params(const sho_params *c)
{
    sho::Params p;
    p.gn_iters = c->gn_iters;
    p.pcg_iters = c->pcg_iters;
    p.lambda_depth = c->lambda_depth;
    p.lambda_flow = c->lambda_flow;
    p.lambda_reg = c->lambda_reg;
    p.lambda_node_translation = c->lambda_node_translation;
    p.damping = c->damping;
    p.rot_cut = c->rot_cut;
    return p;
}

void set_threads(int threads)
{
#ifdef _OPENMP
    omp_set_num_threads(threads > 0 ? threads : 1);
#else
    (void)threads;
#endif
}
} // namespace

extern "C"
{
    int sho_max_threads(void)
    {
#ifdef _OPENMP
        return omp_get_num_procs();
#else
        return 1;
#endif
    }

    // Surfel skinning (Skinner.cu:10-94).  Returns 0.
    int sho_skin(const float *surfel4, int num_surfels, const float *node4, int num_nodes, int K, float r2,
                 uint16_t *knn, float *spatial, float *connect, int threads)
    {
        if (K < 1 || K > 16 || num_nodes < 1)
            return -1;
        set_threads(threads);
        sho::skin_surfels(surfel4, num_surfels, node4, num_nodes, K, r2, knn, spatial, connect);
        return 0;
    }

    // Node graph (NodeGraph.cu:10-118).  Returns 0.
    int sho_node_graph(const float *node4, int num_nodes, int K, float r2, uint16_t *node_knn, float *node_spatial,
                       uint16_t *edges, float *node_connect, int threads)
    {
        if (K < 1 || K > 16 || num_nodes < K)
            return -1;
        set_threads(threads);
        sho::build_node_graph(node4, num_nodes, K, r2, node_knn, node_spatial, edges, node_connect);
        return 0;
    }

    // Full GN solve.  precision: 64 (FP64 oracle) or 32 (FP32 restatement).
    void *sho_solve(const sho_inputs *cin, const sho_params *cp, int precision, int threads)
    {
        Handle *h = new Handle();
        set_threads(threads);
        h->threads = threads;
        const sho::Inputs in = to_inputs(cin);
        const sho::Params p = to_params(cp);
        const auto t0 = std::chrono::steady_clock::now();
        if (precision == 64)
        {
            sho::Solver<double> s(in, p, h->res);
            s.solve();
        }
        else
        {
            sho::Solver<float> s(in, p, h->res);
            s.solve();
        }
        const auto t1 = std::chrono::steady_clock::now();
        h->solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        return h;
    }

    // dtype codes: 1 u8, 2 u16, 4 u32, 8 f64
    int sho_result_get(void *vh, const char *name, const void **ptr, size_t *count, int *dtype)
    {
        Handle *h = static_cast<Handle *>(vh);
        const sho::Result &r = h->res;
        const std::string n(name);
#define SHO_GET(field, code)                                                                                           \
    if (n == #field)                                                                                                   \
    {                                                                                                                  \
        *ptr = r.field.data();                                                                                         \
        *count = r.field.size();                                                                                       \
        *dtype = code;                                                                                                 \
        return 0;                                                                                                      \
    }
        SHO_GET(pixel_pair, 2)
        SHO_GET(pixel_knn, 2)
        SHO_GET(n2t_offset, 4)
        SHO_GET(n2t_terms, 4)
        SHO_GET(half_keys, 4)
        SHO_GET(half_offset, 4)
        SHO_GET(pair_terms, 4)
        SHO_GET(sym_keys, 4)
        SHO_GET(sym_range, 4)
        SHO_GET(row_offset, 4)
        SHO_GET(residual, 8)
        SHO_GET(jacobian, 8)
        SHO_GET(reg_Ti_xj, 8)
        SHO_GET(reg_Tj_xj, 8)
        SHO_GET(reg_valid, 1)
        SHO_GET(diag, 8)
        SHO_GET(rhs, 8)
        SHO_GET(offdiag, 8)
        SHO_GET(diag_inv, 8)
        SHO_GET(pcg_x, 8)
        SHO_GET(cost, 8)
        SHO_GET(dq_iter, 8)
        SHO_GET(twist_l1, 8)
        SHO_GET(rot_cut, 1)
        SHO_GET(dq, 8)
#undef SHO_GET
        if (n == "num_pixels")
        {
            static thread_local uint32_t v;
            v = uint32_t(r.num_pixels);
            *ptr = &v;
            *count = 1;
            *dtype = 4;
            return 0;
        }
        if (n == "solve_ms")
        {
            *ptr = &h->solve_ms;
            *count = 1;
            *dtype = 8;
            return 0;
        }
        return -1;
    }

    void sho_result_free(void *vh) { delete static_cast<Handle *>(vh); }

    // Exposed pieces of the restatement for closed-form / brute-force checks.
    float sho_round_to_half(float f) { return sho::round_to_half(f); }
    uint32_t sho_float2uint_rn(float f) { return sho::float2uint_rn(f); }

    // Closed-form hooks on the DQ algebra (double precision).
    // dq: (q0 w x y z, q1 w x y z); out34: row-major 3x4.
    void sho_dq_se3(const double *dq, double *out34)
    {
        sho::DQ<double> d{{dq[0], dq[1], dq[2], dq[3]}, {dq[4], dq[5], dq[6], dq[7]}};
        const sho::M34<double> m = sho::dq_se3(d);
        for (int a = 0; a < 3; ++a)
        {
            for (int b = 0; b < 3; ++b)
                out34[4 * a + b] = m.R.m[a][b];
        }
        out34[3] = m.t.x;
        out34[7] = m.t.y;
        out34[11] = m.t.z;
    }
    // forward warp in FP64 (precision 64) or FP32 (32); outputs as double
    int sho_forward_warp(const double *warp_wxyz, int K, const float *ref_v, const float *ref_n,
                         const uint16_t *s_knn, const float *s_spatial, const float *s_connect, uint64_t num_surfels,
                         const float *ref_node, const uint16_t *n_knn, const float *n_spatial, const float *n_connect,
                         uint64_t num_nodes, double *live_v, double *live_n, double *live_node, int precision)
    {
        if ((K != 4 && K != 8) || !warp_wxyz)
            return 1;
        if (precision == 32)
            sho::forward_warp<float>(warp_wxyz, K, ref_v, ref_n, s_knn, s_spatial, s_connect, num_surfels, ref_node,
                                     n_knn, n_spatial, n_connect, num_nodes, live_v, live_n, live_node);
        else
            sho::forward_warp<double>(warp_wxyz, K, ref_v, ref_n, s_knn, s_spatial, s_connect, num_surfels, ref_node,
                                      n_knn, n_spatial, n_connect, num_nodes, live_v, live_n, live_node);
        return 0;
    }

    void sho_dq_from_mat34(const double *in34, double *dq)
    {
        sho::M34<double> m;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                m.R.m[a][b] = in34[4 * a + b];
        m.t = {in34[3], in34[7], in34[11]};
        const sho::DQ<double> d = sho::dq_from_mat34(m);
        const double v[8] = {d.q0.w, d.q0.x, d.q0.y, d.q0.z, d.q1.w, d.q1.x, d.q1.y, d.q1.z};
        std::memcpy(dq, v, sizeof(v));
    }
    void sho_twist_exp(const double *xi6, double *out34)
    {
        const sho::M34<double> m = sho::twist_exp<double>({xi6[0], xi6[1], xi6[2]}, {xi6[3], xi6[4], xi6[5]});
        for (int a = 0; a < 3; ++a)
        {
            for (int b = 0; b < 3; ++b)
                out34[4 * a + b] = m.R.m[a][b];
        }
        out34[3] = m.t.x;
        out34[7] = m.t.y;
        out34[11] = m.t.z;
    }
    void sho_invert6(const double *A, double *Ainv) { sho::Solver<double>::invert6(A, Ainv); }

    // BlockPCG<6>::Solve over a block-CSR (sho::block_pcg); precision 64 or 32.
    // x: in (x_init unless zero_init) and out; delta_hist: iters + 1 entries.
    void sho_block_pcg(int nblk, const uint32_t *row_ptr, const uint32_t *col, const double *blocks,
                       const double *inv_diag, const double *b, double *x, int iters, int zero_init, int precision,
                       double *delta_hist)
    {
        std::vector<double> h;
        if (precision == 64)
            sho::block_pcg<double>(nblk, row_ptr, col, blocks, inv_diag, b, x, iters, zero_init != 0, &h);
        else
            sho::block_pcg<float>(nblk, row_ptr, col, blocks, inv_diag, b, x, iters, zero_init != 0, &h);
        if (delta_hist)
            std::copy(h.begin(), h.end(), delta_hist);
    }

    // Evaluate the per-pixel residual/Jacobian and the ARAP records at a given
    // warp field (FP64): dq is N x 8 (q0 wxyz, q1 wxyz).  Outputs: r (P*3),
    // J (P*18), ti/tj (E*3), valid (E).  Returns P, or -1 if r_cap is too small.
    int sho_eval_terms(const sho_inputs *cin, const sho_params *cp, const double *dq, double *r, size_t r_cap,
                       double *J, double *ti, double *tj, uint8_t *valid)
    {
        sho::Result res;
        const sho::Inputs in = to_inputs(cin);
        const sho::Params p = to_params(cp);
        sho::Solver<double> s(in, p, res);
        s.compact_pixels();
        if (size_t(s.P) * 3 > r_cap)
            return -1;
        s.dq.resize(in.num_nodes);
        for (int n = 0; n < in.num_nodes; ++n)
            s.dq[n] = {{dq[8 * n], dq[8 * n + 1], dq[8 * n + 2], dq[8 * n + 3]},
                       {dq[8 * n + 4], dq[8 * n + 5], dq[8 * n + 6], dq[8 * n + 7]}};
        s.dense_terms();
        s.reg_terms();
        std::memcpy(r, s.r.data(), sizeof(double) * s.r.size());
        std::memcpy(J, s.J.data(), sizeof(double) * s.J.size());
        for (int e = 0; e < in.num_edges; ++e)
        {
            ti[3 * e] = s.Ti_xj[e].x;
            ti[3 * e + 1] = s.Ti_xj[e].y;
            ti[3 * e + 2] = s.Ti_xj[e].z;
            tj[3 * e] = s.Tj_xj[e].x;
            tj[3 * e + 1] = s.Tj_xj[e].y;
            tj[3 * e + 2] = s.Tj_xj[e].z;
            valid[e] = s.valid[e];
        }
        return s.P;
    }

    // ---------------------------------------------------- measurement maps
    struct sho_depth_camera
    {
        int width, height;
        float fx, fy, cx, cy, clip_near, clip_far, scale, radius_scale;
    };
    static sho::DepthCamera depth_cam(const sho_depth_camera *c)
    {
        return {c->width, c->height, c->fx, c->fy, c->cx, c->cy, c->clip_near, c->clip_far, c->scale, c->radius_scale};
    }

    // smallest eigenvector of n packed PSD 3x3 matrices (00 01 02 11 12 22)
    void sho_eigen33(const double *cov6, int n, double *vec3, int precision)
    {
        for (int i = 0; i < n; ++i)
        {
            if (precision == 32)
            {
                float c[6];
                for (int k = 0; k < 6; ++k)
                    c[k] = float(cov6[6 * i + k]);
                const sho::V3<float> v = sho::Eig33<float>::smallest(c);
                vec3[3 * i] = v.x;
                vec3[3 * i + 1] = v.y;
                vec3[3 * i + 2] = v.z;
            }
            else
            {
                const sho::V3<double> v = sho::Eig33<double>::smallest(cov6 + 6 * i);
                vec3[3 * i] = v.x;
                vec3[3 * i + 1] = v.y;
                vec3[3 * i + 2] = v.z;
            }
        }
    }

    void sho_filter_depth(const uint16_t *raw, const sho_depth_camera *c, float *out, float *camera_dot,
                          int precision)
    {
        if (precision == 32)
            sho::filter_depth<float>(raw, depth_cam(c), out, camera_dot);
        else
            sho::filter_depth<double>(raw, depth_cam(c), out, camera_dot);
    }

    // filtered depth -> scaled vertex (w*h*4) and normal/radius (w*h*4) maps
    void sho_measure_maps(const float *filtered, const sho_depth_camera *c, double *vertex4, double *normal4,
                          int precision)
    {
        const sho::DepthCamera cam = depth_cam(c);
        int w, h;
        sho::scaled_size(cam, w, h);
        const size_t n = size_t(w) * h * 4;
        if (precision == 32)
        {
            std::vector<float> v(n), nr(n);
            sho::scaled_vertex_map<float>(filtered, cam, v.data());
            sho::normal_radius_map<float>(v.data(), w, h, cam.radius_scale, nr.data());
            for (size_t i = 0; i < n; ++i)
            {
                vertex4[i] = v[i];
                normal4[i] = nr[i];
            }
        }
        else
        {
            sho::scaled_vertex_map<double>(filtered, cam, vertex4);
            sho::normal_radius_map<double>(vertex4, w, h, cam.radius_scale, normal4);
        }
    }

    void sho_normal_radius_map(const float *vertex4, int w, int h, float radius_scale, double *normal4, int precision)
    {
        const size_t n = size_t(w) * h * 4;
        if (precision == 32)
        {
            std::vector<float> nr(n);
            sho::normal_radius_map<float>(vertex4, w, h, radius_scale, nr.data());
            for (size_t i = 0; i < n; ++i)
                normal4[i] = nr[i];
        }
        else
        {
            std::vector<double> v(vertex4, vertex4 + n);
            sho::normal_radius_map<double>(v.data(), w, h, radius_scale, normal4);
        }
    }

    void sho_vertex_confid_map(const uint16_t *raw, int w, int h, float fx, float fy, float cx, float cy,
                               float *vertex4)
    {
        sho::vertex_confid_map<float>(raw, w, h, fx, fy, cx, cy, vertex4);
    }
}
