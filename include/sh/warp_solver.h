/*
 * StarHub warp-field Gauss-Newton solver for AMD Instinct MI355X (gfx950).
 *
 * C-ABI drop-in boundary.  Plain pointers and sizes only; every device pointer
 * is caller-owned HBM that must stay valid for the duration of the call that
 * consumes it (the reference's *4Solver structs are non-owning views in the
 * same way, libs/star/opt/solver_types.h:21-120).  Library-owned outputs stay
 * valid until the next set_inputs / solve on the same handle.
 *
 * Citations are to the StarHub sources (paths relative to the repository root).
 * Status convention: every entry point returns an sh_status; none calls exit()
 * (the reference exits / throws on errors, libs/star/common/containers/safe_call.cuh:54-60).
 */
#ifndef SH_WARP_SOLVER_H
#define SH_WARP_SOLVER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *sh_stream_t; /* hipStream_t; NULL = default stream */

typedef enum sh_status
{
    SH_OK = 0,
    SH_E_ARG = 1,      /* invalid argument / size */
    SH_E_CAPACITY = 2, /* input exceeds the capacity fixed at create time (GBufferArray.h:93-100) */
    SH_E_HIP = 3,      /* HIP runtime error */
    SH_E_NAN = 4,      /* non-finite right-hand side (WarpSolverPCG.cu:19-31 exits instead) */
    SH_E_STATE = 5,    /* call out of order (e.g. solve before set_inputs) */
    SH_E_NOMEM = 6     /* device allocation failed */
} sh_status;

/* Storage of the per-pixel residual/Jacobian stream. */
typedef enum sh_stream_precision
{
    SH_STREAM_FP32 = 0,
    SH_STREAM_FP16 = 1 /* FP16 storage, FP32 accumulation */
} sh_stream_precision;

/* How the PCG applies A = J^T W J + damping. */
typedef enum sh_spmv_mode
{
    SH_SPMV_MATERIALIZED = 0, /* 6x6 block-CSR (JtJMaterializerBinBlockCSR.cu:8-74) */
    SH_SPMV_MATRIX_FREE = 1,  /* J^T (W (J x)) from the term streams (JtJMaterializerSanityCheck.cu:47-79) */
    SH_SPMV_AUTO = 2          /* autotuner picks */
} sh_spmv_mode;

/* Optional behaviour (sh_solver_config.flags). */
typedef enum sh_solver_flags
{
    /* Also build the reference's node-pair -> term index (NodePair2TermsIndex.cu:294-312):
     * unique (i<j) pairs, their sorted term lists and the symmetric keys.  Not needed by
     * the default row-major materialisation; exposed through sh_solver_debug_view. */
    SH_FLAG_REFERENCE_INDEX = 1u,
    /* Materialise off-diagonal blocks pair-major from that index (the reference's
     * traversal, JtJMaterializer.cu:66-175).  Implies SH_FLAG_REFERENCE_INDEX. */
    SH_FLAG_PAIR_MAJOR = 2u,
    /* Run the PCG as two launches per iteration instead of one persistent cooperative launch. */
    SH_FLAG_NO_PERSISTENT_PCG = 4u,
    /* Keep the persistent PCG but stream the blocks from memory each iteration instead of
     * holding them in registers (the variant used when a solve's rows do not fit). */
    SH_FLAG_STREAMING_PCG = 8u,
    /* Record a HIP event before every GN-phase kernel launch on the solve's stream and
     * keep the summed time per phase (sh_solver_phase_times). */
    SH_FLAG_PHASE_TIMING = 16u,
    /* Materialise row-major from node-row segments (each per-pixel record read once per
     * node of the pixel) instead of from pixel tiles (each record read once). */
    SH_FLAG_ROW_SEGMENTS = 32u,
    /* Also evaluate the cost at the final warp field (one more term pass), so the
     * cost history has gn_iterations + 1 entries. */
    SH_FLAG_COST_HISTORY = 64u,
    /* Issue the split / matrix-free PCG launches one by one instead of replaying them
     * from a captured hipGraph. */
    SH_FLAG_NO_GRAPH = 128u,
    /* Register-resident PCG in its pipelined form (Ghysels-Vanroose): one barrier per
     * iteration instead of two, FP64 vectors.  The same iterates as the reference's PCG
     * in exact arithmetic (BlockPCG.cuh:325-483). */
    SH_FLAG_PIPELINED_PCG = 256u,
    /* Pixels per tile of the tile materialisation: 64 or 256 instead of 128 (at most
     * one of the two).  Smaller tiles give more work-groups and shorter slot lists,
     * larger ones fewer slot partials per block.  Tunable; a change takes effect at
     * the next index build. */
    SH_FLAG_TILE_PIX_64 = 512u,
    SH_FLAG_TILE_PIX_256 = 1024u
} sh_solver_flags;

/* Dual quaternion as stored by the reference: two float4 (x, y, z, w),
 * real part q0 first (libs/star/math/DualQuaternion.hpp:206). */
typedef struct sh_dual_quat
{
    float q0[4];
    float q1[4];
} sh_dual_quat;

/* Construction-time capacities and solver constants.  Replaces the
 * ConfigParser singleton + compile-time constants the reference reads
 * (examples/mono_star/common/Constants.cpp:14-34, PenaltyConstants.cu:8-32,
 * libs/star/geometry/constants.h:6-14). */
typedef struct sh_solver_config
{
    uint32_t max_width, max_height; /* solver image capacity (downsampled size) */
    uint32_t max_nodes;             /* <= 4096: node-pair key i*4096+j (solver_encode.h:14) */
    uint32_t knn;                   /* K nearest nodes per surfel / node: 4 or 8 */
    uint32_t gn_iterations;         /* reference default 6 (Constants.cpp:33) */
    uint32_t pcg_iterations;        /* reference default 10 (WarpSolver.cu:303) */
    float lambda_depth;             /* 0.1 */
    float lambda_flow;              /* 0.0 */
    float lambda_reg;               /* 1.0 (2.3 for K=4) */
    float lambda_node_translation;  /* 0.0 */
    float lambda_feature;           /* 1.0; the feature term has a zero Jacobian and no terms here */
    float damping;                  /* 1e-3 added to the JtJ diagonal (PreconditionerRhsBuilder.cu:249-252) */
    uint32_t stream_precision;      /* sh_stream_precision */
    uint32_t spmv_mode;             /* sh_spmv_mode */
    uint32_t flags;                 /* sh_solver_flags */
} sh_solver_config;

/* Pinhole camera of the solver images (ConfigParser.cu:231-256). */
typedef struct sh_camera
{
    uint32_t width, height;
    float fx, fy, cx, cy;
} sh_camera;

/* Measure4Solver (solver_types.h:52-58): measured maps, camera frame,
 * float4 per pixel, row-major width*height, border reads return zero. */
typedef struct sh_measure_maps
{
    const float *vertex_confid; /* float4 */
    const float *normal_radius; /* float4 */
} sh_measure_maps;

/* Render4Solver (solver_types.h:62-68): reference maps in the world frame. */
typedef struct sh_render_maps
{
    const float *reference_vertex; /* float4 */
    const float *reference_normal; /* float4 */
    const uint32_t *index_map;     /* surfel index, 0xffffffff = empty */
} sh_render_maps;

/* Geometry4Solver (libs/star/geometry/surfel/SurfelGeometry.h:76-82). */
typedef struct sh_geometry
{
    const uint16_t *surfel_knn;           /* num_surfels * K */
    const float *surfel_knn_spatial_weight; /* num_surfels * K */
    const float *surfel_knn_connect_weight; /* num_surfels * K */
    uint32_t num_surfels;
} sh_geometry;

/* NodeGraph4Solver (libs/star/geometry/node_graph/NodeGraph.h:89-99). */
typedef struct sh_node_graph
{
    const float *reference_node_coords;    /* float4 * num_nodes */
    const uint16_t *node_graph;            /* ushort3 (i, j, k) * num_edges */
    const float *node_knn_connect_weight;  /* num_nodes * K */
    const uint16_t *node_knn;              /* num_nodes * K (node-motion term) */
    const float *node_knn_spatial_weight;  /* num_nodes * K */
    uint32_t num_nodes;
    uint32_t num_edges;
    float node_radius_square;
} sh_node_graph;

/* NodeFlow4Solver (solver_types.h:108-112): may be NULL pointer (zeros). */
typedef struct sh_node_flow
{
    const float *node_motion_pred; /* float4 * num_nodes: (x, y, z, confidence) */
    uint32_t num_nodes;
} sh_node_flow;

/* OpticalFlow4Solver (solver_types.h:116-120): float2 per pixel, may be NULL. */
typedef struct sh_flow_map
{
    const float *flow;
} sh_flow_map;

/* KeyPoint4Solver (solver_types.h:82-96).  The reference's feature term has a
 * zero Jacobian (PreconditionerRhsBuilder.cu:215-219); matches are accepted
 * for API compatibility and contribute nothing. */
typedef struct sh_keypoints
{
    uint32_t num_matches;
} sh_keypoints;

typedef struct sh_solver sh_solver;

const char *sh_version(void);
const char *sh_status_string(int status);
/* Text of the last HIP runtime error behind an SH_E_HIP status (process-wide). */
const char *sh_last_hip_error(void);

/* --------------------------------------------------------------- one solve */
/* WarpSolver() + AllocateBuffer (WarpSolver.cu:50-103): all capacity is
 * allocated here. */
int sh_solver_create(const sh_solver_config *config, sh_solver **out);
int sh_solver_destroy(sh_solver *solver);

/* WarpSolver::SetSolverInputs (WarpSolver.cu:125-155); cam2world is the
 * row-major 4x4 camera-to-world extrinsic (inverted here, as at :151-154). */
int sh_solver_set_inputs(sh_solver *solver, const sh_camera *camera, const sh_measure_maps *measure,
                         const sh_render_maps *render, const sh_geometry *geometry, const sh_node_graph *node_graph,
                         const sh_node_flow *node_flow, const sh_flow_map *flow, const sh_keypoints *keypoints,
                         const float cam2world[16]);

/* WarpSolver::SolveStreamed (WarpSolver.cu:260-279): index build, then
 * gn_iterations x (terms, JtJ/JtR, PCG, warp update) starting from identity.
 * Work is enqueued on `stream`; the call returns after the stream has been
 * synchronised (OptimizationProcessorWarpSolver.cu:51).  The calling thread
 * waits by polling an event (it stays busy on one core for the index build's
 * two readbacks and the end of the solve) rather than sleeping in a blocking
 * synchronisation, which wakes tens of microseconds late.  SH_E_ARG when a
 * valid index_map entry is not below num_surfels, or a surfel KNN or edge
 * node id is not below num_nodes (the reference reads such ids unchecked). */
int sh_solver_solve(sh_solver *solver, sh_stream_t stream);

/* WarpSolver::SolvedNodeSE3 / OptimizationProcessorWarpSolver::SolvedSE3
 * (WarpSolver.h:45, OptimizationProcessorWarpSolver.h:28-30): device view of
 * num_nodes dual quaternions.  Before the first solve the view holds
 * max_nodes identity transforms, as the reference's SolverIterationData
 * starts (SolverIterationData.cu:78-88, 127-132); afterwards it is the last
 * solve's field (num_nodes of that solve) until the next solve overwrites it. */
int sh_solver_node_se3(const sh_solver *solver, const sh_dual_quat **device_view, uint32_t *num_nodes);

/* Sizes of the last index build: number of valid pixels P, unique node pairs U
 * (upper triangle), node->term entries. */
int sh_solver_sizes(const sh_solver *solver, uint32_t *num_pixels, uint32_t *num_pairs, uint32_t *num_node_terms);

/* Device views of internal arrays for parity checks: returns the device
 * pointer and element count.  Names: pix, pknn, pw, n2t_off, n2t_terms,
 * half_keys, half_off, pair_entries, upper_pos, lower_pos (with
 * SH_FLAG_REFERENCE_INDEX), row_ptr, blk_col, hg (the per-pixel H/g records;
 * on the tile path with FP32 streams the tile kernel builds them without
 * storing them, and they are stored by the next matrix-free apply), dense_cnt, seg_ptr,
 * tile_cbase, tile_sbase, list_off, blk_transpose, reg, blk, dinv, rhs, x,
 * dq0, dq1, delta, rot_cut (uint8, gn_iterations x num_nodes: 1 where the
 * warp update dropped the node's rotation increment because |w|_1 < 1e-4,
 * SolverIterationData.cu:42), twist_l1 (float, gn_iterations x num_nodes:
 * the FP32 |w|_1 each of those decisions was taken on). */
int sh_solver_debug_view(const sh_solver *solver, const char *name, const void **device_ptr, size_t *count);

/* FP64 total weighted cost sum lambda^2 r^2 (dense depth + flow, ARAP) at the start
 * of every GN iteration, plus the final one with SH_FLAG_COST_HISTORY: the GN
 * acceptance residual (the reference sums it in FP32 for logging only,
 * DenseImageDebug.cu:5-27).  Host copy; *count = number of entries; out may
 * be NULL to query the count. */
int sh_solver_cost_history(const sh_solver *solver, double *out, uint32_t capacity, uint32_t *count);

/* ApplySpMVBase<6>::ApplySpMV (libs/star/pcg_solver/ApplySpMVBase.h:19-37):
 * y = A x with the system of the last GN iteration (materialised or matrix-free). */
int sh_solver_apply_jtj(sh_solver *solver, const float *x, float *y, uint32_t matrix_free, sh_stream_t stream);

/* ---------------------------------------------- linear-solver plug-in
 * The reference's BlockPCG<6> over an abstract ApplySpMVBase<6>
 * (libs/star/pcg_solver/BlockPCG.h:14-115, ApplySpMVBase.h:9-47), for a
 * system the caller brings.  Vectors are FP32 device arrays of matrix_size
 * entries (6 per node); dot products, alpha and beta are FP64. */

/* ApplySpMVBase<6> (ApplySpMVBase.h:9-37) as a table of host callbacks that
 * enqueue device work on `stream` and return an sh_status:
 *   apply_spmv     y = A x             (ApplySpMV, :22-24)
 *   init_residual  r = b - A x_init    (InitResidual, :32-36); NULL: the
 *                  solver computes it through apply_spmv
 * matrix_size is MatrixSize() (:19). */
typedef struct sh_spmv_ops
{
    void *user;
    uint32_t matrix_size;
    int (*apply_spmv)(void *user, const float *x, float *y, sh_stream_t stream);
    int (*init_residual)(void *user, const float *x_init, const float *b, float *residual, sh_stream_t stream);
} sh_spmv_ops;

/* ApplySpMVBinBlockCSR<6> (ApplySpMVBinBlockCSR.h:8-52) in this library's own
 * block layout: a caller-owned 6x6 block-CSR with num_block_rows rows,
 * row_ptr[num_block_rows + 1] block offsets, block_col[num_blocks] block
 * columns and blocks[num_blocks * 36] FP32 row-major 6x6 blocks; (i, j) holds
 * A[6i..6i+5][6j..6j+5].  The arrays are borrowed and must outlive the handle.
 * sh_spmv_check verifies the structure on the device (SH_E_ARG when a row
 * range or a column is out of bounds); apply and init_residual trust it. */
typedef struct sh_spmv sh_spmv;
int sh_spmv_create_block_csr(uint32_t num_block_rows, uint32_t num_blocks, const uint32_t *row_ptr,
                             const uint32_t *block_col, const float *blocks, sh_spmv **out);
int sh_spmv_destroy(sh_spmv *spmv);
int sh_spmv_check(const sh_spmv *spmv, sh_stream_t stream);
int sh_spmv_matrix_size(const sh_spmv *spmv, uint32_t *matrix_size);
int sh_spmv_apply(sh_spmv *spmv, const float *x, float *y, sh_stream_t stream);
int sh_spmv_init_residual(sh_spmv *spmv, const float *x_init, const float *b, float *residual, sh_stream_t stream);
/* The callback table of a block-CSR operator (for sh_block_pcg_set_input). */
int sh_spmv_ops_of(sh_spmv *spmv, sh_spmv_ops *ops);
/* The callback table of a solver's own system of its last GN iteration
 * (materialised blocks, or matrix-free when matrix_free != 0; each call
 * synchronises `stream`).  Valid while the solver lives. */
int sh_solver_spmv_ops(sh_solver *solver, uint32_t matrix_free, sh_spmv_ops *ops);

/* BlockPCG<6> (BlockPCG.h:14-115).  sh_block_pcg_create allocates its
 * vectors for max_matrix_size (a multiple of 6), BlockPCG(max_matrix_size)
 * (:50).  set_input is SetSolverInput(inv_diag_blks, spmv_handler, b, x_init,
 * actual_size) (:55-61): inv_diag_blocks holds matrix_size / 6 FP32 6x6
 * inverse diagonal blocks, row-major; x is the caller's solution vector,
 * read as x_init when a solve does not zero-initialise; actual_size 0 takes
 * the operator's matrix_size; SH_E_CAPACITY beyond max_matrix_size.  All
 * pointers are borrowed until the next set_input. */
typedef struct sh_block_pcg sh_block_pcg;
int sh_block_pcg_create(uint32_t max_matrix_size, sh_block_pcg **out);
int sh_block_pcg_destroy(sh_block_pcg *pcg);
int sh_block_pcg_set_input(sh_block_pcg *pcg, const float *inv_diag_blocks, const sh_spmv_ops *spmv, const float *b,
                           float *x, uint32_t actual_size);
/* Solve / SolveNoConvergeCheck (BlockPCG.h:64-65, BlockPCG.cuh:325-349):
 * exactly max_iterations block-Jacobi PCG iterations from x = 0 (zero_init)
 * or from x (r = b - A x through InitResidual); the result is left in x.
 * Returns after synchronising `stream`. */
int sh_block_pcg_solve(sh_block_pcg *pcg, uint32_t max_iterations, int zero_init, sh_stream_t stream);
/* SolveConvergeChecked (BlockPCG.h:68-71, BlockPCG.cuh:351-393): stops after
 * the first iteration with |r.M^-1 r| < epsilon^2 |r0.M^-1 r0| (one host
 * readback per iteration, as in the reference); *iterations = the count run. */
int sh_block_pcg_solve_converge_checked(sh_block_pcg *pcg, uint32_t max_iterations, int zero_init, float epsilon,
                                        sh_stream_t stream, uint32_t *iterations);
/* r.M^-1 r (FP64) before the first iteration and after each iteration of the
 * last solve: *count = iterations + 1; out may be NULL to query the count. */
int sh_block_pcg_delta_history(const sh_block_pcg *pcg, double *out, uint32_t capacity, uint32_t *count);

/* ------------------------------------------------------- input producers */
/* SkinnerKernel (libs/star/geometry/node_graph/Skinner.cu:10-94): K nearest
 * nodes per surfel with normalised Gaussian weights, connect weight 1. */
int sh_skin_surfels(const float *surfel_vertex4, uint32_t num_surfels, const float *node_coords4, uint32_t num_nodes,
                    uint32_t knn, float node_radius_square, uint16_t *surfel_knn, float *spatial_weight,
                    float *connect_weight, sh_stream_t stream);

/* NodeGraph::BuildNodeGraphFromScratch (NodeGraph.cu:10-118, 339-365):
 * FP16-rounded pairwise distances, K-NN (self included), weights, edge list
 * (i, knn[i][k], k) of num_nodes*K ushort3, connect weights 1. */
int sh_build_node_graph(const float *node_coords4, uint32_t num_nodes, uint32_t knn, float node_radius_square,
                        uint16_t *node_knn, float *node_spatial_weight, uint16_t *node_graph,
                        float *node_connect_weight, sh_stream_t stream);

/* ------------------------------------------------- batched uneven solves */
typedef struct sh_batch sh_batch;
int sh_batch_create(const sh_solver_config *config, uint32_t max_solves, sh_batch **out);
int sh_batch_destroy(sh_batch *batch);
int sh_batch_set_inputs(sh_batch *batch, uint32_t slot, const sh_camera *camera, const sh_measure_maps *measure,
                        const sh_render_maps *render, const sh_geometry *geometry, const sh_node_graph *node_graph,
                        const sh_node_flow *node_flow, const sh_flow_map *flow, const sh_keypoints *keypoints,
                        const float cam2world[16]);
/* Solves slots [0, num_solves) together: one launch per GN phase covers all. */
int sh_batch_solve(sh_batch *batch, uint32_t num_solves, sh_stream_t stream);
int sh_batch_node_se3(const sh_batch *batch, uint32_t slot, const sh_dual_quat **device_view, uint32_t *num_nodes);

/* ---------------------------------------------- depth -> measurement maps
 * The step before the solve: SurfelMapInitializer::InitFromRGBDImage's
 * geometry half (libs/star/geometry/geometry_map/SurfelMapInitializer.cu:
 * 213-301), which MeasureProcessorOffline::ProcessFrame runs on every frame
 * (examples/mono_star/measure/MeasureProcessorOffline.cu:81).  Images are
 * linear row-major device arrays; reads outside an image are zero (the
 * reference's border-addressed textures). */
typedef struct sh_depth_camera
{
    uint32_t width, height;    /* raw depth image */
    float fx, fy, cx, cy;      /* raw intrinsics (config "intrinsic") */
    float clip_near, clip_far; /* metres */
    float downsample_scale;    /* solver maps are floor(scale * width) x floor(scale * height) */
    float surfel_radius_scale; /* system "surfel_radius_scale" */
} sh_depth_camera;

/* filterUnreliableDepth (libs/star/img_proc/surfel_reliable_filter.cu:37-191):
 * raw depth [mm] -> filtered depth [m], raw size.  0 where the depth is nearer
 * than clip_near or the 7x7 window normal is missing (<= 24 valid points) or
 * faces away (camera direction . normal > -0.1); 10 beyond clip_far. */
int sh_filter_depth(const uint16_t *depth_mm, const sh_depth_camera *camera, float *filtered_depth_m,
                    sh_stream_t stream);

/* computeRawVertexFromDepth + filterAndScaleVertex + createNormalRadiusMap
 * (SurfelMapInitializer.cu:83-137, 273-343; libs/star/img_proc/normal_map.cu:
 * 32-136): filtered depth -> the solver's Measure4Solver maps at the
 * downsampled size: vertex_confid (float4 x, y, z [m], 1; zero outside
 * [clip_near, clip_far]) and normal_radius (float4 normal, surfel radius [mm]).
 * Nearest downsampling: scaled pixel (x, y) takes raw texel
 * (int(x / scale), int(y / scale)). */
int sh_measure_maps_from_depth(const float *filtered_depth_m, const sh_depth_camera *camera, float *vertex_confid,
                               float *normal_radius, sh_stream_t stream);

/* createNormalRadiusMap (normal_map.cu:32-136) on any float4 vertex map. */
int sh_normal_radius_map(const float *vertex_confid, uint32_t width, uint32_t height, float surfel_radius_scale,
                         float *normal_radius, sh_stream_t stream);

/* createVertexConfidMap (libs/star/img_proc/vertex_map.cu:13-57): raw depth
 * [mm] -> float4 vertex map (x, y, z [m], confidence 1), no filtering. */
int sh_vertex_confid_map(const uint16_t *depth_mm, uint32_t width, uint32_t height, float fx, float fy, float cx,
                         float cy, float *vertex_confid, sh_stream_t stream);

/* ------------------------------------------------- optical flow and node motion
 * VolumeDeformFileFetch::FetchOFImage (libs/star/io/VolumeDeformFileFetch.cpp:
 * 39-52) + createScaledOpticalFlowMap (libs/star/img_proc/optical_flow_map.cu:
 * 5-57): a decoded .of.png (16-bit, 3 channels interleaved as read, the first
 * two (flow + 1000) * 10 of dx, dy in raw pixels; tools/transfer_hdf5.py:42-52)
 * -> the solver's flow map (float2) at floor(scale*width) x floor(scale*height):
 * raw pixel (round-half-even(x / scale), ...), flow = (u16 / 10 - 1000) * scale. */
int sh_flow_map_from_of16(const uint16_t *of16, uint32_t raw_width, uint32_t raw_height, float downsample_scale,
                          float *flow_map2, sh_stream_t stream);

/* ResetNodeMotion + AtomicSumIdealNodeMotionFromOpticalFlowKernel +
 * AverageNodeMotion (libs/star/geometry/node_graph/node_graph_opticalflow.cu:
 * 14-92, 180-245): the node-motion prior (sh_node_flow.node_motion_pred,
 * float4: world motion, weight) from the flow between the previous and the
 * current frame.  Per previous-frame pixel with a surfel s (index_map_prev) and
 * a current depth at the flowed pixel (rgbd w = 1 / depth, point-sampled):
 * the camera-space motion of the back-projected points (dropped when |dz| >
 * 0.05 m), times the spatial weight of s's first KNN slot, rotated by
 * cam2world, summed onto that node with the weight; each node is then divided
 * by its weight sum.  The sums are exact integers of 2^-31 units (the
 * reference's float atomics are order-dependent), so the result is
 * deterministic.  Surfel / node ids out of range contribute nothing.
 * workspace: num_nodes * 32 bytes. */
int sh_node_motion_from_flow(const float *rgbd_prev4, const float *rgbd4, const float *flow_map2,
                             const uint32_t *index_map_prev, uint32_t width, uint32_t height,
                             const uint16_t *surfel_knn, const float *surfel_spatial_weight, uint32_t num_surfels,
                             uint32_t knn, const float cam2world[16], float fx, float fy, float cx, float cy,
                             uint32_t num_nodes, void *workspace, float *node_motion_pred4, sh_stream_t stream);

/* No counterpart in the reference: sh_node_motion_from_flow without the
 * workspace argument, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It allocates nothing: the node sums live in a library-owned
 * device array of SH_NO_ALLOC_MAX_NODES nodes x 32 bytes (128 KiB), one per
 * GPU, held until the process ends and cleared by a kernel at the start of
 * every call.  num_nodes above SH_NO_ALLOC_MAX_NODES: SH_E_CAPACITY.  Every
 * call on one GPU shares that GPU's array, so calls that overlap on one GPU
 * (other streams or threads) race; serialise them.  Same results as
 * sh_node_motion_from_flow, bit for bit. */
#define SH_NO_ALLOC_MAX_NODES 4096u
int sh_node_motion_from_flow_noalloc(const float *rgbd_prev4, const float *rgbd4, const float *flow_map2,
                                     const uint32_t *index_map_prev, uint32_t width, uint32_t height,
                                     const uint16_t *surfel_knn, const float *surfel_spatial, uint32_t num_surfels,
                                     uint32_t knn, const float cam2world[16], float fx, float fy, float cx, float cy,
                                     uint32_t num_nodes, float *node_motion_pred4, sh_stream_t stream);

/* ------------------------------------------------------- solver maps
 * Renderer::DrawSolverMapsWithRecentObservation / ...ConfidentObservation
 * (libs/star/geometry/render/RendererDraw.cu:104-174, shaders/solver_map.vert
 * :21-53, solver_map.frag:1-17): the model surfels rendered as one-pixel
 * points into the solver's reference maps, without OpenGL.  Per surfel:
 * p = world2camera (row-major, host) * (x, y, z, 1); culled when p.z >
 * max_depth or p.z <= 0, or when its confidence (vertex w) is below 1 and
 * |color_time.z - current_time| >= 3 (always, with_recent_observation = 0);
 * color_time4 NULL skips that test.  Pixel (floor(fx p.x/p.z + cx + 0.5),
 * floor(fy p.y/p.z + cy + 0.5)) of camera (the downsampled size and
 * intrinsics); the nearest point wins (depth (p.z/max_depth)/2 + 1/2 < 1),
 * equal depths go to the lower surfel index.  Outputs (width*height, row-major):
 * reference vertex (x, y, z, 1) and normal (x, y, z, 0) of the winner, its
 * index; zero / 0xffffffff where no point lands.  zbuffer: width*height*8
 * bytes of caller workspace. */
int sh_render_solver_maps(const float *reference_vertex4, const float *reference_normal4, const float *color_time4,
                          uint32_t num_surfels, const float world2camera[16], const sh_camera *camera,
                          float max_depth, float current_time, int with_recent_observation, void *zbuffer,
                          float *reference_vertex_map4, float *reference_normal_map4, uint32_t *index_map,
                          sh_stream_t stream);

/* SurfelNodeDeformer::ForwardWarpSurfelsAndNodes (libs/star/geometry/surfel/
 * SurfelNodeDeformer.cu:13-80): live surfels and nodes from the reference ones
 * under a warp field (num_nodes dual quaternions).  Each node DQ is normalised
 * before the K-slot blend (DualQuaternion.hpp:289-306); v' = R v + t,
 * n' = R n, w components kept.  Either set may be empty (count 0). */
int sh_forward_warp(const sh_dual_quat *warp_field, uint32_t num_nodes, uint32_t knn,
                    const float *reference_vertex4, const float *reference_normal4, const uint16_t *surfel_knn,
                    const float *surfel_spatial_weight, const float *surfel_connect_weight, uint32_t num_surfels,
                    const float *reference_node4, const uint16_t *node_knn, const float *node_spatial_weight,
                    const float *node_connect_weight, float *live_vertex4, float *live_normal4, float *live_node4,
                    sh_stream_t stream);

/* ------------------------------------------------------------- serving */
/* No single counterpart in the reference, which solves one frame of one camera
 * per ProcessFrame call (examples/mono_star/app.cu:132-141).  A server takes
 * tracking-solve requests from many camera streams: each stream's solver
 * state (surfel KNN ids and weights, node graph, last warp field) stays
 * resident in server-owned HBM; solves are batched continuously at
 * Gauss-Newton-iteration granularity (requests join at the next iteration
 * boundary as slots free up); frames of one stream are solved in order. */
typedef struct sh_server sh_server;

/* hbm_budget_bytes value: the device's free HBM when the server is created
 * (hipMemGetInfo), less 1/16 of it as headroom. */
#define SH_HBM_BUDGET_FREE (~(uint64_t)0)

typedef struct sh_server_config
{
    sh_solver_config solver;   /* per-solve capacity (the largest request) and solver constants */
    uint32_t max_active;       /* solves in flight at most (continuous-batch width) */
    uint64_t hbm_budget_bytes; /* device bytes the server may hold: solve slots, their tile and partial buffers
                                  (which grow to the largest frame a slot has held) and resident streams;
                                  0 = no limit; SH_HBM_BUDGET_FREE = the free HBM at creation */
} sh_server_config;

typedef struct sh_request_result
{
    uint64_t ticket;
    uint32_t stream_id;
    int status;
    double submit_ms, admit_ms, done_ms; /* server clock, ms since sh_server_create */
} sh_request_result;

typedef struct sh_server_stats
{
    uint64_t slots;           /* solve slots the budget allowed */
    uint64_t streams, streams_refused;
    uint64_t submitted, completed, failed;
    uint64_t steps;           /* GN-iteration launches */
    uint64_t max_batch;       /* most solves in one launch */
    uint64_t slot_iterations; /* sum of batch sizes over steps */
    uint64_t hbm_reserved;    /* bytes held */
    uint64_t hbm_budget;      /* the budget in force (0 = no limit) */
    uint64_t growth_refused;  /* index builds refused because a slot's buffers would pass the budget */
    double now_ms;
} sh_server_stats;

/* Creates the server and its two worker threads.  Fewer than max_active slots
 * are made when the budget cannot hold them; SH_E_CAPACITY when not even one fits.
 * Index builds run on other streams while GN steps run, so the server solves
 * with the split PCG (SH_FLAG_NO_PERSISTENT_PCG: no kernel waits on another
 * group); a served result equals a solve alone with that flag, bit for bit.
 * SH_E_ARG for SH_FLAG_PAIR_MAJOR configurations (their PCG cannot change). */
int sh_server_create(const sh_server_config *config, sh_server **out);
int sh_server_destroy(sh_server *server);
/* Registers a stream and copies its geometry and node graph into resident HBM;
 * SH_E_CAPACITY when the budget cannot hold it (admission by memory budget).
 * A request whose index build would grow its slot's buffers past the budget
 * completes with SH_E_CAPACITY (sh_server_stats.growth_refused).
 * With reference surfels (float4 vertex and normal, both or neither), the
 * server keeps them resident too and forward-warps them and the nodes after
 * every solve of the stream (sh_server_live_surfels). */
int sh_server_add_stream(sh_server *server, uint32_t stream_id, const sh_camera *camera, const sh_geometry *geometry,
                         const sh_node_graph *node_graph, const float cam2world[16], const float *reference_vertex4,
                         const float *reference_normal4);
/* Device views of a stream's live surfels (vertex, normal: num_surfels float4)
 * and live nodes (float4): the reference ones under the last solved warp field. */
int sh_server_live_surfels(sh_server *server, uint32_t stream_id, const float **live_vertex4,
                           const float **live_normal4, const float **live_node4, uint32_t *num_surfels);
/* Queues one frame of a stream.  The maps are caller-owned device memory and must
 * stay valid until the request is done; out_device (may be NULL) receives the
 * solved node dual quaternions. */
int sh_server_submit(sh_server *server, uint32_t stream_id, const sh_measure_maps *measure,
                     const sh_render_maps *render, const sh_flow_map *flow, const sh_node_flow *node_flow,
                     sh_dual_quat *out_device, uint64_t *ticket);
/* Blocks until the request is done; returns its status and timestamps. */
int sh_server_wait(sh_server *server, uint64_t ticket, sh_request_result *result);
/* Device view of a stream's resident warp field (the last solve's result). */
int sh_server_warp_field(sh_server *server, uint32_t stream_id, const sh_dual_quat **device_view,
                         uint32_t *num_nodes);
int sh_server_get_stats(sh_server *server, sh_server_stats *stats);

/* ------------------------------------------------------------ autotuner */
/* No counterpart in the reference (its launch shapes are compile-time).  A
 * tuning is what can change between solves without reallocating. */
typedef struct sh_tuning
{
    uint32_t flags;              /* subset of SH_FLAG_ROW_SEGMENTS | SH_FLAG_STREAMING_PCG | SH_FLAG_NO_PERSISTENT_PCG |
                                    SH_FLAG_PIPELINED_PCG | SH_FLAG_TILE_PIX_64 | SH_FLAG_TILE_PIX_256 */
    uint32_t spmv_mode;          /* sh_spmv_mode */
    uint32_t pcg_rows_per_group; /* register-resident PCG: node rows per work-group, 1..85 (0 = 85) */
    double ms;                   /* autotune: median solve time of the chosen tuning */
} sh_tuning;

/* Times each candidate tuning on the inputs already set (median of `repeats`
 * solves after one warm-up, HIP events on `stream`), keeps the fastest and
 * returns it; a candidate must beat an earlier one by 2% to replace it, so
 * near-ties keep the earlier (default-first) choice.  The candidates are the
 * materialisation paths and PCG variants with their rows per group, then the
 * tile sizes 64 and 256 for a winner on the tile path.  Every candidate
 * solves the same problem to the same tolerance. */
int sh_solver_autotune(sh_solver *solver, sh_stream_t stream, uint32_t repeats, sh_tuning *best);
int sh_solver_set_tuning(sh_solver *solver, const sh_tuning *tuning);
int sh_batch_autotune(sh_batch *batch, uint32_t num_solves, sh_stream_t stream, uint32_t repeats, sh_tuning *best);
int sh_batch_set_tuning(sh_batch *batch, const sh_tuning *tuning);
/* The tuning in use (ms = 0). */
int sh_solver_get_tuning(const sh_solver *solver, sh_tuning *tuning);

/* Persistent-PCG fallback (no counterpart in the reference, whose PCG is one
 * launch per step).  The persistent PCG kernels need every work-group of a
 * solve resident at once and wait for each other at their barriers; kernels
 * that the caller runs on other streams at the same time can hold CUs long
 * enough for a wait to run out.  Each wait polls at most `polls` times
 * (0: the default, 2^22, about 0.1 s), after which the launch drains with
 * meaningless sums.  A solve (or batch step) whose wait ran out is re-run
 * with the split PCG, which has no inter-group waits, and the solver keeps
 * the split PCG afterwards (its tuning then holds SH_FLAG_NO_PERSISTENT_PCG;
 * a later set_tuning or autotune may choose a persistent PCG again; an
 * autotune candidate that ran out is not chosen).  sh_solver_pcg_fallbacks
 * counts the re-runs. */
int sh_solver_set_wait_limit(sh_solver *solver, uint32_t polls);
int sh_solver_pcg_fallbacks(const sh_solver *solver, uint32_t *count);

/* ------------------------------------------------------- measurement hook */
/* No counterpart in the reference (which times nothing on the device).  With
 * SH_FLAG_PHASE_TIMING in the config: the event-to-event time (ms) and the
 * launch count of every GN phase since the last reset, for phases
 * 0 .. num_phases-1 (names from sh_phase_name); reset != 0 zeroes them after
 * the copy.  Phases: dense_terms, reg_terms, rows_seg, rows_combine,
 * diag_rhs, offdiag, dinv, pcg, update, tiles, tiles_combine. */
const char *sh_phase_name(uint32_t phase);
int sh_solver_phase_times(sh_solver *solver, double *ms, uint32_t *calls, uint32_t num_phases, uint32_t reset);
int sh_batch_phase_times(sh_batch *batch, double *ms, uint32_t *calls, uint32_t num_phases, uint32_t reset);
/* The same for a server created with SH_FLAG_PHASE_TIMING in its solver
 * config: the GN-phase kernel time of its steps (a launch covers every solve
 * admitted at that step). */
int sh_server_phase_times(sh_server *server, double *ms, uint32_t *calls, uint32_t num_phases, uint32_t reset);

#ifdef __cplusplus
}
#endif
#endif /* SH_WARP_SOLVER_H */
