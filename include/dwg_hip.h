/*
 * dwg_hip.h — C ABI of the MI355X-native DWG (diffusing winding gradients)
 * reconstruction library (libdwg_hip.so).
 *
 * Plain C: opaque handles, plain pointers, int64 sizes, int status codes.
 * Device pointers ("d_") are HIP device memory owned by the caller; host
 * pointers ("h_") are ordinary host memory.  Point/normal arrays are packed
 * float triples (x,y,z per element, 12-byte stride), the layout of the
 * reference's float3 buffers.  Every call is asynchronous on the context's
 * stream unless it returns a host value.
 *
 * Status: 0 = DWG_OK; otherwise an error code, and dwg_ctx_error() returns
 * the message.  No call exits or throws across the ABI (the reference
 * aborts via gpuErrchk → exit(), src/TurboDWG.cu:3-12).
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef DWG_HIP_H
#define DWG_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DWG_OK 0
#define DWG_ERR_ARG 1       /* invalid argument or size */
#define DWG_ERR_HIP 2       /* HIP runtime error */
#define DWG_ERR_CAPACITY 3  /* output buffer too small; required size returned */
#define DWG_ERR_STATE 4     /* object used in the wrong state */

/* flags for the winding-number sums */
#define DWG_SCREENED 1u   /* screened kernel (TurboLibiGWN.cu:552-585); else plain libigl kernel */
#define DWG_USE_RADII 2u  /* screening radius per source; else 0.0015 for every source */
#define DWG_KAHAN 4u      /* compensated (Kahan) accumulation of per-tile partial sums */
#define DWG_FAR_BF16 8u   /* far-field dipole moments stored and read as bf16 (tree only) */
#define DWG_NEAR_FP64 16u /* screened point terms inside the source's screening radius (r < s)
                             evaluated and summed in FP64 (new; direct and tree sums) */

typedef struct dwg_ctx dwg_ctx;
typedef struct dwg_octree dwg_octree;
typedef struct dwg_knn_index dwg_knn_index;
typedef struct dwg_mlp dwg_mlp;

/* Implicit regular grid: point (i,j,k) = origin + (i,j,k)·h, linear index
 * i + res[0]·(j + res[1]·k) (igl::grid order, grid.cpp:28-42). */
typedef struct dwg_grid {
    float origin[3];
    float h;
    int32_t res[3];
} dwg_grid;

/* ---- context ------------------------------------------------------------ */
/* stream: a hipStream_t, or NULL for the default stream. */
int dwg_ctx_create(int device, void* stream, dwg_ctx** out);
void dwg_ctx_destroy(dwg_ctx* ctx);
const char* dwg_ctx_error(const dwg_ctx* ctx);
int dwg_ctx_sync(dwg_ctx* ctx);
/* The library keeps the device blocks it releases (trees, indices, scratch)
 * for reuse by later calls; this hands every cached block back to HIP
 * (after synchronising each device that has some) and returns the bytes
 * freed in *out_bytes (may be NULL).  No reference counterpart: for
 * long-lived processes that want the memory back between workloads. */
int dwg_cache_release(int64_t* out_bytes);
/* Per-kernel timing with HIP events on the context stream (off by default). */
int dwg_ctx_profile(dwg_ctx* ctx, int enable);
/* Sum of recorded durations (ms) and launch count of kernel `name`; syncs. */
int dwg_ctx_profile_get(dwg_ctx* ctx, const char* name, double* total_ms, int64_t* launches);
int dwg_ctx_profile_reset(dwg_ctx* ctx);
/* Event count `name` recorded while profiling (e.g. "knn_query_centroids:redo",
 * queries that left the tiled kNN path); 0 if none. */
int dwg_ctx_counter_get(dwg_ctx* ctx, const char* name, int64_t* value);

/* ---- winding-number sums ------------------------------------------------ */
/* Direct all-pairs sums w(q) = Σ_j term(p_j − q) over n sources at nq
 * queries; replaces the β ≤ 0 branch of igl::fast_winding_number
 * (fast_winding_number.cpp:313-323) and, with DWG_SCREENED, the per-point
 * term of ScreenedGWNLibiPStoQ (TurboLibiGWN.cu:544-585).
 * d_radii may be NULL unless DWG_USE_RADII.  d_out_grad (nullable, nq×3)
 * receives ∇_q w (new; the reference has no gradient), of the plain or the
 * screened term; a screened term inside its screening radius (ρ = s) has
 * ∇w = −g(s)·A·n. */
int dwg_gwn_direct(dwg_ctx* ctx, const float* d_pts, const float* d_nrm, const float* d_area,
                   const float* d_radii, int64_t n, const float* d_q, int64_t nq, float sigma,
                   uint32_t flags, float* d_out_w, float* d_out_grad);

/* No counterpart in the reference: dwg_gwn_direct for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, kernels and results;
 * the packed sources go to a fixed device array of the library (8 MiB per
 * device, held until the process ends) instead of context buffers, so the
 * call allocates nothing.  n ≤ DWG_NOALLOC_MAX_SOURCES, else DWG_ERR_ARG.
 * Every call on one device shares that array: calls that overlap on a
 * device (other streams or contexts) race. */
#define DWG_NOALLOC_MAX_SOURCES 262144
int dwg_gwn_direct_noalloc(dwg_ctx* ctx, const float* d_pts, const float* d_nrm, const float* d_area,
                           const float* d_radii, int64_t n, const float* d_q, int64_t nq, float sigma,
                           uint32_t flags, float* d_out_w, float* d_out_grad);

/* Octree over sources: octreeDWG split rule (a node with more than
 * leaf_size points and depth < max_depth is split into 8 children,
 * octree.cpp:171-335; the reference uses leaf_size 1) plus the order-0
 * moments CM, R (fast_winding_number.cpp:54-119) and the GPU upload
 * (TurboLibiGWN::PreLibiGWN, TurboLibiGWN.cu:132-231); the whole build runs
 * on the GPU.  h_pts: host doubles (n×3); h_area: host doubles (n) or NULL
 * for the constant 2e-6 (DWGProcess.h:619).  Point order of every other
 * call stays the caller's. */
int dwg_octree_build(dwg_ctx* ctx, const double* h_pts, const double* h_area, int64_t n,
                     int max_depth, int leaf_size, dwg_octree** out);
/* The same from device arrays (d_pts n×3 FP64, d_area n FP64 or NULL). */
int dwg_octree_build_device(dwg_ctx* ctx, const double* d_pts, const double* d_area, int64_t n,
                            int max_depth, int leaf_size, dwg_octree** out);
void dwg_octree_destroy(dwg_octree* tree);
/* GetOctreeLeaves (DWGProcess.h:383-428): centres (device FP64, room for
 * n_leaves×3) and widths (device FP64 or NULL) of the leaf cells holding more
 * than min_num ≥ 0 points, in breadth-first order; the count in *out_count. */
int dwg_octree_leaves(dwg_ctx* ctx, const dwg_octree* tree, int min_num, double* d_centres, double* d_widths,
                      int64_t* out_count);
/* SmoothNormalstoOctreeLeaves + .normalized() (DWGProcess.h:430-472, 537):
 * for each leaf of dwg_octree_leaves' selection (same order), component k
 * of its normal is Σ_j exp(−((P_jk − C_k)/W)²/4)·N_jk over the leaf's
 * points j (C its centre, W its width); the whole n_leaves×3 matrix is then
 * divided by its Frobenius norm (rows are not unit length).  d_pts, d_nrm:
 * the tree's points and their normals, device FP64 n×3, caller order;
 * d_out: device floats, room for n_leaves×3. */
int dwg_octree_leaf_normals(dwg_ctx* ctx, const dwg_octree* tree, int min_num, const double* d_pts,
                            const double* d_nrm, float* d_out, int64_t* out_count);
int dwg_octree_info(const dwg_octree* tree, int64_t* n_nodes, int64_t* n_leaves, int* max_level);
/* The tree's points in tree order (libigl octree's point_indices,
 * octree.cpp:171-335, concatenated over the leaves): d_order (device,
 * room for n) receives the caller index of each point, sorted by octant
 * path with equal paths in caller order; every node's points are one
 * contiguous run of it. */
int dwg_octree_point_order(dwg_ctx* ctx, const dwg_octree* tree, int* d_order);
/* Per-iteration refresh of the far-field moment EC = Σ A·n per node and of
 * the packed sources (TurboLibiGWN::UpdateMassNormals, TurboLibiGWN.cu:1051-1059).
 * d_radii may be NULL. */
int dwg_octree_update(dwg_ctx* ctx, dwg_octree* tree, const float* d_nrm, const float* d_area,
                      const float* d_radii);
/* Re-packs the per-point terms (A·n, A) with other areas (d_area NULL: 2e-6)
 * and keeps the far-field moments EC of the last dwg_octree_update.  The
 * reference's sums take the point areas apart from the node moments
 * (TurboLibiScreenedGWNLaunch's A argument, TurboLibiGWN.cu:912-922);
 * GenerateMesh filters the areas after the moments (TurboDWG.cu:870-875). */
int dwg_octree_set_point_areas(dwg_ctx* ctx, dwg_octree* tree, const float* d_nrm, const float* d_area);
/* Barnes-Hut sums at nq query points (TurboLibiGWN::TurboLibiScreenedGWNLaunch,
 * TurboLibiGWN.cu:912-922; plain: TurboLibiGWNLaunch :900-910). */
int dwg_gwn_tree(dwg_ctx* ctx, const dwg_octree* tree, const float* d_q, int64_t nq, float beta,
                 float sigma, uint32_t flags, float* d_out);
/* Value and gradient ∇_q w of the Barnes-Hut sums at nq query points (new;
 * the reference approximates the gradient direction by marching-cubes face
 * normals): the term set of dwg_gwn_tree's FP32 walk, each term with its
 * exact gradient (far nodes: dipole EC at CM, s = 0.0015).  Flags:
 * DWG_SCREENED; the radii are those of the last dwg_octree_update.
 * d_out_grad (nq×3) may be NULL. */
int dwg_gwn_tree_grad(dwg_ctx* ctx, const dwg_octree* tree, const float* d_q, int64_t nq, float beta,
                      float sigma, uint32_t flags, float* d_out_w, float* d_out_grad);
/* The same sums at the octree's own points, evaluated in tree order (one
 * wave per 64 consecutive tree slots) and written in the caller's point order
 * (the source self-evaluation of TurboScreenedDWGLaunch, TurboDWG.cu:536). */
int dwg_gwn_tree_self(dwg_ctx* ctx, const dwg_octree* tree, float beta, float sigma, uint32_t flags,
                      float* d_out);
/* Barnes-Hut sums at every point of an implicit grid (the reference
 * materialises the grid and calls the launch above, TurboDWG.cu:537). */
int dwg_gwn_tree_grid(dwg_ctx* ctx, const dwg_octree* tree, const dwg_grid* grid, float beta,
                      float sigma, uint32_t flags, float* d_out);
/* Exact number of terms (point terms + dipoles) each query takes (new). */
int dwg_gwn_tree_terms(dwg_ctx* ctx, const dwg_octree* tree, const float* d_q, int64_t nq,
                       float beta, int64_t* d_terms);
int dwg_gwn_tree_grid_terms(dwg_ctx* ctx, const dwg_octree* tree, const dwg_grid* grid,
                            float beta, int64_t* h_total_terms);

/* ---- k nearest neighbours ------------------------------------------------ */
/* Exact kNN under float d² = fma(dz,dz,fma(dy,dy,dx·dx)), results ascending
 * by (d², index) — the order of cukd's FixedCandidateList (knn.h:403-414);
 * the index is the position in the caller's d_pts.  Replaces
 * GenerateKDTreeForDWG + CudaKNNWithKDTreeforDWG (KNN_GWN.cu:183-212, 490-715). */
int dwg_knn_build(dwg_ctx* ctx, const float* d_pts, int64_t n, int k_hint, dwg_knn_index** out);
void dwg_knn_destroy(dwg_knn_index* index);
/* k ≤ 32.  d_skip (nullable): points with d_skip[j] != 0 are not candidates.
 * d_qmask (nullable): queries with d_qmask[i] == 0 are not computed.
 * d_d2 (nullable) receives the distances.  Missing neighbours are -1. */
int dwg_knn_query(dwg_ctx* ctx, const dwg_knn_index* index, const float* d_q, int64_t nq, int k,
                  const uint8_t* d_skip, const uint8_t* d_qmask, int32_t* d_idx, float* d_d2);

/* ---- marching cubes ------------------------------------------------------ */
/* Iso-surface of a grid field (TurboMC::TurboMarchingCubes, TurboMC.cu:572-638).
 * Deterministic order: cubes x-fastest, triangles in table order.  Any of
 * d_tris (T×9), d_cent (T×3), d_nrm (T×3) may be NULL.  *h_count receives
 * T; if T > capacity nothing is written and DWG_ERR_CAPACITY is returned. */
int dwg_marching_cubes(dwg_ctx* ctx, const float* d_field, const dwg_grid* grid, float iso,
                       int64_t capacity, float* d_tris, float* d_cent, float* d_nrm,
                       int64_t* h_count);
/* No counterpart in the reference: the same triangles' centroids and
 * normals (values as dwg_marching_cubes gives them) in one pass and in no
 * fixed order, for callers whose use does not depend on the order (the
 * normal diffusion's fixed-point sums).  *h_count receives T; if
 * T > capacity, DWG_ERR_CAPACITY is returned and the buffers hold an
 * unspecified part of the triangles. */
int dwg_marching_cubes_any_order(dwg_ctx* ctx, const float* d_field, const dwg_grid* grid, float iso,
                                 int64_t capacity, float* d_cent, float* d_nrm, int64_t* h_count);

/* ---- normal diffusion ---------------------------------------------------- */
/* One diffusion update of source normals from triangle face normals
 * (TurboDWG.cu:560-644): 15-NN of every centroid among the sources, sum of
 * face normals per source (exact fixed-point, deterministic), normalise
 * (|acc| < 1e-10 keeps the old normal and counts as zero), and when more
 * than zero_threshold sources are zero, refill them with the normalised sum
 * of their zero_k nearest non-zero sources.  *h_nzero receives the count. */
int dwg_diffuse_normals(dwg_ctx* ctx, const dwg_knn_index* src_index, const float* d_src,
                        int64_t n_src, const float* d_cent, const float* d_tri_nrm, int64_t ntri,
                        int k, int zero_k, int zero_threshold, float* d_nrm, int64_t* h_nzero);
/* r_i = clamp(sqrt(mean_{j∈kNN(i)} |p_i − p_j|²), rmin, rmax), self included
 * (DWGProcess.h:325-343; TurboDWG::UpdateRadii, TurboDWG.cu:674-691). */
int dwg_radii(dwg_ctx* ctx, const dwg_knn_index* index, const float* d_pts, int64_t n, int k,
              float rmin, float rmax, float* d_out);
/* PCA normals (PCL NormalEstimation, DWGProcess.h:152-200): smallest-
 * eigenvalue direction of the covariance of each point's k nearest
 * neighbours (itself included; `index` built over d_pts), turned towards
 * h_viewpoint (the reference uses the bounding-box centre). */
int dwg_pca_normals(dwg_ctx* ctx, const dwg_knn_index* index, const float* d_pts, int64_t n, int k,
                    const float* h_viewpoint, float* d_out);
/* Voronoi point areas (igl::copyleft::cgal::point_areas, point_areas.cpp:
 * 41-205, which DWGProcess.h:345-380 and :790 call with k = 15) without
 * CGAL: the k nearest neighbours of each point (itself first; `index` over
 * d_pts), those whose normal faces the point's own (d_nrm, NULL = all kept),
 * the PCA plane, the Delaunay triangles at the point, its Voronoi share. */
int dwg_point_areas(dwg_ctx* ctx, const dwg_knn_index* index, const float* d_pts, int64_t n, int k,
                    const float* d_nrm, float* d_out);
/* n_i = (p_i − c)/(|p_i − c| + 1e-8) (TurboNormalInit.cu:3-27). */
int dwg_gauss_map(dwg_ctx* ctx, const float* d_pts, int64_t n, const float* h_center,
                  float* d_nrm);
/* n_q = normalise(mean of the k nearest source normals) (LoadNormals,
 * TurboDWG.cu:206-236). */
int dwg_transfer_normals(dwg_ctx* ctx, const dwg_knn_index* src_index, const float* d_src_nrm,
                         const float* d_q, int64_t nq, int k, float* d_out_nrm);
/* Divide the largest ceil(n·threshold) areas by 1000 (OutlierFilter,
 * TurboDWG.cu:991-1009). */
int dwg_outlier_filter(dwg_ctx* ctx, float* d_area, int64_t n, float threshold);
/* Convergence residual of a diffusion step (new; the reference runs a fixed
 * iteration count, TurboDWG.cu:516): mean over the n sources of
 * 1 − |cos ∠(old_i, new_i)|, reduced in FP64 on the GPU; the absolute value
 * ignores the global sign flip of each iteration; a zero normal counts 1. */
int dwg_normal_change(dwg_ctx* ctx, const float* d_old, const float* d_new, int64_t n, double* h_residual);
/* FP64 sum of a float array (the reference's thrust::reduce, TurboDWG.cu:541). */
int dwg_reduce_sum(dwg_ctx* ctx, const float* d_vals, int64_t n, double* h_sum);
/* Axis-aligned bounds of a point array (cukd computeBounds, builder_common.h:76-150). */
int dwg_bounds(dwg_ctx* ctx, const float* d_pts, int64_t n, float* h_lo, float* h_hi);

/* ---- learned normal initialisation --------------------------------------- */
/* New (no reference counterpart; the reference initialises with random or
 * Gauss-map normals, DWGProcess.h:264-281, TurboNormalInit.cu:3-27): a
 * 48→128→128→3 GELU MLP over the 16-nearest-neighbour patch of a point
 * (offsets scaled by the 16th-neighbour distance, distance order) predicts
 * an oriented normal.  BF16 weights and activations, FP32 accumulation on
 * the matrix cores, FP32 master weights and Adam state.  Parameter vector
 * layout: W1[128][48], b1[128], W2[128][128], b2[128], W3[3][128], b3[3]. */
int64_t dwg_mlp_num_params(void);
int dwg_mlp_create(dwg_ctx* ctx, uint64_t seed, dwg_mlp** out);
int dwg_mlp_load(dwg_ctx* ctx, const char* path, dwg_mlp** out);
int dwg_mlp_save(dwg_ctx* ctx, const dwg_mlp* mlp, const char* path);
void dwg_mlp_destroy(dwg_mlp* mlp);
/* host copies of the FP32 parameters / last gradient (either may be NULL) */
int dwg_mlp_get(dwg_ctx* ctx, const dwg_mlp* mlp, float* h_params, float* h_grad);
int dwg_mlp_set(dwg_ctx* ctx, dwg_mlp* mlp, const float* h_params);
/* n×48 BF16 patches (uint16 bit patterns) of the indexed points */
int dwg_mlp_patches(dwg_ctx* ctx, const dwg_knn_index* index, const float* d_pts, int64_t n,
                    uint16_t* d_patches);
/* patches of nq query points d_q over the n_src points d_src of `index`
 * (the first neighbour at distance 0 is dropped) */
int dwg_mlp_patches_from(dwg_ctx* ctx, const dwg_knn_index* index, const float* d_src, int64_t n_src,
                         const float* d_q, int64_t nq, uint16_t* d_patches);
/* Learned initialisation of n point normals: patches over the points (over
 * a 50,000-point hashed subsample for larger clouds), MLP normal line,
 * oriented away from h_center (the Gauss-map hemisphere, TurboNormalInit.cu:3-44). */
int dwg_mlp_init(dwg_ctx* ctx, const dwg_mlp* mlp, const float* d_pts, int64_t n, const float* h_center,
                 float* d_out);
/* n×3 outputs, unit length when normalize != 0 */
int dwg_mlp_forward(dwg_ctx* ctx, const dwg_mlp* mlp, const uint16_t* d_patches, int64_t n, int normalize,
                    float* d_out);
/* loss over the batch and its gradient (kept in the MLP): mode 0 the
 * oriented loss mean(1 − ŷ·t), mode 1 the normal-line loss mean(1 − (ŷ·t)²) */
int dwg_mlp_loss_grad(dwg_ctx* ctx, dwg_mlp* mlp, const uint16_t* d_patches, const float* d_target,
                      int64_t batch, int mode, double* h_loss);
int dwg_mlp_adam(dwg_ctx* ctx, dwg_mlp* mlp, float lr, float beta1, float beta2, float eps);
/* gathers batch rows d_index of (patches, targets) and rotates both by the
 * row-major 3×3 h_rotation (training-time augmentation) */
int dwg_mlp_augment(dwg_ctx* ctx, const uint16_t* d_patches, const float* d_target, const int64_t* d_index,
                    int64_t batch, const float* h_rotation, uint16_t* d_patches_out, float* d_target_out);
/* C[m×n] = A·B, BF16 operands with element strides, FP32 result (row-major
 * C); the matrix-core product every MLP layer uses */
int dwg_gemm_bf16(dwg_ctx* ctx, const uint16_t* d_a, int64_t s_ai, int64_t s_ak, const uint16_t* d_b,
                  int64_t s_bk, int64_t s_bj, float* d_c, int64_t m, int64_t n, int64_t k);

/* ---- whole reconstruction ------------------------------------------------ */
/* Options of DWGLaunch (src/main.cpp:5-15); defaults of src/main.cpp:19-39
 * via dwg_options_default().  Fields after output_mode have no reference
 * flag. */
typedef struct dwg_options {
    int32_t input_normal;        /* 0 random, 1 input normals, 2 random (PCA areas with com_As), 3 Gauss map, 4 learned MLP, 5 PCA */
    int32_t grid_reso;           /* --grid_reso, 400 */
    int32_t final_reso;          /* --final_reso, 500 */
    int32_t kw;                  /* --kw, 10 */
    int32_t final_kw;            /* --final_kw, 10 */
    int32_t normal_kw;           /* --normal_kw, 10 */
    int32_t max_depth;           /* --max_depth, 10 */
    float sigma;                 /* --lambda, 10 */
    int32_t min_num;             /* --min_num, 0 */
    int32_t use_radii;           /* --use_radii, 0 */
    int32_t com_As;              /* --com_As, 0 */
    float As_threshold;          /* --As_thres, 0 */
    float offset_1;              /* --offset_1, 0.05 */
    float offset_2;              /* --offset_2, 0.01 */
    int32_t use_leaves_generate; /* --use_leaves_generate, 0 */
    int32_t total_iter;          /* --total_iter, 40 (at least 1) */
    int32_t test_iter;           /* --test_iter, 999999 */
    int32_t output_mode;         /* --output_mode, 1 */
    int32_t source_mode;         /* 0: octree-leaf centres as the reference, 1: the input points */
    int32_t leaf_size;           /* octree leaf capacity for the sums (1 = octreeDWG) */
    float beta;                  /* Barnes-Hut accuracy, 2.3 (DWGProcess.h:689); ≤ 0: direct sums */
    uint32_t precision;          /* DWG_KAHAN | DWG_FAR_BF16 | DWG_NEAR_FP64 */
    uint64_t seed;               /* random initial normals */
    int32_t normalize_input;     /* 1: centre and scale as ReadPointandNormals (DWGProcess.h:53-97) */
    int32_t verbose;             /* progress lines on stdout */
    int32_t count_terms;         /* 1: count the Barnes-Hut terms of one iteration (untimed) */
    const dwg_mlp* mlp;          /* weights for input_normal 4 */
    int32_t sigma_warm_iters;    /* first iterations at σ = 20 (TurboDWG.cu:526-528), 4 */
    float residual_tol;          /* > 0: stop after an even iteration count whose convergence residual
                                    (dwg_normal_change of that iteration) is below it; 0: total_iter */
} dwg_options;

void dwg_options_default(dwg_options* opt);

typedef struct dwg_result {
    float* normals;        /* in: n×3 buffer or NULL; out: final point normals (normalised frame);
                              with use_leaves_generate 1 only when the sources are the points */
    int32_t want_mesh;     /* in: 1 to receive the final mesh in `mesh` */
    float* mesh;           /* out: tri_count×9 floats (normalised frame) in pinned host
                              memory allocated by the library; release with dwg_free() */
    int64_t tri_count;     /* out: triangles of the final mesh */
    int64_t n_sources;     /* out: sources of the iterations */
    double center[3];      /* out: input normalisation, p_norm = (p − center)/scale */
    double scale;
    double time_setup_s;   /* out: wall time of setup (trees, radii, grid, init) */
    double time_iter_s;    /* out: wall time of the iterations */
    double time_final_s;   /* out: wall time of the final mesh pass */
    double time_total_s;
    double* w_mean;        /* in: total_iter buffer or NULL; out: mean source WN per iteration */
    int64_t* iter_tris;    /* in: total_iter buffer or NULL; out: triangles per iteration */
    int64_t* iter_zeros;   /* in: total_iter buffer or NULL; out: zero normals per iteration */
    int64_t terms_src;     /* out (count_terms): terms of one source evaluation */
    int64_t terms_grid;    /* out (count_terms): terms of one grid evaluation */
    int64_t grid_points;   /* out: points of the iteration grid */
    float* source_normals; /* in: n×3 buffer or NULL; out: the n_sources×3 source normals after the
                              iterations, before the final pass (normalised frame) */
    double* residual;      /* in: total_iter buffer or NULL; out: convergence residual per iteration */
    int32_t iterations;    /* out: iterations run (total_iter unless residual_tol stopped earlier) */
} dwg_result;

/* Full reconstruction of n host points (DWGLaunch → DataprePointCloud →
 * MainIterative, src/main.cpp:5-15, DWGProcess.h:99-808).  h_nrm: input
 * normals for input_normal 1 (else NULL).  out_base (nullable): path prefix
 * for the STL / .xyz outputs as the reference names them. */
int dwg_launch(dwg_ctx* ctx, const float* h_pts, const float* h_nrm, int64_t n, const dwg_options* opt,
               const char* out_base, dwg_result* res);
/* Releases memory the library allocated for a caller (dwg_result.mesh).  Up
 * to two released pinned buffers are kept for reuse by later meshes of up to
 * their size (pinning costs milliseconds per hundred MB). */
void dwg_free(void* p);

#ifdef __cplusplus
}
#endif

#endif /* DWG_HIP_H */
