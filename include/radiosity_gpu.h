/* radiosity_gpu.h -- C ABI of the MI355X radiosity engine (libradgpu.so).
 *
 * Drop-in for the form-factor and shooting path of morris821028/hw-radiosity
 * ("the reference").  The reference is one C++ program with no library API;
 * its seams are the `rad` command line (src/rad.cpp:525-661) and the
 * functions over its global TriStore (include/raycast.h:4-5,
 * include/shade.h:4-5, src/rad.cpp).  Each entry point below names the
 * reference code it replaces.  Plain pointers and sizes only; every call
 * returns a status, and the library never exits the process.
 *
 * Threading: one context per host thread.  A context owns one HIP stream
 * and its device buffers; host arrays belong to the caller.
 */
#ifndef RADIOSITY_GPU_H
#define RADIOSITY_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rad_ctx rad_ctx;

typedef enum {
	RAD_OK = 0,
	RAD_EINVAL = 1,   /* bad option or argument */
	RAD_EIO = 2,      /* file open/read/write failure */
	RAD_ESTATE = 3,   /* call out of order (e.g. shoot before prepare) */
	RAD_EDEVICE = 4   /* HIP runtime error, or no GPU present */
} rad_status;

/* Counter indices for rad_counters(). */
enum {
	RAD_CNT_RAYS = 0,      /* rays traced on the device */
	RAD_CNT_REF_RAYS = 1,  /* rays the reference casts for the same work (CalFF calls reaching RayHitted) */
	RAD_CNT_CELLS = 2,     /* octree leaves visited */
	RAD_CNT_TESTS = 3,     /* ray/triangle tests */
	RAD_CNT_LEVELS = 4,    /* octree levels the reference descends for those leaves */
	RAD_CNT_PRECISION = 5, /* "Precision Exceeded" steps (raycast.cpp:344-350) */
	RAD_CNT_STALLED = 6,   /* rays the reference never finishes; their hit is -2 */
	RAD_CNT_RECEIVERS = 7, /* receiver updates (Shade calls, shade.cpp:435) */
	RAD_CNT_N = 8
};

/* Options: the reference's (rad.cpp:527-528, defaults rad.cpp:17-31):
 *   -debug -adapt_area -sample_area -converge -delta_ff -write_cycle -zip
 *   -triangle -light -o -interactive -jobs -parallel_src
 * plus engine switches:
 *   -fma 0|1          reproduce the reference compiler's FMA contractions (default 1)
 *   -select high|low  tie-break on equal unshot energy: highest index (the
 *                     shipped build, default) or lowest (MaxDeltaRad, rad.cpp:359-386)
 *   -max_tri N        the reference pool size MaxTri (rad.h:10), default 1000000
 *   -old_format 0|1   per-vertex output of the OLD_FORMAT build (rad.cpp:150-159)
 *   -batch N          shooting iterations queued between host checks (default 32)
 *   -kernel_timing 0|1  time every shoot kernel with HIP events (rad_timings)
 *   -graph 0|1        1: replay shooting iterations from a captured HIP
 *                     graph; 0 (default): launch every kernel from the host,
 *                     as -kernel_timing 1 does; results are the same
 *   -quiet 0|1        no log lines
 *   -neighbors device|host  where InitRadiosity's neighbour search runs (default device)
 *   -octree device|host     where BuildTree runs (default device)
 *   -variant v0|v3    v0: the shipped build (default); v3: the reference's
 *                     V3 build, Delta-ff adaptive splitting live
 *                     (shade.cpp:463-525 with PARALLEL, config.h), in the
 *                     receiver order of its -jobs 1 run; one source per
 *                     iteration (-parallel_src 1)
 * Replaces ProcessOption (rad.cpp:525-631) for everything but the input name. */
rad_status rad_create(rad_ctx **out, int argc, const char *const *argv);
void rad_destroy(rad_ctx *ctx);
const char *rad_last_error(const rad_ctx *ctx);

/* Input.  ReadTriangle + BuildTree (InitModelFromFile, rad.cpp:244-252,
 * rad.cpp:47-111, raycast.cpp:216-233).  rad_set_triangles takes n
 * triangles as p9 (x,y,z of 3 vertices), Frgb and Brgb triples.  The
 * octree is built on the device (needs a GPU: RAD_EDEVICE without one)
 * unless the context was created with -octree host; the device starts up
 * while the input is read. */
rad_status rad_load_tri(rad_ctx *ctx, const char *path);
rad_status rad_set_triangles(rad_ctx *ctx, int n, const float *p9, const int *frgb3, const int *brgb3);

/* InitRadiosity + PrePartitionTriangles (rad.cpp:259-356,
 * shade.cpp:556-598).  InitRadiosity's neighbour search (rad.cpp:295-334)
 * runs on the device (needs a GPU: RAD_EDEVICE without one) unless the
 * context was created with -neighbors host; the rest runs on the host. */
rad_status rad_prepare_host(rad_ctx *ctx);
/* Copies the prepared scene and octree to the device and selects the first
 * source.  Needs a GPU. */
rad_status rad_upload(rad_ctx *ctx);

/* Up to max_shots iterations of the shooting loop on the device
 * (parallelRadiosity, rad.cpp:406-451, iterated by DoRadiosity
 * rad.cpp:499-514).  *shots = iterations done, *converged = 1 when no
 * candidate is left. */
rad_status rad_shoot(rad_ctx *ctx, int max_shots, int *shots, int *converged);

/* Whole solve as DoRadiosity (rad.cpp:491-523): prepare, upload, shoot,
 * -write_cycle snapshots <prefix>.<loop>, final <prefix>.fin, log lines.
 * *iterations = the logged "Complete in #Iterations" value (shots + 1). */
rad_status rad_run(rad_ctx *ctx, const char *prefix, int *iterations);

/* PrintOut (rad.cpp:117-242) of the current state to `path`. */
rad_status rad_write_fin(rad_ctx *ctx, const char *path);

/* The WebGL viewer's model file from a .fin, byte-identical to the
 * reference converter's (`tri2json -i fin -o json`,
 * WebGL_demo/model-converter/main.cpp:81-246): vertices moved into a cube
 * of side 10 centred on the origin, normals, colours / 255, indices.  Needs
 * no scene; ctx only carries the error text and -quiet (otherwise the
 * converter's "#Triangle n" and success lines go to stderr).  *triangles
 * (may be NULL) = records converted.  RAD_EIO for an unreadable .fin (the
 * reference aborts) or an unwritable output. */
rad_status rad_fin_to_json(rad_ctx *ctx, const char *fin_path, const char *json_path, int *triangles);

/* State in the reference's 236-byte Triangle layout (rad.h:25-39), slots
 * [0, *trinum); records may be NULL to query *trinum.  g6 (may be NULL)
 * receives G0 and G1 (rad.cpp:105-110). */
rad_status rad_records(rad_ctx *ctx, int *trinum, void *records, float *g6);

/* Octree facts: out3 = {#nodes, #list entries, GridNum}, *min_grid_len =
 * MinGridLen (raycast.cpp:175-232). */
rad_status rad_tree_info(rad_ctx *ctx, int *out3, float *min_grid_len);

/* The octree as the reference's TreeNodeStore / TriListStore hold it after
 * BuildTree (raycast.cpp:9-13, 168-233): out3[0] (rad_tree_info) records of
 * the 60-byte TreeNode layout (rad.h:41-45: g0, g1, sub[8], list) in
 * depth-first preorder, and out3[1] list entries (each leaf's ascending
 * originals followed by -1).  nodes or lists may be NULL. */
rad_status rad_tree_arrays(rad_ctx *ctx, void *nodes, int *lists);

/* RayHitted (raycast.cpp:335-360) for n rays, each {float p[3], float v[3],
 * int ban, int target} (32 bytes), against the uploaded scene.  hit is the
 * returned triangle, -1 for none, -2 where the reference never returns;
 * cells/tests (may be NULL) count leaves visited and triangle tests. */
rad_status rad_cast(rad_ctx *ctx, int n, const void *rays, int *hit, int *cells, int *tests);

/* rad_cast without allocating: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The rays and results are
 * staged through a static device array that each GPU holds until the
 * process ends, so at most RAD_CAST_STATIC_MAX rays per call (RAD_EINVAL
 * beyond), and calls that overlap on one GPU (contexts on other threads)
 * race: serialise them.  Results equal rad_cast's. */
#define RAD_CAST_STATIC_MAX (1 << 18)
rad_status rad_cast_static(rad_ctx *ctx, int n, const void *rays, int *hit, int *cells, int *tests);

/* The per-vertex form factors Shade computes for receiver slot pairs[2i+1]
 * from source slot pairs[2i] (AdaptCalFF(CalFF(...)), shade.cpp:442-450). */
rad_status rad_ff_pairs(rad_ctx *ctx, int n, const int *pairs, float *ff3);

/* Device counters since rad_upload (RAD_CNT_*). */
rad_status rad_counters(rad_ctx *ctx, long long *out);

/* Timings: out[0] host setup ms (load to prepare), out[1] upload ms (HIP
 * start-up excluded),
 * out[2] device ms of the shooting loop, out[3] shooting iterations
 * launched, out[4] device ms of the ray-tracing kernel alone and out[5] the
 * number of such kernels timed (both 0 unless -kernel_timing 1), out[6]
 * device ms of the last rad_cast kernel. */
rad_status rad_timings(rad_ctx *ctx, double *out7);

/* ---- Explicit form-factor matrix: gathering, shooting and adjoint --------
 *
 * The reference has no matrix form of its solver; these entry points solve
 * the radiosity system its progressive shooting converges to,
 *     B = E + diag(rho) F B      (one column per right-hand side),
 * with F[i][j] the form factor from the centroid of patch i to patch j as
 * CalFF computes it (area-weighted point-to-patch term, shade.cpp:107-127,
 * visibility by RayHitted, raycast.cpp:335-360), rows scaled to sum <= 1.
 * Right-hand-side arrays are n x nrhs, row-major (patch-major), in double.
 * A handle owns one HIP stream and the matrix (n rounded up to 256 squared,
 * FP32, plus FP16/BF16/transposed copies made on demand). */
typedef struct rad_lin rad_lin;

enum { RAD_PREC_F32 = 0, RAD_PREC_F16 = 1, RAD_PREC_BF16 = 2 };

/* info[8] written by the solvers */
enum {
	RAD_INFO_OUTER = 0,     /* refinement steps (gather, adjoint) or shots (shoot) */
	RAD_INFO_SWEEPS = 1,    /* low-precision sweeps */
	RAD_INFO_REL_RES = 2,   /* final FP64 relative residual */
	RAD_INFO_MS = 3,        /* solve time, ms */
	RAD_INFO_MS_TARGET = 4, /* time when the error against Bref first met target (-1: never / no Bref) */
	RAD_INFO_ERR = 5,       /* final relative L2 error against Bref (-1: no Bref) */
	RAD_INFO_POWER_E = 6,   /* sum_i area_i E_i over all columns (FP64) */
	RAD_INFO_POWER_B = 7,   /* sum_i area_i B_i (FP64) */
	RAD_INFO_N = 8
};

rad_status rad_lin_create(rad_lin **out, int n);
void rad_lin_destroy(rad_lin *lin);
const char *rad_lin_last_error(const rad_lin *lin);

/* F and patch areas from an uploaded scene (n must equal its slot count;
 * logical slots get zero rows and columns).  F[i][j] is CalFF's value from
 * the centroid of i to the whole patch j (shade.cpp:107-123) times a
 * visibility:
 *   RAD_VIS_SEGMENT    nothing crosses the open segment between the two
 *                      centroids (parameter in (1e-4, 1 - 1e-4), walked from
 *                      the lower slot index to the higher, so symmetric),
 *                      found by a grid walk over complete octree leaf lists;
 *                      needs a scene that pre-partition left unsplit
 *   RAD_VIS_REFERENCE  RayHitted from the centroid returns j
 *                      (raycast.cpp:335-360), blind spots included
 * *ms (may be NULL) = device time of the form-factor kernel.  Replaces the
 * per-pair CalFF of Shade (shade.cpp:435-554) with one n x n evaluation. */
enum { RAD_VIS_SEGMENT = 0, RAD_VIS_REFERENCE = 1 };
rad_status rad_lin_from_scene(rad_lin *lin, rad_ctx *scene, int visibility, double *ms);
rad_status rad_lin_set_matrix(rad_lin *lin, const float *F);   /* n x n row-major */
rad_status rad_lin_get_matrix(rad_lin *lin, float *F);
rad_status rad_lin_set_areas(rad_lin *lin, const double *area); /* power sums and shooting's source order */

/* Picks the MFMA tile (rows per wave rt*16, split-K factor) for nrhs columns
 * by timing every candidate on this matrix; gather does it on first use. */
rad_status rad_lin_autotune(rad_lin *lin, int nrhs, int prec, int *rt, int *splitk);
/* Forces the tile for nrhs columns instead of the autotuner's choice:
 * rt in {1, 2, 4} (rt = 4 with more than 64 columns shares the right-hand
 * sides through LDS), splitk in {1, 2, 4, 8},
 * subject to the matrix size. */
rad_status rad_lin_set_tile(rad_lin *lin, int nrhs, int prec, int rt, int splitk);
/* Average device time of one low-precision product F X, reps launches. */
rad_status rad_lin_gemm_ms(rad_lin *lin, int nrhs, int prec, int reps, double *ms);
/* Y = F X: RAD_PREC_F32 in FP64 arithmetic on the FP32 matrix; FP16/BF16
 * through the MFMA path (X rounded to that precision, FP32 accumulation). */
rad_status rad_lin_multiply(rad_lin *lin, int prec, int nrhs, const double *X, double *Y);

/* Mixed-precision gathering: Jacobi sweeps d <- r + rho o (F d) with F in
 * FP16 or BF16 on MFMA (FP32 accumulation), inside iterative refinement whose
 * residual E + rho o (F x) - x and energy sums are FP64 over the FP32 F.
 * Stops when ||r|| / ||E|| <= tol or after max_outer refinements.  Bref (may
 * be NULL) is a reference solution; info reports the error against it and
 * the time at which it first fell to target.  nrhs 1..192 in one GEMM. */
rad_status rad_lin_gather(rad_lin *lin, int prec, int nrhs, const double *rho, const double *E, double tol,
                          int max_outer, int sweeps, const double *Bref, double target, double *B, double *info);

/* Plain FP32 progressive shooting on the same system, the reference's method
 * (source = largest unshot power area_j * sum_c dB_j, ties to the lowest
 * index as MaxDeltaRad, rad.cpp:359-386; update rho_i F_ij dB_j as
 * CalRadiosity, shade.cpp:244-259), one device launch per shot.  Stops when
 * no unshot power is left, after max_shots, or when the error against Bref
 * (if given) reaches target.  nrhs 1..16. */
rad_status rad_lin_shoot(rad_lin *lin, int nrhs, const double *rho, const double *E, const double *Bref,
                         double target, int max_shots, double *B, double *info);

/* Adjoint of the solve for a loss L(B): solves (I - diag(rho) F)^T lambda =
 * dL/dB (FP16 MFMA sweeps on a transposed copy of F made on the first call,
 * FP64 residual refinement over the FP32 F, no stored iterates) and
 * returns dL/dE = lambda and dL/drho = lambda o (F B), visibility fixed.
 * nrhs 1..4. */
rad_status rad_lin_adjoint(rad_lin *lin, int nrhs, const double *rho, const double *B, const double *dLdB, double tol,
                           int max_outer, int sweeps, double *dLdE, double *dLdrho, double *info);

/* ---- Many scenes of uneven size in one launch ------------------------------
 *
 * count independent systems B_s = E_s + diag(rho_s) F_s B_s with three
 * columns (RGB) each, solved by the mixed-precision gather above with every
 * step of every scene in one kernel (tile lists spanning the scenes, grids
 * sized to the machine, converged scenes dropping out), the whole schedule
 * captured once into a HIP graph and replayed with one launch.  No reference
 * counterpart (the reference solves one scene per process). */
typedef struct rad_batch rad_batch;

rad_status rad_batch_create(rad_batch **out, int count, const int *n);
void rad_batch_destroy(rad_batch *batch);
const char *rad_batch_last_error(const rad_batch *batch);
/* Scene `scene`'s F from an uploaded context with n[scene] slots (see
 * rad_lin_from_scene), or from a host n x n row-major matrix. */
rad_status rad_batch_from_scene(rad_batch *batch, int scene, rad_ctx *ctx, int visibility, double *ms);
rad_status rad_batch_set_matrix(rad_batch *batch, int scene, const float *F);
/* Picks the schedule of the batched sweep product -- rows per wave rt*16
 * (rt in {1, 2, 4}), k per tile kc (1024..8192) and the persistent grid
 * (512..4096 blocks) -- by timing one sweep of every candidate on this
 * batch; solve does it on first use for each precision. */
rad_status rad_batch_autotune(rad_batch *batch, int prec, int *rt, int *kc, int *grid);
/* Forces that schedule instead (rt in {1, 2, 4}, kc in {1024, 2048, 4096,
 * 8192}, grid 64..16384); a later rad_batch_autotune hands the choice back. */
rad_status rad_batch_set_tile(rad_batch *batch, int rt, int kc, int grid);
/* rho, E, B: (sum of n) x 3 row-major, scenes concatenated in order.
 * `outer` refinements of `sweeps` sweeps each, a scene stopping once its
 * FP64 relative residual is <= tol.  use_graph = 1: one graph launch;
 * 0: the same kernels scene after scene (bit-identical results).
 * rel_res[count] = final relative residuals; info4 = {device ms, graph
 * build ms (0 when reused), kernel launches, scenes still above tol}. */
rad_status rad_batch_solve(rad_batch *batch, int prec, int use_graph, const double *rho, const double *E, double tol,
                           int outer, int sweeps, double *B, double *rel_res, double *info4);

#ifdef __cplusplus
}
#endif
#endif
