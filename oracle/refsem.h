/* refsem -- CPU restatement of the reference's radiosity path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle: tests/, the smoke check and
 * the benchmark's CPU-baseline leg load it as the checker.  The engine in
 * hw-radiosity_amd/ never links, loads or calls it.
 *
 * The record layout, the octree, the DDA visibility walk, the form factor,
 * pre-partition, the shooting loop and the .fin writer follow the reference
 * (morris821028/hw-radiosity, "V0" = the shipped include/config.h) so that
 * state dumps compare byte-for-byte with the reference's own TriStore.
 */
#ifndef REFSEM_H
#define REFSEM_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rs_ctx rs_ctx;

/* Counters, indices into rs_counters()'s output. */
enum {
	RS_CNT_RAYS = 0,        /* RayHitted calls (reference-equivalent rays) */
	RS_CNT_CELLS = 1,       /* leaf cells visited */
	RS_CNT_TESTS = 2,       /* TriHitted calls */
	RS_CNT_LEVELS = 3,      /* TreeNodeNum descent levels */
	RS_CNT_PRECISION = 4,   /* "Precision Exceeded" events (raycast.cpp:344-350) */
	RS_CNT_SHADES = 5,      /* Shade calls */
	RS_CNT_ITERS = 6,       /* shots done */
	RS_CNT_CAPPED = 7,      /* rays the reference never finishes (hit = -2) */
	RS_CNT_N = 8
};

/* argv: the reference's rad options (rad.cpp:527-528) without the input file,
 * plus restatement switches:
 *   -fma 0|1          GCC contraction sites on (REF-FMA, default) or off
 *   -select high|low  V0 tie-break (highest index, default) or lowest (V1)
 *   -neighbors exact|hash   O(N^2) scan in loop order, or hashed equivalent (default)
 *   -skip_discarded 0|1     skip the full-source ray that AdaptCalFF discards
 *   -max_tri N        the reference's pool size MaxTri (rad.h:10), default 1000000
 *   -old_format 0|1   per-vertex OLD_FORMAT output (rad.cpp:150-159)
 *   -quiet 0|1        suppress log lines
 * Returns NULL on a parse error. */
rs_ctx *rs_new(int argc, const char **argv);
void rs_free(rs_ctx *);
const char *rs_error(rs_ctx *);

int rs_load_file(rs_ctx *, const char *path);                 /* ReadTriangle */
int rs_load_arrays(rs_ctx *, int n, const float *p9, const int *frgb3, const int *brgb3);
int rs_build_tree(rs_ctx *);                                  /* BuildTree */
int rs_init_radiosity(rs_ctx *);                              /* InitRadiosity */
int rs_prepartition(rs_ctx *);                                /* PrePartitionTriangles */
/* Run up to max_shots shooting iterations; returns shots done, and sets
 * *converged to 1 when an iteration found no candidate. */
int rs_iterate(rs_ctx *, int max_shots, int *converged);
/* Whole DoRadiosity: init, pre-partition, loop with -write_cycle snapshots,
 * final <prefix>.fin.  Returns the logged iteration count (shots + 1). */
int rs_run(rs_ctx *, const char *prefix);
int rs_write_fin(rs_ctx *, const char *path);

int rs_trinum(rs_ctx *);
int rs_record_size(void);
/* Copies trinum raw 236-byte records, and G0/G1 (6 floats) if g != NULL. */
int rs_dump(rs_ctx *, void *records, float *g);
/* Octree facts: out[0]=#nodes, out[1]=#list entries, out[2]=GridNum, and
 * MinGridLen through *min_grid_len. */
int rs_tree_info(rs_ctx *, int *out3, float *min_grid_len);
/* Leaf lookup for the engine's flat table check: leaf node index for
 * integer cell coordinates (gx, gy, gz), by the reference descent. */
int rs_rand(int n, int *out);  /* glibc rand()'s first n outputs (default seed) */
int rs_descend(rs_ctx *, int gx, int gy, int gz);
/* Octree arrays: nodes as {g0[3], g1[3], sub[8], list} (15 x 4 bytes). */
int rs_tree_arrays(rs_ctx *, void *nodes, int *lists);

/* rays: n records {float p[3], float v[3], int ban, int target}; outputs may be NULL. */
int rs_raycast(rs_ctx *, int n, const void *rays, int *hit, int *cells, int *tests);
/* V0 Shade on a zeroed copy of dst with src.deltaB = (1000, 2000, 3000) and
 * dst.Frgb = 255: writes the copy's deltaaccB (9 floats per pair). */
int rs_shade_pairs(rs_ctx *, int n, const int *pairs, float *out9);
/* Per-vertex form factors ff[v] for (src, dst) pairs, as Shade computes them. */
int rs_ff_pairs(rs_ctx *, int n, const int *pairs, float *out3);
/* Explicit gathering matrix: F[i][j] = CalFF's value from the centroid of i
 * to the whole patch j, times its visibility:
 *   RS_VIS_SEGMENT    no other original crosses the open segment between the
 *                     centroids, t in (1e-4, 1 - 1e-4), the segment taken from
 *                     the lower slot index to the higher (grid walk with
 *                     geometric point location; not a reference function)
 *   RS_VIS_REFERENCE  RayHitted from the centroid returns j's logical index
 *   RS_VIS_NONE       always visible (checks)
 *   RS_VIS_BRUTE      RS_VIS_SEGMENT's test against every original (checks)
 * for (i, j) pairs (unscaled)... */
enum { RS_VIS_SEGMENT = 0, RS_VIS_REFERENCE = 1, RS_VIS_NONE = 2, RS_VIS_BRUTE = 3 };
int rs_ff_centroid(rs_ctx *, int vis, int n, const int *pairs, float *out);
/* ...and whole rows (n = trinum entries each), scaled to sum <= 1. */
int rs_ff_rows(rs_ctx *, int vis, int nrows, const int *rows, float *out);
/* Cost sample of InitRadiosity's O(N^2) neighbour pass (rad.cpp:295-334):
 * the reference's comparisons for rows i < rows, OpenMP over j as the
 * reference runs it, nothing written; returns the edge matches. */
long long rs_neighbor_rows(rs_ctx *, int rows);
void rs_counters(rs_ctx *, long long *out);
void rs_reset_counters(rs_ctx *);

#ifdef __cplusplus
}
#endif
#endif
