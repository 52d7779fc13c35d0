// C ABI of the engine (include/radiosity_gpu.h).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/radiosity_gpu.h"
#include "device.h"
#include "fin.h"
#include "viewer_json.h"
#include "batch.h"
#include "linear.h"
#include "octree.h"
#include "scene.h"

using namespace rgpu;

struct rad_ctx {
	HostScene scene;
	FlatOctree tree;
	std::unique_ptr<DeviceEngine> dev;
	FinWriter writer;
	std::string err;
	bool loaded = false, prepared = false, uploaded = false, device_newer = false;
	double ms_setup = 0, ms_upload = 0;
	// device start-up on its own thread while the host prepares (rad_run):
	// the engine it made, or the error; read only after the join
	std::thread starter;
	DeviceEngine *started = nullptr;
	std::string start_err;
	~rad_ctx()
	{
		if (starter.joinable()) starter.join();
		delete started;
	}
};

struct rad_lin {
	std::unique_ptr<rgpu::LinearSystem> sys;
	std::string err;
};

struct rad_batch {
	std::unique_ptr<rgpu::SceneBatch> b;
	std::string err;
};

namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t0)
{
	return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

rad_status fail(rad_ctx *c, rad_status s, const std::string &msg)
{
	c->err = msg;
	return s;
}

rad_status lfail(rad_lin *l, rad_status s, const std::string &msg)
{
	l->err = msg;
	return s;
}

// The device's start-up (HIP runtime, the context's stream, loading the
// device code) on a thread of its own, on the calling thread's current
// device, while the host works; a no-op when the context has a device or a
// start-up under way, or when there is no device.
void start_device(rad_ctx *c)
{
	int device = -1;
	if (c->dev || c->starter.joinable() || (device = DeviceEngine::current_device()) < 0) return;
	c->start_err.clear();
	c->starter = std::thread([c, device] {
		c->started = DeviceEngine::create_on(device, c->start_err);
		if (c->started && c->started->warm(c->start_err)) {
			delete c->started;
			c->started = nullptr;
		}
	});
}

// Waits for a start-up under way; its failure is reported by the first call
// that needs the device.
void join_device(rad_ctx *c)
{
	if (!c->starter.joinable()) return;
	c->starter.join();
	if (c->started) {
		c->dev.reset(c->started);
		c->started = nullptr;
	}
}

// A device for the context: the start-up under way, or a new one.
rad_status need_device(rad_ctx *c)
{
	if (c->dev) return RAD_OK;
	join_device(c);
	if (c->dev) return RAD_OK;
	DeviceEngine *d = DeviceEngine::create(c->err);
	if (!d) return RAD_EDEVICE;
	c->dev.reset(d);
	return RAD_OK;
}

rad_status after_load(rad_ctx *c)
{
	HostScene &s = c->scene;
	c->loaded = c->prepared = c->uploaded = c->device_newer = false;
	for (int k = 0; k < 3; k++)
		if (!std::isfinite(s.G0[k]) || !std::isfinite(s.G1[k]))  // the reference's grid is undefined then
			return fail(c, RAD_EINVAL, "scene coordinates are not finite: the octree grid is undefined");
	if (!s.o.octree_host && need_device(c) != RAD_OK) return RAD_EDEVICE;
	const auto t0 = Clock::now();
	if (s.o.octree_host)
		c->tree = build_octree(s.pool.data(), s.n_orig, s.G0, s.G1, s.o.fma != 0, s.o.jobs > 0 ? std::max(s.o.jobs, 8) : 0);
	else if (c->dev->build_octree(s, c->tree, c->err))
		return RAD_EDEVICE;
	if (c->tree.gt <= 0) return fail(c, RAD_EINVAL, "octree grid too large to tabulate");
	s.log("[INFO] 3D-DDA #TreeNode = %d, #Trilist = %d\n", c->tree.n_nodes, c->tree.n_list);
	c->ms_setup += ms_since(t0);
	c->loaded = true;
	c->prepared = c->uploaded = c->device_newer = false;
	return RAD_OK;
}

rad_status sync_host(rad_ctx *c)
{
	if (c->device_newer) {
		if (c->dev->download(c->scene, c->err)) return RAD_EDEVICE;
		c->device_newer = false;
	}
	return RAD_OK;
}

rad_status write_snapshot(rad_ctx *c, const std::string &path, bool final_file)
{
	rad_status st = sync_host(c);
	if (st != RAD_OK) return st;
	const HostScene &s = c->scene;
	std::string werr;
	if (c->writer.write(path, s.pool.data(), s.trinum, s.o.old_format != 0, werr) != 0) {
		std::fprintf(stderr, "%s\n", werr.c_str());
		if (!final_file) return RAD_OK;  // the reference skips an unwritable snapshot
		// and writes the final result to stdout instead (rad.cpp:131-138)
		if (c->writer.write("/dev/stdout", s.pool.data(), s.trinum, s.o.old_format != 0, werr) != 0)
			return fail(c, RAD_EIO, werr);
		return RAD_OK;
	}
	if (s.o.zip) {
		std::printf("Compress output file %s .\n", path.c_str());
		const std::string cmd = "gzip " + path;
		// rad.cpp:229-232 prints OK when system() returns non-zero
		std::puts(std::system(cmd.c_str()) ? "OK" : "Failed");
	}
	if (s.o.interactive) {
		// rad.cpp:234-240: hand the file to the WebGL demo's script (which
		// converts it with tri2json, main.cpp:248-290) and report its status
		const std::string cmd = "cd ./WebGL_demo && ./run.sh ../" + path + " 2>/dev/null";
		const int ret = std::system(cmd.c_str());
		std::fprintf(stderr, "WebGL: %s, return %d\n", cmd.c_str(), ret);
	}
	return RAD_OK;
}

}  // namespace

extern "C" {

rad_status rad_create(rad_ctx **out, int argc, const char *const *argv)
{
	if (!out) return RAD_EINVAL;
	*out = nullptr;
	std::unique_ptr<rad_ctx> c(new rad_ctx());
	for (int i = 0; i < argc; i++) {
		std::string err;
		if (!parse_option(c->scene.o, argc, argv, i, err)) {
			std::fprintf(stderr, "rad_create: %s\n", err.c_str());
			return RAD_EINVAL;
		}
	}
	*out = c.release();
	return RAD_OK;
}

void rad_destroy(rad_ctx *c) { delete c; }

const char *rad_last_error(const rad_ctx *c) { return c ? c->err.c_str() : "null context"; }

rad_status rad_load_tri(rad_ctx *c, const char *path)
{
	if (!c || !path) return RAD_EINVAL;
	// the device starts while the file is read and the octree built: its
	// start-up belongs with the reference's "Init" (InitModelFromFile,
	// rad.cpp:244-252), not with the solve
	start_device(c);
	const auto t0 = Clock::now();
	c->ms_setup = 0;
	if (c->scene.load_tri(path, c->err)) {
		join_device(c);
		return RAD_EIO;
	}
	c->ms_setup += ms_since(t0);
	const rad_status st = after_load(c);  // joins the start-up for the device octree
	join_device(c);
	return st;
}

rad_status rad_set_triangles(rad_ctx *c, int n, const float *p9, const int *frgb3, const int *brgb3)
{
	if (!c || n < 0 || (n > 0 && (!p9 || !frgb3 || !brgb3))) return RAD_EINVAL;
	start_device(c);  // as rad_load_tri
	const auto t0 = Clock::now();
	c->ms_setup = 0;
	if (c->scene.set_triangles(n, p9, frgb3, brgb3, c->err)) {
		join_device(c);
		return RAD_EINVAL;
	}
	c->ms_setup += ms_since(t0);
	const rad_status st = after_load(c);
	join_device(c);
	return st;
}

rad_status rad_prepare_host(rad_ctx *c)
{
	if (!c) return RAD_EINVAL;
	if (!c->loaded) return fail(c, RAD_ESTATE, "no scene loaded");
	if (c->prepared) return RAD_OK;
	if (c->scene.o.write_cycle <= 0) return fail(c, RAD_EINVAL, "-write_cycle must be positive");
	const bool on_device = c->scene.o.neighbors_host == 0;
	if (on_device && need_device(c) != RAD_OK) return RAD_EDEVICE;  // the neighbour search needs it
	const auto t0 = Clock::now();
	c->scene.init_radiosity(!on_device);
	if (on_device && c->dev->link_neighbors(c->scene, c->err)) return RAD_EDEVICE;
	c->scene.prepartition();
	c->ms_setup += ms_since(t0);
	c->prepared = true;
	return RAD_OK;
}

rad_status rad_upload(rad_ctx *c)
{
	if (!c) return RAD_EINVAL;
	if (!c->prepared) return fail(c, RAD_ESTATE, "rad_prepare_host first");
	if (!c->dev && c->starter.joinable()) {  // started during the host set-up
		join_device(c);
		if (!c->dev) return fail(c, RAD_EDEVICE, c->start_err);
	}
	if (!c->dev) {  // HIP start-up and the context's stream: not part of the upload time
		DeviceEngine *d = DeviceEngine::create(c->err);
		if (!d) return RAD_EDEVICE;
		c->dev.reset(d);
	}
	const auto t0 = Clock::now();
	c->dev->timing = c->scene.o.kernel_timing != 0;
	c->dev->graphs = c->scene.o.graph != 0;
	if (c->dev->upload(c->scene, c->tree, c->err)) return RAD_EDEVICE;
	c->ms_upload = ms_since(t0);
	c->uploaded = true;
	c->device_newer = false;
	return RAD_OK;
}

rad_status rad_shoot(rad_ctx *c, int max_shots, int *shots, int *converged)
{
	if (!c || !shots || !converged || max_shots < 0) return RAD_EINVAL;
	if (!c->uploaded) return fail(c, RAD_ESTATE, "rad_upload first");
	if (c->dev->shoot(max_shots, shots, converged, c->err)) return RAD_EDEVICE;
	if (*shots > 0) c->device_newer = true;
	return RAD_OK;
}

rad_status rad_run(rad_ctx *c, const char *prefix, int *iterations)
{
	if (!c || !prefix) return RAD_EINVAL;
	rad_status st;
	if (!c->uploaded && !c->prepared) start_device(c);  // overlaps InitRadiosity and pre-partition
	if ((st = rad_prepare_host(c)) != RAD_OK) return st;
	if (!c->uploaded && (st = rad_upload(c)) != RAD_OK) return st;
	const HostScene &s = c->scene;
	const int W = s.o.write_cycle;
	int done = 0;  // shots so far
	for (;;) {
		const int next = (done / W + 1) * W;  // loop number of the next snapshot
		int got = 0, conv = 0;
		if ((st = rad_shoot(c, next - done, &got, &conv)) != RAD_OK) return st;
		done += got;
		// the reference announces loop L before trying shot L
		if (done + 1 >= next) s.log("[INFO] Loop %d BEGIN\n", next);
		if (conv) break;
		if ((st = write_snapshot(c, std::string(prefix) + "." + std::to_string(next), false)) != RAD_OK) return st;
		s.log("[INFO] END Loop. There are %i triangles.\n", s.trinum);
	}
	s.log("[INFO] Complete in #Iterations %d\n", done + 1);
	const auto t0 = Clock::now();
	if ((st = write_snapshot(c, std::string(prefix) + ".fin", true)) != RAD_OK) return st;
	s.log("[INFO] Store file took %f sec. time.\n", ms_since(t0) / 1000.0);
	if (iterations) *iterations = done + 1;
	return RAD_OK;
}

rad_status rad_write_fin(rad_ctx *c, const char *path)
{
	if (!c || !path) return RAD_EINVAL;
	if (!c->loaded) return fail(c, RAD_ESTATE, "no scene loaded");
	rad_status st = sync_host(c);
	if (st != RAD_OK) return st;
	if (c->writer.write(path, c->scene.pool.data(), c->scene.trinum, c->scene.o.old_format != 0, c->err)) return RAD_EIO;
	return RAD_OK;
}

rad_status rad_fin_to_json(rad_ctx *c, const char *fin_path, const char *json_path, int *triangles)
{
	if (!c || !fin_path || !json_path) return RAD_EINVAL;
	int n = 0, tiny = 0;
	if (fin_to_json(fin_path, json_path, n, tiny, c->err)) return RAD_EIO;
	if (!c->scene.o.quiet) {
		for (int k = 0; k < tiny; k++) std::fprintf(stderr, "[Warning] Triangle area is too small to show\n");
		std::fprintf(stderr, "#Triangle %d\nModel-Converter Success\n", n);
	}
	if (triangles) *triangles = n;
	return RAD_OK;
}

rad_status rad_records(rad_ctx *c, int *trinum, void *records, float *g6)
{
	if (!c || !trinum) return RAD_EINVAL;
	if (!c->loaded) return fail(c, RAD_ESTATE, "no scene loaded");
	rad_status st = sync_host(c);
	if (st != RAD_OK) return st;
	*trinum = c->scene.trinum;
	if (records) std::memcpy(records, c->scene.pool.data(), sizeof(Patch) * (size_t)c->scene.trinum);
	if (g6) {
		std::memcpy(g6, c->scene.G0, 12);
		std::memcpy(g6 + 3, c->scene.G1, 12);
	}
	return RAD_OK;
}

rad_status rad_tree_info(rad_ctx *c, int *out3, float *mgl)
{
	if (!c || !out3 || !mgl) return RAD_EINVAL;
	if (!c->loaded) return fail(c, RAD_ESTATE, "no scene loaded");
	out3[0] = c->tree.n_nodes;
	out3[1] = c->tree.n_list;
	out3[2] = c->tree.grid_num;
	*mgl = c->tree.min_grid;
	return RAD_OK;
}

rad_status rad_tree_arrays(rad_ctx *c, void *nodes, int *lists)
{
	if (!c) return RAD_EINVAL;
	if (!c->loaded) return fail(c, RAD_ESTATE, "no scene loaded");
	if (nodes) std::memcpy(nodes, c->tree.ref_nodes.data(), c->tree.ref_nodes.size() * sizeof(RefNode));
	if (lists) std::memcpy(lists, c->tree.ref_lists.data(), c->tree.ref_lists.size() * sizeof(int));
	return RAD_OK;
}

rad_status rad_cast(rad_ctx *c, int n, const void *rays, int *hit, int *cells, int *tests)
{
	if (!c || n < 0 || (n > 0 && (!rays || !hit))) return RAD_EINVAL;
	if (!c->uploaded) return fail(c, RAD_ESTATE, "rad_upload first");
	if (c->dev->cast(n, (const RayIn *)rays, hit, cells, tests, c->err)) return RAD_EDEVICE;
	return RAD_OK;
}

rad_status rad_cast_static(rad_ctx *c, int n, const void *rays, int *hit, int *cells, int *tests)
{
	if (!c || n < 0 || (n > 0 && (!rays || !hit))) return RAD_EINVAL;
	if (!c->uploaded) return fail(c, RAD_ESTATE, "rad_upload first");
	if (n > RAD_CAST_STATIC_MAX) {
		c->err = "rad_cast_static: " + std::to_string(n) + " rays exceed its limit of " +
		         std::to_string(RAD_CAST_STATIC_MAX) + " per call";
		return RAD_EINVAL;
	}
	if (c->dev->cast_static(n, (const RayIn *)rays, hit, cells, tests, c->err)) return RAD_EDEVICE;
	return RAD_OK;
}

rad_status rad_ff_pairs(rad_ctx *c, int n, const int *pairs, float *ff3)
{
	if (!c || n < 0 || (n > 0 && (!pairs || !ff3))) return RAD_EINVAL;
	if (!c->uploaded) return fail(c, RAD_ESTATE, "rad_upload first");
	if (c->dev->ff_pairs(n, pairs, ff3, c->err)) return RAD_EDEVICE;
	return RAD_OK;
}

rad_status rad_counters(rad_ctx *c, long long *out)
{
	if (!c || !out) return RAD_EINVAL;
	for (int k = 0; k < RAD_CNT_N; k++) out[k] = 0;
	if (c->dev) c->dev->counters(out);
	return RAD_OK;
}

rad_status rad_timings(rad_ctx *c, double *out7)
{
	if (!c || !out7) return RAD_EINVAL;
	out7[6] = c->dev ? c->dev->cast_ms() : 0;
	out7[0] = c->ms_setup;
	out7[1] = c->ms_upload;
	out7[2] = c->dev ? c->dev->shoot_ms() : 0;
	out7[3] = c->dev ? (double)c->dev->kernel_launches() : 0;
	out7[4] = c->dev ? c->dev->kernel_ms() : 0;
	out7[5] = c->dev ? (double)c->dev->kernel_timed() : 0;
	return RAD_OK;
}

/* ---- explicit form-factor matrix ------------------------------------------ */

rad_status rad_lin_create(rad_lin **out, int n)
{
	if (!out) return RAD_EINVAL;
	*out = nullptr;
	std::unique_ptr<rad_lin> l(new rad_lin());
	if (n <= 0) return RAD_EINVAL;
	std::string err;
	LinearSystem *L = LinearSystem::create(n, err);
	if (!L) {
		std::fprintf(stderr, "rad_lin_create: %s\n", err.c_str());
		return RAD_EDEVICE;
	}
	l->sys.reset(L);
	*out = l.release();
	return RAD_OK;
}

void rad_lin_destroy(rad_lin *l) { delete l; }

const char *rad_lin_last_error(const rad_lin *l) { return l ? l->err.c_str() : "null handle"; }

rad_status rad_lin_from_scene(rad_lin *l, rad_ctx *c, int visibility, double *ms)
{
	if (!l || !c) return RAD_EINVAL;
	if (visibility != RAD_VIS_SEGMENT && visibility != RAD_VIS_REFERENCE)
		return lfail(l, RAD_EINVAL, "unknown visibility mode");
	if (!c->uploaded) return lfail(l, RAD_ESTATE, "rad_upload the scene first");
	const HostScene &s = c->scene;
	if (s.trinum != l->sys->n())
		return lfail(l, RAD_EINVAL, "matrix size " + std::to_string(l->sys->n()) + " != scene slots " +
		                                std::to_string(s.trinum));
	if (c->dev->form_factors(l->sys->matrix32(), s.trinum, l->sys->ld(), visibility, c->tree, s, l->err))
		return RAD_EDEVICE;
	if (ms) *ms = c->dev->cast_ms();
	std::vector<double> area(s.trinum);
	for (int i = 0; i < s.trinum; i++) area[i] = s.pool[i].area;
	if (l->sys->set_areas(area.data(), l->err)) return RAD_EDEVICE;
	return RAD_OK;
}

rad_status rad_lin_set_matrix(rad_lin *l, const float *F)
{
	if (!l || !F) return RAD_EINVAL;
	return l->sys->set_matrix(F, l->err) ? RAD_EDEVICE : RAD_OK;
}

rad_status rad_lin_get_matrix(rad_lin *l, float *F)
{
	if (!l || !F) return RAD_EINVAL;
	return l->sys->get_matrix(F, l->err) ? RAD_EDEVICE : RAD_OK;
}

rad_status rad_lin_set_areas(rad_lin *l, const double *area)
{
	if (!l || !area) return RAD_EINVAL;
	return l->sys->set_areas(area, l->err) ? RAD_EDEVICE : RAD_OK;
}

rad_status rad_lin_autotune(rad_lin *l, int nrhs, int prec, int *rt, int *splitk)
{
	if (!l || nrhs <= 0 || nrhs > 192 || (prec != RAD_PREC_F16 && prec != RAD_PREC_BF16))
		return l ? lfail(l, RAD_EINVAL, "autotune: nrhs 1..192, precision FP16 or BF16") : RAD_EINVAL;
	if (l->sys->autotune(nrhs, prec, l->err)) return RAD_EDEVICE;
	int a, b;
	l->sys->tile_config(nrhs, prec, &a, &b);
	if (rt) *rt = a;
	if (splitk) *splitk = b;
	return RAD_OK;
}

rad_status rad_lin_set_tile(rad_lin *l, int nrhs, int prec, int rt, int splitk)
{
	if (!l) return RAD_EINVAL;
	return l->sys->set_tile(nrhs, prec, rt, splitk, l->err) ? RAD_EINVAL : RAD_OK;
}

rad_status rad_lin_gemm_ms(rad_lin *l, int nrhs, int prec, int reps, double *ms)
{
	if (!l || !ms || nrhs <= 0 || nrhs > 192 || reps <= 0 || (prec != RAD_PREC_F16 && prec != RAD_PREC_BF16))
		return l ? lfail(l, RAD_EINVAL, "gemm_ms: bad arguments") : RAD_EINVAL;
	return l->sys->gemm_ms(nrhs, prec, reps, ms, l->err) ? RAD_EDEVICE : RAD_OK;
}

rad_status rad_lin_multiply(rad_lin *l, int prec, int nrhs, const double *X, double *Y)
{
	if (!l || !X || !Y || nrhs <= 0 || nrhs > 192 || prec < RAD_PREC_F32 || prec > RAD_PREC_BF16)
		return l ? lfail(l, RAD_EINVAL, "multiply: bad arguments") : RAD_EINVAL;
	return l->sys->multiply(prec, nrhs, X, Y, l->err) ? RAD_EDEVICE : RAD_OK;
}

rad_status rad_lin_gather(rad_lin *l, int prec, int nrhs, const double *rho, const double *E, double tol,
                          int max_outer, int sweeps, const double *Bref, double target, double *B, double *info)
{
	if (!l || !rho || !E || !B || nrhs <= 0 || nrhs > 192 || sweeps <= 0 || max_outer < 0 ||
	    (prec != RAD_PREC_F16 && prec != RAD_PREC_BF16))
		return l ? lfail(l, RAD_EINVAL, "gather: bad arguments") : RAD_EINVAL;
	return l->sys->gather(prec, nrhs, rho, E, tol, max_outer, sweeps, Bref, target, B, info, l->err) ? RAD_EDEVICE
	                                                                                                    : RAD_OK;
}

rad_status rad_lin_shoot(rad_lin *l, int nrhs, const double *rho, const double *E, const double *Bref, double target,
                         int max_shots, double *B, double *info)
{
	if (!l || !rho || !E || !B || nrhs <= 0 || nrhs > 16 || max_shots < 0)
		return l ? lfail(l, RAD_EINVAL, "shoot: bad arguments") : RAD_EINVAL;
	return l->sys->shoot(nrhs, rho, E, Bref, target, max_shots, B, info, l->err) ? RAD_EDEVICE : RAD_OK;
}

rad_status rad_lin_adjoint(rad_lin *l, int nrhs, const double *rho, const double *B, const double *dLdB, double tol,
                           int max_outer, int sweeps, double *dLdE, double *dLdrho, double *info)
{
	if (!l || !rho || !B || !dLdB || !dLdE || !dLdrho || nrhs <= 0 || nrhs > 4 || sweeps <= 0 || max_outer < 0)
		return l ? lfail(l, RAD_EINVAL, "adjoint: bad arguments") : RAD_EINVAL;
	return l->sys->adjoint(nrhs, rho, B, dLdB, tol, max_outer, sweeps, dLdE, dLdrho, info, l->err) ? RAD_EDEVICE
	                                                                                                : RAD_OK;
}

/* ---- many scenes ----------------------------------------------------------- */

rad_status rad_batch_create(rad_batch **out, int count, const int *n)
{
	if (!out || count <= 0 || !n) return RAD_EINVAL;
	*out = nullptr;
	std::string err;
	SceneBatch *b = SceneBatch::create(count, n, err);
	if (!b) {
		std::fprintf(stderr, "rad_batch_create: %s\n", err.c_str());
		return err.find("device") != std::string::npos || err.find("hip") != std::string::npos ? RAD_EDEVICE
		                                                                                       : RAD_EINVAL;
	}
	std::unique_ptr<rad_batch> h(new rad_batch());
	h->b.reset(b);
	*out = h.release();
	return RAD_OK;
}

void rad_batch_destroy(rad_batch *b) { delete b; }

const char *rad_batch_last_error(const rad_batch *b) { return b ? b->err.c_str() : "null handle"; }

rad_status rad_batch_from_scene(rad_batch *b, int scene, rad_ctx *c, int visibility, double *ms)
{
	if (!b || !c) return RAD_EINVAL;
	if (scene < 0 || scene >= b->b->count()) {
		b->err = "scene index out of range";
		return RAD_EINVAL;
	}
	if (visibility != RAD_VIS_SEGMENT && visibility != RAD_VIS_REFERENCE) {
		b->err = "unknown visibility mode";
		return RAD_EINVAL;
	}
	if (!c->uploaded) {
		b->err = "rad_upload the scene first";
		return RAD_ESTATE;
	}
	const HostScene &s = c->scene;
	if (s.trinum != b->b->n(scene)) {
		b->err = "scene " + std::to_string(scene) + " has " + std::to_string(b->b->n(scene)) + " patches, the context " +
		         std::to_string(s.trinum);
		return RAD_EINVAL;
	}
	if (c->dev->form_factors(b->b->matrix32(scene), s.trinum, b->b->ld(scene), visibility, c->tree, s, b->err))
		return RAD_EDEVICE;
	if (ms) *ms = c->dev->cast_ms();
	return RAD_OK;
}

rad_status rad_batch_set_matrix(rad_batch *b, int scene, const float *F)
{
	if (!b || !F) return RAD_EINVAL;
	return b->b->set_matrix(scene, F, b->err) ? RAD_EINVAL : RAD_OK;
}

rad_status rad_batch_autotune(rad_batch *b, int prec, int *rt, int *kc, int *grid)
{
	if (!b) return RAD_EINVAL;
	if (prec != RAD_PREC_F16 && prec != RAD_PREC_BF16) {
		b->err = "batch autotune: precision FP16 or BF16";
		return RAD_EINVAL;
	}
	BatchTile t{0, 0, 0};
	if (b->b->autotune(prec, t, b->err)) return RAD_EDEVICE;
	if (rt) *rt = t.rt;
	if (kc) *kc = t.kc;
	if (grid) *grid = t.grid;
	return RAD_OK;
}

rad_status rad_batch_set_tile(rad_batch *b, int rt, int kc, int grid)
{
	if (!b) return RAD_EINVAL;
	return b->b->set_tile(BatchTile{rt, kc, grid}, b->err) ? RAD_EINVAL : RAD_OK;
}

rad_status rad_batch_solve(rad_batch *b, int prec, int use_graph, const double *rho, const double *E, double tol,
                           int outer, int sweeps, double *B, double *rel_res, double *info4)
{
	if (!b || !rho || !E || !B || !rel_res) return RAD_EINVAL;
	if ((prec != RAD_PREC_F16 && prec != RAD_PREC_BF16) || outer < 0 || outer > 64 || sweeps <= 0 || sweeps > 64) {
		b->err = "batch solve: precision FP16/BF16, outer 0..64, sweeps 1..64";
		return RAD_EINVAL;
	}
	return b->b->solve(prec, use_graph != 0, rho, E, tol, outer, sweeps, B, rel_res, info4, b->err) ? RAD_EDEVICE
	                                                                                              : RAD_OK;
}

}  // extern "C"
