// Batched mixed-precision gather over scenes of uneven size (see batch.h).
//
// HBM layout (scene s, n_s patches, ld_s = n_s rounded up to 256):
//   F32   ld_s^2 floats at off[s], row-major (the matrix of record)
//   F16 / BF16  ld_s^2 at off[s], MFMA fragment order, one power-of-two scale
//         for the whole batch (FP16)
//   rows  n_s * 3 at voff[s]: rho, E, residual, scaled residual, correction
//   xT    16 x ld_s at xoff[s]: FP64 iterate (rows 0..2 used), and its
//         low-precision copy in fragment order for the MFMA B operand
//   part64  ceil(ld_s / 2048) x ld_s x 16 at poff[s]: split-K partials of
//         the FP64 residual product
//   part32  ceil(ld_s / kc) x ld_s x 16 at goff[s] (room for kc = 1024):
//         split-K partials of the sweep product
// Work lists, sorted by scene: residual tiles {scene, 256-row block, k range
// of 2048}; sweep-product tiles {scene, 64*rt-row block, k range of kc},
// whose shape (rt, kc) and persistent grid come from the autotuner; element
// tiles {scene, first element} of 256 elements, never straddling a scene, so
// per-scene norms are sums over a fixed range of tile partials.
#include "batch.h"
#include "linear.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

namespace rgpu {
namespace {

#include "lin_common.h"

constexpr int NR = 3;           // right-hand sides per scene (RGB)
constexpr int kKChunk = 2048;   // k per residual tile
constexpr int kPersist = 1024;  // blocks of the element-walking kernels
constexpr int kMinKc = 1024;    // smallest k chunk of a sweep-product tile

struct BMeta {
	int n, ld, nk;  // nk = k chunks of the residual product
	int nkg;        // k chunks of the sweep product (its kc)
	long long off, voff, xoff, poff, goff;
};

// low-precision sweep product, one tile per block per step of the walk: the
// block's 4 waves own 16*RT rows each of the tile's 64*RT, over k-blocks
// [z, w) (kcb = kc / 32 per chunk)
template <typename E, int RT>
__global__ __launch_bounds__(kB) void k_bgemm(const int4 *tiles, int t0, int t1, const BMeta *meta, const int *active,
                                              const E *F16, const E *XT, float *part, int kcb)
{
	for (int t = t0 + blockIdx.x; t < t1; t += gridDim.x) {
		const int4 tl = tiles[t];  // scene, row block, k-blocks [z, w)
		if (!active[tl.x]) continue;
		const BMeta m = meta[tl.x];
		gemm_tile<E, RT, 1>(F16 + m.off, XT + m.xoff, part + m.goff + (size_t)(tl.z / kcb) * m.ld * 16, m.ld, 16,
		                    tl.y * 4 * RT + (threadIdx.x >> 6) * RT, tl.z, tl.w);
	}
}

// FP64 residual product on the vector FP64 units, same tiles (three columns:
// a 16-column FP64 MFMA would spend 13 of them on padding).  Per tile the
// block stages x over the tile's k range (at most kKChunk, a multiple of 256)
// in LDS, column-major, then walks the 256-row block in passes of 32 rows:
// wave w takes kBRows rows, each lane 4 consecutive k per step of 256 (one
// 16-byte load per row, the next step's loads in flight).  The 64 lane
// partials of the 24 sums are combined by a reduce-scatter: the exchanges
// over lane bits 5, 4, 3 halve the set each time (12, 6, 3 shuffles) and
// leave lane l with row l >> 3; bits 2, 1, 0 then sum those 3 in full.
// Writes part[kc][row][c], c < 3 (rows up to ld; F is zero past n).
constexpr int kBRows = 8;
__global__ __launch_bounds__(kB) void k_bfx64v(const int4 *tiles, int t0, int t1, const BMeta *meta, const int *active,
                                               const float *__restrict__ F32, const double *__restrict__ xT,
                                               double *__restrict__ part)
{
	typedef float fv4 __attribute__((ext_vector_type(4)));
	static_assert(kBRows == 8 && NR == 3, "the reduce-scatter assumes 8 rows of 3 columns");
	__shared__ double xs[NR * kKChunk];
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	for (int t = t0 + blockIdx.x; t < t1; t += gridDim.x) {
		const int4 tl = tiles[t];  // scene, 256-row block, k-blocks [z, w)
		if (!active[tl.x]) continue;
		const BMeta m = meta[tl.x];
		const int k0 = tl.z * 32, nkk = (tl.w - tl.z) * 32;
		__syncthreads();  // the previous tile's reads of xs are done
		for (int i = threadIdx.x; i < NR * nkk; i += kB) {
			const int c = i / nkk, kk = i - c * nkk;
			xs[c * kKChunk + kk] = xT[m.xoff + (size_t)c * m.ld + k0 + kk];
		}
		__syncthreads();
		const float *F = F32 + m.off;
		double *P = part + m.poff + (size_t)(tl.z / (kKChunk / 32)) * m.ld * 16;
#pragma unroll 1
		for (int g = 0; g < 256 / (4 * kBRows); g++) {
			const int r0 = tl.y * 256 + (g * 4 + w) * kBRows;
			const float *Fr = F + (size_t)r0 * m.ld + k0 + 4 * lane;
			double v[kBRows * NR];
#pragma unroll
			for (int i = 0; i < kBRows * NR; i++) v[i] = 0.0;
			fv4 a[kBRows], an[kBRows];
#pragma unroll
			for (int r = 0; r < kBRows; r++) a[r] = __builtin_nontemporal_load((const fv4 *)(Fr + (size_t)r * m.ld));
#pragma unroll 1
			for (int st = 0; st < nkk; st += 256) {
				if (st + 256 < nkk) {
#pragma unroll
					for (int r = 0; r < kBRows; r++)
						an[r] = __builtin_nontemporal_load((const fv4 *)(Fr + (size_t)r * m.ld + st + 256));
				}
				double x[NR][4];
#pragma unroll
				for (int c = 0; c < NR; c++)
#pragma unroll
					for (int j = 0; j < 4; j++) x[c][j] = xs[c * kKChunk + st + 4 * lane + j];
#pragma unroll
				for (int j = 0; j < 4; j++)
#pragma unroll
					for (int r = 0; r < kBRows; r++) {
						const double f = (double)a[r][j];
#pragma unroll
						for (int c = 0; c < NR; c++) v[r * NR + c] = __builtin_fma(f, x[c][j], v[r * NR + c]);
					}
#pragma unroll
				for (int r = 0; r < kBRows; r++) a[r] = an[r];
			}
#pragma unroll
			for (int half = kBRows * NR / 2, bit = 32; half >= NR; half >>= 1, bit >>= 1) {
				const bool upper = (lane & bit) != 0;
#pragma unroll
				for (int i = 0; i < half; i++) {
					const double lo = v[i], hi = v[half + i];
					v[i] = (upper ? hi : lo) + __shfl_xor(upper ? lo : hi, bit);
				}
			}
#pragma unroll
			for (int bit = 4; bit > 0; bit >>= 1)
#pragma unroll
				for (int c = 0; c < NR; c++) v[c] += __shfl_xor(v[c], bit);
			if ((lane & 7) == 0) {
				double *o = P + (size_t)(r0 + (lane >> 3)) * 16;
#pragma unroll
				for (int c = 0; c < NR; c++) o[c] = v[c];
			}
		}
	}
}

// r = E + rho o (F x) - x (FP64; part == null: x = 0, r = E); per element
// tile: sum of squares and max |r|
__global__ __launch_bounds__(kB) void k_bresid(const int2 *etiles, int e0, int e1, const BMeta *meta,
                                               const int *active, const double *part, const double *xT,
                                               const double *E, const double *rho, double *r, double *blk)
{
	__shared__ double red[8];
	for (int t = e0 + blockIdx.x; t < e1; t += gridDim.x) {
		const int2 et = etiles[t];
		if (!active[et.x]) continue;
		const BMeta m = meta[et.x];
		const int idx = et.y + threadIdx.x;
		double ss = 0.0, mx = 0.0;
		if (idx < m.n * NR) {
			const int i = idx / NR, c = idx % NR;
			const size_t g = m.voff + idx;
			double v = E[g];
			if (part) {
				double y = 0.0;
				for (int q = 0; q < m.nk; q++) y += part[m.poff + ((size_t)q * m.ld + i) * 16 + c];
				v += rho[g] * y - xT[m.xoff + (size_t)c * m.ld + i];
			}
			r[g] = v;
			ss = v * v;
			mx = fabs(v);
		}
		for (int off = 32; off > 0; off >>= 1) {
			ss += __shfl_xor(ss, off);
			mx = fmax(mx, __shfl_xor(mx, off));
		}
		if ((threadIdx.x & 63) == 0) {
			red[2 * (threadIdx.x >> 6)] = ss;
			red[2 * (threadIdx.x >> 6) + 1] = mx;
		}
		__syncthreads();
		if (threadIdx.x == 0) {
			blk[2 * t] = red[0] + red[2] + red[4] + red[6];
			blk[2 * t + 1] = fmax(fmax(red[1], red[3]), fmax(red[5], red[7]));
		}
		__syncthreads();
	}
}

// per scene (one block each): norms from its element tiles in order;
// relative residual, scale for the next correction, still active?
__global__ void k_bnorm(int sc0, const int *etb, const int *ete, const double *blk, int *active, double *enorm,
                        double *rel, double *scale, double tol, int first)
{
	const int s = sc0 + blockIdx.x;
	if (threadIdx.x != 0 || !active[s]) return;
	double ss = 0.0, mx = 0.0;
	for (int t = etb[s]; t < ete[s]; t++) {
		ss += blk[2 * t];
		mx = fmax(mx, blk[2 * t + 1]);
	}
	if (first) enorm[s] = sqrt(ss);
	const double r = enorm[s] > 0.0 ? sqrt(ss) / enorm[s] : 0.0;
	rel[s] = r;
	scale[s] = mx;
	active[s] = (r > tol && mx > 0.0) ? 1 : 0;
}

// r32 = r / scale; XT <- r32 (fragment order)
template <typename E>
__global__ __launch_bounds__(kB) void k_bscale(const int2 *etiles, int e0, int e1, const BMeta *meta,
                                               const int *active, const double *r, const double *scale, float *r32,
                                               E *XT)
{
	for (int t = e0 + blockIdx.x; t < e1; t += gridDim.x) {
		const int2 et = etiles[t];
		if (!active[et.x]) continue;
		const BMeta m = meta[et.x];
		const int idx = et.y + threadIdx.x;
		if (idx >= m.n * NR) continue;
		const size_t g = m.voff + idx;
		const float v = (float)(r[g] / scale[et.x]);
		r32[g] = v;
		XT[m.xoff + frag_index(idx % NR, idx / NR, m.ld)] = (E)v;
	}
}

// d = r32 + rho o (inv_s * sum part); XT <- d
template <typename E>
__global__ __launch_bounds__(kB) void k_bjacobi(const int2 *etiles, int e0, int e1, const BMeta *meta,
                                                const int *active, const float *part, float inv_s, const float *r32,
                                                const float *rho, float *d, E *XT)
{
	for (int t = e0 + blockIdx.x; t < e1; t += gridDim.x) {
		const int2 et = etiles[t];
		if (!active[et.x]) continue;
		const BMeta m = meta[et.x];
		const int idx = et.y + threadIdx.x;
		if (idx >= m.n * NR) continue;
		const int i = idx / NR, c = idx % NR;
		const size_t g = m.voff + idx;
		float y = 0.f;
		for (int q = 0; q < m.nkg; q++) y += part[m.goff + ((size_t)q * m.ld + i) * 16 + c];
		const float v = r32[g] + rho[g] * (y * inv_s);
		d[g] = v;
		XT[m.xoff + frag_index(c, i, m.ld)] = (E)v;
	}
}

// x += scale * d
__global__ __launch_bounds__(kB) void k_baxpy(const int2 *etiles, int e0, int e1, const BMeta *meta,
                                              const int *active, const double *scale, const float *d, double *xT)
{
	for (int t = e0 + blockIdx.x; t < e1; t += gridDim.x) {
		const int2 et = etiles[t];
		if (!active[et.x]) continue;
		const BMeta m = meta[et.x];
		const int idx = et.y + threadIdx.x;
		if (idx >= m.n * NR) continue;
		xT[m.xoff + (size_t)(idx % NR) * m.ld + idx / NR] += scale[et.x] * (double)d[m.voff + idx];
	}
}

}  // namespace

struct SceneBatch::Impl {
	int count = 0;
	std::vector<int> n, ld, tb, te, gtb, gte, etb, ete;
	std::vector<BMeta> meta;
	size_t fsize = 0, vsize = 0, xsize = 0, psize = 0, gsize = 0;
	int ntiles = 0, ngtiles = 0, netiles = 0, total = 0;
	BatchTile gs{0, 0, 0};  // current sweep-product schedule
	BatchTile best[2] = {{0, 0, 0}, {0, 0, 0}};  // autotuned, per precision
	bool forced = false;    // set_tile overrides the autotuner
	hipStream_t st = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	Buf<float> F32, r32, d32, rho32, part32;
	Buf<_Float16> F16, X16;
	Buf<__bf16> FB, XB;
	Buf<double> E64, rho64, r64, xT, part64, blk, enorm, rel, scale;
	Buf<int> active, d_tb, d_te, d_etb, d_ete;
	Buf<int4> tiles, gtiles;
	Buf<int2> etiles;
	Buf<BMeta> d_meta;
	Buf<unsigned> umax;
	bool have16 = false, haveb = false;
	float s16 = 1.f;
	hipGraphExec_t exec = nullptr;
	hipGraph_t graph = nullptr;
	int g_prec = -1, g_outer = -1, g_sweeps = -1, g_nodes = 0;
	double g_tol = -1.0;

	void drop_graph()
	{
		if (exec) (void)hipGraphExecDestroy(exec);
		if (graph) (void)hipGraphDestroy(graph);
		exec = nullptr;
		graph = nullptr;
		g_prec = -1;
	}
	int ensure_lowp(int prec, std::string &err)
	{
		if ((prec == PREC_F16 && have16) || (prec == PREC_BF16 && haveb)) return 0;
		// a captured graph bakes in the FP16 scale (its sweeps multiply by
		// 1/s16) and reads the copy being rebuilt: capture again
		drop_graph();
		if (prec == PREC_F16) {
			if (umax.alloc(1, err) || F16.alloc(fsize, err)) return -1;
			hipLaunchKernelGGL(k_fill<unsigned>, dim3(1), dim3(kB), 0, st, umax.p, (size_t)1, 0u);
			hipLaunchKernelGGL(k_absmax, dim3(4096), dim3(kB), 0, st, F32.p, fsize, umax.p);
			unsigned bits = 0;
			HIP_OK(hipMemcpyAsync(&bits, umax.p, 4, hipMemcpyDeviceToHost, st), "absmax");
			HIP_OK(hipStreamSynchronize(st), "absmax");
			float m;
			std::memcpy(&m, &bits, 4);
			if (!std::isfinite(m)) {
				err = "matrix has non-finite entries";
				return -1;
			}
			int e = 0;
			if (m > 0.f) e = std::max(-60, std::min(60, (int)std::floor(std::log2(16384.0 / m))));
			s16 = std::ldexp(1.0f, e);
			for (int s = 0; s < count; s++)
				hipLaunchKernelGGL(k_to_frag<_Float16>, dim3(2048), dim3(kB), 0, st, F32.p + meta[s].off, ld[s], s16,
				                   F16.p + meta[s].off);
			have16 = true;
		} else {
			if (FB.alloc(fsize, err)) return -1;
			for (int s = 0; s < count; s++)
				hipLaunchKernelGGL(k_to_frag<__bf16>, dim3(2048), dim3(kB), 0, st, F32.p + meta[s].off, ld[s], 1.0f,
				                   FB.p + meta[s].off);
			haveb = true;
		}
		HIP_OK(hipGetLastError(), "convert");
		return 0;
	}

	static bool valid_tile(const BatchTile &g)
	{
		return (g.rt == 1 || g.rt == 2 || g.rt == 4) &&
		       (g.kc == 1024 || g.kc == 2048 || g.kc == 4096 || g.kc == 8192) && g.grid >= 64 && g.grid <= 16384;
	}
	// rebuilds the sweep-product tile list and chunk counts for schedule g
	int set_sched(const BatchTile &g, std::string &err)
	{
		if (!valid_tile(g)) {
			err = "batch tile: rt in {1, 2, 4}, kc in {1024, 2048, 4096, 8192}, grid 64..16384";
			return -1;
		}
		if (g.rt == gs.rt && g.kc == gs.kc && g.grid == gs.grid) return 0;
		// largest scenes first: the blocks that take the longest tiles start
		// first, which shortens the tail of a launch over the whole batch
		std::vector<int> ord(count);
		for (int s = 0; s < count; s++) ord[s] = s;
		std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return ld[a] > ld[b]; });
		std::vector<int4> tl;
		gtb.assign(count, 0);
		gte.assign(count, 0);
		for (const int s : ord) {
			BMeta &m = meta[s];
			m.nkg = (m.ld + g.kc - 1) / g.kc;
			gtb[s] = (int)tl.size();
			for (int rb = 0; rb < m.ld / (64 * g.rt); rb++)
				for (int q = 0; q < m.nkg; q++) {
					const int kb0 = q * (g.kc / 32), kb1 = std::min(m.ld / 32, kb0 + g.kc / 32);
					tl.push_back(make_int4(s, rb, kb0, kb1));
				}
			gte[s] = (int)tl.size();
		}
		ngtiles = (int)tl.size();
		HIP_OK(hipStreamSynchronize(st), "batch tile");
		if (gtiles.alloc(tl.size(), err)) return -1;
		HIP_OK(hipMemcpyAsync(gtiles.p, tl.data(), tl.size() * sizeof(int4), hipMemcpyHostToDevice, st), "batch tile");
		HIP_OK(hipMemcpyAsync(d_meta.p, meta.data(), count * sizeof(BMeta), hipMemcpyHostToDevice, st), "batch tile");
		HIP_OK(hipStreamSynchronize(st), "batch tile");
		gs = g;
		drop_graph();
		return 0;
	}
	template <typename E> void launch_gemm(const E *Fl, const E *X, int g0, int g1)
	{
		const dim3 grid(std::max(1, std::min(gs.grid, g1 - g0)));
		const int kcb = gs.kc / 32;
		if (gs.rt == 1)
			hipLaunchKernelGGL((k_bgemm<E, 1>), grid, dim3(kB), 0, st, gtiles.p, g0, g1, d_meta.p, active.p, Fl, X,
			                   part32.p, kcb);
		else if (gs.rt == 2)
			hipLaunchKernelGGL((k_bgemm<E, 2>), grid, dim3(kB), 0, st, gtiles.p, g0, g1, d_meta.p, active.p, Fl, X,
			                   part32.p, kcb);
		else
			hipLaunchKernelGGL((k_bgemm<E, 4>), grid, dim3(kB), 0, st, gtiles.p, g0, g1, d_meta.p, active.p, Fl, X,
			                   part32.p, kcb);
	}
	// times one sweep (product + Jacobi update, every scene active) for each
	// candidate schedule on this batch's matrices and keeps the fastest
	template <typename E> int tune(const E *Fl, E *X, float inv, BatchTile &out, std::string &err)
	{
		const size_t V = vsize;
		hipLaunchKernelGGL(k_fill<int>, dim3(grid_for((size_t)count)), dim3(kB), 0, st, active.p, (size_t)count, 1);
		hipLaunchKernelGGL(k_fill<float>, dim3(grid_for(V)), dim3(kB), 0, st, r32.p, V, 0.f);
		hipLaunchKernelGGL(k_fill<float>, dim3(grid_for(V)), dim3(kB), 0, st, rho32.p, V, 0.f);
		hipLaunchKernelGGL(k_fill<E>, dim3(grid_for(xsize)), dim3(kB), 0, st, X, xsize, (E)0.f);
		HIP_OK(hipGetLastError(), "tune");
		const int q0 = etb[0], q1 = ete[count - 1];
		const int ge = std::max(1, std::min(kPersist * 4, q1 - q0));
		double best_ms = 1e300;
		BatchTile pick = gs;
		for (int rt : {1, 2, 4})
			for (int kc : {1024, 2048, 4096, 8192})
				for (int grid : {512, 1024, 2048, 4096}) {
					if (set_sched(BatchTile{rt, kc, grid}, err)) return -1;
					const int g0 = 0, g1 = ngtiles;
					constexpr int kReps = 3;
					for (int r = 0; r <= kReps; r++) {
						if (r == 1) HIP_OK(hipEventRecord(e0, st), "tune");
						launch_gemm<E>(Fl, X, g0, g1);
						hipLaunchKernelGGL(k_bjacobi<E>, dim3(ge), dim3(kB), 0, st, etiles.p, q0, q1, d_meta.p,
						                   active.p, part32.p, inv, r32.p, rho32.p, d32.p, X);
					}
					HIP_OK(hipEventRecord(e1, st), "tune");
					HIP_OK(hipEventSynchronize(e1), "tune");
					float ms = 0.f;
					HIP_OK(hipEventElapsedTime(&ms, e0, e1), "tune");
					if (ms < best_ms) {
						best_ms = ms;
						pick = gs;
					}
				}
		out = pick;
		return set_sched(pick, err);
	}

	// the solve's kernel sequence for scenes [sc0, sc1) (one scene, or all:
	// the sweep-product list is in its own scene order); returns the number
	// of launches
	template <typename E>
	int enqueue(const E *Fl, E *X, int sc0, int sc1, int outer, int sweeps, double tol, float inv)
	{
		int launches = 0;
		const int t0 = tb[sc0], t1 = te[sc1 - 1], q0 = etb[sc0], q1 = ete[sc1 - 1];
		const bool all = sc0 == 0 && sc1 == count;
		const int g0 = all ? 0 : gtb[sc0], g1 = all ? ngtiles : gte[sc0];  // else sc1 = sc0 + 1
		const int gt = std::max(1, std::min(kPersist, t1 - t0)), ge = std::max(1, std::min(kPersist * 4, q1 - q0));
		for (int o = 0; o <= outer; o++) {
			if (o > 0) {
				hipLaunchKernelGGL(k_bfx64v, dim3(gt), dim3(kB), 0, st, tiles.p, t0, t1, d_meta.p, active.p, F32.p, xT.p,
				                   part64.p);
				launches++;
			}
			hipLaunchKernelGGL(k_bresid, dim3(ge), dim3(kB), 0, st, etiles.p, q0, q1, d_meta.p, active.p,
			                   o > 0 ? (const double *)part64.p : (const double *)nullptr, xT.p, E64.p, rho64.p, r64.p,
			                   blk.p);
			hipLaunchKernelGGL(k_bnorm, dim3(sc1 - sc0), dim3(64), 0, st, sc0, d_etb.p, d_ete.p, blk.p, active.p,
			                   enorm.p, rel.p, scale.p, tol, o == 0 ? 1 : 0);
			launches += 2;
			if (o == outer) break;
			hipLaunchKernelGGL(k_bscale<E>, dim3(ge), dim3(kB), 0, st, etiles.p, q0, q1, d_meta.p, active.p, r64.p,
			                   scale.p, r32.p, X);
			launches++;
			for (int m = 0; m < sweeps; m++) {
				launch_gemm<E>(Fl, X, g0, g1);
				hipLaunchKernelGGL(k_bjacobi<E>, dim3(ge), dim3(kB), 0, st, etiles.p, q0, q1, d_meta.p, active.p,
				                   part32.p, inv, r32.p, rho32.p, d32.p, X);
				launches += 2;
			}
			hipLaunchKernelGGL(k_baxpy, dim3(ge), dim3(kB), 0, st, etiles.p, q0, q1, d_meta.p, active.p, scale.p, d32.p,
			                   xT.p);
			launches++;
		}
		return launches;
	}

	template <typename E> int fills(E *X)
	{
		hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(xsize)), dim3(kB), 0, st, xT.p, xsize, 0.0);
		hipLaunchKernelGGL(k_fill<E>, dim3(grid_for(xsize)), dim3(kB), 0, st, X, xsize, (E)0.f);
		hipLaunchKernelGGL(k_fill<int>, dim3(grid_for((size_t)count)), dim3(kB), 0, st, active.p, (size_t)count, 1);
		return 3;
	}
};

SceneBatch::~SceneBatch()
{
	if (!d_) return;
	if (d_->st) (void)hipStreamSynchronize(d_->st);
	d_->drop_graph();
	if (d_->e0) (void)hipEventDestroy(d_->e0);
	if (d_->e1) (void)hipEventDestroy(d_->e1);
	hipStream_t s = d_->st;
	delete d_;
	if (s) (void)hipStreamDestroy(s);
}

SceneBatch *SceneBatch::create(int count, const int *n, std::string &err)
{
	if (count <= 0 || count > 65536 || !n) {
		err = "batch: count must be 1..65536";
		return nullptr;
	}
	for (int s = 0; s < count; s++)
		if (n[s] <= 0 || n[s] > 65536) {
			err = "batch: scene sizes must be 1..65536";
			return nullptr;
		}
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
		err = "no HIP device available";
		return nullptr;
	}
	SceneBatch *B = new SceneBatch();
	B->d_ = new Impl();
	Impl &d = *B->d_;
	d.count = count;
	std::vector<int4> tl;
	std::vector<int2> el;
	for (int s = 0; s < count; s++) {
		BMeta m;
		m.n = n[s];
		m.ld = (n[s] + 255) / 256 * 256;
		m.nk = (m.ld + kKChunk - 1) / kKChunk;
		m.nkg = 0;
		m.off = (long long)d.fsize;
		m.voff = (long long)d.vsize;
		m.xoff = (long long)d.xsize;
		m.poff = (long long)d.psize;
		m.goff = (long long)d.gsize;
		d.fsize += (size_t)m.ld * m.ld;
		d.vsize += (size_t)m.n * NR;
		d.xsize += (size_t)16 * m.ld;
		d.psize += (size_t)m.nk * m.ld * 16;
		d.gsize += (size_t)((m.ld + kMinKc - 1) / kMinKc) * m.ld * 16;
		d.n.push_back(m.n);
		d.ld.push_back(m.ld);
		d.meta.push_back(m);
		d.tb.push_back((int)tl.size());
		for (int rb = 0; rb < m.ld / 256; rb++)
			for (int kc = 0; kc < m.nk; kc++) {
				const int kb0 = kc * (kKChunk / 32), kb1 = std::min(m.ld / 32, kb0 + kKChunk / 32);
				tl.push_back(make_int4(s, rb, kb0, kb1));
			}
		d.te.push_back((int)tl.size());
		d.etb.push_back((int)el.size());
		for (int e = 0; e < m.n * NR; e += kB) el.push_back(make_int2(s, e));
		d.ete.push_back((int)el.size());
		d.total += m.n;
	}
	d.ntiles = (int)tl.size();
	d.netiles = (int)el.size();
	std::string e;
	if (hipStreamCreateWithFlags(&d.st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&d.e0) != hipSuccess ||
	    hipEventCreate(&d.e1) != hipSuccess) {
		err = "stream/event creation failed";
		delete B;
		return nullptr;
	}
	const size_t V = d.vsize;
	if (d.F32.alloc(d.fsize, e) || d.r32.alloc(V, e) || d.d32.alloc(V, e) || d.rho32.alloc(V, e) ||
	    d.part32.alloc(d.gsize, e) || d.E64.alloc(V, e) || d.rho64.alloc(V, e) || d.r64.alloc(V, e) ||
	    d.xT.alloc(d.xsize, e) || d.part64.alloc(d.psize, e) || d.blk.alloc((size_t)d.netiles * 2, e) ||
	    d.enorm.alloc(count, e) || d.rel.alloc(count, e) || d.scale.alloc(count, e) || d.active.alloc(count, e) ||
	    d.d_tb.alloc(count, e) || d.d_te.alloc(count, e) || d.d_etb.alloc(count, e) || d.d_ete.alloc(count, e) ||
	    d.tiles.alloc(tl.size(), e) || d.etiles.alloc(el.size(), e) || d.d_meta.alloc(count, e)) {
		err = e;
		delete B;
		return nullptr;
	}
	if (hipMemsetAsync(d.F32.p, 0, d.fsize * 4, d.st) != hipSuccess ||
	    hipMemcpyAsync(d.tiles.p, tl.data(), tl.size() * sizeof(int4), hipMemcpyHostToDevice, d.st) != hipSuccess ||
	    hipMemcpyAsync(d.etiles.p, el.data(), el.size() * sizeof(int2), hipMemcpyHostToDevice, d.st) != hipSuccess ||
	    hipMemcpyAsync(d.d_meta.p, d.meta.data(), count * sizeof(BMeta), hipMemcpyHostToDevice, d.st) != hipSuccess ||
	    hipMemcpyAsync(d.d_etb.p, d.etb.data(), count * 4, hipMemcpyHostToDevice, d.st) != hipSuccess ||
	    hipMemcpyAsync(d.d_ete.p, d.ete.data(), count * 4, hipMemcpyHostToDevice, d.st) != hipSuccess ||
	    hipStreamSynchronize(d.st) != hipSuccess) {
		err = "batch upload failed";
		delete B;
		return nullptr;
	}
	if (d.set_sched(BatchTile{4, kKChunk, kPersist}, err)) {
		delete B;
		return nullptr;
	}
	return B;
}

int SceneBatch::count() const { return d_->count; }
int SceneBatch::n(int s) const { return d_->n[s]; }
int SceneBatch::ld(int s) const { return d_->ld[s]; }
int SceneBatch::total() const { return d_->total; }

float *SceneBatch::matrix32(int s)
{
	d_->have16 = d_->haveb = false;
	return d_->F32.p + d_->meta[s].off;
}

int SceneBatch::set_matrix(int s, const float *F, std::string &err)
{
	if (s < 0 || s >= d_->count) {
		err = "batch: scene index out of range";
		return -1;
	}
	const int n = d_->n[s], ld = d_->ld[s];
	HIP_OK(hipMemcpy2D(matrix32(s), (size_t)ld * 4, F, (size_t)n * 4, (size_t)n * 4, n, hipMemcpyHostToDevice),
	       "set matrix");
	return 0;
}

int SceneBatch::autotune(int prec, BatchTile &out, std::string &err)
{
	Impl &d = *d_;
	if (prec != PREC_F16 && prec != PREC_BF16) {
		err = "batch: precision must be FP16 or BF16";
		return -1;
	}
	if (d.ensure_lowp(prec, err)) return -1;
	const int pi = prec == PREC_F16 ? 0 : 1;
	if (prec == PREC_F16) {
		if (d.X16.alloc(d.xsize, err) || d.tune<_Float16>(d.F16.p, d.X16.p, 1.0f / d.s16, d.best[pi], err)) return -1;
	} else {
		if (d.XB.alloc(d.xsize, err) || d.tune<__bf16>(d.FB.p, d.XB.p, 1.0f, d.best[pi], err)) return -1;
	}
	d.forced = false;
	out = d.best[pi];
	return 0;
}

int SceneBatch::set_tile(const BatchTile &t, std::string &err)
{
	if (d_->set_sched(t, err)) return -1;
	d_->forced = true;
	return 0;
}

int SceneBatch::solve(int prec, bool graph, const double *rho, const double *E, double tol, int outer, int sweeps,
                      double *Bout, double *rel_res, double *info, std::string &err)
{
	Impl &d = *d_;
	if (prec != PREC_F16 && prec != PREC_BF16) {
		err = "batch: precision must be FP16 or BF16";
		return -1;
	}
	if (outer < 0 || outer > 64 || sweeps <= 0 || sweeps > 64) {
		err = "batch: outer must be 0..64 and sweeps 1..64";
		return -1;
	}
	if (d.ensure_lowp(prec, err)) return -1;
	if (!d.forced) {
		const BatchTile &b = d.best[prec == PREC_F16 ? 0 : 1];
		BatchTile got;
		if (b.rt == 0 ? autotune(prec, got, err) : d.set_sched(b, err)) return -1;
	}
	const size_t V = d.vsize;
	std::vector<float> h(V);
	for (size_t t = 0; t < V; t++) h[t] = (float)rho[t];
	HIP_OK(hipMemcpyAsync(d.rho32.p, h.data(), V * 4, hipMemcpyHostToDevice, d.st), "upload");
	HIP_OK(hipMemcpyAsync(d.rho64.p, rho, V * 8, hipMemcpyHostToDevice, d.st), "upload");
	HIP_OK(hipMemcpyAsync(d.E64.p, E, V * 8, hipMemcpyHostToDevice, d.st), "upload");
	HIP_OK(hipStreamSynchronize(d.st), "upload");
	const float inv = prec == PREC_F16 ? 1.0f / d.s16 : 1.0f;
	double build_ms = 0.0;
	int launches = 0;
	auto run_all = [&](void) -> int {
		int l = 0;
		if (prec == PREC_F16) {
			l += d.fills(d.X16.p);
			l += d.enqueue<_Float16>(d.F16.p, d.X16.p, 0, d.count, outer, sweeps, tol, inv);
		} else {
			l += d.fills(d.XB.p);
			l += d.enqueue<__bf16>(d.FB.p, d.XB.p, 0, d.count, outer, sweeps, tol, inv);
		}
		return l;
	};
	if (prec == PREC_F16 ? d.X16.alloc(d.xsize, err) : d.XB.alloc(d.xsize, err)) return -1;
	if (graph) {
		if (!(d.exec && d.g_prec == prec && d.g_outer == outer && d.g_sweeps == sweeps && d.g_tol == tol)) {
			d.drop_graph();
			const auto c0 = std::chrono::steady_clock::now();
			HIP_OK(hipStreamBeginCapture(d.st, hipStreamCaptureModeThreadLocal), "capture");
			const int l = run_all();
			hipGraph_t g = nullptr;
			const hipError_t ce = hipStreamEndCapture(d.st, &g);
			if (ce != hipSuccess || !g) {
				err = std::string("graph capture: ") + hipGetErrorString(ce);
				return -1;
			}
			d.graph = g;
			HIP_OK(hipGraphInstantiate(&d.exec, d.graph, nullptr, nullptr, 0), "graph instantiate");
			d.g_prec = prec;
			d.g_outer = outer;
			d.g_sweeps = sweeps;
			d.g_tol = tol;
			d.g_nodes = l;
			build_ms = ms_since(c0);
		}
		launches = d.g_nodes;
		HIP_OK(hipEventRecord(d.e0, d.st), "event");
		HIP_OK(hipGraphLaunch(d.exec, d.st), "graph launch");
		HIP_OK(hipEventRecord(d.e1, d.st), "event");
	} else {
		HIP_OK(hipEventRecord(d.e0, d.st), "event");
		launches += prec == PREC_F16 ? d.fills(d.X16.p) : d.fills(d.XB.p);
		for (int s = 0; s < d.count; s++) {
			if (prec == PREC_F16)
				launches += d.enqueue<_Float16>(d.F16.p, d.X16.p, s, s + 1, outer, sweeps, tol, inv);
			else
				launches += d.enqueue<__bf16>(d.FB.p, d.XB.p, s, s + 1, outer, sweeps, tol, inv);
		}
		HIP_OK(hipEventRecord(d.e1, d.st), "event");
	}
	HIP_OK(hipGetLastError(), "batch launch");
	HIP_OK(hipEventSynchronize(d.e1), "batch solve");
	float ms = 0.f;
	HIP_OK(hipEventElapsedTime(&ms, d.e0, d.e1), "event");
	std::vector<double> hx(d.xsize);
	std::vector<int> act(d.count);
	HIP_OK(hipMemcpy(hx.data(), d.xT.p, d.xsize * 8, hipMemcpyDeviceToHost), "download");
	HIP_OK(hipMemcpy(rel_res, d.rel.p, d.count * 8, hipMemcpyDeviceToHost), "download");
	HIP_OK(hipMemcpy(act.data(), d.active.p, d.count * 4, hipMemcpyDeviceToHost), "download");
	int above = 0;
	for (int s = 0; s < d.count; s++) {
		const BMeta &m = d.meta[s];
		for (int i = 0; i < m.n; i++)
			for (int c = 0; c < NR; c++) Bout[(size_t)m.voff + (size_t)i * NR + c] = hx[(size_t)m.xoff + (size_t)c * m.ld + i];
		above += act[s];
	}
	if (info) {
		info[0] = ms;
		info[1] = build_ms;
		info[2] = launches;
		info[3] = above;
	}
	return 0;
}

}  // namespace rgpu
