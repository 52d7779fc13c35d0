// Linear radiosity solvers on an explicit form-factor matrix (see linear.h).
//
// HBM layout
//   F32   ld x ld floats, row-major, zero padded (ld = n rounded up to 256);
//         the matrix of record: FP64 residuals and the FP32 solvers read it.
//   F16 / BF16  s * F in MFMA fragment order, made on demand: the matrix is
//         cut into 16 x 32 tiles, tile (rb, kb) at ((rb * ld/32) + kb) * 512
//         elements, and inside a tile lane l's 8 elements (row rb*16 + l%16,
//         columns kb*32 + 8*(l/16) .. +7) at l*8.  One wave-wide 16-byte load
//         per lane fetches a whole A fragment as 1 KB of contiguous HBM.
//         s is a power of two that lifts the largest |F| to ~2^14 for FP16
//         (form factors of 1e-6 would otherwise be FP16 subnormals) and 1 for
//         BF16; the product is scaled back by 1/s exactly.
//   XT    the low-precision iterate, nrhs (rounded up to 16) x ld, in the
//         same fragment order (it is the MFMA B operand).
//
// MFMA (v_mfma_f32_16x16x32_{f16,bf16}, gfx950): lane l supplies
// A[l%16][8*(l/16)+j] and B[8*(l/16)+j][l%16], j = 0..7, and receives
// D[4*(l/16)+j][l%16], j = 0..3.  A wave owns 16*RT rows and all 16*CT
// right-hand-side columns, so F is read from HBM once per sweep; split-K S
// spreads the K loop over S blocks whose partial tiles the epilogue sums in a
// fixed order (results do not depend on timing).  (RT, S) per (CT, precision)
// come from the autotuner, which times GEMM + epilogue for each candidate.
#include "linear.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

namespace rgpu {
namespace {

#include "lin_common.h"

// part[s][row][col] = sum over the s-th K chunk of A[row][k] X[k][col]
template <typename E, int RT, int CT>
__global__ __launch_bounds__(kB) void k_gemm(const E *__restrict__ A, const E *__restrict__ XT, float *__restrict__ part,
                                             int ld, int rp, int kchunk)
{
	const int w = threadIdx.x >> 6;
	const int s = blockIdx.y, kb0 = s * (kchunk >> 5);
	gemm_tile<E, RT, CT>(A, XT, part + (size_t)s * ld * rp, ld, rp, (blockIdx.x * 4 + w) * RT, kb0, kb0 + (kchunk >> 5));
}

// Wide right-hand sides (NCT = 8 or 12 column tiles) with the operands
// staged by LDS DMA.  A block is eight waves over 256 rows: WC = 1 gives each
// wave 16 RT = 32 rows and every column, WC = 2 gives it 64 rows and half the
// columns, which cuts the block's LDS reads per k-block from 112 to 80 KB (a
// wave with 64 rows and every column would need its accumulators in AGPRs,
// which the register allocator then copies every k-block).  Each k-block's A
// fragments (1 KB each, those of a row group split over its WC waves) and B
// fragments (spread over the waves) go straight from HBM into a ring of NS
// LDS slots (global_load_lds_dwordx4: no registers held while in flight),
// NS - 1 k-blocks ahead of the MFMAs: 64 KB of A per CU in flight, where a
// register pipeline holds 32 KB.  The compiler treats any LDS read as
// dependent on every LDS DMA in flight and would drain them all before it,
// so the slot is read by inline assembly (its own lgkmcnt wait inside;
// early-clobber outputs, so no read overwrites the address of a later one)
// after an explicit vmcnt wait for this k-block's DMAs and a barrier; every
// trip issues the same DMAs per wave (the index is clamped past the chunk's
// end; those slots are never read), so the vmcnt bound holds to the last
// trip.  The second wave on each SIMD (waves 4-7) runs half a k-block
// behind: after each barrier it multiplies the previous k-block while the
// first wave reads, then reads while the first wave multiplies (0.433 ->
// 0.389 ms for 192 columns at n = 32768, where all eight waves had read and
// then multiplied in step).
// cache policy operand of global_load_lds: nt (non-temporal)
constexpr int kCpolNT = 2;

template <int N>
__device__ __forceinline__ void wait_vmcnt()
{
	static_assert(N >= 0 && N < 64, "vmcnt is six bits");
	asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

// RT A fragments (offsets 0, 1024, ...) from ab, then CT B fragments from bb,
// completed before the block ends
#define RGPU_DS1(i, base, off) "ds_read_b128 %" #i ", %" #base " offset:" #off "\n"
template <int RT, int CT, typename V>
__device__ __forceinline__ void read_slot(V (&a)[RT], V (&b)[CT], unsigned ab, unsigned bb)
{
	if constexpr (RT == 2 && CT == 12) {
		asm volatile(RGPU_DS1(0, 14, 0) RGPU_DS1(1, 14, 1024) RGPU_DS1(2, 15, 0) RGPU_DS1(3, 15, 1024)
		             RGPU_DS1(4, 15, 2048) RGPU_DS1(5, 15, 3072) RGPU_DS1(6, 15, 4096) RGPU_DS1(7, 15, 5120)
		             RGPU_DS1(8, 15, 6144) RGPU_DS1(9, 15, 7168) RGPU_DS1(10, 15, 8192) RGPU_DS1(11, 15, 9216)
		             RGPU_DS1(12, 15, 10240) RGPU_DS1(13, 15, 11264) "s_waitcnt lgkmcnt(0)"
		             : "=&v"(a[0]), "=&v"(a[1]), "=&v"(b[0]), "=&v"(b[1]), "=&v"(b[2]), "=&v"(b[3]), "=&v"(b[4]),
		               "=&v"(b[5]), "=&v"(b[6]), "=&v"(b[7]), "=&v"(b[8]), "=&v"(b[9]), "=&v"(b[10]), "=&v"(b[11])
		             : "v"(ab), "v"(bb)
		             : "memory");
	} else if constexpr (RT == 2 && CT == 8) {
		asm volatile(RGPU_DS1(0, 10, 0) RGPU_DS1(1, 10, 1024) RGPU_DS1(2, 11, 0) RGPU_DS1(3, 11, 1024)
		             RGPU_DS1(4, 11, 2048) RGPU_DS1(5, 11, 3072) RGPU_DS1(6, 11, 4096) RGPU_DS1(7, 11, 5120)
		             RGPU_DS1(8, 11, 6144) RGPU_DS1(9, 11, 7168) "s_waitcnt lgkmcnt(0)"
		             : "=&v"(a[0]), "=&v"(a[1]), "=&v"(b[0]), "=&v"(b[1]), "=&v"(b[2]), "=&v"(b[3]), "=&v"(b[4]),
		               "=&v"(b[5]), "=&v"(b[6]), "=&v"(b[7])
		             : "v"(ab), "v"(bb)
		             : "memory");
	} else if constexpr (RT == 4 && CT == 6) {
		asm volatile(RGPU_DS1(0, 10, 0) RGPU_DS1(1, 10, 1024) RGPU_DS1(2, 10, 2048) RGPU_DS1(3, 10, 3072)
		             RGPU_DS1(4, 11, 0) RGPU_DS1(5, 11, 1024) RGPU_DS1(6, 11, 2048) RGPU_DS1(7, 11, 3072)
		             RGPU_DS1(8, 11, 4096) RGPU_DS1(9, 11, 5120) "s_waitcnt lgkmcnt(0)"
		             : "=&v"(a[0]), "=&v"(a[1]), "=&v"(a[2]), "=&v"(a[3]), "=&v"(b[0]), "=&v"(b[1]), "=&v"(b[2]),
		               "=&v"(b[3]), "=&v"(b[4]), "=&v"(b[5])
		             : "v"(ab), "v"(bb)
		             : "memory");
	} else if constexpr (RT == 4 && CT == 4) {
		asm volatile(RGPU_DS1(0, 8, 0) RGPU_DS1(1, 8, 1024) RGPU_DS1(2, 8, 2048) RGPU_DS1(3, 8, 3072)
		             RGPU_DS1(4, 9, 0) RGPU_DS1(5, 9, 1024) RGPU_DS1(6, 9, 2048) RGPU_DS1(7, 9, 3072)
		             "s_waitcnt lgkmcnt(0)"
		             : "=&v"(a[0]), "=&v"(a[1]), "=&v"(a[2]), "=&v"(a[3]), "=&v"(b[0]), "=&v"(b[1]), "=&v"(b[2]),
		               "=&v"(b[3])
		             : "v"(ab), "v"(bb)
		             : "memory");
	} else {
		static_assert(RT < 0, "no slot reader for this tile");
	}
}
#undef RGPU_DS1

template <typename E, int NCT, int RT, int WC>
__global__ __launch_bounds__(512) void k_gemm_dma(const E *__restrict__ A, const E *__restrict__ XT,
                                                  float *__restrict__ part, int ld, int rp, int kchunk)
{
	typedef typename VecOf<E>::type V;
	constexpr int NS = 5, W = 8, WR = W / WC, CT = NCT / WC;
	constexpr int SLOT = (WR * RT + NCT) * 64;  // fragments (16 bytes) per ring slot
	// DMAs per wave and k-block: RT / WC of its row group's A fragments, and
	// BW of the B fragments, one more on waves w < BX
	constexpr int AW = RT / WC, BW = NCT / W, BX = NCT % W;
	static_assert(NCT == 8 || NCT == 12, "8 or 12 column tiles");
	static_assert(WR * RT == 16 && AW * WC == RT && CT * WC == NCT, "256 rows per block, whole splits");
	__shared__ V ring[NS * SLOT];
	const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int wr = w % WR, wc = w / WR;  // this wave's row group and column group
	const int rb0 = (blockIdx.x * WR + wr) * RT;
	// waves 4-7 (the second wave on each SIMD) run half a k-block behind:
	// their MFMAs on the previous k-block overlap the first four waves' LDS
	// reads, and their reads the others' MFMAs
	const bool late = w >= 4;
	const int s = blockIdx.y, kb0 = s * (kchunk >> 5), kb1 = kb0 + (kchunk >> 5);
	const int nkb = ld >> 5;
	const V *Ap = (const V *)A + lane, *Bp = (const V *)XT + lane;
	const bool extra = w < BX;
	const int c0 = w * BW + (extra ? w : BX);  // the first B fragment this wave loads
	f4 acc[RT][CT];
#pragma unroll
	for (int r = 0; r < RT; r++)
#pragma unroll
		for (int c = 0; c < CT; c++) acc[r][c] = f4{0.f, 0.f, 0.f, 0.f};
	auto issue = [&](int kb, int slot) {
		const int k = min(kb, kb1 - 1);
		V *base = ring + slot * SLOT;
		// A streams once: non-temporal (0.389 -> 0.364 ms at 192 columns);
		// B is read by every block and stays in L2
#pragma unroll
		for (int r = wc * AW; r < (wc + 1) * AW; r++)
			__builtin_amdgcn_global_load_lds((const void *)(Ap + ((size_t)(rb0 + r) * nkb + k) * 64),
			                                 (__attribute__((address_space(3))) void *)(base + (wr * RT + r) * 64), 16, 0,
			                                 kCpolNT);
#pragma unroll
		for (int m = 0; m < BW; m++)
			__builtin_amdgcn_global_load_lds((const void *)(Bp + ((size_t)(c0 + m) * nkb + k) * 64),
			                                 (__attribute__((address_space(3))) void *)(base + (WR * RT + c0 + m) * 64), 16, 0, 0);
		if (BX && extra)
			__builtin_amdgcn_global_load_lds((const void *)(Bp + ((size_t)(c0 + BW) * nkb + k) * 64),
			                                 (__attribute__((address_space(3))) void *)(base + (WR * RT + c0 + BW) * 64), 16, 0, 0);
	};
	const unsigned lbase = (unsigned)(uintptr_t)(__attribute__((address_space(3))) void *)ring + lane * 16;
	for (int i = 0; i < NS - 1; i++) issue(kb0 + i, i);
	int slot = 0;
	auto mma = [&](const V (&a)[RT], const V (&b)[CT]) {
#pragma unroll
		for (int r = 0; r < RT; r++)
#pragma unroll
			for (int c = 0; c < CT; c++) acc[r][c] = mfma(a[r], b[c], acc[r][c]);
	};
	// one k-block: its DMAs waited for, the barrier, the refill, then the
	// wave's own reads and MFMAs (body)
	auto step = [&](int kb, auto body) {
		// this k-block's DMAs: NS - 2 k-blocks' DMAs were issued after them
		constexpr int D = AW + BW;
		if (BX && extra) wait_vmcnt<(NS - 2) * (D + 1)>();
		else wait_vmcnt<(NS - 2) * D>();
		__builtin_amdgcn_s_barrier();
		// every wave is past its reads of the previous trip's slot: refill it
		issue(kb + NS - 1, slot == 0 ? NS - 1 : slot - 1);
		body(lbase + (unsigned)(slot * SLOT * 16 + wr * RT * 1024),
		     lbase + (unsigned)(slot * SLOT * 16 + (WR * RT + wc * CT) * 1024));
		slot = slot == NS - 1 ? 0 : slot + 1;
	};
	if (!late) {
		for (int kb = kb0; kb < kb1; kb++)
			step(kb, [&](unsigned ab, unsigned bb) {
				V a[RT], b[CT];
				read_slot<RT, CT>(a, b, ab, bb);
				mma(a, b);
			});
	} else {
		// the previous k-block's MFMAs, then this one's reads, into the
		// other of two fragment sets: no read overwrites an operand of an
		// MFMA issued in the same trip
		V a0[RT], b0[CT], a1[RT], b1[CT];
		step(kb0, [&](unsigned ab, unsigned bb) { read_slot<RT, CT>(a0, b0, ab, bb); });
		for (int kb = kb0 + 1; kb < kb1; kb += 2) {
			step(kb, [&](unsigned ab, unsigned bb) {
				mma(a0, b0);
				read_slot<RT, CT>(a1, b1, ab, bb);
			});
			if (kb + 1 < kb1)
				step(kb + 1, [&](unsigned ab, unsigned bb) {
					mma(a1, b1);
					read_slot<RT, CT>(a0, b0, ab, bb);
				});
		}
		if ((kb1 - kb0) & 1) mma(a0, b0);  // the last k-block went to set 0
		else mma(a1, b1);
	}
	asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // no DMA may outlive the block's LDS
	const int lr = lane & 15;
#pragma unroll
	for (int r = 0; r < RT; r++)
#pragma unroll
		for (int c = 0; c < CT; c++)
#pragma unroll
			for (int j = 0; j < 4; j++) {
				const int row = (rb0 + r) * 16 + (lane >> 4) * 4 + j;
				part[((size_t)s * ld + row) * rp + (wc * CT + c) * 16 + lr] = acc[r][c][j];
			}
}

// d = rhs + rho o (inv_s * sum_s part); XT <- d (fragment order).
// Padding of XT (rows >= nrhs, k >= n) is never written and stays zero.
// RHO_IN (the adjoint's sweep d <- r + F^T (rho o d)): v = rhs + inv_s * sum
// part and XT <- rho o v instead.
template <typename E, bool RHO_IN = false>
__global__ void k_jacobi(const float *part, int S, int ld, int rp, int n, int nrhs, float inv_s, const float *rhs,
                         const float *rho, float *d, E *XT)
{
	const size_t total = (size_t)n * nrhs;
	for (size_t t = blockIdx.x * (size_t)kB + threadIdx.x; t < total; t += (size_t)gridDim.x * kB) {
		const int i = (int)(t / nrhs), c = (int)(t % nrhs);
		float y = 0.f;
		for (int s = 0; s < S; s++) y += part[((size_t)s * ld + i) * rp + c];
		const float v = RHO_IN ? rhs[t] + y * inv_s : rhs[t] + rho[t] * (y * inv_s);
		d[t] = v;
		XT[frag_index(c, i, ld)] = (E)(RHO_IN ? rho[t] * v : v);
	}
}

// Split-K partials part[s][row][col] (FP64) of F X, summed in a fixed order
// by k_resid.
template <int RT, int CT>
__global__ __launch_bounds__(kB) void k_fx64m(const float *__restrict__ F, const double *__restrict__ xT,
                                              double *__restrict__ part, int ld, int rp, int kchunk)
{
	const int w = threadIdx.x >> 6;
	const int s = blockIdx.z;
	fx64_tile<RT, CT>(F, xT, part + (size_t)s * ld * rp, ld, rp, (blockIdx.x * 4 + w) * RT, blockIdx.y * CT,
	                  s * kchunk, (s + 1) * kchunk);
}

// The same FP64 product for at most 4 right-hand sides on the vector FP64
// units, where a 16-column MFMA would spend 13 of its 16 columns on padding:
// wave w of block b owns the kFxRows rows from kFxRows (4 b + w) over the whole K; per
// step each lane takes 4 consecutive k of each row (one 16-byte load per
// row, the next step's loads in flight) against x staged through LDS in
// chunks of kFxChunk k (1024 or 256; 4 columns, FP64), then a butterfly
// sums the lanes.
// Writes part[row][c] (S = 1), c < 4; columns >= nrhs of xT are zero.
constexpr int kFxRows = 8;
template <int kFxChunk>
__global__ __launch_bounds__(kB) void k_fx64v(const float *__restrict__ F, const double *__restrict__ xT,
                                              double *__restrict__ part, int ld, int rp)
{
	typedef float fv4 __attribute__((ext_vector_type(4)));
	__shared__ double xs[kFxChunk * 4];
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	const int r0 = (blockIdx.x * 4 + w) * kFxRows;
	double acc[kFxRows][4];
#pragma unroll
	for (int r = 0; r < kFxRows; r++)
#pragma unroll
		for (int c = 0; c < 4; c++) acc[r][c] = 0.0;
	const float *Fr = F + (size_t)r0 * ld + 4 * lane;
	fv4 a[kFxRows], an[kFxRows];
#pragma unroll
	for (int r = 0; r < kFxRows; r++) a[r] = __builtin_nontemporal_load((const fv4 *)(Fr + (size_t)r * ld));
	for (int k0 = 0; k0 < ld; k0 += kFxChunk) {
		__syncthreads();  // the previous chunk's reads are done
		for (int i = threadIdx.x; i < kFxChunk * 4; i += kB) {
			const int c = i / kFxChunk, kk = i % kFxChunk;
			xs[kk * 4 + c] = xT[(size_t)c * ld + k0 + kk];
		}
		__syncthreads();
#pragma unroll 1
		for (int st = 0; st < kFxChunk; st += 256) {
			const int k = k0 + st;
			if (k + 256 < ld) {
#pragma unroll
				for (int r = 0; r < kFxRows; r++)
					an[r] = __builtin_nontemporal_load((const fv4 *)(Fr + (size_t)r * ld + k + 256));
			}
			const double *xl = xs + (st + 4 * lane) * 4;
#pragma unroll
			for (int j = 0; j < 4; j++) {
				const double x0 = xl[4 * j], x1 = xl[4 * j + 1], x2 = xl[4 * j + 2], x3 = xl[4 * j + 3];
#pragma unroll
				for (int r = 0; r < kFxRows; r++) {
					const double f = (double)a[r][j];
					acc[r][0] = __builtin_fma(f, x0, acc[r][0]);
					acc[r][1] = __builtin_fma(f, x1, acc[r][1]);
					acc[r][2] = __builtin_fma(f, x2, acc[r][2]);
					acc[r][3] = __builtin_fma(f, x3, acc[r][3]);
				}
			}
#pragma unroll
			for (int r = 0; r < kFxRows; r++) a[r] = an[r];
		}
	}
#pragma unroll
	for (int r = 0; r < kFxRows; r++)
#pragma unroll
		for (int c = 0; c < 4; c++) {
			double v = acc[r][c];
			for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
			acc[r][c] = v;
		}
	if (lane < kFxRows * 4) {
		const int r = lane >> 2, c = lane & 3;
		double v = 0.0;
#pragma unroll
		for (int q = 0; q < kFxRows * 4; q++)
			if (q == lane) v = acc[q >> 2][q & 3];
		part[(size_t)(r0 + r) * rp + c] = v;
	}
}

// out[i][c] = base[i][c] + rho[i][c] * y - beta * xT[c][i], y = the split-K
// sum of part (FP64, fixed order); base, rho may be null (0 and 1).  Per
// block: sum of squares and max |out|.
__global__ __launch_bounds__(kB) void k_resid(const double *part, int S, int ld, int rp, int n, int nrhs,
                                              const double *xT, const double *base, const double *rho, double beta,
                                              double *out, double *blk)
{
	const size_t total = (size_t)n * nrhs;
	const size_t t = blockIdx.x * (size_t)kB + threadIdx.x;
	double ss = 0.0, mx = 0.0;
	if (t < total) {
		const int i = (int)(t / nrhs), c = (int)(t % nrhs);
		double y = 0.0;
		for (int q = 0; q < S; q++) y += part[((size_t)q * ld + i) * rp + c];
		const double v = (base ? base[t] : 0.0) + (rho ? rho[t] : 1.0) * y - beta * xT[(size_t)c * ld + i];
		out[t] = v;
		ss = v * v;
		mx = fabs(v);
	}
	for (int off = 32; off > 0; off >>= 1) {
		ss += __shfl_xor(ss, off);
		mx = fmax(mx, __shfl_xor(mx, off));
	}
	__shared__ double r[8];
	if ((threadIdx.x & 63) == 0) {
		r[2 * (threadIdx.x >> 6)] = ss;
		r[2 * (threadIdx.x >> 6) + 1] = mx;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		blk[2 * blockIdx.x] = r[0] + r[2] + r[4] + r[6];
		blk[2 * blockIdx.x + 1] = fmax(fmax(r[1], r[3]), fmax(r[5], r[7]));
	}
}

// Transposed product split over row chunks, for nrhs <= 4:
//   part[s][j][c] = sum over rows i of chunk s of F[i][j] * rho[i][c] * x[i][c]
template <typename T>
__global__ __launch_bounds__(kB) void k_ftx(const float *__restrict__ F, int ld, int n, int nrhs, const T *__restrict__ x,
                                            const T *__restrict__ rho, int rows_per_split, T *part)
{
	// four consecutive columns per lane (one 16-byte load per row; ld is a
	// multiple of 256, so the load stays in the zero-padded row)
	const int j4 = (blockIdx.x * kB + threadIdx.x) * 4;
	const int s = blockIdx.y, i0 = s * rows_per_split, i1 = min(n, i0 + rows_per_split);
	if (j4 >= n) return;
	T acc[4][4];
#pragma unroll
	for (int q = 0; q < 4; q++)
#pragma unroll
		for (int c = 0; c < 4; c++) acc[q][c] = 0;
	constexpr int kRowsUnrolled = sizeof(T) == 8 ? 4 : 1;  // measured: helps the FP64 form only
#pragma unroll kRowsUnrolled
	for (int i = i0; i < i1; i++) {
		const float4 f = *(const float4 *)(F + (size_t)i * ld + j4);
		const float fq[4] = {f.x, f.y, f.z, f.w};
		T w[4];
#pragma unroll
		for (int c = 0; c < 4; c++) w[c] = c < nrhs ? rho[(size_t)i * nrhs + c] * x[(size_t)i * nrhs + c] : (T)0;
#pragma unroll
		for (int q = 0; q < 4; q++)
#pragma unroll
			for (int c = 0; c < 4; c++)
				if (c < nrhs) acc[q][c] += (T)fq[q] * w[c];
	}
#pragma unroll
	for (int q = 0; q < 4; q++)
		if (j4 + q < n)
			for (int c = 0; c < nrhs; c++) part[((size_t)s * n + j4 + q) * nrhs + c] = acc[q][c];
}

// out = base + sum_s part - beta * x; per block sum of squares and max |out|
template <typename T>
__global__ __launch_bounds__(kB) void k_ftx_finish(const T *part, int S, int n, int nrhs, const T *base, const T *x,
                                                   T beta, T *out, double *blk)
{
	const size_t total = (size_t)n * nrhs;
	const size_t t = blockIdx.x * (size_t)kB + threadIdx.x;
	double ss = 0.0, mx = 0.0;
	if (t < total) {
		T y = 0;
		for (int s = 0; s < S; s++) y += part[(size_t)s * total + t];
		const T v = base[t] + y - beta * x[t];
		out[t] = v;
		ss = (double)v * (double)v;
		mx = fabs((double)v);
	}
	for (int off = 32; off > 0; off >>= 1) {
		ss += __shfl_xor(ss, off);
		mx = fmax(mx, __shfl_xor(mx, off));
	}
	__shared__ double r[8];
	if ((threadIdx.x & 63) == 0) {
		r[2 * (threadIdx.x >> 6)] = ss;
		r[2 * (threadIdx.x >> 6) + 1] = mx;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		blk[2 * blockIdx.x] = r[0] + r[2] + r[4] + r[6];
		blk[2 * blockIdx.x + 1] = fmax(fmax(r[1], r[3]), fmax(r[5], r[7]));
	}
}

// r32 = r64 / s; XT (if given) <- r32 in fragment order
template <typename E>
__global__ void k_scale(const double *r64, double inv_s, size_t total, int nrhs, int ld, float *r32, E *XT,
                        const float *rho_in = nullptr)
{
	for (size_t t = blockIdx.x * (size_t)kB + threadIdx.x; t < total; t += (size_t)gridDim.x * kB) {
		const float v = (float)(r64[t] * inv_s);
		r32[t] = v;
		if (XT) XT[frag_index((int)(t % nrhs), (int)(t / nrhs), ld)] = (E)(rho_in ? rho_in[t] * v : v);
	}
}

// xT[c][i] += s * d[i][c]
__global__ void k_axpy_t(double *xT, int ld, const float *d, double s, int n, int nrhs)
{
	const size_t total = (size_t)n * nrhs;
	for (size_t t = blockIdx.x * (size_t)kB + threadIdx.x; t < total; t += (size_t)gridDim.x * kB)
		xT[(size_t)(t % nrhs) * ld + t / nrhs] += s * (double)d[t];
}
__global__ void k_axpy(double *x, const float *d, double s, size_t total)
{
	for (size_t t = blockIdx.x * (size_t)kB + threadIdx.x; t < total; t += (size_t)gridDim.x * kB) x[t] += s * (double)d[t];
}

// per block: sum (x - ref)^2 and sum ref^2 (x transposed [c][ld] or row-major)
__global__ void k_err(const double *x, int transposed, int ld, const double *ref, int n, int nrhs, double *blk)
{
	const size_t total = (size_t)n * nrhs;
	double a = 0.0, b = 0.0;
	for (size_t t = blockIdx.x * (size_t)kB + threadIdx.x; t < total; t += (size_t)gridDim.x * kB) {
		const double xv = transposed ? x[(size_t)(t % nrhs) * ld + t / nrhs] : x[t];
		const double e = xv - ref[t];
		a += e * e;
		b += ref[t] * ref[t];
	}
	for (int off = 32; off > 0; off >>= 1) {
		a += __shfl_xor(a, off);
		b += __shfl_xor(b, off);
	}
	__shared__ double r[8];
	if ((threadIdx.x & 63) == 0) {
		r[2 * (threadIdx.x >> 6)] = a;
		r[2 * (threadIdx.x >> 6) + 1] = b;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		blk[2 * blockIdx.x] = r[0] + r[2] + r[4] + r[6];
		blk[2 * blockIdx.x + 1] = r[1] + r[3] + r[5] + r[7];
	}
}

// ---- plain progressive shooting (FP32) --------------------------------------
// One launch per shot, no atomics and no device-wide fences.  Launch k reads
// the per-block records of launch k-1 (each: the block's largest unshot-power
// key, area_i * sum_c dB_i[c] with ties to the lowest index, and that patch's
// dB row), reduces them to the source j and its dB_j, adds
// rho_i F_ij dB_j to B_i and dB_i for every patch i (the source's own dB
// restarting from 0), and writes its own records for launch k+1.  Records are
// double-buffered by shot parity.  A zero best key means nothing is left to
// shoot: the launch only passes the zero on.
constexpr int kMaxShootRhs = 16;
struct ShotRec {
	unsigned long long key;
	float dB[kMaxShootRhs];
	float pad[2];
};

__device__ __forceinline__ unsigned long long power_key(float p, int j)
{
	// p > 0 here, so its bits order as the value
	return ((unsigned long long)__float_as_uint(p) << 32) | (0xffffffffu - (uint32_t)j);
}

__device__ __forceinline__ unsigned long long wave_max(unsigned long long key)
{
	for (int off = 32; off > 0; off >>= 1) {
		const unsigned long long o = __shfl_xor(key, off);
		key = o > key ? o : key;
	}
	return key;
}

// block maximum of key; the thread holding it writes its dB row to rec
__device__ __forceinline__ void block_record(unsigned long long key, const float *row, int nrhs, ShotRec *rec)
{
	__shared__ unsigned long long wk[4];
	const unsigned long long wm = wave_max(key);
	if ((threadIdx.x & 63) == 0) wk[threadIdx.x >> 6] = wm;
	__syncthreads();
	unsigned long long m = wk[0];
	for (int w = 1; w < 4; w++) m = wk[w] > m ? wk[w] : m;
	if (threadIdx.x == 0) rec->key = m;
	if (m != 0 && key == m)  // keys are unique (the index is in them)
		for (int c = 0; c < nrhs; c++) rec->dB[c] = row[c];
}

// records of the initial dB (launch 0's input)
__global__ __launch_bounds__(kB) void k_lsel(const float *dB, const float *area, int n, int nrhs, ShotRec *rout)
{
	const int i = blockIdx.x * kB + threadIdx.x;
	unsigned long long key = 0;
	float row[kMaxShootRhs];
	if (i < n) {
		float p = 0.f;
		for (int c = 0; c < nrhs; c++) {
			row[c] = dB[(size_t)i * nrhs + c];
			p += row[c];
		}
		p *= area[i];
		if (p > 0.f) key = power_key(p, i);
	}
	block_record(key, row, nrhs, rout + blockIdx.x);
}

__global__ __launch_bounds__(kB) void k_lshoot(const float *__restrict__ Ft, int ld, const float *rho, const float *area,
                                               float *B, float *dB, int n, int nrhs, const ShotRec *rin,
                                               ShotRec *rout, int *shots)
{
	__shared__ unsigned long long wk[4];
	__shared__ int sb;
	__shared__ float src[kMaxShootRhs];
	// the previous launch's records -> source
	unsigned long long k = 0;
	for (int b = threadIdx.x; b < (int)gridDim.x; b += kB) {
		const unsigned long long o = rin[b].key;
		k = o > k ? o : k;
	}
	const unsigned long long wm = wave_max(k);
	if ((threadIdx.x & 63) == 0) wk[threadIdx.x >> 6] = wm;
	__syncthreads();
	unsigned long long best = wk[0];
	for (int w = 1; w < 4; w++) best = wk[w] > best ? wk[w] : best;
	if (best == 0) {  // nothing left to shoot (uniform over the grid)
		if (threadIdx.x == 0) rout[blockIdx.x].key = 0;
		return;
	}
	for (int b = threadIdx.x; b < (int)gridDim.x; b += kB)
		if (rin[b].key == best) sb = b;
	__syncthreads();
	if (threadIdx.x < nrhs) src[threadIdx.x] = rin[sb].dB[threadIdx.x];
	__syncthreads();
	const int j = (int)(0xffffffffu - (uint32_t)best);
	const int i = blockIdx.x * kB + threadIdx.x;
	unsigned long long key = 0;
	float row[kMaxShootRhs];
	if (i < n) {
		const float f = Ft[(size_t)j * ld + i];
		float p = 0.f;
		for (int c = 0; c < nrhs; c++) {
			const size_t t = (size_t)i * nrhs + c;
			const float v = rho[t] * f * src[c];
			B[t] += v;
			const float nd = (i == j ? 0.f : dB[t]) + v;
			dB[t] = nd;
			row[c] = nd;
			p += nd;
		}
		p *= area[i];
		if (p > 0.f) key = power_key(p, i);
	}
	block_record(key, row, nrhs, rout + blockIdx.x);
	if (blockIdx.x == 0 && threadIdx.x == 0) *shots += 1;  // launches on one stream: no race
}

__global__ void k_transpose(const float *F, float *Ft, int ld)
{
	__shared__ float tile[32][33];
	const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
	const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 32 x 8
	for (int r = ty; r < 32; r += 8) tile[r][tx] = F[(size_t)(by + r) * ld + bx + tx];
	__syncthreads();
	for (int r = ty; r < 32; r += 8) Ft[(size_t)(bx + r) * ld + by + tx] = tile[tx][r];
}

const int kCT[] = {1, 2, 4, 8, 12};
int pick_ct(int nrhs)
{
	const int need = (nrhs + 15) / 16;
	for (int c : kCT)
		if (c >= need) return c;
	return -1;
}

template <typename E>
int launch_gemm(int rt, int ct, const E *A, const E *XT, float *part, int ld, int rp, int S, hipStream_t st)
{
	const dim3 grid(ld / (64 * rt), S), blk(kB);
	const int kchunk = ld / S;
#define G(R, C) hipLaunchKernelGGL((k_gemm<E, R, C>), grid, blk, 0, st, A, XT, part, ld, rp, kchunk)
#define GR(R)                 \
	switch (ct) {             \
	case 1: G(R, 1); break;   \
	case 2: G(R, 2); break;   \
	case 4: G(R, 4); break;   \
	case 8: G(R, 8); break;   \
	case 12: G(R, 12); break; \
	default: return -1;       \
	}
	if (rt == 1) {
		GR(1)
	} else if (rt == 2 && ct == 8) {
		hipLaunchKernelGGL((k_gemm_dma<E, 8, 2, 1>), dim3(ld / 256, S), dim3(512), 0, st, A, XT, part, ld, rp, kchunk);
	} else if (rt == 2 && ct == 12) {
		hipLaunchKernelGGL((k_gemm_dma<E, 12, 2, 1>), dim3(ld / 256, S), dim3(512), 0, st, A, XT, part, ld, rp, kchunk);
	} else if (rt == 2) {
		GR(2)
	} else if (rt == 4 && ct <= 4) {
		switch (ct) {
		case 1: G(4, 1); break;
		case 2: G(4, 2); break;
		case 4: G(4, 4); break;
		default: return -1;
		}
	} else if (rt == 4 && ct == 8) {
		hipLaunchKernelGGL((k_gemm_dma<E, 8, 4, 2>), dim3(ld / 256, S), dim3(512), 0, st, A, XT, part, ld, rp, kchunk);
	} else if (rt == 4 && ct == 12) {
		hipLaunchKernelGGL((k_gemm_dma<E, 12, 4, 2>), dim3(ld / 256, S), dim3(512), 0, st, A, XT, part, ld, rp, kchunk);
	} else {
		return -1;
	}
#undef GR
#undef G
	return 0;
}

bool valid_tile(int rt, int ct, int S, int ld)
{
	// k-blocks per pipelined loop trip: 2 * GemmUnroll<ct>::U, 2 * U of the
	// four-wave LDS kernel (8 or 12 column tiles, rt = 4: U = 4 or 2), or 1
	// for the eight-wave LDS DMA kernel (8 or 12 column tiles, rt = 2);
	// rows per block: 64 rt, or 256 for the eight-wave kernel
	const bool dma = ct >= 8 && (rt == 2 || rt == 4);
	const int u2 = dma ? 1 : 2 * (ct <= 4 ? 4 : 2);
	const int rows = dma ? 256 : 64 * rt;
	return ld % rows == 0 && ld % (32 * S) == 0 && (ld / 32 / S) % u2 == 0 && ld / 32 / S >= 1;
}

double power_sum(const std::vector<float> &a, const double *v, int n, int nrhs)
{
	double p = 0.0;
	for (int i = 0; i < n; i++)
		for (int c = 0; c < nrhs; c++) p += (double)a[i] * v[(size_t)i * nrhs + c];
	return p;
}

}  // namespace

// FP64 residual machinery: the caller's xT holds rp x ld doubles with rows
// >= nrhs and columns >= n zero; split-K partials; per-block stats.
struct Fx64 {
	int n = 0, ld = 0, nrhs = 0, rp = 0, rt = 4, ct = 1, S = 1;
	Buf<double> part, blk;
	size_t nblk = 0;
	std::vector<double> hb;

	int setup(int n_, int ld_, int nrhs_, std::string &err)
	{
		n = n_;
		ld = ld_;
		nrhs = nrhs_;
		const int nct = (nrhs + 15) / 16;
		ct = nct >= 4 ? 4 : (nct >= 2 ? 2 : 1);
		rt = ct == 4 ? 2 : 4;
		rp = (nct + ct - 1) / ct * ct * 16;  // whole column groups: every tile the grid covers exists
		const int tiles = (ld / (64 * rt)) * (rp / (16 * ct));
		S = 1;
		if (!vec())
			while (S < 8 && tiles * S < 512 && (ld / (2 * S)) % 32 == 0) S *= 2;
		nblk = ((size_t)n * nrhs + kB - 1) / kB;
		if (part.alloc((size_t)S * ld * rp, err) || blk.alloc(std::max<size_t>(nblk, 8192) * 2, err)) return -1;
		hb.assign(nblk * 2, 0.0);
		return 0;
	}
	// at most 4 columns: the vector-unit kernel (ld is a multiple of 256:
	// whole chunks of 1024 or 256 k, whole blocks of 4 kFxRows rows)
	bool vec() const { return nrhs <= 4 && rp >= 4 && ld % 256 == 0 && ld % (4 * kFxRows) == 0; }
	// out = base + rho o (F x) - beta x; returns (sum of squares, max |out|)
	int run(const float *F, const double *xT, const double *base, const double *rho, double beta, double *out,
	        hipStream_t st, double &ss, double &mx, bool want_stats, std::string &err)
	{
		const dim3 grid(ld / (64 * rt), rp / (16 * ct), S);
		const int kchunk = ld / S;
		if (vec() && ld % 1024 == 0)
			hipLaunchKernelGGL(k_fx64v<1024>, dim3(ld / (4 * kFxRows)), dim3(kB), 0, st, F, xT, part.p, ld, rp);
		else if (vec())
			hipLaunchKernelGGL(k_fx64v<256>, dim3(ld / (4 * kFxRows)), dim3(kB), 0, st, F, xT, part.p, ld, rp);
		else if (ct == 4) hipLaunchKernelGGL((k_fx64m<2, 4>), grid, dim3(kB), 0, st, F, xT, part.p, ld, rp, kchunk);
		else if (ct == 2) hipLaunchKernelGGL((k_fx64m<4, 2>), grid, dim3(kB), 0, st, F, xT, part.p, ld, rp, kchunk);
		else hipLaunchKernelGGL((k_fx64m<4, 1>), grid, dim3(kB), 0, st, F, xT, part.p, ld, rp, kchunk);
		hipLaunchKernelGGL(k_resid, dim3((unsigned)nblk), dim3(kB), 0, st, part.p, S, ld, rp, n, nrhs, xT, base, rho,
		                   beta, out, blk.p);
		HIP_OK(hipGetLastError(), "FP64 product");
		ss = mx = 0.0;
		if (!want_stats) return 0;
		HIP_OK(hipMemcpyAsync(hb.data(), blk.p, nblk * 2 * 8, hipMemcpyDeviceToHost, st), "reduce");
		HIP_OK(hipStreamSynchronize(st), "residual");
		for (size_t b = 0; b < nblk; b++) {
			ss += hb[2 * b];
			mx = std::max(mx, hb[2 * b + 1]);
		}
		return 0;
	}
};

struct LinearSystem::Impl {
	hipStream_t st = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	Buf<float> F32, Ft, area;
	Buf<_Float16> F16, FT16;  // FT16: s16 * F^T in fragment order (the adjoint's sweeps)
	Buf<__bf16> FB16;
	Buf<unsigned> umax;
	float s16 = 1.f, sb16 = 1.f;
	bool have16 = false, haveb16 = false, havet = false, havet16 = false;
	std::map<std::pair<int, int>, std::pair<int, int>> tiles;  // (ct, prec) -> (rt, S)

	// low-precision fragment copy of F for prec (made once per matrix)
	int ensure_lowp(int prec, int ld, std::string &err)
	{
		const size_t npp = (size_t)ld * ld;
		if ((prec == PREC_F16 && have16) || (prec == PREC_BF16 && haveb16)) return 0;
		if (prec == PREC_F16) {
			if (umax.alloc(1, err)) return -1;
			HIP_OK(hipMemsetAsync(umax.p, 0, sizeof(unsigned), st), "memset");
			hipLaunchKernelGGL(k_absmax, dim3(4096), dim3(kB), 0, st, F32.p, npp, umax.p);
			unsigned bits = 0;
			HIP_OK(hipMemcpyAsync(&bits, umax.p, sizeof bits, hipMemcpyDeviceToHost, st), "absmax");
			HIP_OK(hipStreamSynchronize(st), "absmax");
			float m;
			std::memcpy(&m, &bits, sizeof m);
			if (!std::isfinite(m)) {
				err = "matrix has non-finite entries";
				return -1;
			}
			int e = 0;
			if (m > 0.f) e = std::max(-60, std::min(60, (int)std::floor(std::log2(16384.0 / m))));
			const float s = std::ldexp(1.0f, e);
			if (F16.alloc(npp, err)) return -1;
			hipLaunchKernelGGL(k_to_frag<_Float16>, dim3(8192), dim3(kB), 0, st, F32.p, ld, s, F16.p);
			s16 = s;
			have16 = true;
		} else if (prec == PREC_BF16) {
			if (FB16.alloc(npp, err)) return -1;
			hipLaunchKernelGGL(k_to_frag<__bf16>, dim3(8192), dim3(kB), 0, st, F32.p, ld, 1.0f, FB16.p);
			sb16 = 1.f;
			haveb16 = true;
		} else {
			err = "precision must be FP16 or BF16";
			return -1;
		}
		HIP_OK(hipGetLastError(), "convert");
		return 0;
	}
	// FP32 transpose of F (made once per matrix)
	int ensure_t(int ld, std::string &err)
	{
		if (havet) return 0;
		if (Ft.alloc((size_t)ld * ld, err)) return -1;
		hipLaunchKernelGGL(k_transpose, dim3(ld / 32, ld / 32), dim3(kB), 0, st, F32.p, Ft.p, ld);
		HIP_OK(hipGetLastError(), "transpose");
		havet = true;
		return 0;
	}
	// s16 * F^T in FP16 fragment order (made once per matrix)
	int ensure_t16(int ld, std::string &err)
	{
		if (havet16) return 0;
		if (ensure_lowp(PREC_F16, ld, err) || ensure_t(ld, err) || FT16.alloc((size_t)ld * ld, err)) return -1;
		hipLaunchKernelGGL(k_to_frag<_Float16>, dim3(8192), dim3(kB), 0, st, Ft.p, ld, s16, FT16.p);
		HIP_OK(hipGetLastError(), "convert");
		havet16 = true;
		return 0;
	}
};

LinearSystem::~LinearSystem()
{
	if (!d_) return;
	if (d_->st) (void)hipStreamSynchronize(d_->st);
	if (d_->e0) (void)hipEventDestroy(d_->e0);
	if (d_->e1) (void)hipEventDestroy(d_->e1);
	hipStream_t s = d_->st;
	delete d_;
	if (s) (void)hipStreamDestroy(s);
}

LinearSystem *LinearSystem::create(int n, std::string &err)
{
	if (n <= 0 || n > (1 << 17)) {
		err = "linear system size must be 1..131072";
		return nullptr;
	}
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
		err = "no HIP device available";
		return nullptr;
	}
	LinearSystem *L = new LinearSystem();
	L->d_ = new Impl();
	L->n_ = n;
	L->np_ = (n + 255) / 256 * 256;
	const size_t np = L->np_;
	if (hipStreamCreateWithFlags(&L->d_->st, hipStreamNonBlocking) != hipSuccess ||
	    hipEventCreate(&L->d_->e0) != hipSuccess || hipEventCreate(&L->d_->e1) != hipSuccess) {
		err = "stream/event creation failed";
		delete L;
		return nullptr;
	}
	if (L->d_->F32.alloc(np * np, err) || hipMemset(L->d_->F32.p, 0, np * np * sizeof(float)) != hipSuccess) {
		if (err.empty()) err = "matrix clear failed";
		delete L;
		return nullptr;
	}
	std::vector<float> ones(n, 1.0f);
	if (L->d_->area.alloc(n, err) ||
	    hipMemcpy(L->d_->area.p, ones.data(), n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
		if (err.empty()) err = "area upload failed";
		delete L;
		return nullptr;
	}
	return L;
}

float *LinearSystem::matrix32()
{
	d_->have16 = d_->haveb16 = d_->havet = d_->havet16 = false;  // the caller may rewrite F
	return d_->F32.p;
}

void *LinearSystem::stream() const { return (void *)d_->st; }

int LinearSystem::set_matrix(const float *F, std::string &err)
{
	HIP_OK(hipMemcpy2D(matrix32(), (size_t)np_ * 4, F, (size_t)n_ * 4, (size_t)n_ * 4, n_, hipMemcpyHostToDevice),
	       "set matrix");
	return 0;
}

int LinearSystem::get_matrix(float *F, std::string &err)
{
	HIP_OK(hipMemcpy2D(F, (size_t)n_ * 4, d_->F32.p, (size_t)np_ * 4, (size_t)n_ * 4, n_, hipMemcpyDeviceToHost),
	       "get matrix");
	return 0;
}

int LinearSystem::set_areas(const double *area, std::string &err)
{
	std::vector<float> a(n_);
	for (int i = 0; i < n_; i++) a[i] = (float)area[i];
	HIP_OK(hipMemcpy(d_->area.p, a.data(), n_ * sizeof(float), hipMemcpyHostToDevice), "set areas");
	return 0;
}

void LinearSystem::tile_config(int nrhs, int prec, int *rt, int *splitk) const
{
	const int ct = pick_ct(nrhs);
	auto it = d_->tiles.find({ct, prec});
	if (it == d_->tiles.end()) {
		*rt = 0;
		*splitk = 0;
	} else {
		*rt = it->second.first;
		*splitk = it->second.second;
	}
}

int LinearSystem::set_tile(int nrhs, int prec, int rt, int splitk, std::string &err)
{
	const int ct = pick_ct(nrhs);
	if (nrhs <= 0 || ct < 0 || (prec != PREC_F16 && prec != PREC_BF16) || (rt != 1 && rt != 2 && rt != 4) ||
	    (splitk != 1 && splitk != 2 && splitk != 4 && splitk != 8) || !valid_tile(rt, ct, splitk, np_)) {
		err = "set_tile: no kernel for this tile at this size";
		return -1;
	}
	d_->tiles[{ct, prec}] = {rt, splitk};
	return 0;
}

// Times one sweep (GEMM + Jacobi epilogue) for every valid (rows per wave,
// split-K) and keeps the fastest for (nrhs column tiles, precision).
int LinearSystem::autotune(int nrhs, int prec, std::string &err)
{
	Impl &d = *d_;
	const int ct = pick_ct(nrhs);
	if (nrhs <= 0 || ct < 0 || (prec != PREC_F16 && prec != PREC_BF16)) {
		err = "autotune: nrhs must be 1..192 and precision FP16 or BF16";
		return -1;
	}
	const int rp = ct * 16, ld = np_, n = n_;
	if (d.ensure_lowp(prec, ld, err)) return -1;
	const size_t total = (size_t)n * nrhs;
	Buf<_Float16> xt16;
	Buf<__bf16> xtb;
	Buf<float> part, r32, d32, rho;
	if (part.alloc((size_t)8 * ld * rp, err) || r32.alloc(total, err) || d32.alloc(total, err) || rho.alloc(total, err))
		return -1;
	const int g = grid_for(total);
	const size_t xts = (size_t)rp * ld;
	hipLaunchKernelGGL(k_fill<float>, dim3(g), dim3(kB), 0, d.st, r32.p, total, 1.0f);
	hipLaunchKernelGGL(k_fill<float>, dim3(g), dim3(kB), 0, d.st, rho.p, total, 0.5f);
	if (prec == PREC_F16) {
		if (xt16.alloc(xts, err)) return -1;
		hipLaunchKernelGGL(k_fill<_Float16>, dim3(grid_for(xts)), dim3(kB), 0, d.st, xt16.p, xts, (_Float16)0.f);
	} else {
		if (xtb.alloc(xts, err)) return -1;
		hipLaunchKernelGGL(k_fill<__bf16>, dim3(grid_for(xts)), dim3(kB), 0, d.st, xtb.p, xts, (__bf16)0.f);
	}
	auto sweep = [&](int rt, int S) -> int {
		if (prec == PREC_F16) {
			if (launch_gemm<_Float16>(rt, ct, d.F16.p, xt16.p, part.p, ld, rp, S, d.st)) return -1;
			hipLaunchKernelGGL(k_jacobi<_Float16>, dim3(g), dim3(kB), 0, d.st, part.p, S, ld, rp, n, nrhs, 1.0f, r32.p,
			                   rho.p, d32.p, xt16.p);
		} else {
			if (launch_gemm<__bf16>(rt, ct, d.FB16.p, xtb.p, part.p, ld, rp, S, d.st)) return -1;
			hipLaunchKernelGGL(k_jacobi<__bf16>, dim3(g), dim3(kB), 0, d.st, part.p, S, ld, rp, n, nrhs, 1.0f, r32.p,
			                   rho.p, d32.p, xtb.p);
		}
		return 0;
	};
	double best = 1e30;
	int brt = 0, bs = 0;
	for (int rt : {1, 2, 4}) {
		for (int S : {1, 2, 4, 8}) {
			if (!valid_tile(rt, ct, S, ld)) continue;
			if (sweep(rt, S)) continue;  // warm-up
			HIP_OK(hipEventRecord(d.e0, d.st), "event");
			for (int k = 0; k < 3; k++) sweep(rt, S);
			HIP_OK(hipEventRecord(d.e1, d.st), "event");
			HIP_OK(hipEventSynchronize(d.e1), "event");
			float ms = 0;
			HIP_OK(hipEventElapsedTime(&ms, d.e0, d.e1), "event");
			if (ms < best) {
				best = ms;
				brt = rt;
				bs = S;
			}
		}
	}
	HIP_OK(hipGetLastError(), "autotune");
	HIP_OK(hipStreamSynchronize(d.st), "autotune");
	if (!brt) {
		err = "autotune: no valid tile";
		return -1;
	}
	d.tiles[{ct, prec}] = {brt, bs};
	return 0;
}

int LinearSystem::gemm_ms(int nrhs, int prec, int reps, double *ms, std::string &err)
{
	Impl &d = *d_;
	const int ct = pick_ct(nrhs);
	if (nrhs <= 0 || ct < 0 || (prec != PREC_F16 && prec != PREC_BF16) || reps <= 0) {
		err = "gemm_ms: bad arguments";
		return -1;
	}
	if (!d.tiles.count({ct, prec}) && autotune(nrhs, prec, err)) return -1;
	int rt, S;
	tile_config(nrhs, prec, &rt, &S);
	const int rp = ct * 16, ld = np_;
	if (d.ensure_lowp(prec, ld, err)) return -1;
	Buf<_Float16> xt16;
	Buf<__bf16> xtb;
	Buf<float> part;
	const size_t xts = (size_t)rp * ld;
	if (part.alloc((size_t)S * ld * rp, err)) return -1;
	if (prec == PREC_F16) {
		if (xt16.alloc(xts, err)) return -1;
		hipLaunchKernelGGL(k_fill<_Float16>, dim3(grid_for(xts)), dim3(kB), 0, d.st, xt16.p, xts, (_Float16)1.f);
	} else {
		if (xtb.alloc(xts, err)) return -1;
		hipLaunchKernelGGL(k_fill<__bf16>, dim3(grid_for(xts)), dim3(kB), 0, d.st, xtb.p, xts, (__bf16)1.f);
	}
	auto one = [&]() {
		if (prec == PREC_F16) launch_gemm<_Float16>(rt, ct, d.F16.p, xt16.p, part.p, ld, rp, S, d.st);
		else launch_gemm<__bf16>(rt, ct, d.FB16.p, xtb.p, part.p, ld, rp, S, d.st);
	};
	one();
	HIP_OK(hipEventRecord(d.e0, d.st), "event");
	for (int k = 0; k < reps; k++) one();
	HIP_OK(hipEventRecord(d.e1, d.st), "event");
	HIP_OK(hipEventSynchronize(d.e1), "event");
	HIP_OK(hipGetLastError(), "gemm_ms");
	float t = 0;
	HIP_OK(hipEventElapsedTime(&t, d.e0, d.e1), "event");
	*ms = t / reps;
	return 0;
}

int LinearSystem::multiply(int prec, int nrhs, const double *X, double *Y, std::string &err)
{
	Impl &d = *d_;
	const int ct = pick_ct(nrhs);
	if (nrhs <= 0 || ct < 0) {
		err = "multiply: nrhs must be 1..192";
		return -1;
	}
	const int n = n_, ld = np_, rp = ct * 16;
	const size_t total = (size_t)n * nrhs;
	hipStream_t st = d.st;
	if (prec == PREC_F32) {
		Fx64 fx;
		if (fx.setup(n, ld, nrhs, err)) return -1;
		std::vector<double> xt((size_t)fx.rp * ld, 0.0);
		for (int i = 0; i < n; i++)
			for (int c = 0; c < nrhs; c++) xt[(size_t)c * ld + i] = X[(size_t)i * nrhs + c];
		Buf<double> xT, out;
		if (xT.alloc(xt.size(), err) || out.alloc(total, err)) return -1;
		HIP_OK(hipMemcpyAsync(xT.p, xt.data(), xt.size() * 8, hipMemcpyHostToDevice, st), "upload");
		double ss, mx;
		if (fx.run(d.F32.p, xT.p, nullptr, nullptr, 0.0, out.p, st, ss, mx, false, err)) return -1;
		HIP_OK(hipMemcpyAsync(Y, out.p, total * 8, hipMemcpyDeviceToHost, st), "download");
		HIP_OK(hipStreamSynchronize(st), "multiply");
		return 0;
	}
	if (d.ensure_lowp(prec, ld, err)) return -1;
	int rt, S;
	tile_config(nrhs, prec, &rt, &S);
	if (!rt) {
		rt = 1;
		S = 1;
	}
	Buf<float> part, zero, ones, y;
	Buf<double> x64;
	Buf<_Float16> xt16;
	Buf<__bf16> xtb;
	if (part.alloc((size_t)S * ld * rp, err) || zero.alloc(total, err) || ones.alloc(total, err) || y.alloc(total, err) ||
	    x64.alloc(total, err))
		return -1;
	const int g = grid_for(total);
	const size_t xts = (size_t)rp * ld;
	hipLaunchKernelGGL(k_fill<float>, dim3(g), dim3(kB), 0, st, zero.p, total, 0.0f);
	hipLaunchKernelGGL(k_fill<float>, dim3(g), dim3(kB), 0, st, ones.p, total, 1.0f);
	HIP_OK(hipMemcpyAsync(x64.p, X, total * 8, hipMemcpyHostToDevice, st), "upload");
	const float inv = 1.0f / (prec == PREC_F16 ? d.s16 : d.sb16);
	int rc;
	if (prec == PREC_F16) {
		if (xt16.alloc(xts, err)) return -1;
		hipLaunchKernelGGL(k_fill<_Float16>, dim3(grid_for(xts)), dim3(kB), 0, st, xt16.p, xts, (_Float16)0.f);
		hipLaunchKernelGGL(k_scale<_Float16>, dim3(g), dim3(kB), 0, st, x64.p, 1.0, total, nrhs, ld, y.p, xt16.p);
		rc = launch_gemm<_Float16>(rt, ct, d.F16.p, xt16.p, part.p, ld, rp, S, st);
		hipLaunchKernelGGL(k_jacobi<_Float16>, dim3(g), dim3(kB), 0, st, part.p, S, ld, rp, n, nrhs, inv, zero.p, ones.p,
		                   y.p, xt16.p);
	} else {
		if (xtb.alloc(xts, err)) return -1;
		hipLaunchKernelGGL(k_fill<__bf16>, dim3(grid_for(xts)), dim3(kB), 0, st, xtb.p, xts, (__bf16)0.f);
		hipLaunchKernelGGL(k_scale<__bf16>, dim3(g), dim3(kB), 0, st, x64.p, 1.0, total, nrhs, ld, y.p, xtb.p);
		rc = launch_gemm<__bf16>(rt, ct, d.FB16.p, xtb.p, part.p, ld, rp, S, st);
		hipLaunchKernelGGL(k_jacobi<__bf16>, dim3(g), dim3(kB), 0, st, part.p, S, ld, rp, n, nrhs, inv, zero.p, ones.p,
		                   y.p, xtb.p);
	}
	if (rc) {
		err = "multiply: no kernel for this tile";
		return -1;
	}
	HIP_OK(hipGetLastError(), "multiply");
	std::vector<float> hy(total);
	HIP_OK(hipMemcpyAsync(hy.data(), y.p, total * 4, hipMemcpyDeviceToHost, st), "download");
	HIP_OK(hipStreamSynchronize(st), "multiply");
	for (size_t t = 0; t < total; t++) Y[t] = hy[t];
	return 0;
}

int LinearSystem::gather(int prec, int nrhs, const double *rho, const double *E, double tol, int max_outer, int sweeps,
                         const double *Bref, double target, double *B, double *info, std::string &err)
{
	Impl &d = *d_;
	const int ct = pick_ct(nrhs);
	if (nrhs <= 0 || ct < 0) {
		err = "gather: nrhs must be 1..192";
		return -1;
	}
	if (prec != PREC_F16 && prec != PREC_BF16) {
		err = "gather: precision must be FP16 or BF16";
		return -1;
	}
	if (sweeps <= 0 || max_outer < 0) {
		err = "gather: sweeps must be > 0 and max_outer >= 0";
		return -1;
	}
	if (d.ensure_lowp(prec, np_, err)) return -1;
	if (!d.tiles.count({ct, prec}) && autotune(nrhs, prec, err)) return -1;
	int rt, S;
	tile_config(nrhs, prec, &rt, &S);
	const int n = n_, ld = np_, rp = ct * 16;
	const size_t total = (size_t)n * nrhs;
	const float inv = 1.0f / (prec == PREC_F16 ? d.s16 : d.sb16);
	hipStream_t st = d.st;
	Buf<float> rho32, r32, d32, part;
	Buf<double> rho64, E64, xT, r64, blk, Bref64;
	Buf<_Float16> XT16;
	Buf<__bf16> XTb;
	Fx64 fx;
	if (fx.setup(n, ld, nrhs, err)) return -1;
	const size_t xtn = (size_t)fx.rp * ld;
	if (rho32.alloc(total, err) || r32.alloc(total, err) || d32.alloc(total, err) ||
	    part.alloc((size_t)S * ld * rp, err) || rho64.alloc(total, err) || E64.alloc(total, err) ||
	    xT.alloc(xtn, err) || r64.alloc(total, err) || blk.alloc(8192 * 2, err))
		return -1;
	const size_t xts = (size_t)rp * ld;
	if (prec == PREC_F16) {
		if (XT16.alloc(xts, err)) return -1;
		hipLaunchKernelGGL(k_fill<_Float16>, dim3(grid_for(xts)), dim3(kB), 0, st, XT16.p, xts, (_Float16)0.f);
	} else {
		if (XTb.alloc(xts, err)) return -1;
		hipLaunchKernelGGL(k_fill<__bf16>, dim3(grid_for(xts)), dim3(kB), 0, st, XTb.p, xts, (__bf16)0.f);
	}
	std::vector<float> h32(total);
	for (size_t t = 0; t < total; t++) h32[t] = (float)rho[t];
	HIP_OK(hipMemcpyAsync(rho32.p, h32.data(), total * 4, hipMemcpyHostToDevice, st), "upload");
	HIP_OK(hipMemcpyAsync(rho64.p, rho, total * 8, hipMemcpyHostToDevice, st), "upload");
	HIP_OK(hipMemcpyAsync(E64.p, E, total * 8, hipMemcpyHostToDevice, st), "upload");
	hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(xtn)), dim3(kB), 0, st, xT.p, xtn, 0.0);
	if (Bref) {
		if (Bref64.alloc(total, err)) return -1;
		HIP_OK(hipMemcpyAsync(Bref64.p, Bref, total * 8, hipMemcpyHostToDevice, st), "upload");
	}
	HIP_OK(hipStreamSynchronize(st), "upload");
	std::vector<double> hb(8192 * 2);
	const int g = grid_for(total);
	auto err_vs_ref = [&](double &e) -> int {
		hipLaunchKernelGGL(k_err, dim3(g), dim3(kB), 0, st, xT.p, 1, ld, Bref64.p, n, nrhs, blk.p);
		HIP_OK(hipMemcpyAsync(hb.data(), blk.p, (size_t)g * 2 * 8, hipMemcpyDeviceToHost, st), "err");
		HIP_OK(hipStreamSynchronize(st), "err");
		double a = 0.0, b = 0.0;
		for (int k = 0; k < g; k++) {
			a += hb[2 * k];
			b += hb[2 * k + 1];
		}
		e = std::sqrt(a / std::max(b, 1e-300));
		return 0;
	};
	const auto t0 = std::chrono::steady_clock::now();
	double enorm = -1.0, rel = 0.0, ms_target = -1.0, errv = -1.0, check_ms = 0.0;
	int outer = 0, total_sweeps = 0;
	// x = 0 at the start, so the first residual is E itself
	double ss0 = 0.0, mx0 = 0.0;
	for (size_t t = 0; t < total; t++) {
		ss0 += E[t] * E[t];
		mx0 = std::max(mx0, std::fabs(E[t]));
	}
	for (;; outer++) {
		double ss = 0.0, mx = 0.0;
		if (outer == 0) {
			HIP_OK(hipMemcpyAsync(r64.p, E64.p, total * 8, hipMemcpyDeviceToDevice, st), "residual");
			ss = ss0;
			mx = mx0;
		} else if (fx.run(d.F32.p, xT.p, E64.p, rho64.p, 1.0, r64.p, st, ss, mx, true, err)) {
			return -1;
		}
		if (!std::isfinite(ss)) {
			err = "gather diverged (non-finite residual): the iteration needs spectral radius of diag(rho) F below 1";
			return -1;
		}
		if (enorm < 0) enorm = std::sqrt(ss);  // x = 0: the first residual is E
		rel = enorm > 0 ? std::sqrt(ss) / enorm : 0.0;
		if (Bref && outer > 0) {
			const auto c0 = std::chrono::steady_clock::now();
			if (err_vs_ref(errv)) return -1;
			check_ms += ms_since(c0);
			if (ms_target < 0 && errv <= target) ms_target = ms_since(t0) - check_ms;
		}
		if (rel <= tol || outer >= max_outer || mx == 0.0) break;
		const double s = mx;
		if (prec == PREC_F16)
			hipLaunchKernelGGL(k_scale<_Float16>, dim3(g), dim3(kB), 0, st, r64.p, 1.0 / s, total, nrhs, ld, r32.p, XT16.p);
		else
			hipLaunchKernelGGL(k_scale<__bf16>, dim3(g), dim3(kB), 0, st, r64.p, 1.0 / s, total, nrhs, ld, r32.p, XTb.p);
		for (int m = 0; m < sweeps; m++) {
			int rc;
			if (prec == PREC_F16) {
				rc = launch_gemm<_Float16>(rt, ct, d.F16.p, XT16.p, part.p, ld, rp, S, st);
				hipLaunchKernelGGL(k_jacobi<_Float16>, dim3(g), dim3(kB), 0, st, part.p, S, ld, rp, n, nrhs, inv, r32.p,
				                   rho32.p, d32.p, XT16.p);
			} else {
				rc = launch_gemm<__bf16>(rt, ct, d.FB16.p, XTb.p, part.p, ld, rp, S, st);
				hipLaunchKernelGGL(k_jacobi<__bf16>, dim3(g), dim3(kB), 0, st, part.p, S, ld, rp, n, nrhs, inv, r32.p,
				                   rho32.p, d32.p, XTb.p);
			}
			if (rc) {
				err = "gather: no kernel for this tile";
				return -1;
			}
			total_sweeps++;
		}
		hipLaunchKernelGGL(k_axpy_t, dim3(g), dim3(kB), 0, st, xT.p, ld, d32.p, s, n, nrhs);
		HIP_OK(hipGetLastError(), "sweep launch");
	}
	const double ms = ms_since(t0) - check_ms;
	std::vector<double> hx((size_t)nrhs * ld);
	HIP_OK(hipMemcpy(hx.data(), xT.p, hx.size() * 8, hipMemcpyDeviceToHost), "download");  // rows < nrhs
	for (int i = 0; i < n; i++)
		for (int c = 0; c < nrhs; c++) B[(size_t)i * nrhs + c] = hx[(size_t)c * ld + i];
	if (info) {
		std::vector<float> a(n);
		HIP_OK(hipMemcpy(a.data(), d.area.p, n * 4, hipMemcpyDeviceToHost), "area");
		info[INFO_OUTER] = outer;
		info[INFO_SWEEPS] = total_sweeps;
		info[INFO_REL_RES] = rel;
		info[INFO_MS] = ms;
		info[INFO_MS_TARGET] = ms_target;
		info[INFO_ERR] = errv;
		info[INFO_POWER_E] = power_sum(a, E, n, nrhs);
		info[INFO_POWER_B] = power_sum(a, B, n, nrhs);
	}
	return 0;
}

int LinearSystem::shoot(int nrhs, const double *rho, const double *E, const double *Bref, double target, int max_shots,
                        double *B, double *info, std::string &err)
{
	Impl &d = *d_;
	if (nrhs <= 0 || nrhs > kMaxShootRhs) {
		err = "shoot: nrhs must be 1..16";
		return -1;
	}
	if (max_shots < 0) {
		err = "shoot: max_shots must be >= 0";
		return -1;
	}
	const int n = n_, ld = np_;
	hipStream_t st = d.st;
	static_assert(sizeof(ShotRec) % 8 == 0, "record alignment");
	if (d.ensure_t(ld, err)) return -1;
	const size_t total = (size_t)n * nrhs;
	const int nb = (n + kB - 1) / kB;
	Buf<float> rho32, B32, dB32;
	Buf<ShotRec> rec;
	Buf<int> shots_d;
	if (rho32.alloc(total, err) || B32.alloc(total, err) || dB32.alloc(total, err) || rec.alloc(2 * (size_t)nb, err) ||
	    shots_d.alloc(1, err))
		return -1;
	std::vector<float> h(total);
	for (size_t t = 0; t < total; t++) h[t] = (float)rho[t];
	HIP_OK(hipMemcpyAsync(rho32.p, h.data(), total * 4, hipMemcpyHostToDevice, st), "upload");
	std::vector<float> e32(total);
	for (size_t t = 0; t < total; t++) e32[t] = (float)E[t];
	HIP_OK(hipMemcpyAsync(B32.p, e32.data(), total * 4, hipMemcpyHostToDevice, st), "upload");
	HIP_OK(hipMemcpyAsync(dB32.p, e32.data(), total * 4, hipMemcpyHostToDevice, st), "upload");
	hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(kB), 0, st, shots_d.p, (size_t)1, 0);
	HIP_OK(hipStreamSynchronize(st), "upload");
	std::vector<float> hb(total);
	auto err_vs_ref = [&](double &e) -> int {
		HIP_OK(hipMemcpyAsync(hb.data(), B32.p, total * 4, hipMemcpyDeviceToHost, st), "err");
		HIP_OK(hipStreamSynchronize(st), "err");
		double a = 0.0, b = 0.0;
		for (size_t t = 0; t < total; t++) {
			const double x = (double)hb[t] - Bref[t];
			a += x * x;
			b += Bref[t] * Bref[t];
		}
		e = std::sqrt(a / std::max(b, 1e-300));
		return 0;
	};
	// shot k reads the records of shot k - 1 (parity k & 1) and writes the
	// other set
	auto launch_shot = [&](int k) {
		hipLaunchKernelGGL(k_lshoot, dim3(nb), dim3(kB), 0, st, d.Ft.p, ld, rho32.p, d.area.p, B32.p, dB32.p, n, nrhs,
		                   rec.p + (size_t)(k & 1) * nb, rec.p + (size_t)((k + 1) & 1) * nb, shots_d.p);
	};
	// A shot needs the previous one's argmax, so plain shooting is bound by
	// launch latency: runs of kGraphShots shots (an even number, so parities
	// repeat) are captured once into a HIP graph and replayed.
	constexpr int kGraphShots = 64;
	struct GraphGuard {
		hipGraph_t g = nullptr;
		hipGraphExec_t x = nullptr;
		~GraphGuard()
		{
			if (x) (void)hipGraphExecDestroy(x);
			if (g) (void)hipGraphDestroy(g);
		}
	} run;
	HIP_OK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "shot capture");
	for (int k = 0; k < kGraphShots; k++) launch_shot(k);
	HIP_OK(hipStreamEndCapture(st, &run.g), "shot capture");
	HIP_OK(hipGraphInstantiate(&run.x, run.g, nullptr, nullptr, 0), "shot graph");
	const auto t0 = std::chrono::steady_clock::now();
	hipLaunchKernelGGL(k_lsel, dim3(nb), dim3(kB), 0, st, dB32.p, d.area.p, n, nrhs, rec.p);
	double check_ms = 0.0, ms_target = -1.0, errv = -1.0;
	int launched = 0, chunk = 64, done_shots = 0;
	for (;;) {
		const int b = std::min(chunk, max_shots - launched);
		if (b <= 0) break;
		int k = 0;
		if ((launched & 1) == 0)
			for (; k + kGraphShots <= b; k += kGraphShots) HIP_OK(hipGraphLaunch(run.x, st), "shot graph launch");
		for (; k < b; k++) launch_shot(launched + k);
		launched += b;
		HIP_OK(hipGetLastError(), "shoot launch");
		HIP_OK(hipMemcpyAsync(&done_shots, shots_d.p, sizeof(int), hipMemcpyDeviceToHost, st), "state");
		HIP_OK(hipStreamSynchronize(st), "shoot");
		if (Bref) {
			const auto c0 = std::chrono::steady_clock::now();
			if (err_vs_ref(errv)) return -1;
			check_ms += ms_since(c0);
			if (errv <= target) {
				ms_target = ms_since(t0) - check_ms;
				break;
			}
		}
		if (done_shots < launched) break;  // a launch found nothing left to shoot
		chunk = std::min(chunk * 2, 4096);
	}
	const double ms = ms_since(t0) - check_ms;
	HIP_OK(hipMemcpy(hb.data(), B32.p, total * 4, hipMemcpyDeviceToHost), "download");
	for (size_t t = 0; t < total; t++) B[t] = hb[t];
	if (Bref && errv < 0 && err_vs_ref(errv)) return -1;
	if (info) {
		std::vector<float> a(n);
		HIP_OK(hipMemcpy(a.data(), d.area.p, n * 4, hipMemcpyDeviceToHost), "area");
		info[INFO_OUTER] = done_shots;
		info[INFO_SWEEPS] = 0;
		info[INFO_REL_RES] = -1;
		info[INFO_MS] = ms;
		info[INFO_MS_TARGET] = ms_target;
		info[INFO_ERR] = errv;
		info[INFO_POWER_E] = power_sum(a, E, n, nrhs);
		info[INFO_POWER_B] = power_sum(a, B, n, nrhs);
	}
	return 0;
}

int LinearSystem::adjoint(int nrhs, const double *rho, const double *B, const double *g, double tol, int max_outer,
                          int sweeps, double *lambda, double *drho, double *info, std::string &err)
{
	Impl &d = *d_;
	if (nrhs <= 0 || nrhs > 4) {
		err = "adjoint: nrhs must be 1..4";
		return -1;
	}
	if (sweeps <= 0 || max_outer < 0) {
		err = "adjoint: sweeps must be > 0 and max_outer >= 0";
		return -1;
	}
	const int n = n_, ld = np_;
	const size_t total = (size_t)n * nrhs;
	hipStream_t st = d.st;
	// the sweeps multiply by F^T in FP16 on MFMA (the gather's kernels on
	// the transposed fragment copy); the residual stays FP64 over FP32 F
	const int ct = pick_ct(nrhs), rp = ct * 16;
	if (d.ensure_t16(ld, err)) return -1;
	if (!d.tiles.count({ct, PREC_F16}) && autotune(nrhs, PREC_F16, err)) return -1;
	int rt, Sg;
	tile_config(nrhs, PREC_F16, &rt, &Sg);
	const float inv = 1.0f / d.s16;
	const int S = std::max(1, std::min(64, n / 256));
	const int rows_per_split = (n + S - 1) / S;
	const int fin_blocks = (int)((total + kB - 1) / kB);
	Buf<double> rho64, g64, x64, r64, part64, blk, B64T, fb;
	Fx64 fx;
	if (fx.setup(n, ld, nrhs, err)) return -1;
	Buf<float> rho32, r32, l32, part16;
	Buf<_Float16> XT16;
	const size_t xts = (size_t)rp * ld;
	if (rho64.alloc(total, err) || g64.alloc(total, err) || x64.alloc(total, err) || r64.alloc(total, err) ||
	    part64.alloc((size_t)S * total, err) || rho32.alloc(total, err) || r32.alloc(total, err) ||
	    l32.alloc(total, err) || part16.alloc((size_t)Sg * ld * rp, err) || XT16.alloc(xts, err) ||
	    blk.alloc((size_t)std::max(fin_blocks, 8192) * 2, err) || B64T.alloc((size_t)fx.rp * ld, err) ||
	    fb.alloc(total, err))
		return -1;
	hipLaunchKernelGGL(k_fill<_Float16>, dim3(grid_for(xts)), dim3(kB), 0, st, XT16.p, xts, (_Float16)0.f);
	std::vector<float> h(total);
	for (size_t t = 0; t < total; t++) h[t] = (float)rho[t];
	HIP_OK(hipMemcpyAsync(rho32.p, h.data(), total * 4, hipMemcpyHostToDevice, st), "upload");
	HIP_OK(hipMemcpyAsync(rho64.p, rho, total * 8, hipMemcpyHostToDevice, st), "upload");
	HIP_OK(hipMemcpyAsync(g64.p, g, total * 8, hipMemcpyHostToDevice, st), "upload");
	hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(total)), dim3(kB), 0, st, x64.p, total, 0.0);
	const dim3 tgrid(((n + 3) / 4 + kB - 1) / kB, S);  // four columns per lane
	std::vector<double> hb((size_t)fin_blocks * 2);
	HIP_OK(hipStreamSynchronize(st), "upload");
	const auto t0 = std::chrono::steady_clock::now();
	double enorm = -1.0, rel = 0.0;
	int outer = 0, total_sweeps = 0;
	double ss0 = 0.0, mx0 = 0.0;  // lambda = 0 at the start: the first residual is g
	for (size_t t = 0; t < total; t++) {
		ss0 += g[t] * g[t];
		mx0 = std::max(mx0, std::fabs(g[t]));
	}
	for (;; outer++) {
		double ss = 0.0, mx = 0.0;
		if (outer == 0) {
			HIP_OK(hipMemcpyAsync(r64.p, g64.p, total * 8, hipMemcpyDeviceToDevice, st), "residual");
			ss = ss0;
			mx = mx0;
		} else {
			// r = g + F^T (rho o x) - x in FP64
			hipLaunchKernelGGL(k_ftx<double>, tgrid, dim3(kB), 0, st, d.F32.p, ld, n, nrhs, x64.p, rho64.p,
			                   rows_per_split, part64.p);
			hipLaunchKernelGGL(k_ftx_finish<double>, dim3(fin_blocks), dim3(kB), 0, st, part64.p, S, n, nrhs, g64.p,
			                   x64.p, 1.0, r64.p, blk.p);
			HIP_OK(hipGetLastError(), "adjoint residual");
			HIP_OK(hipMemcpyAsync(hb.data(), blk.p, hb.size() * 8, hipMemcpyDeviceToHost, st), "reduce");
			HIP_OK(hipStreamSynchronize(st), "reduce");
			for (int b = 0; b < fin_blocks; b++) {
				ss += hb[2 * b];
				mx = std::max(mx, hb[2 * b + 1]);
			}
		}
		if (!std::isfinite(ss)) {
			err = "adjoint diverged (non-finite residual)";
			return -1;
		}
		if (enorm < 0) enorm = std::sqrt(ss);
		rel = enorm > 0 ? std::sqrt(ss) / enorm : 0.0;
		if (rel <= tol || outer >= max_outer || mx == 0.0) break;
		const int gg = grid_for(total);
		// d <- r (scaled), then sweeps d <- r + F^T (rho o d) with XT = rho o d
		hipLaunchKernelGGL(k_scale<_Float16>, dim3(gg), dim3(kB), 0, st, r64.p, 1.0 / mx, total, nrhs, ld, r32.p,
		                   XT16.p, rho32.p);
		for (int m = 0; m < sweeps; m++) {
			if (launch_gemm<_Float16>(rt, ct, d.FT16.p, XT16.p, part16.p, ld, rp, Sg, st)) {
				err = "adjoint: no kernel for this tile";
				return -1;
			}
			hipLaunchKernelGGL((k_jacobi<_Float16, true>), dim3(gg), dim3(kB), 0, st, part16.p, Sg, ld, rp, n, nrhs,
			                   inv, r32.p, rho32.p, l32.p, XT16.p);
			total_sweeps++;
		}
		hipLaunchKernelGGL(k_axpy, dim3(gg), dim3(kB), 0, st, x64.p, l32.p, mx, total);
		HIP_OK(hipGetLastError(), "adjoint sweep");
	}
	// dL/drho = lambda o (F B), F B in FP64
	std::vector<double> bt((size_t)fx.rp * ld, 0.0);
	for (int i = 0; i < n; i++)
		for (int c = 0; c < nrhs; c++) bt[(size_t)c * ld + i] = B[(size_t)i * nrhs + c];
	HIP_OK(hipMemcpyAsync(B64T.p, bt.data(), bt.size() * 8, hipMemcpyHostToDevice, st), "upload");
	{
		double ss, mx;
		if (fx.run(d.F32.p, B64T.p, nullptr, nullptr, 0.0, fb.p, st, ss, mx, false, err)) return -1;
	}
	std::vector<double> hfb(total);
	HIP_OK(hipMemcpyAsync(lambda, x64.p, total * 8, hipMemcpyDeviceToHost, st), "download");
	HIP_OK(hipMemcpyAsync(hfb.data(), fb.p, total * 8, hipMemcpyDeviceToHost, st), "download");
	HIP_OK(hipStreamSynchronize(st), "adjoint");
	const double ms = ms_since(t0);
	for (size_t t = 0; t < total; t++) drho[t] = lambda[t] * hfb[t];
	if (info) {
		for (int k = 0; k < INFO_N; k++) info[k] = -1;
		info[INFO_OUTER] = outer;
		info[INFO_SWEEPS] = total_sweeps;
		info[INFO_REL_RES] = rel;
		info[INFO_MS] = ms;
	}
	return 0;
}

}  // namespace rgpu
