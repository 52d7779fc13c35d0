// Device building blocks shared by the explicit-matrix solvers (linear.hip)
// and the batched solver (batch.hip): MFMA fragment layout, the
// low-precision GEMM tile, the FP64 MFMA product tile, buffers.
// Included inside an anonymous namespace of each translation unit.
#pragma once

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kB = 256;

__device__ __forceinline__ f4 mfma(h8 a, h8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
__device__ __forceinline__ f4 mfma(b8 a, b8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

template <typename E> struct VecOf;
template <> struct VecOf<_Float16> { typedef h8 type; };
template <> struct VecOf<__bf16> { typedef b8 type; };

// element (row, k) of a fragment-ordered matrix with ld columns
__device__ __forceinline__ size_t frag_index(int row, int k, int ld)
{
	return ((size_t)(row >> 4) * (ld >> 5) + (k >> 5)) * 512 + ((((k & 31) >> 3) << 4) + (row & 15)) * 8 + (k & 7);
}

__global__ void k_absmax(const float *F, size_t count, unsigned *out)
{
	float m = 0.f;
	for (size_t i = blockIdx.x * (size_t)kB + threadIdx.x; i < count; i += (size_t)gridDim.x * kB) m = fmaxf(m, fabsf(F[i]));
	for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
	if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));  // non-negative floats order as their bits
}

// one thread per 8-element fragment row: reads 8 consecutive floats of F32
template <typename E> __global__ void k_to_frag(const float *F, int ld, float s, E *out)
{
	typedef typename VecOf<E>::type V;
	const size_t units = (size_t)ld * ld / 8;
	for (size_t u = blockIdx.x * (size_t)kB + threadIdx.x; u < units; u += (size_t)gridDim.x * kB) {
		const int row = (int)(u / (ld / 8)), k = (int)(u % (ld / 8)) * 8;
		const float4 a = *(const float4 *)(F + (size_t)row * ld + k);
		const float4 b = *(const float4 *)(F + (size_t)row * ld + k + 4);
		V v;
		v[0] = (E)(a.x * s); v[1] = (E)(a.y * s); v[2] = (E)(a.z * s); v[3] = (E)(a.w * s);
		v[4] = (E)(b.x * s); v[5] = (E)(b.y * s); v[6] = (E)(b.z * s); v[7] = (E)(b.w * s);
		*(V *)(out + frag_index(row, k, ld)) = v;
	}
}

template <typename T> __global__ void k_fill(T *p, size_t count, T v)
{
	for (size_t i = blockIdx.x * (size_t)kB + threadIdx.x; i < count; i += (size_t)gridDim.x * kB) p[i] = v;
}

template <typename E, int RT, int CT, int U> struct GemmRegs {
	typedef typename VecOf<E>::type V;
	V a[U][RT], b[U][CT];
};

template <typename E, int RT, int CT, int U>
__device__ __forceinline__ void gemm_load(GemmRegs<E, RT, CT, U> &g, const typename VecOf<E>::type *Ap,
                                          const typename VecOf<E>::type *Bp, int rb0, int nkb, int kb)
{
#pragma unroll
	for (int u = 0; u < U; u++) {
#pragma unroll
		for (int r = 0; r < RT; r++) g.a[u][r] = __builtin_nontemporal_load(Ap + ((size_t)(rb0 + r) * nkb + kb + u) * 64);
#pragma unroll
		for (int c = 0; c < CT; c++) g.b[u][c] = Bp[((size_t)c * nkb + kb + u) * 64];
	}
}

template <typename E, int RT, int CT, int U>
__device__ __forceinline__ void gemm_mma(const GemmRegs<E, RT, CT, U> &g, f4 (&acc)[RT][CT])
{
#pragma unroll
	for (int u = 0; u < U; u++)
#pragma unroll
		for (int r = 0; r < RT; r++)
#pragma unroll
			for (int c = 0; c < CT; c++) acc[r][c] = mfma(g.a[u][r], g.b[u][c], acc[r][c]);
}

template <int CT> struct GemmUnroll { static constexpr int U = CT <= 4 ? 4 : 2; };

// One wave's share of part[s] = A X over k-blocks [kb0, kb1) (32 k each):
// rows 16*rb0 .. 16*(rb0+RT)-1, all 16*CT columns.  A and XT are fragment
// ordered with ld columns; the loop is software pipelined (two register sets
// of U k-blocks, one loading while the other feeds the MFMAs), so kb1 - kb0
// must be a multiple of 2U.
template <typename E, int RT, int CT>
__device__ __forceinline__ void gemm_tile(const E *__restrict__ A, const E *__restrict__ XT, float *__restrict__ part,
                                          int ld, int rp, int rb0, int kb0, int kb1)
{
	typedef typename VecOf<E>::type V;
	constexpr int U = GemmUnroll<CT>::U;
	const int lane = threadIdx.x & 63;
	const int nkb = ld >> 5;
	const V *Ap = (const V *)A + lane;
	const V *Bp = (const V *)XT + lane;
	f4 acc[RT][CT];
#pragma unroll
	for (int r = 0; r < RT; r++)
#pragma unroll
		for (int c = 0; c < CT; c++) acc[r][c] = f4{0.f, 0.f, 0.f, 0.f};
	GemmRegs<E, RT, CT, U> g0, g1;
	gemm_load(g0, Ap, Bp, rb0, nkb, kb0);
	for (int kb = kb0; kb < kb1; kb += 2 * U) {
		gemm_load(g1, Ap, Bp, rb0, nkb, kb + U);
		gemm_mma(g0, acc);
		if (kb + 2 * U < kb1) gemm_load(g0, Ap, Bp, rb0, nkb, kb + 2 * U);
		gemm_mma(g1, acc);
	}
	const int lr = lane & 15;
#pragma unroll
	for (int r = 0; r < RT; r++)
#pragma unroll
		for (int c = 0; c < CT; c++)
#pragma unroll
			for (int j = 0; j < 4; j++) {
				const int row = (rb0 + r) * 16 + (lane >> 4) * 4 + j;
				part[(size_t)row * rp + c * 16 + lr] = acc[r][c][j];
			}
}

// One wave's share of the FP64 product with an FP32 matrix on the FP64 MFMA
// (v_mfma_f64_16x16x4_f64: lane l supplies A[l%16][l/16] and B[l/16][l%16],
// receives D[l/16 + 4*j][l%16], j = 0..3 -- rows interleaved, unlike the
// 16-bit MFMAs): rows 16*rb0 .. 16*(rb0+RT)-1, columns
// 16*cb0 .. 16*(cb0+CT)-1, k in [k0, k1) (a multiple of 32).  Per 32 k each
// lane loads 8 consecutive F entries (converted exactly to FP64) and 8
// consecutive X entries, then runs 8 MFMAs, step s pairing k = kk + 8*(l/16)
// + s in both operands, so every k is summed once.  part is row-major, rp
// columns.
template <int RT, int CT>
__device__ __forceinline__ void fx64_tile(const float *__restrict__ F, const double *__restrict__ xT,
                                          double *__restrict__ part, int ld, int rp, int rb0, int cb0, int k0, int k1)
{
	typedef double d4 __attribute__((ext_vector_type(4)));
	typedef float fv4 __attribute__((ext_vector_type(4)));
	const int lane = threadIdx.x & 63;
	const int lr = lane & 15, g8 = (lane >> 4) * 8;
	d4 acc[RT][CT];
#pragma unroll
	for (int r = 0; r < RT; r++)
#pragma unroll
		for (int c = 0; c < CT; c++) acc[r][c] = d4{0.0, 0.0, 0.0, 0.0};
	for (int kk = k0; kk < k1; kk += 32) {
		float a[RT][8];
		double b[CT][8];
#pragma unroll
		for (int r = 0; r < RT; r++) {
			const fv4 *pa = (const fv4 *)(F + (size_t)((rb0 + r) * 16 + lr) * ld + kk + g8);
			const fv4 u = __builtin_nontemporal_load(pa), v = __builtin_nontemporal_load(pa + 1);
			a[r][0] = u[0]; a[r][1] = u[1]; a[r][2] = u[2]; a[r][3] = u[3];
			a[r][4] = v[0]; a[r][5] = v[1]; a[r][6] = v[2]; a[r][7] = v[3];
		}
#pragma unroll
		for (int c = 0; c < CT; c++) {
			const double2 *pb = (const double2 *)(xT + (size_t)((cb0 + c) * 16 + lr) * ld + kk + g8);
#pragma unroll
			for (int q = 0; q < 4; q++) {
				const double2 t = pb[q];
				b[c][2 * q] = t.x;
				b[c][2 * q + 1] = t.y;
			}
		}
#pragma unroll
		for (int st = 0; st < 8; st++)
#pragma unroll
			for (int r = 0; r < RT; r++)
#pragma unroll
				for (int c = 0; c < CT; c++)
					acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)a[r][st], b[c][st], acc[r][c], 0, 0, 0);
	}
#pragma unroll
	for (int r = 0; r < RT; r++)
#pragma unroll
		for (int c = 0; c < CT; c++)
#pragma unroll
			for (int j = 0; j < 4; j++) {
				const int row = (rb0 + r) * 16 + (lane >> 4) + 4 * j;
				part[(size_t)row * rp + (cb0 + c) * 16 + lr] = acc[r][c][j];
			}
}

#define HIP_OK(expr, what)                                          \
	do {                                                            \
		hipError_t e_ = (expr);                                     \
		if (e_ != hipSuccess) {                                     \
			err = std::string(what) + ": " + hipGetErrorString(e_); \
			return -1;                                              \
		}                                                           \
	} while (0)

template <typename T> struct Buf {
	T *p = nullptr;
	size_t n = 0;
	Buf() = default;
	Buf(const Buf &) = delete;
	Buf &operator=(const Buf &) = delete;
	~Buf()
	{
		if (p) (void)hipFree(p);
	}
	int alloc(size_t count, std::string &err)
	{
		if (count <= n && p) return 0;
		if (p) (void)hipFree(p);
		p = nullptr;
		n = 0;
		if (count == 0) return 0;
		HIP_OK(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
		n = count;
		return 0;
	}
};

int grid_for(size_t total) { return (int)std::min<size_t>((total + kB - 1) / kB, 8192); }

double ms_since(std::chrono::steady_clock::time_point a)
{
	return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
}


