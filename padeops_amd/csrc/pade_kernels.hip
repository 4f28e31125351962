// Compact-difference operator kernels for CDNA4 (gfx950).
//
// One operator application is  out = A^{-1} (B f)  (forward) or
// out = B^T (A^{-T} g)  (adjoint), applied to every pencil of a 3-D
// Fortran-ordered array along one axis.  A is a band matrix of half-width 2
// (cyclic when periodic), B a 9-point stencil with explicit boundary rows.
//
// Fast path (spike_apply): a pencil of n = P*M points is split into P
// partitions of M points.  A workgroup holds 64 pencils; wave p owns partition
// p of all of them, one pencil per lane, with the M values in registers.
// Each lane
//   1. loads its window (M points + 4-point halos) and evaluates the stencil,
//   2. solves its partition's band block with the block's own LU factors
//      (coefficients are wave-uniform -> scalar loads),
//   3. publishes the four "tip" values (first two / last two) through LDS,
//   4. forms the two neighbour tips on each side from the precomputed inverse
//      of the reduced tip system, and
//   5. corrects its partition with the precomputed coupling columns S_p.
// This is an exact partitioned (SPIKE-type) solve: every value of f is read
// once and every value of out written once, so HBM traffic is the algorithmic
// 2 * sizeof(T) bytes per point.
//
// Pencils along y and z are strided in memory: the 64 lanes of a wave are
// 64 neighbouring pencils, so every load is a contiguous 64-element row.
// Pencils along x are contiguous: the wave moves 64 x 16 tiles through LDS so
// that global loads/stores stay coalesced (MODE 1).
//
// Generic path (generic_apply): one lane per pencil with the whole pencil as
// one partition in global memory -- for sizes the register path does not
// cover (n not a multiple of 16, or too large).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pade_kernels.h"

namespace pade {

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int wrap_index(int j, int n, int periodic)
{
    if (j < 0) return periodic ? j + n : 0;
    if (j >= n) return periodic ? j - n : n - 1;
    return j;
}

constexpr int kXTile = 16;       // points per LDS transpose chunk (x pencils)
constexpr int kXPad = 17;        // padded row length of the transpose tile

template <typename T, int M, bool ADJ, int MODE>
__global__ __launch_bounds__(512) void spike_apply(ApplyArgs<T> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *smem = reinterpret_cast<T *>(smem_raw);
    const int lane = threadIdx.x & 63;
    const int p = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int P = a.P;
    const int n = a.n;
    const int pM = p * M;
    const long long q0 = (long long)blockIdx.x * 64;
    const long long q = q0 + lane;
    const bool valid = q < a.npencils;
    const long long base = (q % a.A1) * a.B1 + (q / a.A1) * a.B2;
    const long long st = a.stride;
    const T *__restrict__ tab = a.tab;
    const T *__restrict__ lu = tab + a.off_lu + (long long)p * M * 5;
    const T *__restrict__ S = tab + a.off_S + (long long)p * M * 4;
    const T *__restrict__ Q = tab + a.off_Q + (long long)p * 4 * (4 * P);
    const bool edge_left = !a.periodic && p == 0;
    const bool edge_right = !a.periodic && p == P - 1;

    T *tips = smem;                          // [4P][64]
    T *halo = smem + 4 * P * 64;             // [8P][64]   (adjoint)
    T *xt = smem + (ADJ ? 12 : 4) * P * 64 + p * 64 * kXPad;  // per-wave tile (MODE 1)

    T w[M + 8];

    // ---------------------------------------------------------------- load
    if constexpr (MODE == 0) {
        if constexpr (!ADJ) {
#pragma unroll
            for (int t = 0; t < M + 8; t++) {
                const int j = wrap_index(pM - 4 + t, n, a.periodic);
                w[t] = valid ? a.f[base + (long long)j * st] : T(0);
            }
        } else {
#pragma unroll
            for (int t = 0; t < M; t++) w[t] = valid ? a.f[base + (long long)(pM + t) * st] : T(0);
        }
    } else {
        constexpr int off = ADJ ? 0 : 4;
#pragma unroll
        for (int c = 0; c < M / kXTile; c++) {
#pragma unroll
            for (int s = 0; s < kXTile; s++) {
                const int pen = 4 * s + (lane >> 4);
                const int pt = lane & 15;
                const long long qq = q0 + pen;
                T v = T(0);
                if (qq < a.npencils) v = a.f[qq * n + pM + c * kXTile + pt];
                xt[pen * kXPad + pt] = v;
            }
            wave_lds_sync();
#pragma unroll
            for (int j = 0; j < kXTile; j++) w[off + c * kXTile + j] = xt[lane * kXPad + j];
            wave_lds_sync();
        }
        if constexpr (!ADJ) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int jl = wrap_index(pM - 4 + k, n, a.periodic);
                const int jr = wrap_index(pM + M + k, n, a.periodic);
                w[k] = valid ? a.f[base + jl] : T(0);
                w[M + 4 + k] = valid ? a.f[base + jr] : T(0);
            }
        }
    }

    T lb[kBRowsDev], rb[kBRowsDev];
    if constexpr (!ADJ) {
        // ------------------------------------------------------- B f (stencil)
        if (edge_left) {
            const T *bl = tab + a.off_bl;
#pragma unroll
            for (int r = 0; r < kBRowsDev; r++) {
                T s = T(0);
                if (r < a.nbl) {
#pragma unroll
                    for (int c = 0; c < kBColsDev; c++) s += bl[r * kBColsDev + c] * w[4 + c];
                }
                lb[r] = s;
            }
        }
        if (edge_right) {
            const T *br = tab + a.off_br;
#pragma unroll
            for (int r = 0; r < kBRowsDev; r++) {
                T s = T(0);
                if (r < a.nbr) {
#pragma unroll
                    for (int c = 0; c < kBColsDev; c++) s += br[r * kBColsDev + c] * w[M + 3 - c];
                }
                rb[r] = s;
            }
        }
        T cst[9];
#pragma unroll
        for (int k = 0; k < 9; k++) cst[k] = tab[k];
#pragma unroll
        for (int t = 0; t < M; t++) {
            T s = cst[0] * w[t];
#pragma unroll
            for (int k = 1; k < 9; k++) s += cst[k] * w[t + k];
            w[t] = s;
        }
        if (edge_left) {
#pragma unroll
            for (int r = 0; r < kBRowsDev; r++)
                if (r < a.nbl) w[r] = lb[r];
        }
        if (edge_right) {
#pragma unroll
            for (int r = 0; r < kBRowsDev; r++)
                if (r < a.nbr) w[M - 1 - r] = rb[r];
        }
    }

    // -------------------------------------------------- local band LU solve
#pragma unroll
    for (int t = 1; t < M; t++) {
        T z = w[t] - lu[t * 5 + 0] * w[t - 1];
        if (t >= 2) z -= lu[t * 5 + 1] * w[t - 2];
        w[t] = z;
    }
    w[M - 1] = w[M - 1] * lu[(M - 1) * 5 + 2];
    w[M - 2] = w[M - 2] * lu[(M - 2) * 5 + 2] - lu[(M - 2) * 5 + 3] * w[M - 1];
#pragma unroll
    for (int t = M - 3; t >= 0; t--)
        w[t] = w[t] * lu[t * 5 + 2] - lu[t * 5 + 3] * w[t + 1] - lu[t * 5 + 4] * w[t + 2];

    // ------------------------------------------- reduced system + correction
    if (a.periodic || P > 1) {
        tips[(4 * p + 0) * 64 + lane] = w[0];
        tips[(4 * p + 1) * 64 + lane] = w[1];
        tips[(4 * p + 2) * 64 + lane] = w[M - 2];
        tips[(4 * p + 3) * 64 + lane] = w[M - 1];
        __syncthreads();
        T u0 = T(0), u1 = T(0), u2 = T(0), u3 = T(0);
        const int m = 4 * P;
        for (int j = 0; j < m; j++) {
            const T g = tips[j * 64 + lane];
            u0 += Q[0 * m + j] * g;
            u1 += Q[1 * m + j] * g;
            u2 += Q[2 * m + j] * g;
            u3 += Q[3 * m + j] * g;
        }
#pragma unroll
        for (int t = 0; t < M; t++)
            w[t] -= S[t * 4 + 0] * u0 + S[t * 4 + 1] * u1 + S[t * 4 + 2] * u2 + S[t * 4 + 3] * u3;
    }

    if constexpr (ADJ) {
        // --------------------------------- B^T of the solved values (halo via LDS)
        const bool exch = a.periodic || P > 1;
        if (exch) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                halo[(8 * p + k) * 64 + lane] = w[k];
                halo[(8 * p + 4 + k) * 64 + lane] = w[M - 4 + k];
            }
            __syncthreads();
        }
        int pl = p - 1, pr = p + 1;
        bool hasl = true, hasr = true;
        if (pl < 0) { if (a.periodic) pl += P; else hasl = false; }
        if (pr >= P) { if (a.periodic) pr -= P; else hasr = false; }
#pragma unroll
        for (int t = M - 1; t >= 0; t--) w[t + 4] = w[t];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            w[k] = (exch && hasl) ? halo[(8 * pl + 4 + k) * 64 + lane] : T(0);
            w[M + 4 + k] = (exch && hasr) ? halo[(8 * pr + k) * 64 + lane] : T(0);
        }
        if (edge_left) {
            const T *bl = tab + a.off_bl;
#pragma unroll
            for (int r = 0; r < kBRowsDev; r++) {
                T s = T(0);
                if (r < a.nbl) {
#pragma unroll
                    for (int c = 0; c < kBColsDev; c++) s += bl[r * kBColsDev + c] * w[4 + c];
                }
                lb[r] = s;
            }
        }
        if (edge_right) {
            const T *br = tab + a.off_br;
#pragma unroll
            for (int r = 0; r < kBRowsDev; r++) {
                T s = T(0);
                if (r < a.nbr) {
#pragma unroll
                    for (int c = 0; c < kBColsDev; c++) s += br[r * kBColsDev + c] * w[M + 3 - c];
                }
                rb[r] = s;
            }
        }
        T cst[9];
#pragma unroll
        for (int k = 0; k < 9; k++) cst[k] = tab[k];
#pragma unroll
        for (int t = 0; t < M; t++) {
            T s = cst[0] * w[t];
#pragma unroll
            for (int k = 1; k < 9; k++) s += cst[k] * w[t + k];
            w[t] = s;
        }
        if (edge_left) {
#pragma unroll
            for (int r = 0; r < kBRowsDev; r++)
                if (r < a.nbl) w[r] = lb[r];
        }
        if (edge_right) {
#pragma unroll
            for (int r = 0; r < kBRowsDev; r++)
                if (r < a.nbr) w[M - 1 - r] = rb[r];
        }
    }

    // ---------------------------------------------------------------- store
    if constexpr (MODE == 0) {
        if (valid) {
#pragma unroll
            for (int t = 0; t < M; t++) a.out[base + (long long)(pM + t) * st] = w[t];
        }
    } else {
#pragma unroll
        for (int c = 0; c < M / kXTile; c++) {
#pragma unroll
            for (int j = 0; j < kXTile; j++) xt[lane * kXPad + j] = w[c * kXTile + j];
            wave_lds_sync();
#pragma unroll
            for (int s = 0; s < kXTile; s++) {
                const int pen = 4 * s + (lane >> 4);
                const int pt = lane & 15;
                const long long qq = q0 + pen;
                if (qq < a.npencils) a.out[qq * n + pM + c * kXTile + pt] = xt[pen * kXPad + pt];
            }
            wave_lds_sync();
        }
    }
}

// ------------------------------------------------------------------------
// Generic path: one lane per pencil, whole pencil as one partition, the
// intermediate values kept in the output array (adjoint: in `work`).
template <typename T, bool ADJ>
__global__ __launch_bounds__(256) void generic_apply(ApplyArgs<T> a, T *__restrict__ work)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.npencils) return;
    const long long base = (q % a.A1) * a.B1 + (q / a.A1) * a.B2;
    const long long st = a.stride;
    const int n = a.n;
    const T *__restrict__ tab = a.tab;
    const T *__restrict__ lu = tab + a.off_lu;
    const T *__restrict__ S = tab + a.off_S;
    const T *__restrict__ Q = tab + a.off_Q;
    T *y = ADJ ? work : a.out;

    auto stencil_row = [&](const T *src, int i) -> T {
        if (!a.periodic) {
            if (i < a.nbl) {
                T s = T(0);
                for (int c = 0; c < kBColsDev && c < n; c++) s += tab[a.off_bl + i * kBColsDev + c] * src[base + (long long)c * st];
                return s;
            }
            if (i >= n - a.nbr) {
                const int r = n - 1 - i;
                T s = T(0);
                for (int c = 0; c < kBColsDev && c < n; c++)
                    s += tab[a.off_br + r * kBColsDev + c] * src[base + (long long)(n - 1 - c) * st];
                return s;
            }
        }
        T s = T(0);
        for (int k = 0; k < 9; k++) {
            const T cf = tab[k];
            if (cf == T(0)) continue;
            const int j = wrap_index(i - 4 + k, n, a.periodic);
            s += cf * src[base + (long long)j * st];
        }
        return s;
    };

    if (!ADJ) {
        for (int i = 0; i < n; i++) y[base + (long long)i * st] = stencil_row(a.f, i);
    } else {
        for (int i = 0; i < n; i++) y[base + (long long)i * st] = a.f[base + (long long)i * st];
    }
    // forward sweep
    T zm1 = y[base], zm2 = T(0);
    for (int i = 1; i < n; i++) {
        T z = y[base + (long long)i * st] - lu[i * 5 + 0] * zm1 - lu[i * 5 + 1] * zm2;
        y[base + (long long)i * st] = z;
        zm2 = zm1;
        zm1 = z;
    }
    // backward sweep
    T gp1 = T(0), gp2 = T(0);
    for (int i = n - 1; i >= 0; i--) {
        T g = y[base + (long long)i * st] * lu[i * 5 + 2] - lu[i * 5 + 3] * gp1 - lu[i * 5 + 4] * gp2;
        y[base + (long long)i * st] = g;
        gp2 = gp1;
        gp1 = g;
    }
    if (a.periodic) {
        const T g0 = y[base], g1 = y[base + st], g2 = y[base + (long long)(n - 2) * st],
                g3 = y[base + (long long)(n - 1) * st];
        T u[4];
        for (int c = 0; c < 4; c++) u[c] = Q[c * 4 + 0] * g0 + Q[c * 4 + 1] * g1 + Q[c * 4 + 2] * g2 + Q[c * 4 + 3] * g3;
        for (int i = 0; i < n; i++) {
            const long long o = base + (long long)i * st;
            y[o] -= S[i * 4 + 0] * u[0] + S[i * 4 + 1] * u[1] + S[i * 4 + 2] * u[2] + S[i * 4 + 3] * u[3];
        }
    }
    if (ADJ) {
        for (int i = 0; i < n; i++) a.out[base + (long long)i * st] = stencil_row(work, i);
    }
}

// ------------------------------------------------------------------------
// launchers

template <typename T, int M, bool ADJ, int MODE>
static hipError_t launch_spike_t(const ApplyArgs<T> &a, hipStream_t s)
{
    const long long nblk = (a.npencils + 63) / 64;
    const size_t lds = spike_lds_bytes(sizeof(T), a.P, ADJ, MODE);
    hipLaunchKernelGGL((spike_apply<T, M, ADJ, MODE>), dim3((unsigned)nblk), dim3(64 * a.P), lds, s, a);
    return hipGetLastError();
}

template <typename T, int M>
static hipError_t launch_spike_m(const ApplyArgs<T> &a, bool adj, int mode, hipStream_t s)
{
    if (!adj) return mode == 0 ? launch_spike_t<T, M, false, 0>(a, s) : launch_spike_t<T, M, false, 1>(a, s);
    return mode == 0 ? launch_spike_t<T, M, true, 0>(a, s) : launch_spike_t<T, M, true, 1>(a, s);
}

size_t spike_lds_bytes(size_t tsize, int P, bool adj, int mode)
{
    size_t elems = (size_t)(adj ? 12 : 4) * P * 64;
    if (mode == 1) elems += (size_t)P * 64 * kXPad;
    return elems * tsize;
}

template <typename T>
hipError_t launch_spike(const ApplyArgs<T> &a, int M, bool adj, int mode, hipStream_t s)
{
    switch (M) {
    case 16: return launch_spike_m<T, 16>(a, adj, mode, s);
    case 32: return launch_spike_m<T, 32>(a, adj, mode, s);
    case 64: return launch_spike_m<T, 64>(a, adj, mode, s);
    default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t launch_generic(const ApplyArgs<T> &a, bool adj, T *work, hipStream_t s)
{
    const int bs = 256;
    const long long nblk = (a.npencils + bs - 1) / bs;
    if (adj)
        hipLaunchKernelGGL((generic_apply<T, true>), dim3((unsigned)nblk), dim3(bs), 0, s, a, work);
    else
        hipLaunchKernelGGL((generic_apply<T, false>), dim3((unsigned)nblk), dim3(bs), 0, s, a, work);
    return hipGetLastError();
}

template hipError_t launch_spike<double>(const ApplyArgs<double> &, int, bool, int, hipStream_t);
template hipError_t launch_spike<float>(const ApplyArgs<float> &, int, bool, int, hipStream_t);
template hipError_t launch_generic<double>(const ApplyArgs<double> &, bool, double *, hipStream_t);
template hipError_t launch_generic<float>(const ApplyArgs<float> &, bool, float *, hipStream_t);

}  // namespace pade
