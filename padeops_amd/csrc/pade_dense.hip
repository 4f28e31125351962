// Dense-operator form and pencil re-orientation kernels.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <algorithm>
#include <vector>

#include "pade_dense.h"

namespace pade {

int dense_get(DenseCache &cache, const OpMatrices &m, int key, const DenseMatrix **out)
{
    auto it = cache.mats.find(key);
    if (it != cache.mats.end()) {
        *out = &it->second;
        return 0;
    }
    std::vector<double> D;
    if (int rc = dense_operator(m, D)) return rc;
    DenseMatrix dm;
    dm.n = m.n;
    const size_t nn = (size_t)m.n * m.n;
    std::vector<float> d32(nn);
    std::vector<uint16_t> bf(3 * nn);
    auto to_bf16 = [](float x) -> uint16_t {  // round to nearest even
        uint32_t u;
        std::memcpy(&u, &x, 4);
        const uint32_t lsb = (u >> 16) & 1u;
        u += 0x7fffu + lsb;
        return uint16_t(u >> 16);
    };
    auto from_bf16 = [](uint16_t h) -> double {
        uint32_t u = uint32_t(h) << 16;
        float x;
        std::memcpy(&x, &u, 4);
        return double(x);
    };
    for (size_t k = 0; k < nn; k++) {
        d32[k] = float(D[k]);
        double r = D[k];
        for (int l = 0; l < 3; l++) {
            const uint16_t h = to_bf16(float(r));
            bf[l * nn + k] = h;
            r -= from_bf16(h);
        }
    }
    hipError_t e = hipMalloc(&dm.d32, nn * sizeof(float));
    if (e != hipSuccess) return 1000 + int(e);
    e = hipMalloc(&dm.dbf, 3 * nn * sizeof(uint16_t));
    if (e != hipSuccess) { hipFree(dm.d32); return 1000 + int(e); }
    hipMemcpy(dm.d32, d32.data(), nn * sizeof(float), hipMemcpyHostToDevice);
    hipMemcpy(dm.dbf, bf.data(), 3 * nn * sizeof(uint16_t), hipMemcpyHostToDevice);
    auto res = cache.mats.emplace(key, dm);
    *out = &res.first->second;
    return 0;
}

void dense_release(DenseCache &cache)
{
    for (auto &kv : cache.mats) {
        if (kv.second.d32) hipFree(kv.second.d32);
        if (kv.second.dbf) hipFree(kv.second.dbf);
    }
    cache.mats.clear();
}

hipError_t dense_launch(const DenseMatrix &, int, const void *, int, void *, int, int, int, int,
                        hipStream_t)
{
    return hipErrorNotSupported;
}

// ------------------------------------------------------------------------
// Axis-permuting re-orientation through 32 x 32 LDS tiles.
// dst(perm(i,j,k)) = src(i,j,k); perm 1 swaps i,j; 2 swaps i,k; 3 swaps j,k.
template <typename T>
__global__ __launch_bounds__(256) void permute_tile(const T *__restrict__ src, T *__restrict__ dst,
                                                   long long nu, long long nv, long long nw,
                                                   long long su_v, long long su_w, long long sd_u,
                                                   long long sd_w)
{
    // src(u, v, w) at u + v*su_v + w*su_w  (u contiguous in src)
    // dst(v, u, w) at v + u*sd_u + w*sd_w  (v contiguous in dst)
    __shared__ T tile[32][33];
    const long long u0 = (long long)blockIdx.x * 32, v0 = (long long)blockIdx.y * 32;
    const long long w = blockIdx.z;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const long long u = u0 + tx, v = v0 + r;
        if (u < nu && v < nv) tile[r][tx] = src[u + v * su_v + w * su_w];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const long long v = v0 + tx, u = u0 + r;
        if (u < nu && v < nv) dst[v + u * sd_u + w * sd_w] = tile[tx][r];
    }
}

template <typename T>
__global__ __launch_bounds__(256) void permute_rows(const T *__restrict__ src, T *__restrict__ dst,
                                                   long long n1, long long n2, long long n3)
{
    // perm 3: dst(i, k, j) = src(i, j, k); rows of i stay contiguous
    const long long row = blockIdx.y + (long long)gridDim.y * blockIdx.z;  // j + n2*k
    if (row >= n2 * n3) return;
    const long long j = row % n2, k = row / n2;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n1) dst[i + n1 * (k + n3 * j)] = src[i + n1 * (j + n2 * k)];
}

template <typename T>
static hipError_t permute_t(const T *src, T *dst, long long n1, long long n2, long long n3, int perm,
                            hipStream_t s)
{
    if (perm == 3) {
        const unsigned gx = (unsigned)((n1 + 255) / 256);
        long long rows = n2 * n3;
        unsigned gy = (unsigned)std::min<long long>(rows, 65535);
        unsigned gz = (unsigned)((rows + gy - 1) / gy);
        hipLaunchKernelGGL((permute_rows<T>), dim3(gx, gy, gz), dim3(256), 0, s, src, dst, n1, n2, n3);
        return hipGetLastError();
    }
    if (perm == 1) {
        // u = i (n1), v = j (n2), w = k (n3); dst(j, i, k) at j + i*n2 + k*n1*n2
        dim3 g((unsigned)((n1 + 31) / 32), (unsigned)((n2 + 31) / 32), (unsigned)n3);
        hipLaunchKernelGGL((permute_tile<T>), g, dim3(256), 0, s, src, dst, n1, n2, n3, n1, n1 * n2, n2,
                           n1 * n2);
        return hipGetLastError();
    }
    // perm 2: u = i (n1), v = k (n3), w = j (n2); dst(k, j, i) at k + j*n3 + i*n3*n2
    dim3 g((unsigned)((n1 + 31) / 32), (unsigned)((n3 + 31) / 32), (unsigned)n2);
    hipLaunchKernelGGL((permute_tile<T>), g, dim3(256), 0, s, src, dst, n1, n3, n2, n1 * n2, n1,
                       n3 * n2, n3);
    return hipGetLastError();
}

hipError_t permute_launch(const void *src, void *dst, int elem_size, int n1, int n2, int n3, int perm,
                          hipStream_t s)
{
    if (elem_size == 8)
        return permute_t<double>((const double *)src, (double *)dst, n1, n2, n3, perm, s);
    return permute_t<float>((const float *)src, (float *)dst, n1, n2, n3, perm, s);
}

}  // namespace pade
