// Dense operator form (D = A^{-1} B applied as an MFMA GEMM) and the pencil
// re-orientation kernels (internal to libpadeops_hip).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <vector>

#include "pade_plan.h"

namespace pade {

struct DenseMatrix {
    int n = 0;
    float *d32 = nullptr;       // D, FP32, row-major n x n
    void *dbf = nullptr;        // D as three BF16 limbs, [3][n][n]
};

struct DenseCache {
    std::map<int, DenseMatrix> mats;
};

// Build (once) and return the dense operator for a plan key.
int dense_get(DenseCache &cache, const OpMatrices &m, int key, const DenseMatrix **out);
void dense_release(DenseCache &cache);

hipError_t dense_launch(const DenseMatrix &D, int axis, const void *f, int f_dtype, void *df,
                        int df_dtype, int na, int nb, int prec, hipStream_t s);

hipError_t permute_launch(const void *src, void *dst, int elem_size, int n1, int n2, int n3, int perm,
                          hipStream_t s);

}  // namespace pade
