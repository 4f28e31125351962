// Kernel-side argument block and launchers (internal to libpadeops_hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace pade {

constexpr int kBRowsDev = 8;
constexpr int kBColsDev = 16;

// Device table layout of one plan (element offsets):
//   [0, 16)                stencil (offsets -4..4, 1/dx^p folded in)
//   [off_bl, +8*16)        left special rows
//   [off_br, +8*16)        right special rows (stored reflected)
//   [off_lu, +P*M*5)       per-partition band LU: l1 l2 1/piv u1/piv u2/piv
//   [off_S,  +P*M*4)       per-partition coupling columns
//   [off_Q,  +P*4*4P)      per-partition rows of the inverse reduced system
template <typename T>
struct ApplyArgs {
    const T *f;
    T *out;
    const T *tab;
    long long npencils;
    long long A1, B1, B2;  // pencil q -> base offset (q % A1)*B1 + (q / A1)*B2
    long long stride;      // element distance between consecutive points of a pencil
    int n, M, P, periodic, nbl, nbr;
    int off_bl, off_br, off_lu, off_S, off_Q;
};

size_t spike_lds_bytes(size_t tsize, int P, bool adj, int mode);

template <typename T>
hipError_t launch_spike(const ApplyArgs<T> &a, int M, bool adj, int mode, hipStream_t s);

template <typename T>
hipError_t launch_generic(const ApplyArgs<T> &a, bool adj, T *work, hipStream_t s);

}  // namespace pade
