// Constants and the result store shared by the SpMV kernels.
#pragma once

#include "common.hpp"

namespace gkomi {

// Load-balanced CSR: nominal nnz per workgroup of the partition, and the LDS
// capacity (products and row pointers) of the staged path.
constexpr int64_t kLbChunk = 1024;
constexpr int kLbLdsNnz = 2048;
constexpr int kLbLdsRows = 1024;
// rows up to this length are reduced by one thread from LDS, longer by a wave
constexpr int kLbShortRow = 32;

// c = sum (plain) or c = alpha * sum + beta * c with beta == 0 ignoring the
// old c (reference/matrix/csr_kernels.cpp:96-117: NaN in c does not propagate).
template <bool Adv, typename T>
__device__ __forceinline__ void store_result(T* dst, T sum, const T* alpha, const T* beta)
{
    if constexpr (Adv) {
        const T bt = beta[0];
        const T r = alpha[0] * sum;
        *dst = (bt == T(0)) ? r : r + bt * *dst;
    } else {
        *dst = sum;
    }
}

int check_csr(const gkomi_csr* a);
// lanes per row of the classical kernel: the descriptor's value, or the
// smallest power of two >= the mean row length (capped at 64)
int pick_subwave(const gkomi_csr* a);

template <typename T, bool Adv>
int csr_spmv_launch(hipStream_t stream, const gkomi_csr* a, int64_t nrhs, const T* b, int64_t bs,
                    T* c, int64_t cs, const T* alpha, const T* beta);

}  // namespace gkomi
