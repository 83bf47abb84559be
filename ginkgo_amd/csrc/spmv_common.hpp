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
// Shape of the classical kernel: `sw` lanes per row, `rpg` rows in flight per
// lane group.  gkomi_csr::subwave holds sw in bits 0..7 and rpg in bits 8..15
// (0 = automatic).
struct subwave_shape {
    int sw;
    int rpg;
};

subwave_shape pick_subwave(const gkomi_csr* a, int value_bytes);

inline bool valid_subwave(subwave_shape sh)
{
    return sh.sw >= 1 && sh.sw <= 64 && (sh.sw & (sh.sw - 1)) == 0 &&
           (sh.rpg == 1 || sh.rpg == 2 || sh.rpg == 4 || sh.rpg == 8);
}

template <typename T, bool Adv>
int csr_spmv_launch(hipStream_t stream, const gkomi_csr* a, int64_t nrhs, const T* b, int64_t bs,
                    T* c, int64_t cs, const T* alpha, const T* beta);

}  // namespace gkomi
