// Batched BiCGSTAB and CG for gfx950: thousands of small independent systems.
//
// One workgroup solves one system at a time, or several small ones side by
// side; the systems are size-sorted (largest first) and cut into size bins,
// one launch per bin or one launch for the whole batch.  Three kernel
// families:
//  * batch_solver_kernel: every work vector in LDS, the matrix (values,
//    16-bit local column indices, row pointers) in LDS when it fits, else
//    read from global memory (L2 resident for the working set of the active
//    workgroups); each workgroup pulls its systems from an atomic queue, so
//    uneven sizes balance themselves;
//  * batch_reg_kernel (Csr formats): the work vectors in registers, one or
//    two rows per thread, only the gathered vectors in LDS beside the matrix
//    or with the matrix rows in registers too; workgroup b takes system b of
//    the bin first and uses the queue only when the bin holds more systems
//    than the grid has workgroups;
//  * batch_packed_kernel (uneven Csr, systems of up to 512 rows, rows of up
//    to 16 entries): the whole batch in one launch of 256-thread workgroups
//    over one queue of work items, largest first, an item being one system
//    of more than 256 rows or 2, 4 or 8 smaller ones side by side.
// The whole iteration runs inside the kernel, with the reference's
// per-system semantics and logging of iterations and residual norm: BiCGSTAB
// (reference/solver/batch_bicgstab_kernels.hpp:180-354: no breakdown guards,
// the half-step convergence exit) or CG (reference/solver/
// batch_cg_kernels.hpp:31-221: the implicit residual norm sqrt|r.z|).
//
// Storage formats:
//  * uneven: a block-diagonal CSR, system k owning global rows
//    [row_off[k], row_off[k+1]) with column indices local to the system;
//  * uniform Csr: batch::matrix::Csr of the reference (include/ginkgo/core/
//    matrix/batch_csr.hpp:29-31): one shared pattern of n rows, values per item;
//  * uniform Ell: batch::matrix::Ell (batch_ell.hpp): shared column-major
//    column indices, -1 after a row's last entry, values per item;
//  * uniform Dense: batch::matrix::Dense (batch_dense.hpp): row-major values
//    per item (staged transposed in LDS, so that the lanes of a wave read
//    consecutive words).
#include <mutex>
#include <vector>

#include "common.hpp"

namespace gkomi {

namespace {

// Sums a and b over the workgroup; every thread receives both sums.  One
// barrier: the caller alternates between two `lds` buffers of 2 * waves
// values (the barrier of the reduction in between orders every read of one
// buffer before its next write).
template <typename T, int NThreads>
__device__ __forceinline__ void block_sum2(T& a, T& b, T* lds)
{
    constexpr int kW = NThreads / kWave;
    a = wave_total_last(a);
    b = wave_total_last(b);
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    if (lane == kWave - 1) {
        lds[wave] = a;
        lds[kW + wave] = b;
    }
    __syncthreads();
    T s0 = lds[0], s1 = lds[kW];
#pragma unroll
    for (int w = 1; w < kW; ++w) {
        s0 += lds[w];
        s1 += lds[kW + w];
    }
    a = s0;
    b = s1;
}

template <typename T, int NThreads>
__device__ __forceinline__ T block_sum1(T a, T* lds)
{
    constexpr int kW = NThreads / kWave;
    a = wave_total_last(a);
    if (threadIdx.x % kWave == kWave - 1) lds[threadIdx.x / kWave] = a;
    __syncthreads();
    T s = lds[0];
#pragma unroll
    for (int w = 1; w < kW; ++w) s += lds[w];
    return s;
}

// group_sum over aligned groups of g lanes, g a power of two <= 64 known
// only at run time (uniform over the workgroup)
template <typename T>
__device__ __forceinline__ T group_sum_rt(T v, int g)
{
    if (g >= 2) v += dpp<0xB1>(v);
    if (g >= 4) v += dpp<0x4E>(v);
    if (g >= 8) v += dpp<0x141>(v);
    if (g >= 16) v += dpp<0x140>(v);
    if (g >= 32) v += swz16(v);
    if (g >= 64) v += __shfl_xor(v, 32, 64);
    return v;
}

struct sys_view {
    int32_t n;
    int32_t nnz;
    int64_t row0;        // first global row (vector offset)
    int64_t nnz0;        // first value of the system
    const int32_t* rp;   // global row pointers of the system (uneven) or shared
    const int32_t* ci;   // column indices (local)
    int32_t rp_base;     // subtracted from rp entries to get local offsets
};

}  // namespace

// How the kernels read the matrix of a batch.
enum batch_fmt : int { kFmtUneven = 0, kFmtCsr = 1, kFmtEll = 2, kFmtDense = 3 };

// The solver run per system.
enum batch_method { kBatchBicgstab = 0, kBatchCg = 1 };

// Work vectors per system: BiCGSTAB r, r_hat, p, p_hat, v, s, s_hat, t, x,
// prec; CG r, z, p, Ap, x, prec.
__host__ __device__ constexpr int batch_vectors(int method)
{
    return method == kBatchCg ? 6 : 10;
}

// LDS layout per system slot (n = lds_rows, z = lds_nnz, V = batch_vectors):
//   T vec[V][n]    the work vectors, x and prec last
//   T vals[z]      (MatLds; Dense transposed)
//   int32 rp[n+1]  (MatLds, Csr formats)
//   uint16 ci[z]   (MatLds, Csr and Ell formats; Ell padding 0xFFFF)
template <typename T, int NThreads, bool MatLds, int Fmt, int Method>
__global__ void __launch_bounds__(NThreads)
    batch_solver_kernel(int64_t num_systems, const int32_t* __restrict__ order,
                          const int64_t* __restrict__ row_off,
                          const int64_t* __restrict__ item_off, const int32_t* __restrict__ rp,
                          const int32_t* __restrict__ ci, const T* __restrict__ vals,
                          int64_t uni_n, int64_t uni_nnz, int32_t ell_width,
                          const T* __restrict__ b, T* __restrict__ x, int32_t max_iters,
                          int32_t relative, double tol_d, int32_t jacobi,
                          int32_t* __restrict__ iters_out, T* __restrict__ res_out,
                          unsigned* __restrict__ queue, int64_t lds_rows, int64_t lds_nnz,
                          int32_t subwave)
{
    constexpr bool kCsr = Fmt == kFmtUneven || Fmt == kFmtCsr;
    extern __shared__ __align__(16) unsigned char smem[];
    // two reduction buffers, used in turn (block_sum2)
    __shared__ T red2[4 * (NThreads / kWave)];
    int red_turn = 0;
    auto red = [&]() {
        red_turn ^= 1;
        return red2 + red_turn * 2 * (NThreads / kWave);
    };
    __shared__ int64_t slot;
    const int tid = threadIdx.x;
    T* vec = reinterpret_cast<T*>(smem);
    const int64_t L = lds_rows;
    constexpr int kVec = batch_vectors(Method);
    T* r = vec;
    T* xl = vec + (kVec - 2) * L;
    T* prec = vec + (kVec - 1) * L;
    T* lv = vec + kVec * L;
    int32_t* lrp = reinterpret_cast<int32_t*>(lv + (MatLds ? lds_nnz : 0));
    uint16_t* lci = reinterpret_cast<uint16_t*>(lrp + (MatLds && kCsr ? L + 1 : 0));
    const T tol = static_cast<T>(tol_d);

    while (true) {
        __syncthreads();
        if (tid == 0) slot = static_cast<int64_t>(atomicAdd(queue, 1u));
        __syncthreads();
        const int64_t k = slot;
        if (k >= num_systems) break;
        const int64_t sys = order ? order[k] : k;
        sys_view sv;
        // Ell slots per row of this system
        int32_t ew = ell_width;
        if constexpr (Fmt == kFmtEll || Fmt == kFmtDense) {
            if (row_off != nullptr) {
                // uneven Ell / Dense: the system's rows from row_off, its
                // stored entries (and Ell column indices) from item_off
                sv.row0 = row_off[sys];
                sv.n = static_cast<int32_t>(row_off[sys + 1] - sv.row0);
                sv.nnz0 = item_off[sys];
                sv.nnz = static_cast<int32_t>(item_off[sys + 1] - sv.nnz0);
                sv.rp = rp;
                sv.ci = Fmt == kFmtEll ? ci + sv.nnz0 : ci;
                sv.rp_base = 0;
                ew = sv.n > 0 ? sv.nnz / sv.n : 0;
            } else {
                sv.n = static_cast<int32_t>(uni_n);
                sv.nnz = static_cast<int32_t>(uni_nnz);
                sv.row0 = sys * uni_n;
                sv.nnz0 = sys * uni_nnz;
                sv.rp = rp;
                sv.ci = ci;
                sv.rp_base = 0;
            }
        } else if constexpr (Fmt != kFmtUneven) {
            sv.n = static_cast<int32_t>(uni_n);
            sv.nnz = static_cast<int32_t>(uni_nnz);
            sv.row0 = sys * uni_n;
            sv.nnz0 = sys * uni_nnz;
            sv.rp = rp;
            sv.ci = ci;
            sv.rp_base = 0;
        } else {
            sv.row0 = row_off[sys];
            sv.n = static_cast<int32_t>(row_off[sys + 1] - sv.row0);
            sv.rp = rp + sv.row0;
            sv.rp_base = sv.rp[0];
            sv.nnz = sv.rp[sv.n] - sv.rp_base;
            sv.nnz0 = sv.rp_base;
            sv.ci = ci + sv.nnz0;
        }
        const int32_t n = sv.n;
        const T* gv = vals + sv.nnz0;
        const T* gb = b + sv.row0;
        T* gx = x + sv.row0;
        // stage the matrix
        if constexpr (MatLds) {
            if constexpr (Fmt == kFmtDense) {
                // transposed: entry (i, j) at j * n + i
                for (int e = tid; e < sv.nnz; e += NThreads) {
                    const int i = e / n;
                    const int j = e - i * n;
                    lv[j * n + i] = gv[e];
                }
            } else if constexpr (Fmt == kFmtEll) {
                for (int e = tid; e < sv.nnz; e += NThreads) {
                    lv[e] = gv[e];
                    const int32_t c = sv.ci[e];
                    lci[e] = c < 0 ? uint16_t{0xFFFF} : static_cast<uint16_t>(c);
                }
            } else {
                for (int i = tid; i < sv.nnz; i += NThreads) {
                    lv[i] = gv[i];
                    lci[i] = static_cast<uint16_t>(sv.ci[i]);
                }
                for (int i = tid; i <= n; i += NThreads) lrp[i] = sv.rp[i] - sv.rp_base;
            }
        }
        __syncthreads();
        auto row_begin = [&](int row) {
            if constexpr (MatLds) return lrp[row];
            return sv.rp[row] - sv.rp_base;
        };
        auto row_end = [&](int row) {
            if constexpr (MatLds) return lrp[row + 1];
            return sv.rp[row + 1] - sv.rp_base;
        };
        auto val = [&](int64_t e) {
            if constexpr (MatLds) return lv[e];
            return gv[e];
        };
        // column of stored entry e; -1 for Ell padding
        auto col = [&](int64_t e) -> int32_t {
            if constexpr (MatLds) {
                const int32_t c = lci[e];
                if constexpr (Fmt == kFmtEll) return c == 0xFFFF ? -1 : c;
                return c;
            }
            return sv.ci[e];
        };
        // sum_j a_ij in_j in the order of the reference's simple_apply of
        // the format (common/cuda_hip/matrix/batch_{csr,ell,dense}_kernels.hpp)
        auto row_dot = [&](int row, const T* in) -> T {
            T sum = 0;
            if constexpr (Fmt == kFmtDense) {
                for (int j = 0; j < n; ++j) {
                    const T a = MatLds ? lv[static_cast<int64_t>(j) * n + row]
                                       : gv[static_cast<int64_t>(row) * n + j];
                    sum += a * in[j];
                }
            } else if constexpr (Fmt == kFmtEll) {
                for (int j = 0; j < ew; ++j) {
                    const int64_t e = row + static_cast<int64_t>(j) * n;
                    const int32_t c = col(e);
                    if (c < 0) break;
                    sum += val(e) * in[c];
                }
            } else {
                const int32_t e1 = row_end(row);
                for (int32_t e = row_begin(row); e < e1; ++e) sum += val(e) * in[col(e)];
            }
            return sum;
        };
        // out = A in: Csr rows by groups of `subwave` lanes reading
        // consecutive entries (the lanes of a group stride over the row,
        // then a DPP group sum), the other formats a thread per row.  The
        // thread that stores out[row] also calls post(row, out[row]), so that
        // dot products with the result need no further pass.
        auto spmv_post = [&](const T* in, T* out, auto&& post) {
            if constexpr (kCsr) {
                if (subwave > 1) {
                    const int l = tid & (subwave - 1);
                    for (int row = tid / subwave; row < n; row += NThreads / subwave) {
                        T sum = 0;
                        const int32_t e1 = row_end(row);
                        for (int32_t e = row_begin(row) + l; e < e1; e += subwave) {
                            sum += val(e) * in[col(e)];
                        }
                        sum = group_sum_rt(sum, subwave);
                        if (l == 0) {
                            out[row] = sum;
                            post(row, sum);
                        }
                    }
                    return;
                }
            }
            for (int row = tid; row < n; row += NThreads) {
                const T sum = row_dot(row, in);
                out[row] = sum;
                post(row, sum);
            }
        };
        auto spmv = [&](const T* in, T* out) { spmv_post(in, out, [](int, T) {}); };
        // the scalar Jacobi entry of a row, per format as the reference's
        // ScalarJacobi::generate (common/cuda_hip/preconditioner/
        // batch_scalar_jacobi.hpp): Csr and Dense 1 / a_ii with a zero or
        // missing diagonal giving 1; Ell 1 / a_ii (no zero guard), the row
        // ending at the first column index below its slot number
        auto jacobi_of = [&](int row) -> T {
            T d = T(1);
            if constexpr (Fmt == kFmtDense) {
                const T a = MatLds ? lv[static_cast<int64_t>(row) * n + row]
                                   : gv[static_cast<int64_t>(row) * n + row];
                if (a != T(0)) d = T(1) / a;
            } else if constexpr (Fmt == kFmtEll) {
                for (int j = 0; j < ew; ++j) {
                    const int64_t e = row + static_cast<int64_t>(j) * n;
                    const int32_t c = col(e);
                    if (c < j) break;
                    if (c == row) d = T(1) / val(e);
                }
            } else {
                const int32_t e1 = row_end(row);
                for (int32_t e = row_begin(row); e < e1; ++e) {
                    if (col(e) == row) {
                        if (val(e) != T(0)) d = T(1) / val(e);
                        break;
                    }
                }
            }
            return d;
        };
        // prec.generate: batch scalar Jacobi (reference/preconditioner/
        // batch_scalar_jacobi.hpp:79-92) or identity
        T bb = 0, unused = 0;
        for (int row = tid; row < n; row += NThreads) {
            prec[row] = jacobi ? jacobi_of(row) : T(1);
            const T bi = gb[row];
            xl[row] = gx[row];
            r[row] = bi;
            bb += bi * bi;
        }
        block_sum2<T, NThreads>(bb, unused, red());
        const T rhs_norm = sqrt(bb);
        __syncthreads();
        const T limit = relative ? tol * rhs_norm : tol;
        int32_t iter = 0;
        T res = 0;
        if constexpr (Method == kBatchCg) {
            // reference/solver/batch_cg_kernels.hpp:31-221 (batch_entry_cg_impl):
            // r = b - A x, p = z = Ap = 0, rho_old = 1; per iteration
            // z = M r, rho = r.z, the implicit norm sqrt|rho| against the
            // tolerance (iterations counted from 1), p = z + rho/rho_old p
            // (p = z when rho_old == 0), Ap = A p, x += rho/(p.Ap) p,
            // r -= rho/(p.Ap) Ap.
            T* z = vec + L;
            T* p = vec + 2 * L;
            T* ap = vec + 3 * L;
            spmv(xl, ap);
            __syncthreads();
            // z = M r and r.z of the first iteration are formed here, later
            // ones with the update of x and r
            T rz = 0;
            for (int row = tid; row < n; row += NThreads) {
                const T ri = r[row] - ap[row];
                r[row] = ri;
                p[row] = T(0);
                ap[row] = T(0);
                const T zi = prec[row] * ri;
                z[row] = zi;
                rz += ri * zi;
            }
            T rho_old = 1;
            while (true) {
                const T rho = block_sum1<T, NThreads>(rz, red());
                ++iter;
                res = sqrt(rho < T(0) ? -rho : rho);
                if (iter >= max_iters || res <= limit) break;
                const bool restart = rho_old == T(0);
                const T beta = restart ? T(0) : rho / rho_old;
                for (int row = tid; row < n; row += NThreads) {
                    p[row] = restart ? z[row] : z[row] + beta * p[row];
                }
                __syncthreads();
                T pap = 0;
                spmv_post(p, ap, [&](int row, T v) { pap += p[row] * v; });
                pap = block_sum1<T, NThreads>(pap, red());
                const T temp = rho / pap;
                rz = 0;
                for (int row = tid; row < n; row += NThreads) {
                    xl[row] += temp * p[row];
                    const T ri = r[row] - temp * ap[row];
                    r[row] = ri;
                    const T zi = prec[row] * ri;
                    z[row] = zi;
                    rz += ri * zi;
                }
                rho_old = rho;
            }
        } else {
        T* rh = vec + L;
        T* p = vec + 2 * L;
        T* ph = vec + 3 * L;
        T* v = vec + 4 * L;
        T* s = vec + 5 * L;
        T* sh = vec + 6 * L;
        T* t = vec + 7 * L;
        // r = b - A x, ||r||, r_hat = r, p = p_hat = v = 0
        T rr = 0;
        spmv(xl, v);
        __syncthreads();
        for (int row = tid; row < n; row += NThreads) {
            const T ri = r[row] - v[row];
            r[row] = ri;
            rh[row] = ri;
            p[row] = T(0);
            ph[row] = T(0);
            v[row] = T(0);
            rr += ri * ri;
        }
        T rho_new = rr;  // r_hat . r with r_hat = r
        block_sum2<T, NThreads>(rr, rho_new, red());
        res = sqrt(rr);
        T rho_old = 1, omega = 1, alpha = 1;
        for (; iter < max_iters; ++iter) {
            if (res <= limit) break;
            // p = r + beta (p - omega v), p_hat = M p
            const T beta = (rho_new / rho_old) * (alpha / omega);
            for (int row = tid; row < n; row += NThreads) {
                const T pi = r[row] + beta * (p[row] - omega * v[row]);
                p[row] = pi;
                ph[row] = prec[row] * pi;
            }
            __syncthreads();
            T rv = 0;
            spmv_post(ph, v, [&](int row, T val_) { rv += rh[row] * val_; });
            rv = block_sum1<T, NThreads>(rv, red());
            alpha = rho_new / rv;
            T ss = 0;
            for (int row = tid; row < n; row += NThreads) {
                const T si = r[row] - alpha * v[row];
                s[row] = si;
                sh[row] = prec[row] * si;
                ss += si * si;
            }
            ss = block_sum1<T, NThreads>(ss, red());
            res = sqrt(ss);
            if (res <= limit) {
                for (int row = tid; row < n; row += NThreads) xl[row] += alpha * ph[row];
                break;
            }
            T ts = 0, tt = 0;
            spmv_post(sh, t, [&](int row, T val_) {
                ts += val_ * s[row];
                tt += val_ * val_;
            });
            block_sum2<T, NThreads>(ts, tt, red());
            omega = ts / tt;
            T r2 = 0, rhr = 0;
            for (int row = tid; row < n; row += NThreads) {
                // one explicit contraction, the same in both kernels
                xl[row] += fma(alpha, ph[row], omega * sh[row]);
                const T ri = s[row] - omega * t[row];
                r[row] = ri;
                r2 += ri * ri;
                rhr += rh[row] * ri;
            }
            block_sum2<T, NThreads>(r2, rhr, red());
            res = sqrt(r2);
            rho_old = rho_new;
            rho_new = rhr;
        }
        }
        __syncthreads();
        for (int row = tid; row < n; row += NThreads) gx[row] = xl[row];
        if (tid == 0) {
            iters_out[sys] = iter;
            res_out[sys] = res;
        }
    }
    // the last workgroup out resets the queue for the next launch
    if (tid == 0) {
        const unsigned d = atomicAdd(queue + 1, 1u);
        if (d == gridDim.x - 1) {
            queue[0] = 0;
            queue[1] = 0;
        }
    }
}

// The same solvers with the work vectors in registers (Csr formats).  Thread
// t owns rows t, t + NT, ..., t + (R - 1) NT of the system and keeps every
// work vector's entries of those rows in registers; only the vectors the
// products gather -- p_hat and s_hat for BiCGSTAB, p for CG -- go to LDS,
// beside the system's matrix (values, 16-bit local columns, row pointers).
// At about 10 bytes per stored entry plus 16 (FP64) or 8 (FP32) bytes per
// row, a 512-row system of 9 entries per row takes 54 KB instead of the
// 86 KB of the LDS-vector kernel, and the vector updates touch no LDS.  A
// product sums each row by one thread in the stored order, as the
// reference's batch Csr product does (common/cuda_hip/matrix/
// batch_csr_kernels.hpp).  The iteration is the LDS-vector kernel's: the
// reference's batch BiCGSTAB (no breakdown guards, the half-step exit) or
// batch CG (implicit norm sqrt|r.z|, iterations counted from 1).
//
// With K > 0 the matrix rows stay in registers as well: the thread holding
// row i keeps its first K stored entries (values and columns), reads any
// further ones from global memory in the products, and LDS holds only the
// gathered vectors -- one LDS access per entry in a product instead of three.
//
// LDS layout per system slot (n = lds_rows, z = lds_nnz):
//   T in0[n], T in1[n] (BiCGSTAB only), then for K == 0
//   T vals[z], int32 rp[n+1], uint16 ci[z]
template <typename T, int NT, int R, int K, int Fmt, int Method>
__global__ void __launch_bounds__(NT)
    batch_reg_kernel(int64_t num_systems, const int32_t* __restrict__ order,
                     const int64_t* __restrict__ row_off, const int32_t* __restrict__ rp,
                     const int32_t* __restrict__ ci, const T* __restrict__ vals, int64_t uni_n,
                     int64_t uni_nnz, const T* __restrict__ b, T* __restrict__ x,
                     int32_t max_iters, int32_t relative, double tol_d, int32_t jacobi,
                     int32_t* __restrict__ iters_out, T* __restrict__ res_out,
                     unsigned* __restrict__ queue, int64_t lds_rows, int64_t lds_nnz,
                     int32_t dynamic)
{
    static_assert(Fmt == kFmtUneven || Fmt == kFmtCsr, "register kernel: Csr formats");
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ T red2[4 * (NT / kWave)];
    int red_turn = 0;
    auto red = [&]() {
        red_turn ^= 1;
        return red2 + red_turn * 2 * (NT / kWave);
    };
    __shared__ int64_t slot;
    const int tid = threadIdx.x;
    const int64_t L = lds_rows;
    constexpr int kIn = Method == kBatchCg ? 1 : 2;
    // K > 0: each gathered vector has one more slot that stays zero (index
    // kZero); a register slot past its row's end points there with value 0,
    // so a product slot is one fused multiply-add, without a compare or a
    // select.  In FP64 the gathered vectors sit in static LDS at a fixed
    // stride and a register slot holds its column's byte offset in the
    // first vector: the same register addresses the second through the LDS
    // instruction's immediate offset, so a slot needs no decoding.  In FP32
    // (measured faster that way) they sit in dynamic LDS, L + 1 values
    // each, and the columns are packed two per register.
    constexpr bool kOff = sizeof(T) >= 8;
    constexpr bool kStatic = K > 0 && kOff;
    constexpr int kStride = R * NT + 1;
    __shared__ __align__(16) T stat_in[kStatic ? kIn : 1][kStatic ? kStride : 1];
    T* in0 = kStatic ? &stat_in[0][0] : reinterpret_cast<T*>(smem);
    T* in1 = kStatic ? &stat_in[kIn - 1][0] : in0 + L + (K > 0 ? 1 : 0);
    const int32_t kZero = kStatic ? R * NT : static_cast<int32_t>(L);
    T* lv = in0 + kIn * L;
    int32_t* lrp = reinterpret_cast<int32_t*>(lv + lds_nnz);
    uint16_t* lci = reinterpret_cast<uint16_t*>(lrp + L + 1);
    const T tol = static_cast<T>(tol_d);
    if constexpr (K > 0) {
        if (tid == 0) {
            in0[kZero] = T(0);
            in1[kZero] = T(0);
        }
    }

    // the first system of a workgroup is its block index; further ones come
    // from the queue (numbered after the grid) only when the launch has more
    // systems than workgroups, so a launch with one system per workgroup
    // touches no atomic counter
    bool first = true;
    while (true) {
        __syncthreads();
        if (tid == 0) {
            slot = first ? static_cast<int64_t>(blockIdx.x)
                         : (dynamic ? static_cast<int64_t>(gridDim.x) +
                                          static_cast<int64_t>(atomicAdd(queue, 1u))
                                    : num_systems);
        }
        first = false;
        __syncthreads();
        const int64_t k = slot;
        if (k >= num_systems) break;
        const int64_t sys = order ? order[k] : k;
        int32_t n;
        int64_t row0, nnz0;
        const int32_t* srp;
        int32_t rp_base, nnz;
        if constexpr (Fmt == kFmtCsr) {
            n = static_cast<int32_t>(uni_n);
            nnz = static_cast<int32_t>(uni_nnz);
            row0 = sys * uni_n;
            nnz0 = sys * uni_nnz;
            srp = rp;
            rp_base = 0;
        } else {
            row0 = row_off[sys];
            n = static_cast<int32_t>(row_off[sys + 1] - row0);
            srp = rp + row0;
            rp_base = srp[0];
            nnz = srp[n] - rp_base;
            nnz0 = rp_base;
        }
        const T* gv = vals + nnz0;
        const int32_t* gci = ci + (Fmt == kFmtCsr ? 0 : nnz0);
        const T* gb = b + row0;
        T* gx = x + row0;
        // K > 0: the own rows' first K entries in registers; in FP64 each
        // column as its byte offset in a gathered vector (no decoding in the
        // products), in FP32 the columns packed two per register (entry q in
        // bits 16 (q & 1) ..; fewer registers, which keep the occupancy)
        constexpr int KR = K > 0 ? K : 1;
        constexpr int KM = kOff ? KR : (KR + 1) / 2;
        T mv[R][KR];
        uint32_t mc[R][KM];
        int32_t mb[R], ml[R];
        auto col_is = [&](int j, int q, int col) {
            if constexpr (kOff) {
                return mc[j][q] == static_cast<uint32_t>(col * static_cast<int>(sizeof(T)));
            } else {
                return static_cast<int>((mc[j][q / 2] >> (16 * (q & 1))) & 0xFFFFu) == col;
            }
        };
        if constexpr (K > 0) {
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const int row = tid + j * NT;
                mb[j] = row < n ? srp[row] - rp_base : 0;
                ml[j] = row < n ? srp[row + 1] - rp_base - mb[j] : 0;
#pragma unroll
                for (int q = 0; q < KM; ++q) mc[j][q] = 0;
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    const bool in_row = q < ml[j];
                    mv[j][q] = in_row ? gv[mb[j] + q] : T(0);
                    const uint32_t c = static_cast<uint32_t>(in_row ? gci[mb[j] + q] : kZero);
                    if constexpr (kOff) {
                        mc[j][q] = c * static_cast<uint32_t>(sizeof(T));
                    } else {
                        mc[j][q / 2] |= c << (16 * (q & 1));
                    }
                }
            }
        } else {
            for (int i = tid; i < nnz; i += NT) {
                lv[i] = gv[i];
                lci[i] = static_cast<uint16_t>(gci[i]);
            }
            for (int i = tid; i <= n; i += NT) lrp[i] = srp[i] - rp_base;
        }
        // own rows: prec, x, b
        T prec[R], xr[R], r[R];
        T bb = 0, unused = 0;
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int row = tid + j * NT;
            const bool own = row < n;
            xr[j] = own ? gx[row] : T(0);
            r[j] = own ? gb[row] : T(0);
            bb += r[j] * r[j];
            if (own) in0[row] = xr[j];
        }
        __syncthreads();
        // the scalar Jacobi of the reference's batch Csr (1 for a zero or
        // missing diagonal), or the identity
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int row = tid + j * NT;
            T d = T(1);
            if (jacobi && row < n) {
                if constexpr (K > 0) {
                    // the first stored diagonal entry, as the stored-order scan
                    bool found = false;
#pragma unroll
                    for (int q = 0; q < K; ++q) {
                        if (!found && q < ml[j] && col_is(j, q, row)) {
                            found = true;
                            if (mv[j][q] != T(0)) d = T(1) / mv[j][q];
                        }
                    }
                    for (int32_t e = mb[j] + K; !found && e < mb[j] + ml[j]; ++e) {
                        if (gci[e] == row) {
                            found = true;
                            if (gv[e] != T(0)) d = T(1) / gv[e];
                        }
                    }
                } else {
                    const int32_t e1 = lrp[row + 1];
                    for (int32_t e = lrp[row]; e < e1; ++e) {
                        if (lci[e] == row) {
                            if (lv[e] != T(0)) d = T(1) / lv[e];
                            break;
                        }
                    }
                }
            }
            prec[j] = d;
        }
        // out = A in at the own rows, each row summed in stored order
        auto spmv = [&](const T* in, T out[R]) {
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const int row = tid + j * NT;
                T sum = 0;
                if constexpr (K > 0) {
                    // a slot past the row's end reads the zero slot with
                    // value 0: the row sum is the stored-order sum of the
                    // row's own entries
                    if constexpr (kOff) {
                        const char* base = reinterpret_cast<const char*>(in);
#pragma unroll
                        for (int q = 0; q < K; ++q) {
                            sum = fma(mv[j][q], *reinterpret_cast<const T*>(base + mc[j][q]),
                                      sum);
                        }
                    } else {
#pragma unroll
                        for (int qq = 0; qq < KM; ++qq) {
                            // the empty asm keeps the packed word opaque, so
                            // the compiler decodes it here instead of holding
                            // every entry's LDS address in a register
                            uint32_t w = mc[j][qq];
                            asm volatile("" : "+v"(w));
                            sum = fma(mv[j][2 * qq], in[w & 0xFFFFu], sum);
                            if (2 * qq + 1 < K) sum = fma(mv[j][2 * qq + 1], in[w >> 16], sum);
                        }
                    }
                    for (int32_t e = mb[j] + K; e < mb[j] + ml[j]; ++e) sum += gv[e] * in[gci[e]];
                } else if (row < n) {
                    const int32_t e1 = lrp[row + 1];
                    for (int32_t e = lrp[row]; e < e1; ++e) sum += lv[e] * in[lci[e]];
                }
                out[j] = sum;
            }
        };
        block_sum2<T, NT>(bb, unused, red());
        const T rhs_norm = sqrt(bb);
        const T limit = relative ? tol * rhs_norm : tol;
        int32_t iter = 0;
        T res = 0;
        if constexpr (Method == kBatchCg) {
            T z[R], p[R], ap[R];
            spmv(in0, ap);
            T rz = 0;
#pragma unroll
            for (int j = 0; j < R; ++j) {
                r[j] -= ap[j];
                p[j] = T(0);
                z[j] = prec[j] * r[j];
                rz += r[j] * z[j];
            }
            T rho_old = 1;
            while (true) {
                // the first block sum also orders the reads of x in in0
                // (the product above) before p overwrites it
                const T rho = block_sum1<T, NT>(rz, red());
                ++iter;
                res = sqrt(rho < T(0) ? -rho : rho);
                if (iter >= max_iters || res <= limit) break;
                const bool restart = rho_old == T(0);
                const T beta = restart ? T(0) : rho / rho_old;
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    const int row = tid + j * NT;
                    p[j] = restart ? z[j] : z[j] + beta * p[j];
                    if (row < n) in0[row] = p[j];
                }
                __syncthreads();
                spmv(in0, ap);
                T pap = 0;
#pragma unroll
                for (int j = 0; j < R; ++j) pap += p[j] * ap[j];
                pap = block_sum1<T, NT>(pap, red());
                const T temp = rho / pap;
                rz = 0;
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    xr[j] += temp * p[j];
                    r[j] -= temp * ap[j];
                    z[j] = prec[j] * r[j];
                    rz += r[j] * z[j];
                }
                rho_old = rho;
            }
        } else {
            // p_hat = M p and s_hat = M s are formed again where x is
            // updated (the same products, so the same values) instead of
            // being held in registers across the products
            T rh[R], p[R], v[R], s[R], t[R];
            spmv(in0, v);
            T rr = 0;
#pragma unroll
            for (int j = 0; j < R; ++j) {
                r[j] -= v[j];
                rh[j] = r[j];
                p[j] = T(0);
                v[j] = T(0);
                rr += r[j] * r[j];
            }
            T rho_new = rr;
            // also orders the reads of x in in0 before p_hat overwrites it
            block_sum2<T, NT>(rr, rho_new, red());
            res = sqrt(rr);
            T rho_old = 1, omega = 1, alpha = 1;
            for (; iter < max_iters; ++iter) {
                if (res <= limit) break;
                const T beta = (rho_new / rho_old) * (alpha / omega);
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    const int row = tid + j * NT;
                    p[j] = r[j] + beta * (p[j] - omega * v[j]);
                    if (row < n) in0[row] = prec[j] * p[j];
                }
                __syncthreads();
                spmv(in0, v);
                T rv = 0;
#pragma unroll
                for (int j = 0; j < R; ++j) rv += rh[j] * v[j];
                rv = block_sum1<T, NT>(rv, red());
                alpha = rho_new / rv;
                T ss = 0;
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    const int row = tid + j * NT;
                    s[j] = r[j] - alpha * v[j];
                    if (row < n) in1[row] = prec[j] * s[j];
                    ss += s[j] * s[j];
                }
                // publishes s_hat for the product below as well
                ss = block_sum1<T, NT>(ss, red());
                res = sqrt(ss);
                if (res <= limit) {
#pragma unroll
                    for (int j = 0; j < R; ++j) xr[j] += alpha * (prec[j] * p[j]);
                    break;
                }
                spmv(in1, t);
                T ts = 0, tt = 0;
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    ts += t[j] * s[j];
                    tt += t[j] * t[j];
                }
                block_sum2<T, NT>(ts, tt, red());
                omega = ts / tt;
                T r2 = 0, rhr = 0;
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    xr[j] += fma(alpha, prec[j] * p[j], omega * (prec[j] * s[j]));
                    r[j] = s[j] - omega * t[j];
                    r2 += r[j] * r[j];
                    rhr += rh[j] * r[j];
                }
                block_sum2<T, NT>(r2, rhr, red());
                res = sqrt(r2);
                rho_old = rho_new;
                rho_new = rhr;
            }
        }
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int row = tid + j * NT;
            if (row < n) gx[row] = xr[j];
        }
        if (tid == 0) {
            iters_out[sys] = iter;
            res_out[sys] = res;
        }
    }
    // the last workgroup out resets the queue for the next launch
    if (dynamic && tid == 0) {
        const unsigned d = atomicAdd(queue + 1, 1u);
        if (d == gridDim.x - 1) {
            queue[0] = 0;
            queue[1] = 0;
        }
    }
}

// ----------------------------------------------------- one-launch kernel ----

// Work items of the one-launch kernel.  Bin q of the plan holds the items
// [item_begin[q], item_begin[q+1]); item i of bin q is the systems
// k = sys_begin[q] + (i - item_begin[q]) * (256 >> shift[q]) + g of the size
// order, g = 0 .. (256 >> shift[q]) - 1, those with k < sys_begin[q+1], one
// per group of 2^shift[q] threads (each system has at most 2 * 2^shift[q]
// rows).
struct batch_items {
    int32_t num_bins;
    int32_t shift[GKOMI_BATCH_MAX_BINS];
    int64_t item_begin[GKOMI_BATCH_MAX_BINS + 1];
    int64_t sys_begin[GKOMI_BATCH_MAX_BINS + 1];
};

// Sums a and b over each group of L = 2^shift threads (L = 32, 64, 128 or
// 256); every thread receives its group's sums.  A wave's rows of 16 lanes
// are reduced by DPP, then row_bcast:15 leaves each half-wave's total in
// lanes 31 and 63 and, for groups of a wave or more, row_bcast:31 the wave's
// total in lane 63; those lanes write one partial each, and after one barrier
// a group adds its partials in order (for L = 256 the order of block_sum2).
// The caller alternates between two `lds` buffers of 16 values.
template <typename T>
__device__ __forceinline__ void group_sum2_lds(T& a, T& b, T* lds, int shift)
{
    const int tid = threadIdx.x;
    a = group_sum<16>(a);
    b = group_sum<16>(b);
    a += dpp_rows<0x142, 0xA>(a);
    b += dpp_rows<0x142, 0xA>(b);
    const int pshift = shift < 6 ? 5 : 6;  // threads per partial: 32 or 64
    if (pshift == 6) {
        a += dpp_rows<0x143, 0xC>(a);
        b += dpp_rows<0x143, 0xC>(b);
    }
    if ((tid & ((1 << pshift) - 1)) == (1 << pshift) - 1) {
        lds[tid >> pshift] = a;
        lds[8 + (tid >> pshift)] = b;
    }
    __syncthreads();
    const int per = 1 << (shift - pshift);
    const int first = (tid >> shift) * per;
    T s0 = lds[first], s1 = lds[8 + first];
    for (int w = 1; w < per; ++w) {
        s0 += lds[first + w];
        s1 += lds[8 + first + w];
    }
    a = s0;
    b = s1;
}

template <typename T>
__device__ __forceinline__ T group_sum1_lds(T a, T* lds, int shift)
{
    const int tid = threadIdx.x;
    a = group_sum<16>(a);
    a += dpp_rows<0x142, 0xA>(a);
    const int pshift = shift < 6 ? 5 : 6;
    if (pshift == 6) a += dpp_rows<0x143, 0xC>(a);
    if ((tid & ((1 << pshift) - 1)) == (1 << pshift) - 1) lds[tid >> pshift] = a;
    __syncthreads();
    const int per = 1 << (shift - pshift);
    const int first = (tid >> shift) * per;
    T s = lds[first];
    for (int w = 1; w < per; ++w) s += lds[first + w];
    return s;
}

// The whole uneven Csr batch in one persistent launch: 256 threads, two rows
// per thread, the matrix rows (up to K entries, local columns packed two per
// register) and the work vectors in registers, only the gathered vectors in
// LDS.  A workgroup takes its first work item by block index and further
// ones from one queue ordered largest first, so the last work of the launch
// is its smallest systems and no size bin waits for another's tail.  An item
// is one system of 257..512 rows or 256 / L systems of up to 2 L rows (L =
// 128, 64, 32 threads each); the systems of one item iterate side by side,
// every group stepping its own system with the semantics of the register
// kernel (BiCGSTAB or CG, iteration counts and residual norms logged per
// system) and idling once it has stopped, until every group of the item has.
template <typename T, int K, int Method>
__global__ void __launch_bounds__(256)
    batch_packed_kernel(int64_t num_items, batch_items items, const int32_t* __restrict__ order,
                        const int64_t* __restrict__ row_off, const int32_t* __restrict__ rp,
                        const int32_t* __restrict__ ci, const T* __restrict__ vals,
                        const T* __restrict__ b, T* __restrict__ x, int32_t max_iters,
                        int32_t relative, double tol_d, int32_t jacobi,
                        int32_t* __restrict__ iters_out, T* __restrict__ res_out,
                        unsigned* __restrict__ queue, int32_t dynamic)
{
    constexpr int NT = 256;
    constexpr int R = 2;
    constexpr int kIn = Method == kBatchCg ? 1 : 2;
    // the gathered vectors (p_hat, s_hat; CG p), and in FP32 also x, prec
    // and (BiCGSTAB) r_hat, p and v: a thread reads back only its own rows of
    // those, so they take no barrier, and the registers they free raise the
    // FP32 kernel from four to five waves per SIMD.  FP64 keeps them in
    // registers (measured faster at its three waves per SIMD).
    constexpr bool kLdsVec = sizeof(T) == 4;
    constexpr int kVec = kIn + (kLdsVec ? (Method == kBatchCg ? 2 : 5) : 0);
    // one more slot per vector, index R * NT, that stays zero: a register slot
    // past its row's end points there with value 0, so the product is one
    // fused multiply-add per slot, without a compare or a select.  A register
    // slot holds its column's byte offset in the first vector (the group's
    // rows included); the second vector is read through the LDS
    // instruction's immediate offset.
    __shared__ __align__(16) T in_s[kVec][R * NT + 1];
    __shared__ T red2[2 * 16];
    __shared__ int32_t go_flag[8];
    __shared__ int64_t slot;
    int red_turn = 0;
    auto red = [&]() {
        red_turn ^= 1;
        return red2 + red_turn * 16;
    };
    const int tid = threadIdx.x;
    const T tol = static_cast<T>(tol_d);
    if (tid == 0) {
        in_s[0][R * NT] = T(0);
        in_s[kIn - 1][R * NT] = T(0);
    }

    bool first = true;
    while (true) {
        __syncthreads();
        if (tid == 0) {
            slot = first ? static_cast<int64_t>(blockIdx.x)
                         : (dynamic ? static_cast<int64_t>(gridDim.x) +
                                          static_cast<int64_t>(atomicAdd(queue, 1u))
                                    : num_items);
        }
        first = false;
        __syncthreads();
        const int64_t item = slot;
        if (item >= num_items) break;
        int bin = 0;
        while (bin + 1 < items.num_bins && item >= items.item_begin[bin + 1]) ++bin;
        const int shift = items.shift[bin];
        const int L = 1 << shift;
        const int G = NT >> shift;
        const int g = tid >> shift;
        const int lt = tid & (L - 1);
        const int64_t k =
            items.sys_begin[bin] + (item - items.item_begin[bin]) * G + g;
        const bool valid = k < items.sys_begin[bin + 1];
        const int64_t sys = valid ? (order ? static_cast<int64_t>(order[k]) : k) : 0;
        const int64_t row0 = valid ? row_off[sys] : 0;
        const int32_t n = valid ? static_cast<int32_t>(row_off[sys + 1] - row0) : 0;
        const int32_t* srp = rp + row0;
        const int32_t rp_base = valid ? srp[0] : 0;
        const T* gv = vals + rp_base;
        const int32_t* gci = ci + rp_base;
        const T* gb = b + row0;
        T* gx = x + row0;
        T* in0 = in_s[0] + g * (R * L);
        T* in1 = in_s[kIn - 1] + g * (R * L);
        // own-row vectors: LDS rows (kLdsVec) or registers
        T* lx = in_s[kLdsVec ? kIn : 0] + g * (R * L);
        T* lprec = in_s[kLdsVec ? kIn + 1 : 0] + g * (R * L);
        T* lrh = in_s[kLdsVec ? kVec - 1 : 0] + g * (R * L);
        T* lp = in_s[kLdsVec && kIn + 2 < kVec ? kIn + 2 : 0] + g * (R * L);
        T* lv = in_s[kLdsVec && kIn + 3 < kVec ? kIn + 3 : 0] + g * (R * L);
        T x_r[R], prec_r[R], rh_r[R], p_r[R], v_r[R];
        auto own_row = [&](T* lds_rows, T* regs, int j) -> T& {
            if constexpr (kLdsVec) {
                return lds_rows[lt + j * L];
            } else {
                return regs[j];
            }
        };
        auto xr = [&](int j) -> T& { return own_row(lx, x_r, j); };
        auto prec = [&](int j) -> T& { return own_row(lprec, prec_r, j); };

        // the own rows' first K entries: in FP64 each column as its byte
        // offset from the start of a vector of in_s (the group's rows start
        // at first_col); in FP32 the group-local columns packed two per
        // register, read through the group's vector pointers (measured
        // faster in FP32)
        constexpr bool kOff = sizeof(T) >= 8;
        constexpr int KM = kOff ? K : (K + 1) / 2;
        T mv[R][K];
        uint32_t mc[R][KM];
        int32_t mb[R], ml[R];
        const int32_t first_col = kOff ? g * (R * L) : 0;
        // the zero slot: index R * NT of in_s, relative to the columns' origin
        const int32_t zero_col = R * NT - (kOff ? 0 : g * (R * L));
        auto col_is = [&](int j, int q, int col) {
            if constexpr (kOff) {
                return mc[j][q] == static_cast<uint32_t>(col * static_cast<int>(sizeof(T)));
            } else {
                return static_cast<int>((mc[j][q / 2] >> (16 * (q & 1))) & 0xFFFFu) == col;
            }
        };
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int row = lt + j * L;
            mb[j] = row < n ? srp[row] - rp_base : 0;
            ml[j] = row < n ? srp[row + 1] - rp_base - mb[j] : 0;
#pragma unroll
            for (int q = 0; q < KM; ++q) mc[j][q] = 0;
#pragma unroll
            for (int q = 0; q < K; ++q) {
                const bool in_row = q < ml[j];
                mv[j][q] = in_row ? gv[mb[j] + q] : T(0);
                const uint32_t c =
                    static_cast<uint32_t>(in_row ? first_col + gci[mb[j] + q] : zero_col);
                if constexpr (kOff) {
                    mc[j][q] = c * static_cast<uint32_t>(sizeof(T));
                } else {
                    mc[j][q / 2] |= c << (16 * (q & 1));
                }
            }
        }
        T r[R];
        T bb = 0;
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int row = lt + j * L;
            const bool own = row < n;
            const T xj = own ? gx[row] : T(0);
            r[j] = own ? gb[row] : T(0);
            bb += r[j] * r[j];
            xr(j) = xj;
            if (own) in0[row] = xj;
        }
        __syncthreads();
        // the scalar Jacobi of the reference's batch Csr: 1 / the first
        // stored diagonal entry, 1 for a zero or missing one
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int row = lt + j * L;
            T d = T(1);
            if (jacobi && row < n) {
                bool found = false;
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    if (!found && q < ml[j] && col_is(j, q, first_col + row)) {
                        found = true;
                        if (mv[j][q] != T(0)) d = T(1) / mv[j][q];
                    }
                }
                for (int32_t e = mb[j] + K; !found && e < mb[j] + ml[j]; ++e) {
                    if (gci[e] == row) {
                        found = true;
                        if (gv[e] != T(0)) d = T(1) / gv[e];
                    }
                }
            }
            prec(j) = d;
        }
        // out = A in at the own rows, each row summed in stored order, for
        // in = vector `which` of in_s (0 or kIn - 1)
        auto spmv = [&](int which, T out[R]) {
            const T* vec = in_s[which];
            const char* base = reinterpret_cast<const char*>(vec);
            const T* in = vec + g * (R * L);
#pragma unroll
            for (int j = 0; j < R; ++j) {
                T sum = 0;
                if constexpr (kOff) {
#pragma unroll
                    for (int q = 0; q < K; ++q) {
                        sum = fma(mv[j][q], *reinterpret_cast<const T*>(base + mc[j][q]), sum);
                    }
                } else {
#pragma unroll
                    for (int qq = 0; qq < KM; ++qq) {
                        uint32_t w = mc[j][qq];
                        asm volatile("" : "+v"(w));
                        sum = fma(mv[j][2 * qq], in[w & 0xFFFFu], sum);
                        if (2 * qq + 1 < K) sum = fma(mv[j][2 * qq + 1], in[w >> 16], sum);
                    }
                }
                for (int32_t e = mb[j] + K; e < mb[j] + ml[j]; ++e) sum += gv[e] * in[gci[e]];
                out[j] = sum;
            }
        };
        // every group's flag, published by the barrier that follows; true
        // while any group of the item still iterates
        auto any_going = [&](bool go) {
            if (lt == 0) go_flag[g] = go ? 1 : 0;
            __syncthreads();
            int32_t any = 0;
            for (int h = 0; h < G; ++h) any |= go_flag[h];
            return any != 0;
        };
        bb = group_sum1_lds(bb, red(), shift);
        const T rhs_norm = sqrt(bb);
        const T limit = relative ? tol * rhs_norm : tol;
        int32_t iter = 0;
        T res = 0;
        bool go = valid;
        if constexpr (Method == kBatchCg) {
            T z[R], p[R], ap[R];
            spmv(0, ap);
            T rz = 0;
#pragma unroll
            for (int j = 0; j < R; ++j) {
                r[j] -= ap[j];
                p[j] = T(0);
                z[j] = prec(j) * r[j];
                rz += r[j] * z[j];
            }
            T rho_old = 1;
            while (true) {
                // also orders the reads of x in in0 before p overwrites it
                const T rho = group_sum1_lds(go ? rz : T(0), red(), shift);
                if (go) {
                    ++iter;
                    res = sqrt(rho < T(0) ? -rho : rho);
                    if (iter >= max_iters || res <= limit) go = false;
                }
                if (go) {
                    const bool restart = rho_old == T(0);
                    const T beta = restart ? T(0) : rho / rho_old;
#pragma unroll
                    for (int j = 0; j < R; ++j) {
                        const int row = lt + j * L;
                        p[j] = restart ? z[j] : z[j] + beta * p[j];
                        if (row < n) in0[row] = p[j];
                    }
                }
                if (!any_going(go)) break;
                T pap = 0;
                if (go) {
                    spmv(0, ap);
#pragma unroll
                    for (int j = 0; j < R; ++j) pap += p[j] * ap[j];
                }
                pap = group_sum1_lds(pap, red(), shift);
                if (go) {
                    const T temp = rho / pap;
                    rz = 0;
#pragma unroll
                    for (int j = 0; j < R; ++j) {
                        xr(j) += temp * p[j];
                        r[j] -= temp * ap[j];
                        z[j] = prec(j) * r[j];
                        rz += r[j] * z[j];
                    }
                    rho_old = rho;
                }
            }
        } else {
            T s[R], t[R];
            auto rh = [&](int j) -> T& { return own_row(lrh, rh_r, j); };
            auto p = [&](int j) -> T& { return own_row(lp, p_r, j); };
            auto v = [&](int j) -> T& { return own_row(lv, v_r, j); };
            spmv(0, t);
            T rr = 0;
#pragma unroll
            for (int j = 0; j < R; ++j) {
                r[j] -= t[j];
                rh(j) = r[j];
                p(j) = T(0);
                v(j) = T(0);
                s[j] = T(0);
                rr += r[j] * r[j];
            }
            // also orders the reads of x in in0 before p_hat overwrites it
            rr = group_sum1_lds(rr, red(), shift);
            T rho_new = rr;
            res = sqrt(rr);
            T rho_old = 1, omega = 1, alpha = 1;
            while (true) {
                if (go && (iter >= max_iters || res <= limit)) go = false;
                if (go) {
                    const T beta = (rho_new / rho_old) * (alpha / omega);
#pragma unroll
                    for (int j = 0; j < R; ++j) {
                        const int row = lt + j * L;
                        p(j) = r[j] + beta * (p(j) - omega * v(j));
                        if (row < n) in0[row] = prec(j) * p(j);
                    }
                }
                if (!any_going(go)) break;
                T rv = 0;
                if (go) {
                    spmv(0, t);
#pragma unroll
                    for (int j = 0; j < R; ++j) {
                        v(j) = t[j];
                        rv += rh(j) * t[j];
                    }
                }
                rv = group_sum1_lds(rv, red(), shift);
                T ss = 0;
                if (go) {
                    alpha = rho_new / rv;
#pragma unroll
                    for (int j = 0; j < R; ++j) {
                        const int row = lt + j * L;
                        s[j] = r[j] - alpha * v(j);
                        if (row < n) in1[row] = prec(j) * s[j];
                        ss += s[j] * s[j];
                    }
                }
                // publishes s_hat for the product below as well
                ss = group_sum1_lds(ss, red(), shift);
                if (go) {
                    res = sqrt(ss);
                    if (res <= limit) {
                        // the half-step exit: x += alpha p_hat, the
                        // iteration not counted
#pragma unroll
                        for (int j = 0; j < R; ++j) xr(j) += alpha * (prec(j) * p(j));
                        go = false;
                    }
                }
                T ts = 0, tt = 0;
                if (go) {
                    spmv(kIn - 1, t);
#pragma unroll
                    for (int j = 0; j < R; ++j) {
                        ts += t[j] * s[j];
                        tt += t[j] * t[j];
                    }
                }
                group_sum2_lds(ts, tt, red(), shift);
                T r2 = 0, rhr = 0;
                if (go) {
                    omega = ts / tt;
#pragma unroll
                    for (int j = 0; j < R; ++j) {
                        const T pj = prec(j);
                        xr(j) += fma(alpha, pj * p(j), omega * (pj * s[j]));
                        r[j] = s[j] - omega * t[j];
                        r2 += r[j] * r[j];
                        rhr += rh(j) * r[j];
                    }
                }
                group_sum2_lds(r2, rhr, red(), shift);
                if (go) {
                    res = sqrt(r2);
                    rho_old = rho_new;
                    rho_new = rhr;
                    ++iter;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int row = lt + j * L;
            if (row < n) gx[row] = xr(j);
        }
        if (valid && lt == 0) {
            iters_out[sys] = iter;
            res_out[sys] = res;
        }
    }
    // the last workgroup out resets the queue for the next launch
    if (dynamic && tid == 0) {
        const unsigned d = atomicAdd(queue + 1, 1u);
        if (d == gridDim.x - 1) {
            queue[0] = 0;
            queue[1] = 0;
        }
    }
}

// ------------------------------------------------------------ launcher ----

constexpr int kBatchMaxBins = GKOMI_BATCH_MAX_BINS;
constexpr size_t kBatchQueueBytes = 64;  // per bin: queue + done counters

size_t batch_lds_bytes(int vbytes, int64_t rows, int64_t nnz, bool mat_lds,
                       int method = kBatchBicgstab, int fmt = kFmtCsr)
{
    size_t b = static_cast<size_t>(vbytes) * batch_vectors(method) * rows;
    if (mat_lds) {
        b += static_cast<size_t>(vbytes) * nnz;
        if (fmt == kFmtUneven || fmt == kFmtCsr) b += 4 * (rows + 1);
        if (fmt != kFmtDense) b += 2 * nnz;
    }
    return (b + 15) / 16 * 16;
}

// Lanes per row of the Csr product: the plan's value, or the largest power
// of two up to 16 that leaves each lane more than four entries of an average
// row of the bin (measured on 20,000 uneven systems of 9 and 16 entries per
// row: 2 lanes per row beat 1, 4 and 8).
int32_t batch_subwave(int32_t requested, int64_t rows, int64_t nnz)
{
    if (requested > 0) return requested;
    const int64_t avg = rows > 0 ? nnz / rows : 0;
    int32_t sw = 1;
    while (sw < 16 && int64_t{8} * sw < avg) sw *= 2;
    return sw;
}

// LDS of the register-vector kernel: the gathered vectors (two for
// BiCGSTAB, one for CG), values, row pointers and 16-bit columns.
size_t batch_reg_lds_bytes(int vbytes, int64_t rows, int64_t nnz, int method)
{
    const int vecs = method == kBatchCg ? 1 : 2;
    const size_t b = static_cast<size_t>(vbytes) * (vecs * rows + nnz) + 4 * (rows + 1) + 2 * nnz;
    return (b + 15) / 16 * 16;
}

int fmt_of(const gkomi_batch_csr* a)
{
    if (a->format == GKOMI_BATCH_ELL) return kFmtEll;
    if (a->format == GKOMI_BATCH_DENSE) return kFmtDense;
    return a->row_offsets ? kFmtUneven : kFmtCsr;
}

template <typename T, int NT, bool MatLds, int Fmt, int Method>
int launch_batch(hipStream_t stream, int grid, size_t lds, int64_t count, const int32_t* order,
                 const gkomi_batch_csr* a, const T* b, T* x, const gkomi_batch_settings* st,
                 int32_t* iters, T* res, unsigned* queue, int64_t lds_rows, int64_t lds_nnz,
                 int32_t sw)
{
    auto kern = batch_solver_kernel<T, NT, MatLds, Fmt, Method>;
    if (lds > 64 * 1024) {
        GKOMI_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                            hipFuncAttributeMaxDynamicSharedMemorySize,
                                            static_cast<int>(lds)));
    }
    kern<<<grid, NT, lds, stream>>>(count, order, a->row_offsets, a->item_offsets, a->row_ptrs,
                                    a->col_idxs,
                                    static_cast<const T*>(a->values), a->uniform_rows,
                                    a->uniform_nnz, a->ell_width, b, x, st->max_iterations,
                                    st->tolerance_type,
                                    st->tolerance, st->preconditioner, iters, res, queue,
                                    lds_rows, lds_nnz, sw);
    GKOMI_CHECK_LAUNCH("batch_solver_kernel");
    return GKOMI_SUCCESS;
}

template <typename T, bool MatLds, int Fmt, int Method>
int launch_batch_nt(int nt, hipStream_t stream, int grid, size_t lds, int64_t count,
                    const int32_t* order, const gkomi_batch_csr* a, const T* b, T* x,
                    const gkomi_batch_settings* st, int32_t* iters, T* res, unsigned* queue,
                    int64_t lds_rows, int64_t lds_nnz, int32_t sw)
{
    switch (nt) {
    case 64:
        return launch_batch<T, 64, MatLds, Fmt, Method>(stream, grid, lds, count, order, a, b, x, st,
                                                    iters, res, queue, lds_rows, lds_nnz, sw);
    case 128:
        return launch_batch<T, 128, MatLds, Fmt, Method>(stream, grid, lds, count, order, a, b, x,
                                                     st, iters, res, queue, lds_rows, lds_nnz, sw);
    case 256:
        return launch_batch<T, 256, MatLds, Fmt, Method>(stream, grid, lds, count, order, a, b, x,
                                                     st, iters, res, queue, lds_rows, lds_nnz, sw);
    case 512:
        return launch_batch<T, 512, MatLds, Fmt, Method>(stream, grid, lds, count, order, a, b, x,
                                                     st, iters, res, queue, lds_rows, lds_nnz, sw);
    default:
        return fail(GKOMI_INVALID_ARGUMENT, "batch: block size must be 64, 128, 256 or 512");
    }
}

int device_lds_limit()
{
    int dev = 0;
    int v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        v <= 0) {
        return 64 * 1024;
    }
    return v;
}

// The matrix part of a descriptor (solvers and products).
int check_batch_matrix(const gkomi_batch_csr* a)
{
    GKOMI_REQUIRE(a, GKOMI_INVALID_ARGUMENT, "batch: null descriptor");
    GKOMI_REQUIRE(a->num_systems >= 0 && a->max_rows >= 0 && a->max_nnz >= 0 &&
                      a->uniform_cols >= 0,
                  GKOMI_INVALID_ARGUMENT, "batch: negative sizes");
    GKOMI_REQUIRE(a->format == GKOMI_BATCH_CSR || a->format == GKOMI_BATCH_ELL ||
                      a->format == GKOMI_BATCH_DENSE,
                  GKOMI_INVALID_ARGUMENT, "batch: format must be CSR, ELL or DENSE");
    const bool uniform = a->row_offsets == nullptr;
    GKOMI_REQUIRE(uniform || a->format == GKOMI_BATCH_CSR || a->item_offsets != nullptr,
                  GKOMI_INVALID_ARGUMENT,
                  "batch: uneven Ell and Dense batches need item_offsets");
    GKOMI_REQUIRE(!uniform || (a->uniform_rows == a->max_rows && a->uniform_nnz == a->max_nnz),
                  GKOMI_INVALID_ARGUMENT, "batch: uniform batch needs max_rows/max_nnz = n/nnz");
    GKOMI_REQUIRE(uniform || a->uniform_cols == 0, GKOMI_INVALID_ARGUMENT,
                  "batch: uneven systems are square (uniform_cols must be 0)");
    GKOMI_REQUIRE(a->max_rows <= 65535, GKOMI_NOT_SUPPORTED,
                  "batch: systems are limited to 65535 rows (16-bit local indices)");
    GKOMI_REQUIRE(a->max_nnz < (int64_t{1} << 31), GKOMI_NOT_SUPPORTED,
                  "batch: at most 2^31 - 1 stored values per system");
    const int64_t cols = a->uniform_cols ? a->uniform_cols : a->uniform_rows;
    if (!uniform && a->format != GKOMI_BATCH_CSR) {
        GKOMI_REQUIRE(a->num_systems == 0 || a->max_nnz == 0 ||
                          (a->values && (a->format == GKOMI_BATCH_DENSE || a->col_idxs)),
                      GKOMI_INVALID_ARGUMENT, "batch: null column indices or values");
    } else if (a->format == GKOMI_BATCH_ELL) {
        GKOMI_REQUIRE(a->ell_width >= 0 && a->uniform_nnz == a->uniform_rows * a->ell_width,
                      GKOMI_INVALID_ARGUMENT,
                      "batch Ell: uniform_nnz must be uniform_rows * ell_width");
        GKOMI_REQUIRE(a->num_systems == 0 || a->uniform_nnz == 0 || (a->col_idxs && a->values),
                      GKOMI_INVALID_ARGUMENT, "batch Ell: null column indices or values");
    } else if (a->format == GKOMI_BATCH_DENSE) {
        GKOMI_REQUIRE(a->uniform_nnz == a->uniform_rows * cols, GKOMI_INVALID_ARGUMENT,
                      "batch Dense: uniform_nnz must be uniform_rows * uniform_cols");
        GKOMI_REQUIRE(a->num_systems == 0 || a->uniform_nnz == 0 || a->values,
                      GKOMI_INVALID_ARGUMENT, "batch Dense: null values");
    } else {
        GKOMI_REQUIRE(a->num_systems == 0 || (a->row_ptrs && a->col_idxs && a->values),
                      GKOMI_INVALID_ARGUMENT, "batch: null matrix arrays");
    }
    return GKOMI_SUCCESS;
}

int check_batch(const gkomi_batch_csr* a, const gkomi_batch_settings* st)
{
    GKOMI_REQUIRE(a && st, GKOMI_INVALID_ARGUMENT, "batch: null descriptor or settings");
    if (int rc = check_batch_matrix(a)) return rc;
    GKOMI_REQUIRE(a->uniform_cols == 0 || a->uniform_cols == a->uniform_rows,
                  GKOMI_DIMENSION_MISMATCH, "batch solver: the systems must be square");
    GKOMI_REQUIRE(st->tolerance_type == 0 || st->tolerance_type == 1, GKOMI_INVALID_ARGUMENT,
                  "batch: tolerance_type must be absolute (0) or relative (1)");
    GKOMI_REQUIRE(st->preconditioner == 0 || st->preconditioner == 1, GKOMI_NOT_SUPPORTED,
                  "batch: preconditioner must be identity (0) or scalar Jacobi (1)");
    GKOMI_REQUIRE(st->max_iterations >= 0, GKOMI_INVALID_ARGUMENT, "batch: negative max_iterations");
    return GKOMI_SUCCESS;
}

// Workgroups of `kern` resident per CU with `lds` bytes of dynamic LDS
// (cached per kernel and LDS size).
int resident_per_cu(const void* kern, int nt, size_t lds)
{
    static std::mutex mtx;
    static std::vector<std::pair<std::pair<const void*, size_t>, int>> cache;
    std::lock_guard<std::mutex> guard(mtx);
    for (const auto& e : cache) {
        if (e.first.first == kern && e.first.second == lds) return e.second;
    }
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, nt, lds) != hipSuccess ||
        nb < 1) {
        (void)hipGetLastError();
        nb = 1;
    }
    cache.push_back({{kern, lds}, nb});
    return nb;
}

template <typename T, int NT, int R, int K, int Fmt, int Method>
int launch_reg(hipStream_t stream, int grid, size_t lds, int64_t count, const int32_t* order,
               const gkomi_batch_csr* a, const T* b, T* x, const gkomi_batch_settings* st,
               int32_t* iters, T* res, unsigned* queue, int64_t lds_rows, int64_t lds_nnz)
{
    auto kern = batch_reg_kernel<T, NT, R, K, Fmt, Method>;
    if (lds > 64 * 1024) {
        GKOMI_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                            hipFuncAttributeMaxDynamicSharedMemorySize,
                                            static_cast<int>(lds)));
    }
    // grid = 0: every workgroup resident at once (occupancy of this kernel)
    if (grid <= 0) {
        const int64_t res_wg =
            static_cast<int64_t>(num_cus()) *
            resident_per_cu(reinterpret_cast<const void*>(kern), NT, lds);
        grid = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(count, res_wg)));
    }
    const int32_t dynamic = count > grid ? 1 : 0;
    kern<<<grid, NT, lds, stream>>>(count, order, a->row_offsets, a->row_ptrs, a->col_idxs,
                                    static_cast<const T*>(a->values), a->uniform_rows,
                                    a->uniform_nnz, b, x, st->max_iterations, st->tolerance_type,
                                    st->tolerance, st->preconditioner, iters, res, queue,
                                    lds_rows, lds_nnz, dynamic);
    GKOMI_CHECK_LAUNCH("batch_reg_kernel");
    return GKOMI_SUCCESS;
}

// Rows per thread of the register kernel for a bin: 1 or 2 (0 = does not apply).
int reg_rows_per_thread(int nt, int64_t rows)
{
    if (rows <= nt) return 1;
    if (rows <= 2 * int64_t{nt}) return 2;
    return 0;
}

// Entries per row the register kernel keeps in registers (10, 12, 16 or 32;
// 0 = the matrix in LDS) for `nt` threads with `rpt` rows each, a bin whose
// longest row has `max_row_nnz` entries (0 = unknown) and the plan's
// matrix_in_lds (1 = LDS; 0 = registers, longer rows finishing from global
// memory; -1 = registers when every row fits).  Two rows per thread keep at
// most 16 entries each, and only with 256 or 512 threads.  The smallest
// shape that holds the longest row leaves the most registers free, so the
// most workgroups resident per CU.
int reg_row_entries(int nt, int rpt, int64_t max_row_nnz, int mat)
{
    if (mat == 1 || !(rpt == 1 || (rpt == 2 && nt >= 256))) return 0;
    const int cap = rpt == 1 ? 32 : 16;
    if (max_row_nnz > 0 && max_row_nnz <= 10) return 10;
    if (max_row_nnz > 0 && max_row_nnz <= 12) return 12;
    if (max_row_nnz > 0 && max_row_nnz <= 16) return 16;
    if (max_row_nnz > 0 && max_row_nnz <= cap) return cap;
    return mat == 0 ? cap : 0;
}

template <typename T, int Fmt, int Method>
int launch_reg_shape(int nt, int rpt, int kreg, hipStream_t s, int grid, size_t lds, int64_t cnt,
                     const int32_t* order, const gkomi_batch_csr* a, const T* b, T* x,
                     const gkomi_batch_settings* st, int32_t* iters, T* res, unsigned* q,
                     int64_t rows, int64_t nnz)
{
#define GKOMI_REG_CASE(NTV, RV, KV)                                                            \
    if (nt == NTV && rpt == RV && kreg == KV)                                                  \
        return launch_reg<T, NTV, RV, KV, Fmt, Method>(s, grid, lds, cnt, order, a, b, x, st,   \
                                                       iters, res, q, rows, nnz);
#define GKOMI_REG_SHAPE(KV)                                                                    \
    GKOMI_REG_CASE(64, 1, KV)                                                                  \
    GKOMI_REG_CASE(128, 1, KV)                                                                 \
    GKOMI_REG_CASE(256, 1, KV)                                                                 \
    GKOMI_REG_CASE(512, 1, KV)
    GKOMI_REG_SHAPE(0)
    GKOMI_REG_CASE(64, 2, 0)
    GKOMI_REG_CASE(128, 2, 0)
    GKOMI_REG_CASE(256, 2, 0)
    GKOMI_REG_CASE(512, 2, 0)
    if constexpr (Fmt == kFmtUneven) {
        GKOMI_REG_SHAPE(10)
        GKOMI_REG_SHAPE(12)
        GKOMI_REG_SHAPE(16)
        GKOMI_REG_SHAPE(32)
        GKOMI_REG_CASE(256, 2, 10)
        GKOMI_REG_CASE(512, 2, 10)
        GKOMI_REG_CASE(256, 2, 12)
        GKOMI_REG_CASE(512, 2, 12)
        GKOMI_REG_CASE(256, 2, 16)
        GKOMI_REG_CASE(512, 2, 16)
    }
#undef GKOMI_REG_SHAPE
#undef GKOMI_REG_CASE
    return fail(GKOMI_INVALID_ARGUMENT,
                "batch: no register kernel for this block size, rows per thread and row length");
}

template <typename T, bool MatLds, int Method>
int launch_batch_fmt(int fmt, int nt, hipStream_t s, int grid, size_t lds, int64_t cnt,
                     const int32_t* order, const gkomi_batch_csr* a, const T* b, T* x,
                     const gkomi_batch_settings* st, int32_t* iters, T* res, unsigned* q,
                     int64_t rows, int64_t nnz, int32_t sw)
{
    switch (fmt) {
    case kFmtUneven:
        return launch_batch_nt<T, MatLds, kFmtUneven, Method>(nt, s, grid, lds, cnt, order, a, b,
                                                              x, st, iters, res, q, rows, nnz, sw);
    case kFmtCsr:
        return launch_batch_nt<T, MatLds, kFmtCsr, Method>(nt, s, grid, lds, cnt, order, a, b, x,
                                                           st, iters, res, q, rows, nnz, sw);
    case kFmtEll:
        return launch_batch_nt<T, MatLds, kFmtEll, Method>(nt, s, grid, lds, cnt, order, a, b, x,
                                                           st, iters, res, q, rows, nnz, sw);
    default:
        return launch_batch_nt<T, MatLds, kFmtDense, Method>(nt, s, grid, lds, cnt, order, a, b,
                                                             x, st, iters, res, q, rows, nnz, sw);
    }
}

// The one-launch form (batch_packed_kernel) applies to an uneven Csr batch of
// systems of at most 512 rows whose plan knows every bin's longest row and no
// row is longer than 16 entries; *kreg gets the register entries per row.
bool packed_applies(const gkomi_batch_csr* a, const gkomi_batch_plan* plan, int* kreg)
{
    if (fmt_of(a) != kFmtUneven || a->max_rows > 512) return false;
    if (plan->order == nullptr && plan->num_bins > 1) return false;
    int64_t longest = 0;
    for (int q = 0; q < plan->num_bins; ++q) {
        if (plan->bin_begin[q + 1] == plan->bin_begin[q]) continue;
        if (plan->bin_max_row_nnz[q] <= 0 || plan->bin_max_rows[q] > 512) return false;
        longest = std::max(longest, plan->bin_max_row_nnz[q]);
    }
    if (longest > 16) return false;
    *kreg = longest <= 10 ? 10 : (longest <= 12 ? 12 : 16);
    return true;
}

template <typename T, int K, int Method>
int launch_packed_k(hipStream_t s, const gkomi_batch_csr* a, const gkomi_batch_settings* st,
                    const gkomi_batch_plan* plan, const T* b, T* x, int32_t* iters, T* res,
                    unsigned* queue)
{
    batch_items it{};
    it.num_bins = plan->num_bins;
    int64_t total = 0;
    for (int q = 0; q < plan->num_bins; ++q) {
        const int64_t cnt = plan->bin_begin[q + 1] - plan->bin_begin[q];
        const int64_t half = (plan->bin_max_rows[q] + 1) / 2;
        int shift = 5;
        while ((int64_t{1} << shift) < half) ++shift;
        GKOMI_REQUIRE(shift <= 8, GKOMI_INVALID_ARGUMENT,
                      "batch (one launch): a bin's systems exceed 512 rows");
        it.shift[q] = shift;
        it.item_begin[q] = total;
        it.sys_begin[q] = plan->bin_begin[q];
        const int64_t per_item = 256 >> shift;
        total += (cnt + per_item - 1) / per_item;
    }
    it.item_begin[plan->num_bins] = total;
    it.sys_begin[plan->num_bins] = plan->bin_begin[plan->num_bins];
    if (total == 0) return GKOMI_SUCCESS;
    auto kern = batch_packed_kernel<T, K, Method>;
    const int64_t resident =
        static_cast<int64_t>(num_cus()) * resident_per_cu(reinterpret_cast<const void*>(kern), 256, 0);
    const int grid = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(total, resident)));
    const int32_t dynamic = total > grid ? 1 : 0;
    kern<<<grid, 256, 0, s>>>(total, it, plan->order, a->row_offsets, a->row_ptrs, a->col_idxs,
                              static_cast<const T*>(a->values), b, x, st->max_iterations,
                              st->tolerance_type, st->tolerance, st->preconditioner, iters, res,
                              queue, dynamic);
    GKOMI_CHECK_LAUNCH("batch_packed_kernel");
    return GKOMI_SUCCESS;
}

template <typename T, int Method>
int launch_packed(hipStream_t s, int kreg, const gkomi_batch_csr* a,
                  const gkomi_batch_settings* st, const gkomi_batch_plan* plan, const T* b, T* x,
                  int32_t* iters, T* res, unsigned* queue)
{
    switch (kreg) {
    case 10:
        return launch_packed_k<T, 10, Method>(s, a, st, plan, b, x, iters, res, queue);
    case 12:
        return launch_packed_k<T, 12, Method>(s, a, st, plan, b, x, iters, res, queue);
    default:
        return launch_packed_k<T, 16, Method>(s, a, st, plan, b, x, iters, res, queue);
    }
}

template <typename T, int Method>
int batch_solve_impl(const gkomi_ctx* ctx, const gkomi_batch_csr* a,
                     const gkomi_batch_settings* st, const gkomi_batch_plan* plan, const T* b,
                     T* x, int32_t* iters, T* res, void* ws, size_t wsb)
{
    if (int rc = enter(ctx)) return rc;
    if (int rc = check_batch(a, st)) return rc;
    if (a->num_systems == 0) return GKOMI_SUCCESS;
    GKOMI_REQUIRE(b && x && iters && res, GKOMI_INVALID_ARGUMENT, "batch: null vectors");
    GKOMI_REQUIRE(ws && wsb >= gkomi_batch_workspace_bytes(a), GKOMI_INVALID_ARGUMENT,
                  "batch: workspace too small (zero-filled, gkomi_batch_workspace_bytes)");
    gkomi_batch_plan single{};
    if (plan == nullptr) {
        single.num_bins = 1;
        single.bin_begin[0] = 0;
        single.bin_begin[1] = a->num_systems;
        single.bin_max_rows[0] = a->max_rows;
        single.bin_max_nnz[0] = a->max_nnz;
        single.block_size[0] = a->max_rows > 256 ? 256 : (a->max_rows > 64 ? 128 : 64);
        single.matrix_in_lds[0] = -1;
        single.blocks_per_cu[0] = 0;
        plan = &single;
    }
    GKOMI_REQUIRE(plan->num_bins >= 1 && plan->num_bins <= GKOMI_BATCH_MAX_BINS,
                  GKOMI_INVALID_ARGUMENT, "batch: bad number of bins");
    GKOMI_REQUIRE(plan->bin_begin[0] == 0 && plan->bin_begin[plan->num_bins] == a->num_systems,
                  GKOMI_INVALID_ARGUMENT, "batch: bins must cover all systems");
    const int fmt = fmt_of(a);
    const int lds_limit = device_lds_limit();
    hipStream_t s = stream_of(ctx);
    unsigned* queues = static_cast<unsigned*>(ws);
    GKOMI_REQUIRE(plan->single_launch == 0 || plan->single_launch == 1, GKOMI_INVALID_ARGUMENT,
                  "batch: single_launch must be 0 or 1");
    if (plan->single_launch == 1) {
        int kreg = 0;
        GKOMI_REQUIRE(packed_applies(a, plan, &kreg), GKOMI_NOT_SUPPORTED,
                      "batch: one launch needs an uneven Csr batch of systems of at most 512 "
                      "rows, an order, and every bin's longest row known and at most 16 entries");
        return launch_packed<T, Method>(s, kreg, a, st, plan, b, x, iters, res, queues);
    }
    for (int bin = 0; bin < plan->num_bins; ++bin) {
        const int64_t beg = plan->bin_begin[bin];
        const int64_t cnt = plan->bin_begin[bin + 1] - beg;
        if (cnt <= 0) continue;
        const int64_t rows = plan->bin_max_rows[bin];
        const int64_t nnz = plan->bin_max_nnz[bin];
        GKOMI_REQUIRE(rows <= a->max_rows && nnz <= a->max_nnz, GKOMI_INVALID_ARGUMENT,
                      "batch: bin maxima exceed the batch maxima");
        const int32_t* order_bin = plan->order ? plan->order + beg : nullptr;
        GKOMI_REQUIRE(plan->order != nullptr || plan->num_bins == 1, GKOMI_INVALID_ARGUMENT,
                      "batch: several bins need an order");
        unsigned* qb = queues + bin * (kBatchQueueBytes / sizeof(unsigned));
        {
            // work vectors in registers (Csr formats, up to two rows per
            // thread, the matrix in LDS): the default where it applies
            const int want = plan->vectors[bin];
            GKOMI_REQUIRE(want >= 0 && want <= 2, GKOMI_INVALID_ARGUMENT,
                          "batch: vectors must be 0 (automatic), 1 (LDS) or 2 (registers)");
            const int nt = plan->block_size[bin];
            const int rpt = reg_rows_per_thread(nt, rows);
            // matrix rows in registers (uneven Csr): matrix_in_lds 0, or -1
            // when the bin's longest row is known and short enough
            const int kreg = fmt == kFmtUneven
                ? reg_row_entries(nt, rpt, plan->bin_max_row_nnz[bin], plan->matrix_in_lds[bin])
                : 0;
            // the rows in registers: the gathered vectors, each with its zero
            // slot (static LDS in FP64)
            const size_t rlds =
                kreg > 0 ? (sizeof(T) >= 8 ? 0
                                           : (sizeof(T) * (Method == kBatchCg ? 1 : 2) * (rows + 1) +
                                              15) / 16 * 16)
                         : batch_reg_lds_bytes(sizeof(T), rows, nnz, Method);
            const bool applies = (fmt == kFmtUneven || fmt == kFmtCsr) && rpt > 0 &&
                                 rlds <= static_cast<size_t>(lds_limit);
            GKOMI_REQUIRE(want != 2 || applies, GKOMI_NOT_SUPPORTED,
                          "batch: register work vectors need a Csr batch, at most two rows "
                          "per thread and the matrix in LDS");
            if (want == 2 || (want == 0 && applies)) {
                // the plan's workgroups per CU, or 0: as many as are resident
                const int bpc = plan->blocks_per_cu[bin];
                const int grid = bpc > 0 ? static_cast<int>(std::max<int64_t>(
                                               1, std::min<int64_t>(
                                                      cnt, static_cast<int64_t>(num_cus()) * bpc)))
                                         : 0;
                const int rc = fmt == kFmtUneven
                    ? launch_reg_shape<T, kFmtUneven, Method>(nt, rpt, kreg, s, grid, rlds, cnt,
                                                              order_bin, a, b, x, st, iters, res,
                                                              qb, rows, nnz)
                    : launch_reg_shape<T, kFmtCsr, Method>(nt, rpt, kreg, s, grid, rlds, cnt,
                                                           order_bin, a, b, x, st, iters, res, qb,
                                                           rows, nnz);
                if (rc) return rc;
                continue;
            }
        }
        int mat = plan->matrix_in_lds[bin];
        if (mat < 0) {
            mat = batch_lds_bytes(sizeof(T), rows, nnz, true, Method, fmt) <=
                  static_cast<size_t>(lds_limit);
        }
        const size_t lds = batch_lds_bytes(sizeof(T), rows, nnz, mat != 0, Method, fmt);
        GKOMI_REQUIRE(lds <= static_cast<size_t>(lds_limit), GKOMI_NOT_SUPPORTED,
                      "batch: the work vectors of the largest system exceed LDS");
        int bpc = plan->blocks_per_cu[bin];
        if (bpc <= 0) {
            bpc = std::max(1, std::min(8, static_cast<int>((160 * 1024) / std::max<size_t>(lds, 1))));
        }
        const int grid = static_cast<int>(
            std::max<int64_t>(1, std::min<int64_t>(cnt, static_cast<int64_t>(num_cus()) * bpc)));
        const int32_t* order = plan->order ? plan->order + beg : nullptr;
        GKOMI_REQUIRE(plan->order != nullptr || plan->num_bins == 1, GKOMI_INVALID_ARGUMENT,
                      "batch: several bins need an order");
        unsigned* q = queues + bin * (kBatchQueueBytes / sizeof(unsigned));
        const int nt = plan->block_size[bin];
        const int32_t sw = batch_subwave(plan->subwave[bin], rows, nnz);
        GKOMI_REQUIRE(sw >= 1 && sw <= 16 && (sw & (sw - 1)) == 0, GKOMI_INVALID_ARGUMENT,
                      "batch: subwave must be 0 (automatic), 1, 2, 4, 8 or 16");
        const int rc = mat ? launch_batch_fmt<T, true, Method>(fmt, nt, s, grid, lds, cnt, order,
                                                               a, b, x, st, iters, res, q, rows,
                                                               nnz, sw)
                           : launch_batch_fmt<T, false, Method>(fmt, nt, s, grid, lds, cnt, order,
                                                                a, b, x, st, iters, res, q, rows,
                                                                nnz, sw);
        if (rc) return rc;
    }
    return GKOMI_SUCCESS;
}


// ------------------------------------------------------ batched product ----

// x_k = A_k b_k, or with Adv x_k = sum_j (alpha_k a_ij) b_j + beta_k x_k, per
// system k (workgroups stride over the systems).  Csr and Ell run a thread
// per row with the reference's sequential order over the row; Dense runs a
// wave per row with coalesced loads of the row and a wave reduction (the
// reference reduces a warp-sized group per row as well).
template <typename T, int Fmt, bool Adv>
__global__ void __launch_bounds__(kBlock)
    batch_spmv_kernel(int64_t num_systems, const int64_t* __restrict__ row_off,
                      const int64_t* __restrict__ item_off,
                      const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                      const T* __restrict__ vals, int64_t uni_rows, int64_t uni_cols,
                      int64_t uni_nnz, int32_t ell_width, const T* __restrict__ alpha,
                      const T* __restrict__ b, const T* __restrict__ beta, T* __restrict__ x)
{
    for (int64_t sys = blockIdx.x; sys < num_systems; sys += gridDim.x) {
        int64_t rows, cols, row0, col0, nnz0;
        const int32_t* irp = rp;
        const int32_t* ici = ci;
        int32_t base = 0;
        int32_t ew = ell_width;
        if ((Fmt == kFmtEll || Fmt == kFmtDense) && row_off != nullptr) {
            row0 = row_off[sys];
            rows = row_off[sys + 1] - row0;
            cols = rows;
            col0 = row0;
            nnz0 = item_off[sys];
            if (Fmt == kFmtEll) ici = ci + nnz0;
            ew = rows > 0 ? static_cast<int32_t>((item_off[sys + 1] - nnz0) / rows) : 0;
        } else if constexpr (Fmt == kFmtUneven) {
            row0 = row_off[sys];
            rows = row_off[sys + 1] - row0;
            cols = rows;
            col0 = row0;
            irp = rp + row0;
            base = irp[0];
            nnz0 = base;
            ici = ci + nnz0;
        } else {
            rows = uni_rows;
            cols = uni_cols;
            row0 = sys * uni_rows;
            col0 = sys * uni_cols;
            nnz0 = sys * uni_nnz;
        }
        const T* v = vals + nnz0;
        const T* in = b + col0;
        T* out = x + row0;
        const T a = Adv ? alpha[sys] : T(1);
        const T bt = Adv ? beta[sys] : T(0);
        if constexpr (Fmt == kFmtDense) {
            const int lane = threadIdx.x % kWave;
            for (int64_t row = threadIdx.x / kWave; row < rows; row += kWavesPerBlock) {
                T part = 0;
                for (int64_t j = lane; j < cols; j += kWave) {
                    if constexpr (Adv) {
                        part += a * v[row * cols + j] * in[j];
                    } else {
                        part += v[row * cols + j] * in[j];
                    }
                }
                part = wave_sum(part);
                if (lane == 0) out[row] = Adv ? part + bt * out[row] : part;
            }
        } else {
            for (int64_t row = threadIdx.x; row < rows; row += kBlock) {
                T sum = 0;
                if constexpr (Fmt == kFmtEll) {
                    for (int j = 0; j < ew; ++j) {
                        const int64_t e = row + static_cast<int64_t>(j) * rows;
                        const int32_t c = ici[e];
                        if (c < 0) break;
                        if constexpr (Adv) {
                            sum += a * v[e] * in[c];
                        } else {
                            sum += v[e] * in[c];
                        }
                    }
                } else {
                    const int32_t e1 = irp[row + 1] - base;
                    for (int32_t e = irp[row] - base; e < e1; ++e) {
                        if constexpr (Adv) {
                            sum += a * v[e] * in[ici[e]];
                        } else {
                            sum += v[e] * in[ici[e]];
                        }
                    }
                }
                out[row] = Adv ? sum + bt * out[row] : sum;
            }
        }
    }
}

template <typename T, int Fmt>
int batch_spmv_launch(hipStream_t s, const gkomi_batch_csr* a, const T* alpha, const T* b,
                      const T* beta, T* x)
{
    const unsigned grid =
        static_cast<unsigned>(std::min<int64_t>(a->num_systems, int64_t{1} << 20));
    const int64_t cols = a->uniform_cols ? a->uniform_cols : a->uniform_rows;
    const T* v = static_cast<const T*>(a->values);
    if (alpha) {
        batch_spmv_kernel<T, Fmt, true><<<grid, kBlock, 0, s>>>(
            a->num_systems, a->row_offsets, a->item_offsets, a->row_ptrs, a->col_idxs, v,
            a->uniform_rows, cols,
            a->uniform_nnz, a->ell_width, alpha, b, beta, x);
    } else {
        batch_spmv_kernel<T, Fmt, false><<<grid, kBlock, 0, s>>>(
            a->num_systems, a->row_offsets, a->item_offsets, a->row_ptrs, a->col_idxs, v,
            a->uniform_rows, cols,
            a->uniform_nnz, a->ell_width, alpha, b, beta, x);
    }
    GKOMI_CHECK_LAUNCH("batch_spmv_kernel");
    return GKOMI_SUCCESS;
}

template <typename T>
int batch_spmv_impl(const gkomi_ctx* ctx, const gkomi_batch_csr* a, const T* alpha, const T* b,
                    const T* beta, T* x)
{
    if (int rc = enter(ctx)) return rc;
    if (int rc = check_batch_matrix(a)) return rc;
    GKOMI_REQUIRE((alpha == nullptr) == (beta == nullptr), GKOMI_INVALID_ARGUMENT,
                  "batch spmv: alpha and beta must both be given or both be NULL");
    if (a->num_systems == 0 || a->max_rows == 0) return GKOMI_SUCCESS;
    GKOMI_REQUIRE(b && x, GKOMI_INVALID_ARGUMENT, "batch spmv: null vector");
    hipStream_t s = stream_of(ctx);
    switch (fmt_of(a)) {
    case kFmtUneven:
        return batch_spmv_launch<T, kFmtUneven>(s, a, alpha, b, beta, x);
    case kFmtCsr:
        return batch_spmv_launch<T, kFmtCsr>(s, a, alpha, b, beta, x);
    case kFmtEll:
        return batch_spmv_launch<T, kFmtEll>(s, a, alpha, b, beta, x);
    default:
        return batch_spmv_launch<T, kFmtDense>(s, a, alpha, b, beta, x);
    }
}

template <typename T, int Method>
int batch_tune_impl(const gkomi_ctx* ctx, const gkomi_batch_csr* a,
                    const gkomi_batch_settings* st, gkomi_batch_plan* plan, const T* b,
                    const T* x, T* xs, int32_t* iters, T* res, void* ws, size_t wsb)
{
    if (int rc = enter(ctx)) return rc;
    if (int rc = check_batch(a, st)) return rc;
    GKOMI_REQUIRE(plan && x && xs, GKOMI_INVALID_ARGUMENT, "batch tune: null plan or vectors");
    if (a->num_systems == 0) return GKOMI_SUCCESS;
    const int64_t total_rows =
        a->row_offsets ? -1 : a->num_systems * a->uniform_rows;  // uneven: read below
    int64_t rows_all = total_rows;
    if (rows_all < 0) {
        GKOMI_CHECK_HIP(hipMemcpy(&rows_all, a->row_offsets + a->num_systems, sizeof(int64_t),
                                  hipMemcpyDeviceToHost));
    }
    const int lds_limit = device_lds_limit();
    hipStream_t s = stream_of(ctx);
    hipEvent_t e0, e1;
    GKOMI_CHECK_HIP(hipEventCreate(&e0));
    GKOMI_CHECK_HIP(hipEventCreate(&e1));
    int rc = GKOMI_SUCCESS;
    for (int bin = 0; bin < plan->num_bins && rc == GKOMI_SUCCESS; ++bin) {
        gkomi_batch_plan one = *plan;
        one.num_bins = 1;
        one.single_launch = 0;
        one.order = plan->order ? plan->order + plan->bin_begin[bin] : nullptr;
        const int64_t cnt = plan->bin_begin[bin + 1] - plan->bin_begin[bin];
        if (cnt <= 0) continue;
        one.bin_begin[0] = 0;
        one.bin_begin[1] = cnt;
        one.bin_max_rows[0] = plan->bin_max_rows[bin];
        one.bin_max_nnz[0] = plan->bin_max_nnz[bin];
        one.blocks_per_cu[0] = plan->blocks_per_cu[bin];
        gkomi_batch_csr sub = *a;
        sub.num_systems = cnt;
        float best = 1e30f;
        int best_nt = 256, best_mat = 0, best_vec = 1;
        const bool fits = batch_lds_bytes(sizeof(T), one.bin_max_rows[0], one.bin_max_nnz[0],
                                          true, Method, fmt_of(a)) <=
                          static_cast<size_t>(lds_limit);
        const int fmt = fmt_of(a);
        const bool reg_fits = (fmt == kFmtUneven || fmt == kFmtCsr) &&
                              batch_reg_lds_bytes(sizeof(T), one.bin_max_rows[0],
                                                  one.bin_max_nnz[0], Method) <=
                                  static_cast<size_t>(lds_limit);
        one.bin_max_row_nnz[0] = plan->bin_max_row_nnz[bin];
        for (int nt : {64, 128, 256, 512}) {
            // candidates: register work vectors with the matrix rows in
            // registers (uneven Csr, rows of known length) and in LDS, then
            // LDS work vectors with the matrix in LDS and in global memory
            for (int cand = 3; cand >= 0; --cand) {
                const int vec = cand >= 2 ? 2 : 1;
                const int mat = cand == 3 ? 0 : (cand == 2 ? 1 : cand);
                const int rpt = reg_rows_per_thread(nt, one.bin_max_rows[0]);
                if (cand == 3 && !(fmt == kFmtUneven && rpt > 0 &&
                                   reg_row_entries(nt, rpt, one.bin_max_row_nnz[0], -1) > 0)) {
                    continue;
                }
                if (cand == 2 && !(reg_fits && rpt > 0)) continue;
                if (vec == 1 && mat == 1 && !fits) continue;
                one.block_size[0] = nt;
                one.matrix_in_lds[0] = mat;
                one.vectors[0] = vec;
                // sub-batch of this bin only: order is rebased, so the
                // descriptor keeps the full arrays but counts cnt systems
                GKOMI_CHECK_HIP(hipMemcpyAsync(xs, x, sizeof(T) * rows_all,
                                               hipMemcpyDeviceToDevice, s));
                GKOMI_CHECK_HIP(hipEventRecord(e0, s));
                rc = batch_solve_impl<T, Method>(ctx, &sub, st, &one, b, xs, iters, res, ws,
                                                 wsb);
                if (rc) break;
                GKOMI_CHECK_HIP(hipEventRecord(e1, s));
                GKOMI_CHECK_HIP(hipEventSynchronize(e1));
                float ms = 0;
                GKOMI_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) {
                    best = ms;
                    best_nt = nt;
                    best_mat = mat;
                    best_vec = vec;
                }
            }
            if (rc) break;
        }
        plan->block_size[bin] = best_nt;
        plan->matrix_in_lds[bin] = best_mat;
        plan->vectors[bin] = best_vec;
    }
    // the tuned per-bin launches against one launch over the whole batch:
    // the two forms alternate three times and each keeps its best time (the
    // two are often within a few percent)
    int kreg = 0;
    plan->single_launch = 0;
    if (rc == GKOMI_SUCCESS && packed_applies(a, plan, &kreg)) {
        float best[2] = {1e30f, 1e30f};
        for (int rep = 0; rep < 3 && rc == GKOMI_SUCCESS; ++rep) {
            for (int single : {0, 1}) {
                gkomi_batch_plan trial = *plan;
                trial.single_launch = single;
                GKOMI_CHECK_HIP(hipMemcpyAsync(xs, x, sizeof(T) * rows_all,
                                               hipMemcpyDeviceToDevice, s));
                GKOMI_CHECK_HIP(hipEventRecord(e0, s));
                rc = batch_solve_impl<T, Method>(ctx, a, st, &trial, b, xs, iters, res, ws,
                                                 wsb);
                if (rc) break;
                GKOMI_CHECK_HIP(hipEventRecord(e1, s));
                GKOMI_CHECK_HIP(hipEventSynchronize(e1));
                float ms = 0;
                GKOMI_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
                best[single] = std::min(best[single], ms);
            }
        }
        plan->single_launch = best[1] < best[0] ? 1 : 0;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

}  // namespace gkomi

using namespace gkomi;

extern "C" {

size_t gkomi_batch_workspace_bytes(const gkomi_batch_csr* a)
{
    (void)a;
    return kBatchQueueBytes * kBatchMaxBins;
}

size_t gkomi_batch_lds_bytes(int32_t value_bytes, int64_t rows, int64_t nnz,
                             int32_t matrix_in_lds)
{
    return batch_lds_bytes(value_bytes, rows, nnz, matrix_in_lds != 0);
}

size_t gkomi_batch_cg_lds_bytes(int32_t value_bytes, int64_t rows, int64_t nnz,
                                int32_t matrix_in_lds)
{
    return batch_lds_bytes(value_bytes, rows, nnz, matrix_in_lds != 0, kBatchCg);
}

int gkomi_batch_bicgstab_solve_f64(const gkomi_ctx* ctx, const gkomi_batch_csr* a,
                                   const gkomi_batch_settings* st, const gkomi_batch_plan* plan,
                                   const double* b, double* x, int32_t* iters, double* res,
                                   void* ws, size_t wsb)
{
    return batch_solve_impl<double, kBatchBicgstab>(ctx, a, st, plan, b, x, iters, res,
                                                            ws, wsb);
}

int gkomi_batch_bicgstab_solve_f32(const gkomi_ctx* ctx, const gkomi_batch_csr* a,
                                   const gkomi_batch_settings* st, const gkomi_batch_plan* plan,
                                   const float* b, float* x, int32_t* iters, float* res,
                                   void* ws, size_t wsb)
{
    return batch_solve_impl<float, kBatchBicgstab>(ctx, a, st, plan, b, x, iters, res,
                                                            ws, wsb);
}

int gkomi_batch_bicgstab_tune_f64(const gkomi_ctx* ctx, const gkomi_batch_csr* a,
                                  const gkomi_batch_settings* st, gkomi_batch_plan* plan,
                                  const double* b, const double* x, double* xs, int32_t* iters,
                                  double* res, void* ws, size_t wsb)
{
    return batch_tune_impl<double, kBatchBicgstab>(ctx, a, st, plan, b, x, xs, iters,
                                                           res, ws, wsb);
}

int gkomi_batch_bicgstab_tune_f32(const gkomi_ctx* ctx, const gkomi_batch_csr* a,
                                  const gkomi_batch_settings* st, gkomi_batch_plan* plan,
                                  const float* b, const float* x, float* xs, int32_t* iters,
                                  float* res, void* ws, size_t wsb)
{
    return batch_tune_impl<float, kBatchBicgstab>(ctx, a, st, plan, b, x, xs, iters,
                                                           res, ws, wsb);
}

int gkomi_batch_spmv_f64(const gkomi_ctx* ctx, const gkomi_batch_csr* a, const double* alpha,
                         const double* b, const double* beta, double* x)
{
    return batch_spmv_impl(ctx, a, alpha, b, beta, x);
}

int gkomi_batch_spmv_f32(const gkomi_ctx* ctx, const gkomi_batch_csr* a, const float* alpha,
                         const float* b, const float* beta, float* x)
{
    return batch_spmv_impl(ctx, a, alpha, b, beta, x);
}

#define GKOMI_BATCH_CG_ENTRY(T, S)                                                            \
    int gkomi_batch_cg_solve_##S(const gkomi_ctx* ctx, const gkomi_batch_csr* a,              \
                                 const gkomi_batch_settings* st, const gkomi_batch_plan* plan, \
                                 const T* b, T* x, int32_t* iters, T* res, void* ws,           \
                                 size_t wsb)                                                   \
    {                                                                                         \
        return batch_solve_impl<T, kBatchCg>(ctx, a, st, plan, b, x, iters, res, ws, wsb);    \
    }                                                                                         \
    int gkomi_batch_cg_tune_##S(const gkomi_ctx* ctx, const gkomi_batch_csr* a,               \
                                const gkomi_batch_settings* st, gkomi_batch_plan* plan,       \
                                const T* b, const T* x, T* xs, int32_t* iters, T* res,        \
                                void* ws, size_t wsb)                                         \
    {                                                                                         \
        return batch_tune_impl<T, kBatchCg>(ctx, a, st, plan, b, x, xs, iters, res, ws, wsb); \
    }

GKOMI_BATCH_CG_ENTRY(double, f64)
GKOMI_BATCH_CG_ENTRY(float, f32)
#undef GKOMI_BATCH_CG_ENTRY


// ------------------------------------------------ allocation-free solve ----
// gkomi_batch_bicgstab_solve_noalloc_f64 keeps the work-queue counters of
// every size bin in this file-scope device array instead of a workspace
// argument.  Each GPU holds its own copy for the life of the process.  In
// host code the array's name is only a host-side stand-in: the device
// address is resolved once per device by hipGetSymbolAddress (with that
// device current) and kept here; only that address reaches kernels.
}  // extern "C"

namespace gkomi {

constexpr size_t kBatchScratchBytes = 4096;
static __device__ __align__(256) unsigned char scratch[kBatchScratchBytes];

// The current device's address of `scratch`, resolved on the first call on
// that device.
int batch_scratch_address(void** out)
{
    static std::mutex mtx;
    static std::vector<void*> addresses;
    int dev = 0;
    GKOMI_CHECK_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> guard(mtx);
    if (static_cast<int>(addresses.size()) <= dev) addresses.resize(dev + 1, nullptr);
    if (addresses[dev] == nullptr) {
        void* ptr = nullptr;
        GKOMI_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        GKOMI_REQUIRE(ptr != nullptr, GKOMI_HIP_ERROR,
                      "batch: no device address for the allocation-free scratch");
        addresses[dev] = ptr;
    }
    *out = addresses[dev];
    return GKOMI_SUCCESS;
}

}  // namespace gkomi

extern "C" {

int gkomi_batch_bicgstab_solve_noalloc_f64(const gkomi_ctx* ctx, const gkomi_batch_csr* a,
                                           const gkomi_batch_settings* st,
                                           const gkomi_batch_plan* plan, const double* b,
                                           double* x, int32_t* iters, double* res)
{
    if (int rc = enter(ctx)) return rc;
    if (int rc = check_batch(a, st)) return rc;
    const size_t need = gkomi_batch_workspace_bytes(a);
    GKOMI_REQUIRE(need <= kBatchScratchBytes, GKOMI_NOT_SUPPORTED,
                  "batch (allocation-free): the work queues need more than the 4 KiB of "
                  "device scratch; use gkomi_batch_bicgstab_solve_f64 with a workspace");
    void* ws = nullptr;
    if (int rc = batch_scratch_address(&ws)) return rc;
    // the queues start at zero: cleared by a kernel (no memset node under
    // graph capture), whatever an earlier call left
    if (int rc = gkomi_fill_zero(ctx, ws, need)) return rc;
    return batch_solve_impl<double, kBatchBicgstab>(ctx, a, st, plan, b, x, iters, res, ws,
                                                    need);
}

}  // extern "C"
