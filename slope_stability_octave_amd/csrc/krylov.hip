// Sparse and dense linear algebra of the Krylov solvers and the AMG V-cycle
// on gfx950, for double and float operands:
//   node-block CSR SpMV (dim x dim blocks), scalar CSR SpMV, masked BLAS-1,
//   tall-skinny multi-dot V'u and update u += V h, node-block Jacobi,
//   Chebyshev smoothing steps, dense coarse-level product, precision
//   conversion.
//
// Reductions are two-stage with a fixed grid (deterministic: no atomics), a
// 64-lane wavefront reduction by cross-lane shuffles then LDS across waves,
// always accumulated in double.  Results of reductions stay on the device so
// that a caller can chain them without host round trips.
#include <mutex>
#include <unordered_map>

#include "linalg.h"
#include "ssr_common.h"

namespace {

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block-wide sum of kRedThreads values; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double *sh)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < kRedThreads / 64; ++k) s += sh[k];
    __syncthreads();
    return s;
}

// values left per lane after the halving steps from distance o down to 1
__host__ __device__ constexpr int group_final_count(int n, int o)
{
    return o == 0 ? n : group_final_count(n % 2 == 0 ? n / 2 : n, o / 2);
}

// Sum of N values over a group of G lanes by halving exchanges: at distance
// O a lane keeps half of its values (the half chosen by its bit O) and adds
// the partner's copy of that half; once the count is odd the remaining
// distances exchange everything.  Every value is summed over the same pairs,
// in the same order, as the full xor butterfly (lane + partner at each
// distance), so the sums are the same bits, for fewer shuffles, and the
// final values end up spread over the writer lanes.  This is distance O, then
// the smaller ones.  w[0..N) are the lane's current values; on return
// w[0..group_final_count(N, O)) hold its kept sums, whose flat indices start
// at base.  After a step that cannot halve, only the lane with bit O clear
// writes (writer).
template <int N, int O, int G, class T, int NA>
__device__ __forceinline__ void group_halving_sum(T (&w)[NA], int l, int &base, bool &writer)
{
    if constexpr (O > 0) {
        if constexpr (N % 2 == 0) {
            constexpr int H = N / 2;
            const bool hi = (l & O) != 0;
#pragma unroll
            for (int i = 0; i < H; ++i) {
                const T send = hi ? w[i] : w[H + i];
                const T keep = hi ? w[H + i] : w[i];
                w[i] = keep + __shfl_xor(send, O, G);
            }
            if (hi) base += H;
            group_halving_sum<H, O / 2, G, T>(w, l, base, writer);
        } else {
#pragma unroll
            for (int i = 0; i < N; ++i) w[i] += __shfl_xor(w[i], O, G);
            if (l & O) writer = false;
            group_halving_sum<N, O / 2, G, T>(w, l, base, writer);
        }
    }
}

// ---------------------------------------------------------------------------
// SpMV y = alpha K x + beta z on node-block CSR.  A group of G lanes handles
// one block row: lanes stride over the row's blocks, each forming dim partial
// outputs, then the group reduces with shuffles.  Block values row-major.
// ---------------------------------------------------------------------------
// U blocks per lane per pass (U = 2: lane l takes blocks b and b + G, and
// issues the loads of both before the products, which hides more of the
// row_ptr -> col -> x latency chain; measured on the 3D slope operator at
// 1.49M unknowns: 232 us against 253 us for U = 1, G = 32).  Each row's sum
// is formed in the same order for a given (G, U).
//
// MIRROR (symmetric operators whose blocks (a, b) and (b, a) are transposes
// bit for bit, as K_r is by construction): a block left of the diagonal
// (diag[r] is the diagonal block's index) is read as the transpose of its
// mirror tblk[b] in the upper half of row col[b], so the kernel streams only
// the blocks on and above the diagonal from memory; the mirrors were read
// by the rows just before (the matrix is banded) and come from the caches.
// The products and their order are those of the plain form: the same bits.
//
// Measured on the ~1.5M-unknown slope operator (14.4M blocks; rows reach up
// to 287,044 nodes back, 6,041 on average): 287 us against 232 us for the
// plain form in double, 203 against 134 us in float.  The mirrors come from
// all over the matrix and from other dies' caches, each lane fetching its
// own 72-byte block, so the form is selectable but off by default.
template <int DIM, int G, class T, int U = 1, bool MIRROR = false>
__global__ __launch_bounds__(256) void k_spmv_bsr(int64_t n_rows, const int32_t *__restrict__ row_ptr,
                                                  const int32_t *__restrict__ col, const T *__restrict__ val,
                                                  const T *__restrict__ x, T *y, double alpha, double beta,
                                                  const T *z, const int32_t *__restrict__ diag,
                                                  const int32_t *__restrict__ tblk)
{
    constexpr int DD = DIM * DIM;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r = gid / G;
    const int l = (int)(gid % G);
    T acc[DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i) acc[i] = T(0);
    if (r < n_rows) {
        const int32_t b0 = row_ptr[r], b1 = row_ptr[r + 1];
        const int32_t bd = MIRROR ? diag[r] : b0;
        // entry e = i * DIM + j of the block, or of the transposed mirror
        auto at = [](bool t, int e) { return t ? (e % DIM) * DIM + e / DIM : e; };
        if (U == 1) {
            for (int32_t b = b0 + l; b < b1; b += G) {
                const int64_t c = col[b];
                const bool t = MIRROR && b < bd;
                T xv[DIM];
#pragma unroll
                for (int j = 0; j < DIM; ++j) xv[j] = x[c * DIM + j];
                const T *v = val + (t ? (int64_t)tblk[b] : (int64_t)b) * DD;
#pragma unroll
                for (int i = 0; i < DIM; ++i)
#pragma unroll
                    for (int j = 0; j < DIM; ++j) acc[i] += v[at(t, i * DIM + j)] * xv[j];
            }
        } else {
            for (int32_t b = b0 + l; b < b1; b += 2 * G) {
                const int32_t bb = b + G;
                const bool two = bb < b1;
                const int64_t c0 = col[b];
                const int64_t c1 = two ? (int64_t)col[bb] : c0;
                const bool t0 = MIRROR && b < bd, t1 = MIRROR && two && bb < bd;
                const int64_t p0 = t0 ? (int64_t)tblk[b] : (int64_t)b;
                const int64_t p1 = t1 ? (int64_t)tblk[bb] : (int64_t)bb;
                T v0[DD], v1[DD], x0[DIM], x1[DIM];
#pragma unroll
                for (int e = 0; e < DD; ++e) v0[e] = val[p0 * DD + at(t0, e)];
#pragma unroll
                for (int e = 0; e < DD; ++e) v1[e] = two ? val[p1 * DD + at(t1, e)] : T(0);
#pragma unroll
                for (int j = 0; j < DIM; ++j) x0[j] = x[c0 * DIM + j];
#pragma unroll
                for (int j = 0; j < DIM; ++j) x1[j] = x[c1 * DIM + j];
#pragma unroll
                for (int i = 0; i < DIM; ++i)
#pragma unroll
                    for (int j = 0; j < DIM; ++j) acc[i] += v0[i * DIM + j] * x0[j] + v1[i * DIM + j] * x1[j];
            }
        }
    }
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) acc[i] += __shfl_xor(acc[i], o, G);
    if (r < n_rows && l == 0) {
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            const T o = T(alpha) * acc[i];
            y[r * DIM + i] = beta == 0.0 ? o : o + T(beta) * z[r * DIM + i];
        }
    }
}

// The same product with the block values staged through LDS.  Lanes that
// each read a whole 72-byte block issue nine 8-byte loads whose addresses
// stride by the block size, so every load instruction touches about as many
// cache lines as it has lanes and the vector L1 serves each line nine times.
// Here the one wavefront of a workgroup first copies the value span of its
// rows (contiguous in the block array) into LDS with 16-byte loads that
// consecutive lanes issue for consecutive addresses, then forms the same
// products as k_spmv_bsr from LDS, in the same order (the same bits).  A span
// longer than CAP blocks is read from global memory directly.
// The products of one lane of k_spmv_bsr over a block row, reading block b
// at v[(b - vb0) * dim^2] (the lane's sum in k_spmv_bsr's order).
template <int DIM, int G, class T, int U>
__device__ __forceinline__ void bsr_row_products(int32_t b0, int32_t b1, int l, const int32_t *__restrict__ col,
                                                 const T *__restrict__ x, const T *__restrict__ v, int32_t vb0,
                                                 T acc[DIM])
{
    constexpr int DD = DIM * DIM;
    if (U == 1) {
        for (int32_t b = b0 + l; b < b1; b += G) {
            const int64_t c = col[b];
            T xv[DIM];
#pragma unroll
            for (int j = 0; j < DIM; ++j) xv[j] = x[c * DIM + j];
            const T *vv = v + (int64_t)(b - vb0) * DD;
#pragma unroll
            for (int i = 0; i < DIM; ++i)
#pragma unroll
                for (int j = 0; j < DIM; ++j) acc[i] += vv[i * DIM + j] * xv[j];
        }
    } else {
        for (int32_t b = b0 + l; b < b1; b += 2 * G) {
            const int32_t bb = b + G;
            const bool two = bb < b1;
            const int64_t c0 = col[b];
            const int64_t c1 = two ? (int64_t)col[bb] : c0;
            T v0[DD], v1[DD], x0[DIM], x1[DIM];
#pragma unroll
            for (int e = 0; e < DD; ++e) v0[e] = v[(int64_t)(b - vb0) * DD + e];
#pragma unroll
            for (int e = 0; e < DD; ++e) v1[e] = two ? v[(int64_t)(bb - vb0) * DD + e] : T(0);
#pragma unroll
            for (int j = 0; j < DIM; ++j) x0[j] = x[c0 * DIM + j];
#pragma unroll
            for (int j = 0; j < DIM; ++j) x1[j] = x[c1 * DIM + j];
#pragma unroll
            for (int i = 0; i < DIM; ++i)
#pragma unroll
                for (int j = 0; j < DIM; ++j) acc[i] += v0[i * DIM + j] * x0[j] + v1[i * DIM + j] * x1[j];
        }
    }
}

template <int DIM, int G, class T, int U, int CAP>
__global__ __launch_bounds__(64) void k_spmv_bsr_lds(int64_t n_rows, const int32_t *__restrict__ row_ptr,
                                                     const int32_t *__restrict__ col, const T *__restrict__ val,
                                                     const T *__restrict__ x, T *y, double alpha, double beta,
                                                     const T *z)
{
    constexpr int DD = DIM * DIM, RW = 64 / G;
    constexpr int BYTES = CAP * DD * (int)sizeof(T);
    __shared__ __align__(16) unsigned char sh[BYTES + 16];
    const int lane = threadIdx.x;
    const int64_t r_first = (int64_t)blockIdx.x * RW;
    const int64_t r = r_first + lane / G;
    const int l = lane % G;
    const int64_t r_last = r_first + RW < n_rows ? r_first + RW : n_rows;
    const int32_t bw0 = row_ptr[r_first], bw1 = row_ptr[r_last];
    const bool staged = bw1 - bw0 <= CAP;
    int mis = 0;
    if (staged) {
        const unsigned char *base = (const unsigned char *)val;
        const int64_t s0 = (int64_t)bw0 * DD * (int64_t)sizeof(T), s1 = (int64_t)bw1 * DD * (int64_t)sizeof(T);
        const int64_t a0 = s0 & ~(int64_t)15;
        mis = (int)(s0 - a0);
        const int n16 = (int)((s1 - a0) >> 4);
        typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
        // every load of the span is issued before the first LDS write
        constexpr int NCH = (BYTES + 16 + 16 * 64 - 1) / (16 * 64);
        u32x4 buf[NCH];
        const u32x4 *src = (const u32x4 *)(base + a0);
#pragma unroll
        for (int t = 0; t < NCH; ++t) {
            const int c = lane + 64 * t;
            if (c < n16) buf[t] = __builtin_nontemporal_load(src + c);
        }
#pragma unroll
        for (int t = 0; t < NCH; ++t) {
            const int c = lane + 64 * t;
            if (c < n16) *(u32x4 *)(sh + 16 * c) = buf[t];
        }
        const int tail = (int)((s1 - a0) & 15);
        if (lane < tail / 4)
            *(uint32_t *)(sh + 16 * n16 + 4 * lane) = *(const uint32_t *)(base + a0 + 16 * n16 + 4 * lane);
    }
    __syncthreads();  // one wavefront per workgroup: its LDS writes are visible to all its lanes
    T acc[DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i) acc[i] = T(0);
    if (r < n_rows) {
        const int32_t b0 = row_ptr[r], b1 = row_ptr[r + 1];
        // staged block b sits at LDS offset (b - bw0) blocks: small positive
        // offsets from the array (an LDS pointer displaced by -bw0 blocks
        // wraps its 32-bit offset and leaves the LDS aperture when added back)
        if (staged)
            bsr_row_products<DIM, G, T, U>(b0, b1, l, col, x, (const T *)(sh + mis), bw0, acc);
        else
            bsr_row_products<DIM, G, T, U>(b0, b1, l, col, x, val, 0, acc);
    }
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) acc[i] += __shfl_xor(acc[i], o, G);
    if (r < n_rows && l == 0) {
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            const T o = T(alpha) * acc[i];
            y[r * DIM + i] = beta == 0.0 ? o : o + T(beta) * z[r * DIM + i];
        }
    }
}

// The same product with one wavefront per block row and lanes over the
// row's values: lane l owns entry q = l mod dim^2 of block l / dim^2 in each
// pass, so every pass reads 64 / dim^2 consecutive blocks (63 of 64 lanes in
// 3D) as one contiguous span, and each lane keeps a single partial sum of row
// entry q / dim.  Three masked wavefront sums give the dim outputs.
template <class T>
__device__ __forceinline__ T wave_sum_t(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <int DIM, class T>
__global__ __launch_bounds__(256) void k_spmv_bsr_w(int64_t n_rows, const int32_t *__restrict__ row_ptr,
                                                    const int32_t *__restrict__ col, const T *__restrict__ val,
                                                    const T *__restrict__ x, T *y, double alpha, double beta,
                                                    const T *z)
{
    constexpr int DD = DIM * DIM, BPW = 64 / DD, ACT = BPW * DD;
    const int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int l = threadIdx.x & 63;
    const int lb = l / DD, q = l - lb * DD, i = q / DIM, j = q - i * DIM;
    T acc = T(0);
    if (r < n_rows && l < ACT) {
        const int32_t b0 = row_ptr[r], b1 = row_ptr[r + 1];
        for (int32_t b = b0 + lb; b < b1; b += BPW) acc += val[(int64_t)b * DD + q] * x[(int64_t)col[b] * DIM + j];
    }
    T o[DIM];
#pragma unroll
    for (int ii = 0; ii < DIM; ++ii) o[ii] = wave_sum_t<T>(i == ii ? acc : T(0));
    if (r < n_rows && l < DIM) {
        T v = o[0];
#pragma unroll
        for (int ii = 1; ii < DIM; ++ii)
            if (l == ii) v = o[ii];
        const T a = T(alpha) * v;
        y[r * DIM + l] = beta == 0.0 ? a : a + T(beta) * z[r * DIM + l];
    }
}

// Block CSR with dense RB x CB blocks (row-major), the coarse AMG operators
// and transfers whose unknowns come in whole aggregates (6 rigid-body modes in
// 3D, 3 in 2D) and nodes: one column index per block instead of one per
// entry, 8.1 instead of 12 bytes per FP64 value of a 6 x 6 block.  G lanes per
// block row as in k_spmv_bsr.
// PF: a lane loads its next block (clamped to the row's last) together with
// the current one and adds the two blocks' products one after the other, as
// the plain loop does (the same bits); the plain loop's loads wait for the
// previous block's.
template <int RB, int CB, int G, class T, bool HALVE = true, bool PF = false>
__global__ __launch_bounds__(256) void k_spmv_brc(int64_t n_rows, const int32_t *__restrict__ row_ptr,
                                                  const int32_t *__restrict__ col, const T *__restrict__ val,
                                                  const T *__restrict__ x, T *y, double alpha, double beta,
                                                  const T *z)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r = gid / G;
    const int l = (int)(gid % G);
    T acc[RB];
#pragma unroll
    for (int i = 0; i < RB; ++i) acc[i] = T(0);
    if (r < n_rows) {
        const int32_t b0 = row_ptr[r], b1 = row_ptr[r + 1];
        if constexpr (PF) {
            for (int32_t b = b0 + l; b < b1; b += 2 * G) {
                const int32_t bb = b + G < b1 ? b + G : b;
                const int64_t c0 = col[b], c1 = col[bb];
                T x0[CB], x1[CB], v0[RB * CB], v1[RB * CB];
#pragma unroll
                for (int j = 0; j < CB; ++j) x0[j] = x[c0 * CB + j];
#pragma unroll
                for (int j = 0; j < CB; ++j) x1[j] = x[c1 * CB + j];
#pragma unroll
                for (int e = 0; e < RB * CB; ++e) v0[e] = val[(int64_t)b * RB * CB + e];
#pragma unroll
                for (int e = 0; e < RB * CB; ++e) v1[e] = val[(int64_t)bb * RB * CB + e];
#pragma unroll
                for (int i = 0; i < RB; ++i)
#pragma unroll
                    for (int j = 0; j < CB; ++j) acc[i] += v0[i * CB + j] * x0[j];
                if (b + G < b1) {
#pragma unroll
                    for (int i = 0; i < RB; ++i)
#pragma unroll
                        for (int j = 0; j < CB; ++j) acc[i] += v1[i * CB + j] * x1[j];
                }
            }
        } else {
            for (int32_t b = b0 + l; b < b1; b += G) {
                const int64_t c = col[b];
                T xv[CB];
#pragma unroll
                for (int j = 0; j < CB; ++j) xv[j] = x[c * CB + j];
                const T *v = val + (int64_t)b * RB * CB;
#pragma unroll
                for (int i = 0; i < RB; ++i)
#pragma unroll
                    for (int j = 0; j < CB; ++j) acc[i] += v[i * CB + j] * xv[j];
            }
        }
    }
    if (!HALVE) {
        // the full xor butterfly on every row, lane 0 writing
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
            for (int o = G / 2; o > 0; o >>= 1) acc[i] += __shfl_xor(acc[i], o, G);
        if (r < n_rows && l == 0) {
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                const T o = T(alpha) * acc[i];
                y[r * RB + i] = beta == 0.0 ? o : o + T(beta) * z[r * RB + i];
            }
        }
        return;
    }
    // the lanes' partials summed by halving exchanges (group_halving_sum: the
    // butterfly's bits; 6 rows over 64 lanes: 18 shuffles instead of 36)
    int base = 0;
    bool writer = true;
    group_halving_sum<RB, G / 2, G, T>(acc, l, base, writer);
    constexpr int NFIN = group_final_count(RB, G / 2);
    if (r < n_rows && writer) {
#pragma unroll
        for (int t = 0; t < NFIN; ++t) {
            const int i = base + t;
            const T o = T(alpha) * acc[t];
            y[r * RB + i] = beta == 0.0 ? o : o + T(beta) * z[r * RB + i];
        }
    }
}

// Scalar CSR, G lanes per row (rows of ~10-300 entries).
template <int G, class T>
__global__ __launch_bounds__(256) void k_spmv_csr(int64_t n_rows, const int64_t *__restrict__ rp,
                                                  const int32_t *__restrict__ ci, const T *__restrict__ v,
                                                  const T *__restrict__ x, T *y, double alpha, double beta,
                                                  const T *z)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r = gid / G;
    const int l = (int)(gid % G);
    T s = T(0);
    if (r < n_rows)
        for (int64_t k = rp[r] + l; k < rp[r + 1]; k += G) s += v[k] * x[ci[k]];
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
    if (r < n_rows && l == 0) y[r] = T(alpha) * s + (beta == 0.0 ? T(0) : T(beta) * z[r]);
}

// partial[b] = sum over the block's share of a[i]*b[i] (times mask if given)
template <class T>
__global__ __launch_bounds__(kRedThreads) void k_dot_partial(int64_t n, const T *__restrict__ a,
                                                             const T *__restrict__ b,
                                                             const double *__restrict__ m,
                                                             double *__restrict__ part)
{
    __shared__ double sh[kRedThreads / 64];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double t = (double)a[i] * (double)b[i];
        s += m ? m[i] * t : t;
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// part[b] = largest u(1)^2 + ... + u(dim)^2 over the block's share of the
// nodes of U (dim x n_n node-major); k_max_partials then takes the largest
// partial and its square root (sqrt is monotone: the same as the largest root)
template <int DIM>
__global__ __launch_bounds__(kRedThreads) void k_node_sq_max_partial(int64_t n_n, const double *__restrict__ U,
                                                                     double *__restrict__ part)
{
#pragma clang fp contract(off)
    __shared__ double sh[kRedThreads];
    double m = 0.0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n_n; v += (int64_t)gridDim.x * blockDim.x) {
        double q = U[v * DIM] * U[v * DIM];
#pragma unroll
        for (int c = 1; c < DIM; ++c) q = q + U[v * DIM + c] * U[v * DIM + c];
        m = q > m ? q : m;
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int o = kRedThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o && sh[threadIdx.x + o] > sh[threadIdx.x]) sh[threadIdx.x] = sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(kRedThreads) void k_sqrt_max_partials(int nb, const double *__restrict__ part,
                                                                   double *__restrict__ out)
{
    __shared__ double sh[kRedThreads];
    double m = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) m = part[i] > m ? part[i] : m;
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int o = kRedThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o && sh[threadIdx.x + o] > sh[threadIdx.x]) sh[threadIdx.x] = sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sqrt(sh[0]);
}

// out[c] = sum_b part[c*nb + b] for c < k (one block per output)
__global__ __launch_bounds__(kRedThreads) void k_reduce_partials(int nb, int k, const double *__restrict__ part,
                                                                 double *__restrict__ out)
{
    __shared__ double sh[kRedThreads / 64];
    const int c = blockIdx.x;
    if (c >= k) return;
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s += part[(int64_t)c * nb + i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[c] = s;
}

// Sums of N per-lane values over the wavefront (N a power of two <= 64) in
// N-1 + 6-log2(N) shuffles: log2(N) halving exchanges (lanes xor 32, 16, ...)
// leave each lane with one of the N values summed over its exchange partners,
// selected by its high lane bits; the remaining xor steps finish the sum.
// Lane l then holds the total of value md_index<N>(l).
template <int N>
__device__ __forceinline__ int md_index(int l)
{
    int c = 0;
#pragma unroll
    for (int m = 32, b = N / 2; b > 0; m >>= 1, b >>= 1) c += ((l & m) ? b : 0);
    return c;
}

template <int N>
__device__ __forceinline__ double wave_sum_n(double (&s)[N], int l)
{
#pragma unroll
    for (int m = 32, h = N / 2; h > 0; m >>= 1, h >>= 1) {
        const bool up = (l & m) != 0;
#pragma unroll
        for (int j = 0; j < h; ++j) {
            const double send = up ? s[j] : s[j + h];
            s[j] = (up ? s[j + h] : s[j]) + __shfl_xor(send, m, 64);
        }
    }
    double v = s[0];
#pragma unroll
    for (int m = 32 / N; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// Multi-dot h = V(:,0:k)' u, columns stride ld, k <= 128.  The grid is
// (row blocks) x (8-column groups): every block takes RPT rows per thread of
// one column group, so all groups run at once instead of one after another
// (the latency of a small multi-dot is that of one pass).  Each wavefront
// reduces its 8 column sums by one shuffle transposition, and one barrier
// combines the four wavefronts into this block's partial for each column.
constexpr int kMdCols = 8, kMdBlocks = 8192;

//
// PF: all of a pass's loads (RPT rows of u and of the column group) are
// issued before the first product, with clamped indices instead of guards,
// and the products then added in the plain form's order (the same bits); the
// plain form's guarded loads wait one after another.
template <class T, int RPT, bool PF = false>
__global__ __launch_bounds__(kRedThreads) void k_multidot_partial(int64_t n, int k, const T *__restrict__ V,
                                                                  int64_t ld, const T *__restrict__ u,
                                                                  double *__restrict__ part)
{
    __shared__ double sh[kRedThreads / 64][kMdCols];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int ci = md_index<kMdCols>(l);
    const int c0 = blockIdx.y * kMdCols;
    const int nc = k - c0 < kMdCols ? k - c0 : kMdCols;
    double s[kMdCols];
#pragma unroll
    for (int c = 0; c < kMdCols; ++c) s[c] = 0.0;
    const int64_t span = (int64_t)RPT * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * span; base < n; base += (int64_t)gridDim.x * span) {
        if constexpr (PF) {
            T uu[RPT], vv[RPT][kMdCols];
#pragma unroll
            for (int rr = 0; rr < RPT; ++rr) {
                const int64_t i = base + (int64_t)rr * blockDim.x + threadIdx.x;
                const int64_t ic = i < n ? i : n - 1;
                uu[rr] = u[ic];
#pragma unroll
                for (int c = 0; c < kMdCols; ++c) {
                    const int cc = c < nc ? c0 + c : c0;
                    vv[rr][c] = V[(int64_t)cc * ld + ic];
                }
            }
#pragma unroll
            for (int rr = 0; rr < RPT; ++rr) {
                const int64_t i = base + (int64_t)rr * blockDim.x + threadIdx.x;
                if (i < n) {
                    const double ui = (double)uu[rr];
#pragma unroll
                    for (int c = 0; c < kMdCols; ++c)
                        if (c < nc) s[c] += (double)vv[rr][c] * ui;
                }
            }
        } else {
#pragma unroll
            for (int rr = 0; rr < RPT; ++rr) {
                const int64_t i = base + (int64_t)rr * blockDim.x + threadIdx.x;
                if (i < n) {
                    const double ui = (double)u[i];
#pragma unroll
                    for (int c = 0; c < kMdCols; ++c)
                        if (c < nc) s[c] += (double)V[(int64_t)(c0 + c) * ld + i] * ui;
                }
            }
        }
    }
    const double t = wave_sum_n<kMdCols>(s, l);
    if ((l & (64 / kMdCols - 1)) == 0 && ci < nc) sh[w][ci] = t;
    __syncthreads();
    if (threadIdx.x < nc) {
        double a = 0.0;
#pragma unroll
        for (int q = 0; q < kRedThreads / 64; ++q) a += sh[q][threadIdx.x];
        part[(int64_t)(c0 + threadIdx.x) * gridDim.x + blockIdx.x] = a;
    }
}

// u = alpha*u + beta * V(:,0:k) h   (h on device, k <= 128)
// DIRECT: the coefficients are read where they are used (wave-uniform loads
// through the scalar cache) instead of staged in LDS behind a barrier, so
// the column loads start at once; the same sums (the same bits).
template <class TV, class TU, bool DIRECT = false>
__global__ __launch_bounds__(256) void k_gemv_update(int64_t n, int k, const TV *__restrict__ V, int64_t ld,
                                                     const double *__restrict__ h, double alpha, double beta,
                                                     TU *__restrict__ u)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    if constexpr (DIRECT) {
        if (i >= n) return;
        for (int c = 0; c < k; ++c) s += (double)V[(int64_t)c * ld + i] * h[c];
    } else {
        __shared__ double sh[128];
        for (int c = threadIdx.x; c < k; c += blockDim.x) sh[c] = h[c];
        __syncthreads();
        if (i >= n) return;
        for (int c = 0; c < k; ++c) s += (double)V[(int64_t)c * ld + i] * sh[c];
    }
    u[i] = (TU)((alpha == 0.0 ? 0.0 : alpha * (double)u[i]) + beta * s);
}

// ---------------------------------------------------------------------------
// Multi-vector SpMV Y(:,s) = K X(:,s) for s < nv <= kSpmmNV: every block of
// K is read once for all nv vectors (the basis products of the deflation
// re-orthogonalisation).  Same lane grouping as k_spmv_bsr.
// ---------------------------------------------------------------------------
constexpr int kSpmmNV = 8;

template <int DIM, int G, class T, bool HALVE = true>
__global__ __launch_bounds__(256) void k_spmm_bsr(int64_t n_rows, const int32_t *__restrict__ row_ptr,
                                                  const int32_t *__restrict__ col, const T *__restrict__ val,
                                                  const T *__restrict__ X, int64_t ldx, T *__restrict__ Y,
                                                  int64_t ldy, int nv)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r = gid / G;
    const int l = (int)(gid % G);
    T acc[kSpmmNV][DIM];
#pragma unroll
    for (int s = 0; s < kSpmmNV; ++s)
#pragma unroll
        for (int i = 0; i < DIM; ++i) acc[s][i] = T(0);
    if (r < n_rows) {
        const int32_t b0 = row_ptr[r], b1 = row_ptr[r + 1];
        for (int32_t b = b0 + l; b < b1; b += G) {
            const int64_t c = col[b];
            T v[DIM * DIM];
#pragma unroll
            for (int e = 0; e < DIM * DIM; ++e) v[e] = val[(int64_t)b * DIM * DIM + e];
#pragma unroll
            for (int s = 0; s < kSpmmNV; ++s) {
                if (s < nv) {
                    T xv[DIM];
#pragma unroll
                    for (int j = 0; j < DIM; ++j) xv[j] = X[s * ldx + c * DIM + j];
#pragma unroll
                    for (int i = 0; i < DIM; ++i)
#pragma unroll
                        for (int j = 0; j < DIM; ++j) acc[s][i] += v[i * DIM + j] * xv[j];
                }
            }
        }
    }
    // the G lanes' partials of all kSpmmNV*DIM values summed by halving
    // exchanges (group_halving_sum: the butterfly's bits; 3D, 16 lanes: 24
    // shuffles instead of 96)
    if (!HALVE) {
        // the full xor butterfly on every value, lane 0 writing the row
#pragma unroll
        for (int s = 0; s < kSpmmNV; ++s)
#pragma unroll
            for (int i = 0; i < DIM; ++i)
#pragma unroll
                for (int o = G / 2; o > 0; o >>= 1) acc[s][i] += __shfl_xor(acc[s][i], o, G);
        if (r < n_rows && l == 0) {
#pragma unroll
            for (int s = 0; s < kSpmmNV; ++s)
                if (s < nv)
#pragma unroll
                    for (int i = 0; i < DIM; ++i) Y[s * ldy + r * DIM + i] = acc[s][i];
        }
        return;
    }
    constexpr int NVAL = kSpmmNV * DIM;
    T w[NVAL];
#pragma unroll
    for (int s = 0; s < kSpmmNV; ++s)
#pragma unroll
        for (int i = 0; i < DIM; ++i) w[s * DIM + i] = acc[s][i];
    int base = 0;      // flat index of the lane's first kept value
    bool writer = true;
    group_halving_sum<NVAL, G / 2, G, T>(w, l, base, writer);
    constexpr int NFIN = group_final_count(NVAL, G / 2);
    if (r < n_rows && writer) {
#pragma unroll
        for (int t = 0; t < NFIN; ++t) {
            const int f = base + t, s = f / DIM, i = f - (f / DIM) * DIM;
            if (s < nv) Y[s * ldy + r * DIM + i] = w[t];
        }
    }
}

// ---------------------------------------------------------------------------
// Tall-skinny dense products of the deflation basis (double, column-major):
//   k_gram_partial: partial X'Y over one row chunk for one 64 x 64 tile of
//     the k x k result (tile pair = blockIdx.x, chunk = blockIdx.y); the
//     chunks are then summed in order by k_sum_chunks (deterministic).
//   k_gemm_tall:    Y = X C for X n x k, C k x m (one 64-row x 64-column
//     output tile per block).
// 256 threads as 16 x 16, 4 x 4 outputs each; operands staged through LDS in
// 32-deep slabs read along the columns (coalesced), the next slab held in
// registers while the current one is multiplied.
// ---------------------------------------------------------------------------
constexpr int kTile = 64, kSlab = 32;

__global__ __launch_bounds__(256) void k_gram_partial(int64_t n, int k, int tiles, int64_t rows_per_chunk,
                                                      const double *__restrict__ X, const double *__restrict__ Y,
                                                      double *__restrict__ part, int upper_only)
{
    constexpr int PER = kSlab * kTile / 256;  // slab elements per thread and operand
    __shared__ double xs[kSlab][kTile + 1];
    __shared__ double ys[kSlab][kTile + 1];
    if (upper_only && blockIdx.x % tiles > blockIdx.x / tiles) return;  // mirrored by k_sum_chunks
    const int i0 = (blockIdx.x % tiles) * kTile, j0 = (blockIdx.x / tiles) * kTile;
    const int64_t r0 = (int64_t)blockIdx.y * rows_per_chunk;
    const int64_t r1 = r0 + rows_per_chunk < n ? r0 + rows_per_chunk : n;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[4][4];
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[p][q] = 0.0;
    // the next slab is loaded into registers while the current one is used
    double px[PER], py[PER];
    auto fetch = [&](int64_t rb) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int e = threadIdx.x + 256 * u;
            const int rr = e % kSlab, cc = e / kSlab;
            const int64_t r = rb + rr;
            px[u] = (r < r1 && i0 + cc < k) ? X[(int64_t)(i0 + cc) * n + r] : 0.0;
            py[u] = (r < r1 && j0 + cc < k) ? Y[(int64_t)(j0 + cc) * n + r] : 0.0;
        }
    };
    if (r0 < r1) fetch(r0);
    for (int64_t rb = r0; rb < r1; rb += kSlab) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int e = threadIdx.x + 256 * u;
            xs[e % kSlab][e / kSlab] = px[u];
            ys[e % kSlab][e / kSlab] = py[u];
        }
        __syncthreads();
        if (rb + kSlab < r1) fetch(rb + kSlab);
#pragma unroll 4
        for (int rr = 0; rr < kSlab; ++rr) {
            double a[4], b[4];
#pragma unroll
            for (int p = 0; p < 4; ++p) a[p] = xs[rr][tx + 16 * p];
#pragma unroll
            for (int q = 0; q < 4; ++q) b[q] = ys[rr][ty + 16 * q];
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[p][q] += a[p] * b[q];
        }
    }
    double *out = part + (int64_t)blockIdx.y * k * k;
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = i0 + tx + 16 * p, j = j0 + ty + 16 * q;
            if (i < k && j < k) out[(int64_t)j * k + i] = acc[p][q];
        }
}

// out[e] = sum over c < nc of part[c*len + e], in chunk order
// (k x k result; with upper_only the tiles below the diagonal were skipped
// and their entries are read from the transposed position)
__global__ void k_sum_chunks(int nc, int k, const double *__restrict__ part, double *__restrict__ out, int upper_only)
{
    const int64_t len = (int64_t)k * k;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= len) return;
    const int i = (int)(e % k), j = (int)(e / k);
    const int64_t src = (upper_only && i / kTile > j / kTile) ? (int64_t)i * k + j : e;
    // eight interleaved partial sums keep eight loads in flight (a single
    // running sum waits on every load); combined in a fixed order
    double s[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int c = 0;
    for (; c + 8 <= nc; c += 8)
#pragma unroll
        for (int u = 0; u < 8; ++u) s[u] += part[(int64_t)(c + u) * len + src];
    for (int u = 0; c < nc; ++c, ++u) s[u] += part[(int64_t)c * len + src];
    out[e] = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}

__global__ __launch_bounds__(256) void k_gemm_tall(int64_t n, int k, int m, const double *__restrict__ X,
                                                   const double *__restrict__ C, double *__restrict__ Y,
                                                   const int *__restrict__ qstart, float *__restrict__ Y32)
{
    constexpr int PER = kSlab * kTile / 256;
    __shared__ double xs[kSlab][kTile + 1];  // [q][row]
    __shared__ double cs[kSlab][kTile + 1];  // [q][column]
    const int64_t r0 = (int64_t)blockIdx.x * kTile;
    const int c0 = blockIdx.y * kTile;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[4][4];
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[p][q] = 0.0;
    double px[PER], pc[PER];
    auto fetch = [&](int q0) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int e = threadIdx.x + 256 * u;
            const int rr = e % kTile, qq = e / kTile;
            px[u] = (r0 + rr < n && q0 + qq < k) ? X[(int64_t)(q0 + qq) * n + r0 + rr] : 0.0;
            const int q2 = e % kSlab, cc = e / kSlab;
            pc[u] = (q0 + q2 < k && c0 + cc < m) ? C[(int64_t)(c0 + cc) * k + q0 + q2] : 0.0;
        }
    };
    // rows of C above qstart[column tile] are zero in every column of the tile
    const int qs = qstart ? (qstart[blockIdx.y] / kSlab) * kSlab : 0;
    if (qs < k) fetch(qs);
    for (int q0 = qs; q0 < k; q0 += kSlab) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int e = threadIdx.x + 256 * u;
            xs[e / kTile][e % kTile] = px[u];
            cs[e % kSlab][e / kSlab] = pc[u];
        }
        __syncthreads();
        if (q0 + kSlab < k) fetch(q0 + kSlab);
#pragma unroll 4
        for (int qq = 0; qq < kSlab; ++qq) {
            double a[4], b[4];
#pragma unroll
            for (int p = 0; p < 4; ++p) a[p] = xs[qq][tx + 16 * p];
#pragma unroll
            for (int q = 0; q < 4; ++q) b[q] = cs[qq][ty + 16 * q];
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[p][q] += a[p] * b[q];
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = c0 + ty + 16 * q;
        if (c >= m) continue;
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const int64_t r = r0 + tx + 16 * p;
            if (r < n) {
                Y[(int64_t)c * n + r] = acc[p][q];
                if (Y32) Y32[(int64_t)c * n + r] = (float)acc[p][q];
            }
        }
    }
}

template <class T>
__global__ void k_axpby(int64_t n, double a, const T *__restrict__ x, double b, T *__restrict__ y)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = T(a) * x[i] + (b == 0.0 ? T(0) : T(b) * y[i]);
}

template <class T>
__global__ void k_lincomb(int64_t n, double a, const T *__restrict__ x, double b, const T *__restrict__ y,
                          T *__restrict__ z)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = T(a) * x[i] + T(b) * y[i];
}

template <class T>
__global__ void k_mul(int64_t n, const double *__restrict__ m, T *__restrict__ x)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] *= T(m[i]);
}

// x = v (write-only; also the graph-safe way to clear a buffer)
template <class T>
__global__ void k_fill(int64_t n, double v, T *__restrict__ x)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = T(v);
}

template <class TI, class TO>
__global__ void k_convert(int64_t n, double a, const TI *__restrict__ x, TO *__restrict__ y)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (TO)(a * (double)x[i]);
}

template <class TI, class TO>
__global__ void k_axpy(int64_t n, double a, const TI *__restrict__ x, TO *__restrict__ y)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (TO)((double)y[i] + a * (double)x[i]);
}

template <class T>
__global__ void k_scale_inv_sqrt(int64_t n, const double *__restrict__ nrm2, double tiny, T *__restrict__ u)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double nr = sqrt(nrm2[0] > 0.0 ? nrm2[0] : 0.0);
    u[i] = nr > tiny ? (T)((double)u[i] / nr) : T(0);
}

// Node-block Jacobi: invert each diagonal block (masked rows/cols give identity).
template <int DIM, class T>
__global__ void k_block_inverse(int64_t n_n, const int32_t *__restrict__ diag, const double *__restrict__ val,
                                T *__restrict__ inv)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_n) return;
    const double *a = val + (int64_t)diag[r] * DIM * DIM;
    T *o = inv + r * DIM * DIM;
    if (DIM == 2) {
        const double det = a[0] * a[3] - a[1] * a[2];
        o[0] = a[3] / det; o[1] = -a[1] / det; o[2] = -a[2] / det; o[3] = a[0] / det;
    } else {
        const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
        const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
        const double id = 1.0 / det;
        o[0] = c00 * id; o[1] = (a[2] * a[7] - a[1] * a[8]) * id; o[2] = (a[1] * a[5] - a[2] * a[4]) * id;
        o[3] = c01 * id; o[4] = (a[0] * a[8] - a[2] * a[6]) * id; o[5] = (a[2] * a[3] - a[0] * a[5]) * id;
        o[6] = c02 * id; o[7] = (a[1] * a[6] - a[0] * a[7]) * id; o[8] = (a[0] * a[4] - a[1] * a[3]) * id;
    }
}

template <int DIM, class T>
__global__ void k_block_apply(int64_t n_n, const T *__restrict__ inv, const T *__restrict__ x,
                              const double *__restrict__ mask, T *__restrict__ y)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_n) return;
    const T *o = inv + r * DIM * DIM;
    T xv[DIM];
#pragma unroll
    for (int j = 0; j < DIM; ++j) xv[j] = x[r * DIM + j];
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        T s = T(0);
#pragma unroll
        for (int j = 0; j < DIM; ++j) s += o[i * DIM + j] * xv[j];
        y[r * DIM + i] = mask ? T(mask[r * DIM + i]) * s : s;
    }
}

// Chebyshev: d = c3*dinv.*r ; x = d
template <class T>
__global__ void k_cheb_first(int64_t n, double c3, const T *__restrict__ dinv, const T *__restrict__ r,
                             T *__restrict__ d, T *__restrict__ x)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T dn = T(c3) * dinv[i] * r[i];
    d[i] = dn;
    x[i] = dn;
}

// Chebyshev: d = c2*d + c3*dinv.*r ; x += d
template <class T>
__global__ void k_cheb_step(int64_t n, double c2, double c3, const T *__restrict__ dinv, const T *__restrict__ r,
                            T *__restrict__ d, T *__restrict__ x)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T dn = c2 == 0.0 ? T(c3) * dinv[i] * r[i] : T(c2) * d[i] + T(c3) * dinv[i] * r[i];
    d[i] = dn;
    x[i] += dn;
}

// Chebyshev with node-block scaling: d = c2 d + c3 Binv r ; x = d or x += d
template <int DIM, class T>
__global__ void k_cheb_block(int64_t n_n, double c2, double c3, const T *__restrict__ Binv, const T *__restrict__ r,
                             T *__restrict__ d, T *__restrict__ x, int first)
{
    const int64_t nd = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (nd >= n_n) return;
    const T *B = Binv + nd * DIM * DIM;
    T rv[DIM];
#pragma unroll
    for (int j = 0; j < DIM; ++j) rv[j] = r[nd * DIM + j];
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        T s = T(0);
#pragma unroll
        for (int j = 0; j < DIM; ++j) s += B[i * DIM + j] * rv[j];
        const int64_t q = nd * DIM + i;
        const T dn = (first || c2 == 0.0) ? T(c3) * s : T(c2) * d[q] + T(c3) * s;
        d[q] = dn;
        x[q] = first ? dn : x[q] + dn;
    }
}

// k_cheb_block with one thread per unknown q = dim*node + i: the thread
// reads row i of the node's block and the node's three residual entries and
// forms the same sum and update as k_cheb_block's row i (the same bits), and
// consecutive lanes read and write consecutive entries of r, d and x.
template <int DIM, class T>
__global__ void k_cheb_block_dof(int64_t n_n, double c2, double c3, const T *__restrict__ Binv,
                                 const T *__restrict__ r, T *__restrict__ d, T *__restrict__ x, int first)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_n * DIM) return;
    const int64_t nd = q / DIM;
    const int i = (int)(q - nd * DIM);
    const T *B = Binv + nd * DIM * DIM + i * DIM;
    T s = T(0);
#pragma unroll
    for (int j = 0; j < DIM; ++j) s += B[j] * r[nd * DIM + j];
    const T dn = (first || c2 == 0.0) ? T(c3) * s : T(c2) * d[q] + T(c3) * s;
    d[q] = dn;
    x[q] = first ? dn : x[q] + dn;
}

// y = M x for a dense n x n row-major matrix (one row per wavefront)
template <class T>
__global__ __launch_bounds__(256) void k_dense_gemv(int n, const T *__restrict__ M, const T *__restrict__ x,
                                                    T *__restrict__ y)
{
    const int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int l = threadIdx.x & 63;
    double s = 0.0;
    if (w < n)
        for (int j = l; j < n; j += 64) s += (double)M[(int64_t)w * n + j] * (double)x[j];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (w < n && l == 0) y[w] = (T)s;
}

// element-wise z = a .* b
__global__ void k_ewmul(int64_t n, const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ z)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = a[i] * b[i];
}

// Reduction scratch per stream, so that threads running independent
// problems on their own streams never share partial sums.
struct RedScratch {
    double *part = nullptr;  // kMdBlocks * 128 partial sums
    double *red = nullptr;   // 128 reduced values
};
std::mutex g_scr_mu;
std::unordered_map<hipStream_t, RedScratch> g_scr;

int red_scratch(RedScratch **out)
{
    const hipStream_t st = ssr_stream();
    std::lock_guard<std::mutex> lk(g_scr_mu);
    RedScratch &r = g_scr[st];  // element references survive rehashing
    if (!r.part) SSR_HIP_CHECK(hipMalloc((void **)&r.part, sizeof(double) * kMdBlocks * 128));
    if (!r.red) SSR_HIP_CHECK(hipMalloc((void **)&r.red, sizeof(double) * 128));
    *out = &r;
    return SSR_OK;
}

// Partial sums of ssr_dev_multidot_noalloc: kMdBlocks x 128 doubles (8 MiB),
// held by every device the library runs on until the process ends.  Calls
// that overlap on one device share it and race.
constexpr int kNoAllocMaxCols = 128;
static __device__ __align__(256) double scratch[(size_t)kMdBlocks * kNoAllocMaxCols];

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
double *g_noalloc_ptr[kMaxDevices] = {};

// device address of `scratch` on the current device (looked up on the first
// call there)
int noalloc_scratch(double **out)
{
    int dev = 0;
    SSR_HIP_CHECK(hipGetDevice(&dev));
    SSR_REQUIRE(dev >= 0 && dev < kMaxDevices, "multidot_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            ssr_set_error("multidot_noalloc: hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
            return SSR_ERR_HIP;
        }
        g_noalloc_ptr[dev] = (double *)p;
    }
    *out = g_noalloc_ptr[dev];
    return SSR_OK;
}

unsigned red_blocks(int64_t n)
{
    int64_t b = (n + kRedThreads - 1) / kRedThreads;
    return (unsigned)(b < 1 ? 1 : (b > kRedBlocks ? kRedBlocks : b));
}

#define SSR_LAUNCH1(K_, n_, ...)                                                                      \
    do {                                                                                              \
        if ((n_) == 0) return SSR_OK;                                                                 \
        hipLaunchKernelGGL(K_, dim3(ssr_blocks((n_), 256)), dim3(256), 0, ssr_stream(), __VA_ARGS__); \
        SSR_HIP_CHECK(hipGetLastError());                                                             \
        return SSR_OK;                                                                                \
    } while (0)

}  // namespace

namespace ssr {

template <class T>
int spmv_bsr(int dim, int64_t n_rows, const int32_t *row_ptr, const int32_t *col, const T *val, const T *x, T *y,
             double alpha, double beta, const T *z, const int32_t *diag, const int32_t *tblk)
{
    if (n_rows == 0) return SSR_OK;
    if (!z) z = y;
    const bool mirror = diag && tblk && ssr_tuning(SSR_TUNE_SPMV_MIRROR);
    int G = ssr_tuning(SSR_TUNE_SPMV_GROUP);
    int U = ssr_tuning(SSR_TUNE_SPMV_UNROLL);
    if (G == 0) {
        // measured best for P2 node-block rows (3D: 27 or 65 blocks, 28 on
        // average; 2D: about 11)
        G = dim == 3 ? 16 : 4;
        if (U == 0) U = dim == 3 ? 2 : 1;
    }
    if (U == 0) U = 1;
    const int64_t threads = n_rows * G;
    const dim3 grid(ssr_blocks(threads, 256)), blk(256);
    if (!mirror && ssr_tuning(SSR_TUNE_SPMV_STAGE) &&
        ((dim == 3 && G == 16 && U == 2) || (dim == 2 && G == 4 && U == 1))) {
        const dim3 gw(ssr_blocks(threads, 64)), bw(64);
        if (dim == 3)
            hipLaunchKernelGGL((k_spmv_bsr_lds<3, 16, T, 2, 192>), gw, bw, 0, ssr_stream(), n_rows, row_ptr, col, val,
                               x, y, alpha, beta, z);
        else
            hipLaunchKernelGGL((k_spmv_bsr_lds<2, 4, T, 1, 320>), gw, bw, 0, ssr_stream(), n_rows, row_ptr, col, val,
                               x, y, alpha, beta, z);
        SSR_HIP_CHECK(hipGetLastError());
        return SSR_OK;
    }
#define SPMV_M(D_, G_, U_, M_)                                                                                  \
    hipLaunchKernelGGL((k_spmv_bsr<D_, G_, T, U_, M_>), grid, blk, 0, ssr_stream(), n_rows, row_ptr, col, val, x, y, \
                       alpha, beta, z, diag, tblk)
#define SPMV(D_, G_)                              \
    do {                                          \
        if (U == 2) SPMV_M(D_, G_, 2, false);     \
        else SPMV_M(D_, G_, 1, false);            \
    } while (0)
    // the mirror form exists for the default shapes (other shapes: plain)
    if (mirror && dim == 3 && G == 16 && U == 2) SPMV_M(3, 16, 2, true);
    else if (mirror && dim == 2 && G == 4 && U == 1) SPMV_M(2, 4, 1, true);
    else if (G == 64) {
        const dim3 gw(ssr_blocks(n_rows * 64, 256));
        if (dim == 3)
            hipLaunchKernelGGL((k_spmv_bsr_w<3, T>), gw, blk, 0, ssr_stream(), n_rows, row_ptr, col, val, x, y, alpha,
                               beta, z);
        else
            hipLaunchKernelGGL((k_spmv_bsr_w<2, T>), gw, blk, 0, ssr_stream(), n_rows, row_ptr, col, val, x, y, alpha,
                               beta, z);
    } else if (dim == 3) {
        if (G == 4) SPMV(3, 4);
        else if (G == 16) SPMV(3, 16);
        else if (G == 32) SPMV(3, 32);
        else SPMV(3, 8);
    } else {
        if (G == 4) SPMV(2, 4);
        else if (G == 16) SPMV(2, 16);
        else if (G == 32) SPMV(2, 32);
        else SPMV(2, 8);
    }
#undef SPMV
#undef SPMV_M
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

template <class T>
int spmv_csr(int64_t n_rows, const int64_t *rp, const int32_t *ci, const T *v, const T *x, T *y, double alpha,
             double beta, const T *z, int64_t nnz)
{
    if (n_rows == 0) return SSR_OK;
    if (!z) z = y;
    // lanes per row from the average row length (8 when unknown): the
    // Galerkin coarse levels and the restriction have rows of hundreds
    const int64_t avg = nnz < 0 ? 24 : nnz / n_rows;
    const dim3 blk(256);
#define CSR(G_) hipLaunchKernelGGL((k_spmv_csr<G_, T>), dim3(ssr_blocks(n_rows * G_, 256)), blk, 0, ssr_stream(), \
                                   n_rows, rp, ci, v, x, y, alpha, beta, z)
    if (avg <= 12) CSR(4);
    else if (avg <= 48) CSR(8);
    else if (avg <= 96) CSR(16);
    else if (avg <= 192) CSR(32);
    else CSR(64);
#undef CSR
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

template <class T>
int spmv_brc(int rb, int cb, int64_t n_rows, const int32_t *rp, const int32_t *col, const T *val, const T *x, T *y,
             double alpha, double beta, const T *z, int64_t nblk)
{
    if (n_rows == 0) return SSR_OK;
    if (!z) z = y;
    const int64_t avg = nblk / n_rows;
    int G = avg <= 6 ? 4 : avg <= 12 ? 8 : avg <= 40 ? 16 : avg <= 96 ? 32 : 64;
    // few long rows (the first coarse AMG level: some 10^4 aggregate rows of
    // about 70 blocks): more lanes per row until the grid fills the device
    while (G < 64 && n_rows * G < (int64_t)ssr_cu_count() * 2048 && avg > G / 2) G *= 2;
    const dim3 blk(256);
    const bool halve = ssr_tuning(SSR_TUNE_HALVING_SUMS) != 0;
    // the prefetching form pays for the transfers' 3 x 6 and 6 x 3 blocks in
    // float (51.9 -> 44.5 and 46.9 -> 39.2 us on the 1.5M slope) but not in
    // double (94 -> 111 and 86 -> 101 us) nor for the 6 x 6 coarse operator
    // (float 11.5 -> 16.0 us): the second block doubles the value registers
    const bool pf = ssr_tuning(SSR_TUNE_BRC_PREFETCH) != 0 && rb != cb && sizeof(T) == 4;
#define BRC_G(RB_, CB_, G_)                                                                                        \
    do {                                                                                                           \
        const dim3 grid(ssr_blocks(n_rows * G_, 256));                                                             \
        if (halve && pf)                                                                                           \
            hipLaunchKernelGGL((k_spmv_brc<RB_, CB_, G_, T, true, true>), grid, blk, 0, ssr_stream(), n_rows, rp, col, \
                               val, x, y, alpha, beta, z);                                                         \
        else if (halve)                                                                                            \
            hipLaunchKernelGGL((k_spmv_brc<RB_, CB_, G_, T, true>), grid, blk, 0, ssr_stream(), n_rows, rp, col, val, \
                               x, y, alpha, beta, z);                                                              \
        else                                                                                                       \
            hipLaunchKernelGGL((k_spmv_brc<RB_, CB_, G_, T, false>), grid, blk, 0, ssr_stream(), n_rows, rp, col,    \
                               val, x, y, alpha, beta, z);                                                         \
    } while (0)
#define BRC(RB_, CB_)                                                                                              \
    do {                                                                                                           \
        if (G == 4) BRC_G(RB_, CB_, 4);                                                                            \
        else if (G == 8) BRC_G(RB_, CB_, 8);                                                                       \
        else if (G == 16) BRC_G(RB_, CB_, 16);                                                                     \
        else if (G == 32) BRC_G(RB_, CB_, 32);                                                                     \
        else BRC_G(RB_, CB_, 64);                                                                                  \
    } while (0)
    if (rb == 6 && cb == 6) BRC(6, 6);
    else if (rb == 3 && cb == 6) BRC(3, 6);
    else if (rb == 6 && cb == 3) BRC(6, 3);
    else if (rb == 3 && cb == 3) BRC(3, 3);
    else if (rb == 2 && cb == 3) BRC(2, 3);
    else if (rb == 3 && cb == 2) BRC(3, 2);
    else {
        ssr_set_error("spmv_brc: block shape %d x %d not supported", rb, cb);
        return SSR_ERR_ARG;
    }
#undef BRC
#undef BRC_G
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

bool brc_supported(int rb, int cb)
{
    return (rb == 6 && cb == 6) || (rb == 3 && cb == 6) || (rb == 6 && cb == 3) || (rb == 3 && cb == 3) ||
           (rb == 2 && cb == 3) || (rb == 3 && cb == 2);
}

template <class T>
int dot(int64_t n, const T *a, const T *b, const double *mask, double *result)
{
    RedScratch *sc = nullptr;
    int rc = red_scratch(&sc);
    if (rc) return rc;
    const unsigned nb = red_blocks(n);
    hipLaunchKernelGGL((k_dot_partial<T>), dim3(nb), dim3(kRedThreads), 0, ssr_stream(), n, a, b, mask, sc->part);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kRedThreads), 0, ssr_stream(), (int)nb, 1, sc->part, result);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

// h = V(:,0:k)' u with the partial sums in `part` (kMdBlocks * k doubles at
// most; every entry the reduction reads is written first, so `part` needs no
// clearing).
template <class T>
int multidot_into(int64_t n, int k, const T *V, int64_t ld, const T *u, double *h, double *part)
{
    SSR_REQUIRE(k >= 0 && k <= 128, "multidot: k must be in [0,128]");
    SSR_REQUIRE(n >= 0, "multidot: n must be >= 0");
    if (k == 0) return SSR_OK;
    // rows per thread: 4 for long columns, 2 otherwise (measured at 21k,
    // 84k and 711k rows)
    const int rpt = n >= 262144 ? 4 : 2;
    const int64_t nbl = (n + (int64_t)kRedThreads * rpt - 1) / ((int64_t)kRedThreads * rpt);
    const unsigned nb = (unsigned)(nbl < 1 ? 1 : (nbl > kMdBlocks ? kMdBlocks : nbl));
    const dim3 grid(nb, (unsigned)((k + kMdCols - 1) / kMdCols));
    const bool pf = ssr_tuning(SSR_TUNE_KRYLOV_PREFETCH) != 0;
    if (rpt == 4 && pf)
        hipLaunchKernelGGL((k_multidot_partial<T, 4, true>), grid, dim3(kRedThreads), 0, ssr_stream(), n, k, V, ld,
                           u, part);
    else if (rpt == 4)
        hipLaunchKernelGGL((k_multidot_partial<T, 4>), grid, dim3(kRedThreads), 0, ssr_stream(), n, k, V, ld, u,
                           part);
    else if (pf)
        hipLaunchKernelGGL((k_multidot_partial<T, 2, true>), grid, dim3(kRedThreads), 0, ssr_stream(), n, k, V, ld,
                           u, part);
    else
        hipLaunchKernelGGL((k_multidot_partial<T, 2>), grid, dim3(kRedThreads), 0, ssr_stream(), n, k, V, ld, u,
                           part);
    hipLaunchKernelGGL(k_reduce_partials, dim3(k), dim3(kRedThreads), 0, ssr_stream(), (int)nb, k, part, h);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

template <class T>
int multidot(int64_t n, int k, const T *V, int64_t ld, const T *u, double *h)
{
    SSR_REQUIRE(k >= 0 && k <= 128, "multidot: k must be in [0,128]");
    if (k == 0) return SSR_OK;
    RedScratch *sc = nullptr;
    int rc = red_scratch(&sc);
    if (rc) return rc;
    return multidot_into<T>(n, k, V, ld, u, h, sc->part);
}

template <class TV, class TU>
int gemv_update(int64_t n, int k, const TV *V, int64_t ld, const double *h, double alpha, double beta, TU *u)
{
    SSR_REQUIRE(k >= 0 && k <= 128, "gemv_update: k must be in [0,128]");
    // the direct form pays with double columns (46.9 -> 43.5 us on the 1.5M
    // slope) but not with float ones (24.1 -> 30.7 us)
    if (sizeof(TV) == 8 && ssr_tuning(SSR_TUNE_KRYLOV_PREFETCH))
        SSR_LAUNCH1((k_gemv_update<TV, TU, true>), n, n, k, V, ld, h, alpha, beta, u);
    SSR_LAUNCH1((k_gemv_update<TV, TU>), n, n, k, V, ld, h, alpha, beta, u);
}

template <class T>
int axpby(int64_t n, double a, const T *x, double b, T *y)
{
    SSR_LAUNCH1(k_axpby<T>, n, n, a, x, b, y);
}

template <class T>
int lincomb(int64_t n, double a, const T *x, double b, const T *y, T *z)
{
    SSR_LAUNCH1(k_lincomb<T>, n, n, a, x, b, y, z);
}

template <class T>
int mul(int64_t n, const double *m, T *x)
{
    SSR_LAUNCH1(k_mul<T>, n, n, m, x);
}

template <class T>
int fill(int64_t n, double v, T *x)
{
    SSR_LAUNCH1(k_fill<T>, n, n, v, x);
}

template <class TI, class TO>
int convert(int64_t n, double a, const TI *x, TO *y)
{
    SSR_LAUNCH1((k_convert<TI, TO>), n, n, a, x, y);
}

template <class TI, class TO>
int axpy(int64_t n, double a, const TI *x, TO *y)
{
    SSR_LAUNCH1((k_axpy<TI, TO>), n, n, a, x, y);
}

template <class T>
int scale_inv_sqrt(int64_t n, const double *nrm2, double tiny, T *u)
{
    SSR_LAUNCH1(k_scale_inv_sqrt<T>, n, n, nrm2, tiny, u);
}

template <class T>
int block_jacobi_setup(int dim, int64_t n_n, const int32_t *diag, const double *val, T *inv)
{
    if (dim == 3) SSR_LAUNCH1((k_block_inverse<3, T>), n_n, n_n, diag, val, inv);
    SSR_LAUNCH1((k_block_inverse<2, T>), n_n, n_n, diag, val, inv);
}

template <class T>
int block_jacobi_apply(int dim, int64_t n_n, const T *inv, const T *x, const double *mask, T *y)
{
    if (dim == 3) SSR_LAUNCH1((k_block_apply<3, T>), n_n, n_n, inv, x, mask, y);
    SSR_LAUNCH1((k_block_apply<2, T>), n_n, n_n, inv, x, mask, y);
}

template <class T>
int cheb_first(int64_t n, double c3, const T *dinv, const T *r, T *d, T *x)
{
    SSR_LAUNCH1(k_cheb_first<T>, n, n, c3, dinv, r, d, x);
}

template <class T>
int cheb_step(int64_t n, double c2, double c3, const T *dinv, const T *r, T *d, T *x)
{
    SSR_LAUNCH1(k_cheb_step<T>, n, n, c2, c3, dinv, r, d, x);
}

template <class T>
int cheb_block_step(int dim, int64_t n_n, double c2, double c3, const T *Binv, const T *r, T *d, T *x, int first)
{
    if (ssr_tuning(SSR_TUNE_CHEB_PER_DOF)) {
        if (dim == 3) SSR_LAUNCH1((k_cheb_block_dof<3, T>), n_n * 3, n_n, c2, c3, Binv, r, d, x, first);
        SSR_LAUNCH1((k_cheb_block_dof<2, T>), n_n * 2, n_n, c2, c3, Binv, r, d, x, first);
    }
    if (dim == 3) SSR_LAUNCH1((k_cheb_block<3, T>), n_n, n_n, c2, c3, Binv, r, d, x, first);
    SSR_LAUNCH1((k_cheb_block<2, T>), n_n, n_n, c2, c3, Binv, r, d, x, first);
}

template <class T>
int dense_gemv(int n, const T *M, const T *x, T *y)
{
    if (n == 0) return SSR_OK;
    hipLaunchKernelGGL(k_dense_gemv<T>, dim3(ssr_blocks((int64_t)n * 64, 256)), dim3(256), 0, ssr_stream(), n, M, x,
                       y);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

template <class T>
int spmm_bsr(int dim, int64_t n_rows, const int32_t *row_ptr, const int32_t *col, const T *val, const T *X,
             int64_t ldx, T *Y, int64_t ldy, int nv)
{
    SSR_REQUIRE(dim == 2 || dim == 3, "spmm_bsr: dim must be 2 or 3");
    SSR_REQUIRE(nv >= 0 && nv <= kSpmmNV, "spmm_bsr: nv must be in [0,8]");
    SSR_REQUIRE(ldx >= n_rows * dim && ldy >= n_rows * dim, "spmm_bsr: leading dimension too small");
    if (n_rows == 0 || nv == 0) return SSR_OK;
    const int G = ssr_tuning(SSR_TUNE_SPMM_GROUP) > 0 ? ssr_tuning(SSR_TUNE_SPMM_GROUP) : (dim == 3 ? 16 : 4);
    const dim3 grid(ssr_blocks(n_rows * G, 256)), blk(256);
#define SPMM(D_, G_)                                                                                          \
    do {                                                                                                      \
        if (ssr_tuning(SSR_TUNE_HALVING_SUMS))                                                                 \
            hipLaunchKernelGGL((k_spmm_bsr<D_, G_, T, true>), grid, blk, 0, ssr_stream(), n_rows, row_ptr, col, val, \
                               X, ldx, Y, ldy, nv);                                                           \
        else                                                                                                  \
            hipLaunchKernelGGL((k_spmm_bsr<D_, G_, T, false>), grid, blk, 0, ssr_stream(), n_rows, row_ptr, col,     \
                               val, X, ldx, Y, ldy, nv);                                                      \
    } while (0)
    if (dim == 3) {
        if (G == 4) SPMM(3, 4);
        else if (G == 8) SPMM(3, 8);
        else if (G == 32) SPMM(3, 32);
        else SPMM(3, 16);
    } else {
        if (G == 8) SPMM(2, 8);
        else if (G == 16) SPMM(2, 16);
        else if (G == 32) SPMM(2, 32);
        else SPMM(2, 4);
    }
#undef SPMM
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

namespace {
// row chunks of the Gram product: about 2048 blocks in all, each chunk a
// multiple of the slab depth and at least 8 slabs deep
void gram_chunks(int64_t n, int k, int *tiles, int64_t *rows_per_chunk, int *chunks)
{
    const int t = (k + kTile - 1) / kTile;
    const int64_t want = (2048 + (int64_t)t * t - 1) / ((int64_t)t * t);
    int64_t rpc = (n + want - 1) / want;
    if (rpc < 8 * kSlab) rpc = 8 * kSlab;
    rpc = (rpc + kSlab - 1) / kSlab * kSlab;
    *tiles = t;
    *rows_per_chunk = rpc;
    *chunks = (int)((n + rpc - 1) / rpc);
}
}  // namespace

int64_t gram_work_len(int64_t n, int k)
{
    if (n <= 0 || k <= 0) return 0;
    int t, nc;
    int64_t rpc;
    gram_chunks(n, k, &t, &rpc, &nc);
    return (int64_t)nc * k * k;
}

int gram(int64_t n, int k, const double *X, const double *Y, double *G, double *work, int64_t work_len,
         int symmetric)
{
    SSR_REQUIRE(k >= 0 && k <= 16384, "gram: k out of range");
    if (k == 0) return SSR_OK;
    const int64_t kk = (int64_t)k * k;
    if (n == 0) return fill<double>(kk, 0.0, G);
    int t, nc;
    int64_t rpc;
    gram_chunks(n, k, &t, &rpc, &nc);
    SSR_REQUIRE(work && work_len >= (int64_t)nc * kk, "gram: workspace too small");
    SSR_REQUIRE((int64_t)t * t <= 0x7fffffff && nc <= 65535, "gram: grid too large");
    hipLaunchKernelGGL(k_gram_partial, dim3((unsigned)(t * t), (unsigned)nc), dim3(256), 0, ssr_stream(), n, k, t,
                       rpc, X, Y, work, symmetric ? 1 : 0);
    hipLaunchKernelGGL(k_sum_chunks, dim3(ssr_blocks(kk, 256)), dim3(256), 0, ssr_stream(), nc, k, work, G,
                       symmetric ? 1 : 0);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

int gemm_tall(int64_t n, int k, int m, const double *X, const double *C, double *Y, const int *qstart, float *Y32)
{
    SSR_REQUIRE(k >= 0 && m >= 0 && m <= 65535 * kTile, "gemm_tall: sizes out of range");
    if (n == 0 || m == 0) return SSR_OK;
    if (k == 0) {
        int rc = fill<double>(n * m, 0.0, Y);
        return (rc || !Y32) ? rc : fill<float>(n * m, 0.0, Y32);
    }
    const int64_t rb = (n + kTile - 1) / kTile;
    SSR_REQUIRE(rb <= 0x7fffffff, "gemm_tall: too many rows");
    hipLaunchKernelGGL(k_gemm_tall, dim3((unsigned)rb, (unsigned)((m + kTile - 1) / kTile)), dim3(256), 0,
                       ssr_stream(), n, k, m, X, C, Y, qstart, Y32);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

#define SSR_INST(T)                                                                                                  \
    template int spmm_bsr<T>(int, int64_t, const int32_t *, const int32_t *, const T *, const T *, int64_t, T *,      \
                             int64_t, int);                                                                          \
    template int spmv_bsr<T>(int, int64_t, const int32_t *, const int32_t *, const T *, const T *, T *, double,       \
                             double, const T *, const int32_t *, const int32_t *);                                   \
    template int spmv_csr<T>(int64_t, const int64_t *, const int32_t *, const T *, const T *, T *, double, double,   \
                             const T *, int64_t);                                                                    \
    template int spmv_brc<T>(int, int, int64_t, const int32_t *, const int32_t *, const T *, const T *, T *, double,  \
                             double, const T *, int64_t);                                                            \
    template int dot<T>(int64_t, const T *, const T *, const double *, double *);                                    \
    template int multidot<T>(int64_t, int, const T *, int64_t, const T *, double *);                                 \
    template int axpby<T>(int64_t, double, const T *, double, T *);                                                   \
    template int lincomb<T>(int64_t, double, const T *, double, const T *, T *);                                      \
    template int mul<T>(int64_t, const double *, T *);                                                               \
    template int fill<T>(int64_t, double, T *);                                                                      \
    template int scale_inv_sqrt<T>(int64_t, const double *, double, T *);                                            \
    template int block_jacobi_setup<T>(int, int64_t, const int32_t *, const double *, T *);                          \
    template int block_jacobi_apply<T>(int, int64_t, const T *, const T *, const double *, T *);                     \
    template int cheb_first<T>(int64_t, double, const T *, const T *, T *, T *);                                      \
    template int cheb_step<T>(int64_t, double, double, const T *, const T *, T *, T *);                               \
    template int dense_gemv<T>(int, const T *, const T *, T *);                                                       \
    template int cheb_block_step<T>(int, int64_t, double, double, const T *, const T *, T *, T *, int);
SSR_INST(double)
SSR_INST(float)
#undef SSR_INST
template int gemv_update<double, double>(int64_t, int, const double *, int64_t, const double *, double, double,
                                         double *);
template int gemv_update<float, float>(int64_t, int, const float *, int64_t, const double *, double, double, float *);
template int gemv_update<float, double>(int64_t, int, const float *, int64_t, const double *, double, double,
                                        double *);
template int convert<double, float>(int64_t, double, const double *, float *);
template int convert<float, double>(int64_t, double, const float *, double *);
template int convert<double, double>(int64_t, double, const double *, double *);
template int axpy<float, double>(int64_t, double, const float *, double *);
template int axpy<double, double>(int64_t, double, const double *, double *);

}  // namespace ssr

// ---------------------------------------------------------------------------
// C ABI (double, device pointers)
// ---------------------------------------------------------------------------
extern "C" {

int ssr_dev_spmv_bsr(int dim, int64_t n_rows, const int32_t *row_ptr, const int32_t *col, const double *val,
                     const double *x, double *y, double alpha, double beta)
{
    return ssr::spmv_bsr<double>(dim, n_rows, row_ptr, col, val, x, y, alpha, beta);
}

int ssr_dev_spmv_bsr_f32(int dim, int64_t n_rows, const int32_t *row_ptr, const int32_t *col, const float *val,
                         const float *x, float *y, double alpha, double beta)
{
    return ssr::spmv_bsr<float>(dim, n_rows, row_ptr, col, val, x, y, alpha, beta);
}

int ssr_dev_spmv_bsr_sym(int dim, int64_t n_rows, const int32_t *row_ptr, const int32_t *col, const int32_t *diag,
                         const int32_t *tblk, const double *val, const double *x, double *y, double alpha, double beta)
{
    SSR_REQUIRE(diag && tblk, "spmv_bsr_sym: diag and tblk are required");
    return ssr::spmv_bsr<double>(dim, n_rows, row_ptr, col, val, x, y, alpha, beta, nullptr, diag, tblk);
}

int ssr_dev_spmv_bsr_sym_f32(int dim, int64_t n_rows, const int32_t *row_ptr, const int32_t *col,
                             const int32_t *diag, const int32_t *tblk, const float *val, const float *x, float *y,
                             double alpha, double beta)
{
    SSR_REQUIRE(diag && tblk, "spmv_bsr_sym_f32: diag and tblk are required");
    return ssr::spmv_bsr<float>(dim, n_rows, row_ptr, col, val, x, y, alpha, beta, nullptr, diag, tblk);
}

int ssr_dev_convert_f64_f32(int64_t n, const double *x, float *y) { return ssr::convert<double, float>(n, 1.0, x, y); }

// result[0] = sum_i a_i b_i (m_i) written to DEVICE memory
int ssr_dev_dot(int64_t n, const double *a, const double *b, const double *mask, double *result)
{
    return ssr::dot<double>(n, a, b, mask, result);
}

// host-returning dot (synchronises)
int ssr_dot(int64_t n, const double *a, const double *b, const double *mask, double *host_result)
{
    RedScratch *sc = nullptr;
    int rc = red_scratch(&sc);
    if (rc) return rc;
    if ((rc = ssr::dot<double>(n, a, b, mask, sc->red))) return rc;
    SSR_HIP_CHECK(hipMemcpyAsync(host_result, sc->red, sizeof(double), hipMemcpyDeviceToHost, ssr_stream()));
    SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));
    return SSR_OK;
}

// max over nodes of the Euclidean norm of U's node vector (U dim x n_n
// node-major), returned to the host (synchronises): the U_max of the
// continuation histories (SSR_indirect_continuation.m:92-93,182)
int ssr_node_norm_max(int dim, int64_t n_n, const double *U, double *host_result)
{
    SSR_REQUIRE((dim == 2 || dim == 3) && n_n >= 0 && host_result && (n_n == 0 || U),
                "node_norm_max: bad arguments");
    RedScratch *sc = nullptr;
    int rc = red_scratch(&sc);
    if (rc) return rc;
    const unsigned nb = red_blocks(n_n);
    if (dim == 3)
        hipLaunchKernelGGL(k_node_sq_max_partial<3>, dim3(nb), dim3(kRedThreads), 0, ssr_stream(), n_n, U, sc->part);
    else
        hipLaunchKernelGGL(k_node_sq_max_partial<2>, dim3(nb), dim3(kRedThreads), 0, ssr_stream(), n_n, U, sc->part);
    hipLaunchKernelGGL(k_sqrt_max_partials, dim3(1), dim3(kRedThreads), 0, ssr_stream(), (int)nb, sc->part, sc->red);
    SSR_HIP_CHECK(hipGetLastError());
    SSR_HIP_CHECK(hipMemcpyAsync(host_result, sc->red, sizeof(double), hipMemcpyDeviceToHost, ssr_stream()));
    SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));
    return SSR_OK;
}

// h (device, k values) = V(:,0:k)' u
int ssr_dev_multidot(int64_t n, int k, const double *V, int64_t ld, const double *u, double *h)
{
    return ssr::multidot<double>(n, k, V, ld, u, h);
}

// The same multi-dot with its partial sums in the file-scope `scratch` array
// instead of library-allocated memory: no hipMalloc, hipFree or allocating
// library call on any path.  The array's device address is looked up once
// per device (with that device current) and kept per device; the name
// `scratch` is a host-side stand-in and is never passed to a kernel.
int ssr_dev_multidot_noalloc(int64_t n, int k, const double *V, int64_t ld, const double *u, double *h)
{
    SSR_REQUIRE(k >= 0 && k <= kNoAllocMaxCols,
                "multidot_noalloc: k = %d exceeds the %d columns of the static scratch (%d x %d doubles)", k,
                kNoAllocMaxCols, kMdBlocks, kNoAllocMaxCols);
    SSR_REQUIRE(n >= 0 && (k == 0 || (V && u && h)), "multidot_noalloc: bad arguments");
    if (k == 0) return SSR_OK;
    double *part = nullptr;
    int rc = noalloc_scratch(&part);
    if (rc) return rc;
    return ssr::multidot_into<double>(n, k, V, ld, u, h, part);
}

int ssr_dev_gemv_update(int64_t n, int k, const double *V, int64_t ld, const double *h, double alpha, double beta,
                        double *u)
{
    return ssr::gemv_update<double, double>(n, k, V, ld, h, alpha, beta, u);
}

// the float forms of the two (float basis and vector, double coefficients
// and sums): what the precision-32 Krylov iterations run
int ssr_dev_multidot_f32(int64_t n, int k, const float *V, int64_t ld, const float *u, double *h)
{
    return ssr::multidot<float>(n, k, V, ld, u, h);
}

int ssr_dev_gemv_update_f32(int64_t n, int k, const float *V, int64_t ld, const double *h, double alpha, double beta,
                            float *u)
{
    return ssr::gemv_update<float, float>(n, k, V, ld, h, alpha, beta, u);
}

int ssr_dev_axpby(int64_t n, double a, const double *x, double b, double *y) { return ssr::axpby<double>(n, a, x, b, y); }

int ssr_dev_lincomb(int64_t n, double a, const double *x, double b, const double *y, double *z)
{
    return ssr::lincomb<double>(n, a, x, b, y, z);
}

int ssr_dev_fill(int64_t n, double v, double *x) { return ssr::fill<double>(n, v, x); }

int ssr_dev_mul(int64_t n, const double *m, double *x) { return ssr::mul<double>(n, m, x); }

int ssr_dev_block_jacobi_setup(int dim, int64_t n_n, const int32_t *diag, const double *val, double *inv)
{
    return ssr::block_jacobi_setup<double>(dim, n_n, diag, val, inv);
}

int ssr_dev_block_jacobi_apply(int dim, int64_t n_n, const double *inv, const double *x, const double *mask,
                               double *y)
{
    return ssr::block_jacobi_apply<double>(dim, n_n, inv, x, mask, y);
}

int ssr_dev_csr_spmv(int64_t n_rows, const int64_t *rp, const int32_t *ci, const double *v, const double *x,
                     double *y, double alpha, double beta)
{
    return ssr::spmv_csr<double>(n_rows, rp, ci, v, x, y, alpha, beta);
}

int ssr_dev_ewmul(int64_t n, const double *a, const double *b, double *z)
{
    SSR_LAUNCH1(k_ewmul, n, n, a, b, z);
}

int ssr_dev_cheb_step(int64_t n, double c2, double c3, const double *dinv, const double *r, double *d, double *x)
{
    return ssr::cheb_step<double>(n, c2, c3, dinv, r, d, x);
}

int ssr_dev_dense_gemv(int n, const double *M, const double *x, double *y) { return ssr::dense_gemv<double>(n, M, x, y); }

}  // extern "C"
