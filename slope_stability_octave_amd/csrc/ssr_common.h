// Shared helpers for the SSR hot-path HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#define SSR_OK 0
#define SSR_ERR_ARG 1
#define SSR_ERR_HIP 2
#define SSR_ERR_ALLOC 3
#define SSR_ERR_STATE 4

extern "C" const char *ssr_last_error(void);
void ssr_set_error(const char *fmt, ...);

#define SSR_HIP_CHECK(expr)                                                              \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess) {                                                          \
            ssr_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, \
                          __LINE__);                                                     \
            return SSR_ERR_HIP;                                                          \
        }                                                                                \
    } while (0)

#define SSR_REQUIRE(cond, ...)          \
    do {                                \
        if (!(cond)) {                  \
            ssr_set_error(__VA_ARGS__); \
            return SSR_ERR_ARG;         \
        }                               \
    } while (0)

// Grid size for a 1-D launch over n items; callers never pass n <= 0.
static inline unsigned ssr_blocks(int64_t n, int threads)
{
    int64_t b = (n + threads - 1) / threads;
    return (unsigned)(b < 1 ? 1 : b);
}

// Cached device buffers (devpool.cpp): blocks freed here are kept for later
// requests of a similar size instead of being returned to the runtime.
int ssr_pool_malloc(void **out, size_t bytes);
void ssr_pool_free(void *p);

hipStream_t ssr_stream(void);
// divert ssr_stream() of the calling thread to `s` while capturing a graph
void ssr_capture_stream_push(hipStream_t s);
void ssr_capture_stream_pop(void);
// true once any host thread has set a stream of its own (ssr_set_thread_stream)
bool ssr_thread_streams_used(void);
int ssr_cu_count(void);  // compute units of the current device (cached)

// Preconditioner interface used by the Krylov solvers (solver.cpp): block
// Jacobi and smoothed-aggregation AMG implement it.
struct ssr_precond {
    virtual ~ssr_precond() {}
    // operator values (device); K32 is the float copy when the solver runs
    // its Krylov iterations in float (else nullptr)
    virtual int setup(const double *Kvals, const float *K32) = 0;
    virtual int apply(const double *x, double *y) = 0;  // y = M^{-1} x (device)
    virtual int apply(const float *x, float *y) = 0;    // same in float
    // the next setup rebuilds whatever it would otherwise keep
    virtual void invalidate() {}
    // the values of the next setups are symmetric bit for bit (K_r)
    virtual void set_symmetric(bool) {}
};

// Kernel-variant / launch-shape knobs (set by the autotuner or the caller).
#define SSR_TUNE_CONSTITUTIVE_VARIANT 0 /* 3D return-mapping kernel form, see constitutive.hip */
#define SSR_TUNE_SPMV_GROUP 1           /* lanes per block row in the node-block SpMV: 4, 8, 16, 32; 64 =
                                           one wavefront per row, lanes over the values; 0 = by dimension */
#define SSR_TUNE_PERSIST_WAVES 2        /* wavefronts per compute unit of persistent grids (0: auto) */
#define SSR_TUNE_AMG_GRAPH 3            /* 1: replay the V-cycle as a captured HIP graph, 0: eager launches */
#define SSR_TUNE_AMG_BLOCK_SMOOTH 4     /* 1: node-block Jacobi scaling in the finest-level Chebyshev smoother */
#define SSR_TUNE_ASSEMBLY_FORM 5        /* 0: element blocks (and element-node forces), then one
                                           deterministic gather per block (per node);
                                           1: FP64 atomic adds of the element contributions */
#define SSR_TUNE_SPMM_GROUP 6           /* lanes per block row of the multi-vector SpMV (0: by dimension) */
#define SSR_TUNE_SPMV_UNROLL 7          /* blocks per lane per pass of the node-block SpMV: 1 or 2 (0: by dimension) */
#define SSR_TUNE_ELEMENT_KERNEL 8       /* element pair blocks: 0 = two pairs of one column node per lane
                                           (k_assemble_pairs2; 2D as 2),
                                           1 = D*B shared through LDS per point (k_assemble),
                                           2 = per-lane products from LDS-staged point data (k_assemble_pairs) */
#define SSR_TUNE_AMG_COARSE_BLOCKS 9     /* 1: coarse AMG operators and transfers as block CSR over whole nodes
                                           (default), 0: scalar CSR */
#define SSR_TUNE_AMG_DEVICE_SETUP 10     /* 1: build the AMG hierarchy on the device (default), 0: on the host */
#define SSR_TUNE_SPMV_STAGE 11          /* 1: node-block SpMV with the block values staged through LDS,
                                           0: read from global memory by each lane (default) */
#define SSR_TUNE_HALVING_SUMS 12        /* lane-group sums of the multi-vector and block-CSR SpMVs: 1 = halving
                                           exchanges (default), 0 = full butterfly per value; the same bits */
#define SSR_TUNE_CHEB_PER_DOF 13        /* finest-level Chebyshev block step: 1 = one thread per unknown (default),
                                           0 = one thread per node; the same bits */
#define SSR_TUNE_SPMV_MIRROR 14         /* 1: node-block SpMV of a symmetric operator reads the blocks on and
                                           above the diagonal only, the others as transposed mirrors; the
                                           same bits; 0 (default): every block */
#define SSR_TUNE_CONSTITUTIVE_2D 15     /* 2D return mapping with DS: 1 = one wavefront per block with the
                                           strains read and S, DS written through LDS as contiguous spans
                                           (default), 0 = one lane per point, direct accesses (the form for
                                           S alone); the same bits */
#define SSR_TUNE_NEWTON_GRAPH 16        /* 1: the Newton loops replay their element phases (strain, stress and
                                           tangent, internal force; K_r assembly) from captured HIP graphs
                                           (default), 0: eager launches; the same bits */
#define SSR_TUNE_KRYLOV_PREFETCH 17     /* 1: the Krylov multi-dot issues a pass's loads before its products,
                                           and the double basis update reads its coefficients directly
                                           (default), 0: guarded loads in turn, coefficients through LDS;
                                           the same bits */
#define SSR_TUNE_BRC_PREFETCH 18        /* 1: the float block-CSR SpMV of the AMG transfers (rectangular blocks)
                                           loads a lane's next block with its current one (default), 0: in
                                           turn; the same bits */
#define SSR_TUNE_COUNT 19
int ssr_tuning(int key);
