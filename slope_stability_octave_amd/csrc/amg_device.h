// Smoothed-aggregation hierarchy built on the device (amg_device.hip) and the
// aggregation shared with the host build (amg.cpp).
#pragma once

#include <stdint.h>

#include <vector>

namespace ssr_amg {

// Greedy aggregation of a node graph given by its strong neighbours (sp/sc:
// CSR, neighbours of node i in sc[sp[i]..sp[i+1]) in the order they were
// found).  Pass 1: a node whose strong neighbours are all unaggregated roots
// an aggregate with them; pass 2: every other node joins the aggregate of its
// first aggregated strong neighbour; pass 3: nodes still left with strong
// neighbours root new aggregates.  agg[i] = -1 for nodes without strong
// neighbours.  Sequential: the result depends only on the graph.
std::vector<int32_t> aggregate_graph(int64_t nn, const std::vector<int64_t> &sp, const std::vector<int32_t> &sc,
                                     int32_t &n_agg);

// Block CSR on the device with rb x cb row-major blocks (ownership passes to
// whoever adopts it).
struct DevBlocks {
    int rb = 0, cb = 0;
    int64_t nbr = 0, nblk = 0;
    int32_t *brp = nullptr, *bcol = nullptr, *brow = nullptr;
    double *bval = nullptr;
    void release();
};

struct DevLevelOut {
    int rb = 0;           // unknowns per node
    int64_t n_nodes = 0;  // block rows of the level operator
    DevBlocks A;          // level operator (empty on level 0: the fine operator is the caller's)
    int32_t *diag = nullptr;
    double *dinv = nullptr;  // rb * n_nodes
    double rho = 1.0, rho_up = 1.0;
    DevBlocks P, R;  // prolongation from / restriction to the next level (empty on the coarsest)
    void release();
};

struct DevOptions {
    double theta = 0.0;
    int max_levels = 12;
    int64_t max_coarse = 600;
};

// Status of amg_device_build besides SSR_OK / errors: an aggregate lost one of
// the near-null-space vectors (its coarse unknowns would not come in whole
// blocks); the caller builds on the host instead.
constexpr int kNotUniform = 100;

// Build every level below the fine operator (dim x dim node blocks: d_rp,
// d_col, d_brow, d_diag, d_val on the device, h_rp / h_col the host copies of
// the pattern) with the near-null space d_B (n x k column-major, zero rows at
// Dirichlet DOFs).  out[0] describes the fine level (dinv, rho, P, R), out[l]
// for l >= 1 the coarse operators.
int amg_device_build(int dim, int64_t n_nodes, const int32_t *d_rp, const int32_t *d_col, const int32_t *d_brow,
                     const int32_t *d_diag, const double *d_val, const std::vector<int32_t> &h_rp,
                     const std::vector<int32_t> &h_col, const double *d_B, int k, const DevOptions &opt,
                     std::vector<DevLevelOut> &out, double *phase_s);

// max over node rows of sum_j |(D_B^-1 A)_ij| with D_B the diagonal blocks:
// the Gershgorin bound of the node-block Jacobi smoother on the fine level
int fine_block_gershgorin(int dim, int64_t n_nodes, const int32_t *d_rp, const int32_t *d_diag, const double *d_val,
                          double *out);

}  // namespace ssr_amg
