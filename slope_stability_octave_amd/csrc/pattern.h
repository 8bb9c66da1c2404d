// The fixed-pattern record shared by the host builder (pattern.cpp) and the
// device builder (setup.hip).
#pragma once

#include <stdint.h>

#include <vector>

struct ssr_pattern {
    int dim = 0;
    int64_t n_n = 0, n_e = 0, n_p = 0;
    int64_t nnzb = 0;      // number of blocks
    int64_t n_free = 0;    // free DOFs
    int64_t nnz_free = 0;  // entries of K(Q,Q)
    int64_t n_ublk = 0;    // blocks on or above the block diagonal
    // true when the tables were built on the device (setup.hip): the host
    // vectors other than `mask` stay empty and ssr_pattern_get downloads
    bool device_built = false;
    // host copies
    std::vector<int32_t> row_ptr, col, emap, tblk, diag, brow;
    std::vector<int32_t> ublk;  // blocks on or above the block diagonal (row <= column), row-major order
    std::vector<int64_t> ref_pos;
    std::vector<uint8_t> mask;  // dim*n_n, 1 = free
    // device copies
    int32_t *d_row_ptr = nullptr, *d_col = nullptr, *d_emap = nullptr, *d_tblk = nullptr,
            *d_diag = nullptr, *d_elem = nullptr, *d_brow = nullptr, *d_ublk = nullptr;
    int64_t *d_ref_pos = nullptr;
    double *d_mask = nullptr;  // dim*n_n, 1.0 free / 0.0 Dirichlet
    // gather-form assembly: the element blocks that sum into each block, in
    // element order, as code = (e * n_pair + pair) * 2 + transposed
    std::vector<int32_t> cptr, ccode;
    int32_t *d_cptr = nullptr, *d_ccode = nullptr;
    // slot form: every element pair block has one slot in the contribution
    // list of its upper block (row <= column); lists in ublk order, elements
    // ascending inside a list.  pslot: n_e * n_pair slot indices, uptr: list
    // bounds per upper block
    std::vector<int32_t> uptr, pslot;
    int32_t *d_uptr = nullptr, *d_pslot = nullptr;
    mutable double *d_keblk = nullptr;  // n_e * n_pair * dim * dim element blocks (allocated on first use)
    // gather-form internal force: the element nodes (code = e * n_p + a) that
    // sum into each node, in element order
    std::vector<int32_t> nptr, ncode;
    int32_t *d_nptr = nullptr, *d_ncode = nullptr;
    mutable double *d_fe = nullptr;  // n_e * n_p * dim element-node forces (allocated on first use)
};
