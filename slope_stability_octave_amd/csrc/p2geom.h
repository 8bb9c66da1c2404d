// P2 reference elements and the per-point geometry of the physical
// gradients, shared by the setup kernels (which store DPhi and WEIGHT) and the
// element kernels that compute the gradients of a point where they need them
// instead of reading stored ones.  Both go through the functions below, whose
// arithmetic is the reference's, in its order, with FP contraction off in
// each function body, so that a gradient computed here is the same bits
// wherever it is computed.
//
//  * Jacobian as the sum over the element's nodes of coordinate times
//    reference derivative, determinant and inverse by cofactors, physical
//    gradient as Jinv * reference gradient: elastic_stiffness_matrix_3D.m:60-96
//    and elastic_stiffness_matrix_2D.m:50-80.
//  * P2 derivatives: local_basis_volume_3D.m:75-106, local_basis_volume_2D.m:44-47.
//  * Quadrature: quadrature_volume_3D.m:39-42, quadrature_volume_2D.m:28-33.
//
// Coordinates are node-major (dim values per node); elem lists each
// element's n_p nodes contiguously.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace p2geom {

// 11-point rule on the reference tetrahedron
static __constant__ double kXi3[3][11] = {
    {1.0 / 4, 0.0714285714285714, 0.785714285714286, 0.0714285714285714, 0.0714285714285714, 0.399403576166799,
     0.100596423833201, 0.100596423833201, 0.399403576166799, 0.399403576166799, 0.100596423833201},
    {1.0 / 4, 0.0714285714285714, 0.0714285714285714, 0.785714285714286, 0.0714285714285714, 0.100596423833201,
     0.399403576166799, 0.100596423833201, 0.399403576166799, 0.100596423833201, 0.399403576166799},
    {1.0 / 4, 0.0714285714285714, 0.0714285714285714, 0.0714285714285714, 0.785714285714286, 0.100596423833201,
     0.100596423833201, 0.399403576166799, 0.100596423833201, 0.399403576166799, 0.399403576166799}};
static __constant__ double kWF3[11] = {-0.013155555555555, 0.007622222222222, 0.007622222222222, 0.007622222222222,
                                       0.007622222222222,  0.024888888888888, 0.024888888888888, 0.024888888888888,
                                       0.024888888888888,  0.024888888888888, 0.024888888888888};
// 7-point rule on the reference triangle
static __constant__ double kXi2[2][7] = {
    {0.1012865073235, 0.7974269853531, 0.1012865073235, 0.4701420641051, 0.4701420641051, 0.0597158717898, 1.0 / 3},
    {0.1012865073235, 0.1012865073235, 0.7974269853531, 0.0597158717898, 0.4701420641051, 0.4701420641051, 1.0 / 3}};
static __constant__ double kWF2[7] = {0.1259391805448 / 2, 0.1259391805448 / 2, 0.1259391805448 / 2,
                                      0.1323941527885 / 2, 0.1323941527885 / 2, 0.1323941527885 / 2, 0.225 / 2};

// reference derivatives of the ten P2 tetrahedron functions at point q
__device__ inline void tet_dhat(int q, double D1[10], double D2[10], double D3[10])
{
#pragma clang fp contract(off)
    const double x1 = kXi3[0][q], x2 = kXi3[1][q], x3 = kXi3[2][q];
    const double x0 = 1 - x1 - x2 - x3;
    D1[0] = -4 * x0 + 1; D1[1] = 4 * x1 - 1; D1[2] = 0;  D1[3] = 0;
    D1[4] = 4 * (x0 - x1); D1[5] = 4 * x2; D1[6] = -4 * x2; D1[7] = 4 * x3; D1[8] = 0; D1[9] = -4 * x3;
    D2[0] = -4 * x0 + 1; D2[1] = 0; D2[2] = 4 * x2 - 1; D2[3] = 0;
    D2[4] = -4 * x1; D2[5] = 4 * x1; D2[6] = 4 * (x0 - x2); D2[7] = 0; D2[8] = 4 * x3; D2[9] = -4 * x3;
    D3[0] = -4 * x0 + 1; D3[1] = 0; D3[2] = 0; D3[3] = 4 * x3 - 1;
    D3[4] = -4 * x1; D3[5] = 0; D3[6] = -4 * x2; D3[7] = 4 * x1; D3[8] = 4 * x2; D3[9] = 4 * (x0 - x3);
}

// reference derivatives of the six P2 triangle functions at point q
__device__ inline void tri_dhat(int q, double D1[6], double D2[6])
{
#pragma clang fp contract(off)
    const double x1 = kXi2[0][q], x2 = kXi2[1][q];
    const double x0 = 1 - x1 - x2;
    D1[0] = -4 * x0 + 1; D1[1] = 4 * x1 - 1; D1[2] = 0; D1[3] = 4 * x2; D1[4] = -4 * x2; D1[5] = 4 * (x0 - x1);
    D2[0] = -4 * x0 + 1; D2[1] = 0; D2[2] = 4 * x2 - 1; D2[3] = 4 * x1; D2[4] = 4 * (x0 - x2); D2[5] = -4 * x1;
}

// Inverse Jacobian (row-major iv[r*3+c]) and determinant at one point of a
// tetrahedron whose ten nodes are en[0..9].
__device__ inline void tet_jinv(const double *__restrict__ coord, const int32_t *__restrict__ en, const double H1[10],
                                const double H2[10], const double H3[10], double iv[9], double &DET)
{
#pragma clang fp contract(off)
    double J11 = 0, J12 = 0, J13 = 0, J21 = 0, J22 = 0, J23 = 0, J31 = 0, J32 = 0, J33 = 0;
#pragma unroll
    for (int a = 0; a < 10; ++a) {
        const int64_t v = en[a];
        const double X1 = coord[3 * v], X2 = coord[3 * v + 1], X3 = coord[3 * v + 2];
        if (a == 0) {
            J11 = X1 * H1[a]; J12 = X2 * H1[a]; J13 = X3 * H1[a];
            J21 = X1 * H2[a]; J22 = X2 * H2[a]; J23 = X3 * H2[a];
            J31 = X1 * H3[a]; J32 = X2 * H3[a]; J33 = X3 * H3[a];
        } else {
            J11 += X1 * H1[a]; J12 += X2 * H1[a]; J13 += X3 * H1[a];
            J21 += X1 * H2[a]; J22 += X2 * H2[a]; J23 += X3 * H2[a];
            J31 += X1 * H3[a]; J32 += X2 * H3[a]; J33 += X3 * H3[a];
        }
    }
    DET = J11 * (J22 * J33 - J32 * J23) - J12 * (J21 * J33 - J23 * J31) + J13 * (J21 * J32 - J22 * J31);
    iv[0] = (J22 * J33 - J23 * J32) / DET;
    iv[1] = -(J12 * J33 - J13 * J32) / DET;
    iv[2] = (J12 * J23 - J13 * J22) / DET;
    iv[3] = -(J21 * J33 - J23 * J31) / DET;
    iv[4] = (J11 * J33 - J13 * J31) / DET;
    iv[5] = -(J11 * J23 - J13 * J21) / DET;
    iv[6] = (J21 * J32 - J22 * J31) / DET;
    iv[7] = -(J11 * J32 - J12 * J31) / DET;
    iv[8] = (J11 * J22 - J12 * J21) / DET;
}

// Inverse Jacobian (iv[r*2+c]) and determinant at one point of a triangle
// whose six nodes are en[0..5].
__device__ inline void tri_jinv(const double *__restrict__ coord, const int32_t *__restrict__ en, const double H1[6],
                                const double H2[6], double iv[4], double &DET)
{
#pragma clang fp contract(off)
    double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        const int64_t v = en[a];
        const double X1 = coord[2 * v], X2 = coord[2 * v + 1];
        if (a == 0) {
            J11 = X1 * H1[a]; J12 = X2 * H1[a]; J21 = X1 * H2[a]; J22 = X2 * H2[a];
        } else {
            J11 += X1 * H1[a]; J12 += X2 * H1[a]; J21 += X1 * H2[a]; J22 += X2 * H2[a];
        }
    }
    DET = J11 * J22 - J12 * J21;
    iv[0] = J22 / DET;
    iv[1] = -J12 / DET;
    iv[2] = -J21 / DET;
    iv[3] = J11 / DET;
}

// physical gradient component r of basis function a: row r of Jinv times the
// reference gradient
__device__ inline double tet_grad(const double iv[9], const double H1[10], const double H2[10], const double H3[10],
                                  int r, int a)
{
#pragma clang fp contract(off)
    return iv[3 * r] * H1[a] + iv[3 * r + 1] * H2[a] + iv[3 * r + 2] * H3[a];
}

__device__ inline double tri_grad(const double iv[4], const double H1[6], const double H2[6], int r, int a)
{
#pragma clang fp contract(off)
    return iv[2 * r] * H1[a] + iv[2 * r + 1] * H2[a];
}

// Gradients of every basis function of one point (G[r][a]).
template <int DIM, int NP>
__device__ inline void point_gradients(const double *__restrict__ coord, const int32_t *__restrict__ en, int q,
                                       double (&G)[DIM][NP])
{
    if constexpr (DIM == 3) {
        double H1[10], H2[10], H3[10], iv[9], det;
        tet_dhat(q, H1, H2, H3);
        tet_jinv(coord, en, H1, H2, H3, iv, det);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int a = 0; a < 10; ++a) G[r][a] = tet_grad(iv, H1, H2, H3, r, a);
    } else {
        double H1[6], H2[6], iv[4], det;
        tri_dhat(q, H1, H2);
        tri_jinv(coord, en, H1, H2, iv, det);
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int a = 0; a < 6; ++a) G[r][a] = tri_grad(iv, H1, H2, r, a);
    }
}

}  // namespace p2geom
