// scn_node.h - SCN node algebra on the device (FP32).
//
// Port numbering follows the reference (TlmSolver.cpp:20-29):
//   p = axis*4 + side*2 + pol, side 0 = '-' face, POL[axis][pol] = E axis
//   POL = {{y,z},{x,z},{x,y}}
// so the 12 ports of a node are
//   0 x-Ey  1 x-Ez  2 x+Ey  3 x+Ez   4 y-Ex  5 y-Ez  6 y+Ex  7 y+Ez
//   8 z-Ex  9 z-Ey 10 z+Ex 11 z+Ey
//
// The expressions below are the reference's loops (VjOf TlmSolver.cpp:46-52,
// ZIof :37-42, ScatterCell :59-76, ScatterCellDiel :86-113) written out per
// port with the same association order, so that with -ffp-contract=off the
// device produces the same FP32 bits as the reference CPU path.
#pragma once

#include <hip/hip_runtime.h>

namespace scn {

// E-axis of (axis, pol): POL table of TlmSolver.cpp:23.
__host__ __device__ constexpr int pol_axis(int axis, int pol)
{
    return axis == 0 ? (pol == 0 ? 1 : 2) : axis == 1 ? (pol == 0 ? 0 : 2) : (pol == 0 ? 0 : 1);
}

// portOf(axis, side, polAxis), TlmSolver.cpp:25-29.
__host__ __device__ constexpr int port_of(int axis, int side, int polAxis)
{
    return axis * 4 + side * 2 + (polAxis == pol_axis(axis, 0) ? 0 : 1);
}

// Node voltage sums for E_x, E_y, E_z: the four ports polarised along j,
// summed ((a + b) + c) + d as in VjOf.
__device__ __forceinline__ float sum_x(const float v[12]) { return ((v[4] + v[6]) + v[8]) + v[10]; }
__device__ __forceinline__ float sum_y(const float v[12]) { return ((v[0] + v[2]) + v[9]) + v[11]; }
__device__ __forceinline__ float sum_z(const float v[12]) { return ((v[1] + v[3]) + v[5]) + v[7]; }

// Loop circulations ZI_k, ((a - b) - c) + d as in ZIof.
__device__ __forceinline__ float zi_x(const float v[12]) { return 0.5f * (((v[5] - v[7]) - v[9]) + v[11]); }
__device__ __forceinline__ float zi_y(const float v[12]) { return 0.5f * (((v[8] - v[10]) - v[1]) + v[3]); }
__device__ __forceinline__ float zi_z(const float v[12]) { return 0.5f * (((v[0] - v[2]) - v[4]) + v[6]); }

// VjOf for a given E axis.
__device__ __forceinline__ float vj(const float v[12], int j)
{
    return 0.5f * (j == 0 ? sum_x(v) : j == 1 ? sum_y(v) : sum_z(v));
}
// ZIof for a given circulation axis.
__device__ __forceinline__ float zi(const float v[12], int k)
{
    return k == 0 ? zi_x(v) : k == 1 ? zi_y(v) : zi_z(v);
}

// Reflected pulses from node voltages (Vx, Vy, Vz) and circulations:
// out(i,s,j) = (V_j + sg*ZI_k) - in(i,1-s,j) with sg = +-1 from the cyclic
// order of (i, j) and the side (TlmSolver.cpp:72-74).
__device__ __forceinline__ void reflect(const float v[12], float Vx, float Vy, float Vz, float Zx,
                                        float Zy, float Zz, float o[12])
{
    o[0] = (Vy - Zz) - v[2];
    o[1] = (Vz + Zy) - v[3];
    o[2] = (Vy + Zz) - v[0];
    o[3] = (Vz - Zy) - v[1];
    o[4] = (Vx + Zz) - v[6];
    o[5] = (Vz - Zx) - v[7];
    o[6] = (Vx - Zz) - v[4];
    o[7] = (Vz + Zx) - v[5];
    o[8] = (Vx - Zy) - v[10];
    o[9] = (Vy + Zx) - v[11];
    o[10] = (Vx + Zy) - v[8];
    o[11] = (Vy - Zx) - v[9];
}

// Air node (ScatterCell).
__device__ __forceinline__ void scatter_air(const float v[12], float o[12])
{
    float Vx = 0.5f * sum_x(v), Vy = 0.5f * sum_y(v), Vz = 0.5f * sum_z(v);
    reflect(v, Vx, Vy, Vz, zi_x(v), zi_y(v), zi_z(v), o);
}

// Stub-loaded lossy node (ScatterCellDiel): V_j = 2(sum4 + yh*vs_j)/den,
// den = (4 + yh) + gh precomputed per material, stub vs_j <- V_j - vs_j.
__device__ __forceinline__ void scatter_diel(const float v[12], float vs[3], float yh, float den,
                                             float o[12])
{
    float Vx = (2.0f * (sum_x(v) + yh * vs[0])) / den;
    float Vy = (2.0f * (sum_y(v) + yh * vs[1])) / den;
    float Vz = (2.0f * (sum_z(v) + yh * vs[2])) / den;
    reflect(v, Vx, Vy, Vz, zi_x(v), zi_y(v), zi_z(v), o);
    vs[0] = Vx - vs[0];
    vs[1] = Vy - vs[1];
    vs[2] = Vz - vs[2];
}

} // namespace scn
