//---------------------------------------------------------------------------
// ntff_oracle.cpp - CPU restatement of the reference's near-to-far-field
// transform from the Huygens box DFTs.
//
// TEST INFRASTRUCTURE ONLY: the checker of the device far field. Only tests/
// and bench.py's CPU baseline leg load it (through liboracle_tlm.so).
//
// Restates ComputeFarFieldFromHuygens (TlmSolver.cpp:805-916): per-face
// equivalent currents J' = n x (eta0 H), M = -(n x E) on the tangential axes
// POL[axis] (TlmSolver.cpp:23), the radiation sums with exp(+j k0 r.r'),
// U = |L_phi + N'_theta|^2 + |L_theta - N'_phi|^2 on the (theta, phi) grid,
// and the trapezoid-rule directivity. Face positions and outward signs as
// TlmSolver::buildHuygens sets them (TlmSolver.cpp:323-334): cellCenter minus
// the domain centre. The sums run in face order, accumulated in float (the
// reference) or double (ACC = double, the precision reference); the
// trigonometry is the C library's float sin/cos in both.
// Pinned in tests/test_oracle_ntff.py by closed forms (a current element's
// sin^2 pattern and directivity 1.5, a two-element array factor, a Huygens
// element's (1 + cos)^2 cardioid per face orientation).
//---------------------------------------------------------------------------
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

const float kPi = (float)M_PI;
const float kC0 = 299792458.0f;
const int POL3[3][2] = {{1, 2}, {0, 2}, {0, 1}};

struct FaceCur
{
    float pos[3];
    float jr[3], ji[3], mr[3], mi[3];
};

template <class ACC>
void far_field(int nh, const float *pos, const int32_t *axis, const int32_t *sign,
               const float *dft, float f0, int nTheta, int nPhi, float *U, float *summary)
{
    const float k0 = 2.0f * kPi * f0 / kC0;
    std::vector<FaceCur> fc(nh);
    for (int f = 0; f < nh; ++f)
    {
        const float *a = dft + (size_t)f * 8;
        const int ax = axis[f];
        const int t1 = POL3[ax][0], t2 = POL3[ax][1];
        const float sa = (ax == 1) ? -1.0f : 1.0f;
        const float s = (float)sign[f];
        FaceCur &q = fc[f];
        for (int c = 0; c < 3; ++c)
        {
            q.pos[c] = pos[3 * f + c];
            q.jr[c] = q.ji[c] = q.mr[c] = q.mi[c] = 0.0f;
        }
        q.jr[t2] = s * sa * a[4];
        q.ji[t2] = s * sa * a[5];
        q.jr[t1] = -s * sa * a[6];
        q.ji[t1] = -s * sa * a[7];
        q.mr[t2] = -s * sa * a[0];
        q.mi[t2] = -s * sa * a[1];
        q.mr[t1] = s * sa * a[2];
        q.mi[t1] = s * sa * a[3];
    }
    for (int ti = 0; ti < nTheta; ++ti)
    {
        const float th = kPi * ti / (nTheta - 1);
        const float st = std::sin(th), ct = std::cos(th);
        for (int pi = 0; pi < nPhi; ++pi)
        {
            const float ph = 2.0f * kPi * pi / (nPhi - 1);
            const float cp = std::cos(ph), sp = std::sin(ph);
            const float rh[3] = {st * cp, st * sp, ct};
            const float thh[3] = {ct * cp, ct * sp, -st};
            const float phh[3] = {-sp, cp, 0.0f};
            ACC NtR = 0, NtI = 0, NpR = 0, NpI = 0, LtR = 0, LtI = 0, LpR = 0, LpI = 0;
            for (int f = 0; f < nh; ++f)
            {
                const FaceCur &q = fc[f];
                const float pha = k0 * (rh[0] * q.pos[0] + rh[1] * q.pos[1] + rh[2] * q.pos[2]);
                const float c = std::cos(pha), s = std::sin(pha);
                const float jtR = q.jr[0] * thh[0] + q.jr[1] * thh[1] + q.jr[2] * thh[2];
                const float jtI = q.ji[0] * thh[0] + q.ji[1] * thh[1] + q.ji[2] * thh[2];
                const float jpR = q.jr[0] * phh[0] + q.jr[1] * phh[1];
                const float jpI = q.ji[0] * phh[0] + q.ji[1] * phh[1];
                const float mtR = q.mr[0] * thh[0] + q.mr[1] * thh[1] + q.mr[2] * thh[2];
                const float mtI = q.mi[0] * thh[0] + q.mi[1] * thh[1] + q.mi[2] * thh[2];
                const float mpR = q.mr[0] * phh[0] + q.mr[1] * phh[1];
                const float mpI = q.mi[0] * phh[0] + q.mi[1] * phh[1];
                NtR += (ACC)(jtR * c - jtI * s);
                NtI += (ACC)(jtR * s + jtI * c);
                NpR += (ACC)(jpR * c - jpI * s);
                NpI += (ACC)(jpR * s + jpI * c);
                LtR += (ACC)(mtR * c - mtI * s);
                LtI += (ACC)(mtR * s + mtI * c);
                LpR += (ACC)(mpR * c - mpI * s);
                LpI += (ACC)(mpR * s + mpI * c);
            }
            const ACC aR = LpR + NtR, aI = LpI + NtI;
            const ACC bR = LtR - NpR, bI = LtI - NpI;
            U[(size_t)ti * nPhi + pi] = (float)(aR * aR + aI * aI + bR * bR + bI * bI);
        }
    }
    // directivity (TlmSolver.cpp:893-913)
    float uMax = 0.0f;
    int iMax = 0;
    double prad = 0.0;
    const float dTh = kPi / (nTheta - 1);
    const float dPh = 2.0f * kPi / (nPhi - 1);
    for (int ti = 0; ti < nTheta; ++ti)
    {
        const float th = kPi * ti / (nTheta - 1);
        const float wTh = (ti == 0 || ti == nTheta - 1) ? 0.5f : 1.0f;
        for (int pi = 0; pi < nPhi; ++pi)
        {
            const float wPh = (pi == 0 || pi == nPhi - 1) ? 0.5f : 1.0f;
            const float u = U[(size_t)ti * nPhi + pi];
            if (u > uMax)
            {
                uMax = u;
                iMax = ti * nPhi + pi;
            }
            prad += (double)u * std::sin(th) * wTh * wPh * dTh * dPh;
        }
    }
    summary[0] = uMax;
    summary[1] = (prad > 0) ? (float)(4.0 * M_PI * uMax / prad) : 0.0f;
    summary[2] = 180.0f * (iMax / nPhi) / (nTheta - 1);
    summary[3] = 360.0f * (iMax % nPhi) / (nPhi - 1);
}

} // namespace

extern "C" {

// U: nTheta*nPhi (theta-major); summary: uMax, directivity, peak theta and
// phi in degrees. Returns 0, or -1 when there is no data (the reference's
// false).
int ot_far_field(int nh, const float *pos, const int32_t *axis, const int32_t *sign,
                 const float *dft, int samples, float f0, int nTheta, int nPhi, int fp64,
                 float *U, float *summary)
{
    if (nh <= 0 || samples == 0 || nTheta < 2 || nPhi < 2)
        return -1;
    if (fp64)
        far_field<double>(nh, pos, axis, sign, dft, f0, nTheta, nPhi, U, summary);
    else
        far_field<float>(nh, pos, axis, sign, dft, f0, nTheta, nPhi, U, summary);
    return 0;
}

// HuyFace::pos and outSign of Huygens cells (TlmSolver.cpp:323-334):
// cellCenter(i,j,k) - (origin + (nx dl, ny dl, nz dl) * 0.5), and the side
// of the box the face is on along its axis (low half of the grid: -1).
void ot_huygens_geometry(int nx, int ny, int nz, const float *origin, float dl, int n,
                         const int64_t *cells, const int32_t *axis, float *pos, int32_t *sign)
{
    const float dc[3] = {origin[0] + (nx * dl) * 0.5f, origin[1] + (ny * dl) * 0.5f,
                         origin[2] + (nz * dl) * 0.5f};
    const int nn[3] = {nx, ny, nz};
    for (int f = 0; f < n; ++f)
    {
        const int64_t c = cells[f];
        const int co[3] = {(int)(c % nx), (int)((c / nx) % ny), (int)(c / ((int64_t)nx * ny))};
        for (int q = 0; q < 3; ++q)
            pos[3 * f + q] = (origin[q] + (co[q] + 0.5f) * dl) - dc[q];
        const int a = axis[f];
        sign[f] = (2 * co[a] + 1 < nn[a]) ? -1 : 1;
    }
}

} // extern "C"
