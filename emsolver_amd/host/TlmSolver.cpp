//---------------------------------------------------------------------------
// TlmSolver.cpp - facade of the reference's TlmSolver over the HIP step.
// Scene set-up, monitors' host state and post-processing live here; the
// time loop is RunGpuTlm (GpuTlm.cpp).
//---------------------------------------------------------------------------
#include "TlmSolver.h"

#include "GpuTlm.h"
#include "GpuTlmHip.h"

#include <algorithm>
#include <cmath>

namespace {
const float kC0 = 299792458.0f;   // TlmSolver.cpp:15
const float kEta0 = 376.730313f;  // TlmSolver.cpp:16
} // namespace

void BuildPecSurfaceFaces(const VoxelGridSpec &g, const std::vector<uint8_t> &mat, uint8_t matPec,
                          std::vector<SurfaceFace> &out)
{
    out.clear();
    if (matPec != MAT_PEC || mat.size() != (size_t)g.nx * g.ny * g.nz)
        return;
    const int64_t n = tlm_build_pec_faces(g.nx, g.ny, g.nz, mat.data(), nullptr, nullptr, nullptr, 0);
    std::vector<int64_t> cell(n);
    std::vector<int32_t> axis(n), sign(n);
    tlm_build_pec_faces(g.nx, g.ny, g.nz, mat.data(), cell.data(), axis.data(), sign.data(), n);
    out.resize(n);
    const long long dirs[3] = {1, g.nx, (long long)g.nx * g.ny};
    for (int64_t f = 0; f < n; ++f)
    {
        // the face sits half a cell from the PEC cell's centre towards the air cell
        const long long pecCell = cell[f] - sign[f] * dirs[axis[f]];
        const int i = (int)(pecCell % g.nx), j = (int)((pecCell / g.nx) % g.ny),
                  k = (int)(pecCell / ((long long)g.nx * g.ny));
        Vec3 cc = g.cellCenter(i, j, k);
        cc.set(axis[f], cc[axis[f]] + sign[f] * g.dl * 0.5f);
        out[f] = {(int)cell[f], cc, (int8_t)axis[f], (int8_t)sign[f]};
    }
}

TlmSolver::TlmSolver() {}
TlmSolver::~TlmSolver() {}

// TlmSolver::setup semantics (TlmSolver.cpp:259-304)
void TlmSolver::setup(const VoxelGridSpec &grid, std::vector<uint8_t> materials,
                      std::vector<MatProps> table, const TlmConfig &cfg)
{
    g = grid;
    config = cfg;
    mat = std::move(materials);
    matTable = std::move(table);
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    V.assign(n * 12, 0.0f);
    dt = g.dl / (2.0f * kC0);
    yhat.clear();
    ghat.clear();
    for (const auto &mp : matTable)
    {
        yhat.push_back(4.0f * (std::max(1.0f, mp.epsr) - 1.0f));
        ghat.push_back(std::max(0.0f, mp.sigma) * g.dl * kEta0);
    }
    const bool anyDiel = std::any_of(mat.begin(), mat.end(), [](uint8_t m) { return m >= MAT_DIEL0; });
    if (anyDiel)
        Vs.assign(n * 3, 0.0f);
    else
        Vs.clear();
    sources.clear();
    portList.clear();
    probes.clear();
    dftFreqs.clear();
    frames.clear();
    energySteps.clear();
    energyVals.clear();
    curStep = 0;
    finished = false;
    stopFlag = false;
    dftSamples = 0;
    huySamples = 0;
    usedGpu = false;
    gpuMsg.clear();
    buildSurfaceFaces();
    buildHuygens();
    jsInstant.assign(surfFaces.size(), 0.0f);
    jsDftRe.assign(surfFaces.size() * 2, 0.0f);
    jsDftIm.assign(surfFaces.size() * 2, 0.0f);
    jsDftMag.clear();
}

void TlmSolver::buildSurfaceFaces()
{
    BuildPecSurfaceFaces(g, mat, MAT_PEC, surfFaces);
}

// Huygens box (TlmSolver.cpp:313-359): inset huyOffset cells, AIR cells only
void TlmSolver::buildHuygens()
{
    huyFaces.clear();
    huyDft.clear();
    const int64_t n = tlm_build_huygens(g.nx, g.ny, g.nz, mat.data(), config.huyOffset, nullptr,
                                        nullptr, nullptr, 0);
    if (n <= 0)
        return;
    std::vector<int64_t> cell(n);
    std::vector<int32_t> axis(n), sgn(n);
    tlm_build_huygens(g.nx, g.ny, g.nz, mat.data(), config.huyOffset, cell.data(), axis.data(),
                      sgn.data(), n);
    const Vec3 domCenter = g.origin + Vec3(g.nx * g.dl, g.ny * g.dl, g.nz * g.dl) * 0.5f;
    huyFaces.resize(n);
    for (int64_t f = 0; f < n; ++f)
    {
        const int i = (int)(cell[f] % g.nx), j = (int)((cell[f] / g.nx) % g.ny),
                  k = (int)(cell[f] / ((long long)g.nx * g.ny));
        huyFaces[f] = {(int)cell[f], g.cellCenter(i, j, k) - domCenter, (int8_t)axis[f],
                       (int8_t)sgn[f]};
    }
    huyDft.assign(huyFaces.size() * 8, 0.0f);
}

// addPort keeps AIR cells only (TlmSolver.cpp:362-374)
int TlmSolver::addPort(const std::vector<size_t> &cells, int polAxis, float amp)
{
    TlmPort p;
    p.polAxis = polAxis;
    p.amp = amp;
    for (size_t c : cells)
        if (c < mat.size() && mat[c] == MAT_AIR)
            p.cells.push_back(c);
    p.vRec.reserve(config.totalSteps);
    p.iRec.reserve(config.totalSteps);
    portList.push_back(std::move(p));
    return (int)portList.size() - 1;
}

int TlmSolver::addProbe(const Vec3 &worldPos)
{
    int i, j, k;
    g.cellOf(worldPos, i, j, k);
    if (!g.inGrid(i, j, k))
        return -1;
    FieldProbe pr;
    pr.cell = g.cellIndex(i, j, k);
    pr.pos = worldPos;
    pr.dftRe.assign(3 * dftFreqs.size(), 0.0);
    pr.dftIm.assign(3 * dftFreqs.size(), 0.0);
    probes.push_back(std::move(pr));
    return (int)probes.size() - 1;
}

void TlmSolver::setProbeDft(const std::vector<double> &freqsHz)
{
    dftFreqs = freqsHz;
    for (auto &p : probes)
    {
        p.dftRe.assign(3 * dftFreqs.size(), 0.0);
        p.dftIm.assign(3 * dftFreqs.size(), 0.0);
    }
}

// plane-wave soft source on the AIR cells of one plane (TlmSolver.cpp:394-407)
void TlmSolver::addPlaneWave(int propAxis, int planeIndex, int polAxis, float amp)
{
    for (int k = 0; k < g.nz; ++k)
        for (int j = 0; j < g.ny; ++j)
            for (int i = 0; i < g.nx; ++i)
            {
                const int co[3] = {i, j, k};
                if (co[propAxis] != planeIndex)
                    continue;
                const size_t c = g.cellIndex(i, j, k);
                if (mat[c] == MAT_AIR)
                    sources.push_back({c, polAxis, amp});
            }
}

void TlmSolver::addSoftSource(size_t cell, int polAxis, float amp)
{
    if (cell < mat.size() && polAxis >= 0 && polAxis < 3)
        sources.push_back({cell, polAxis, amp});
}

float TlmSolver::waveformValue(int n) const
{
    float sv = 0.0f;
    tlm_schedule_tables((int)config.waveform, config.f0, dt, n, 1, config.gdWidthSteps,
                        config.gdDelaySteps, &sv, nullptr, nullptr);
    return sv;
}

void TlmSolver::run()
{
    running = true;
    usedGpu = RunGpuTlm(*this, gpuMsg);
    finalizeDft();
    finished = true;
    running = false;
}

// jsDftMag = (2/N) sqrt(|J1|^2 + |J2|^2) (TlmSolver.cpp:467-482)
void TlmSolver::finalizeDft()
{
    std::lock_guard<std::mutex> lk(vizMutex);
    jsDftMag.assign(surfFaces.size(), 0.0f);
    if (dftSamples > 0)
    {
        const float norm = 2.0f / dftSamples;
        for (size_t f = 0; f < surfFaces.size(); ++f)
        {
            const float r1 = jsDftRe[f * 2], i1 = jsDftIm[f * 2];
            const float r2 = jsDftRe[f * 2 + 1], i2 = jsDftIm[f * 2 + 1];
            jsDftMag[f] = norm * std::sqrt(r1 * r1 + i1 * i1 + r2 * r2 + i2 * i2);
        }
    }
}

void TlmSolver::getJsInstant(std::vector<float> &out)
{
    std::lock_guard<std::mutex> lk(vizMutex);
    out = jsInstant;
}

bool TlmSolver::getJsDft(std::vector<float> &out)
{
    std::lock_guard<std::mutex> lk(vizMutex);
    if (jsDftMag.empty())
        return false;
    out = jsDftMag;
    return true;
}

void TlmSolver::setFieldPlane(int axis, int index)
{
    std::lock_guard<std::mutex> lk(vizMutex);
    planeIndex = index;
    planeAxis = axis;
}

bool TlmSolver::getFieldPlane(std::vector<float> &out, int &axis, int &index, int &n1, int &n2)
{
    std::lock_guard<std::mutex> lk(vizMutex);
    if (planeAxis < 0 || planeBuf.empty())
        return false;
    out = planeBuf;
    axis = planeAxis;
    index = planeIndex;
    n1 = planeN1;
    n2 = planeN2;
    return true;
}

int TlmSolver::frameCount()
{
    std::lock_guard<std::mutex> lk(vizMutex);
    return (int)frames.size();
}

bool TlmSolver::getFrame(int idx, VizFrame &out)
{
    std::lock_guard<std::mutex> lk(vizMutex);
    if (idx < 0 || idx >= (int)frames.size())
        return false;
    out = frames[idx];
    return true;
}

void TlmSolver::getEnergyHistory(std::vector<int> &steps, std::vector<float> &vals)
{
    std::lock_guard<std::mutex> lk(vizMutex);
    steps = energySteps;
    vals = energyVals;
}

bool TlmSolver::computeFarField(FarFieldData &out, int nTheta, int nPhi)
{
    return ComputeFarFieldFromHuygens(huyFaces, huyDft, huySamples, config.f0, nullptr, out, nTheta,
                                      nPhi);
}

float TlmSolver::ScatterUnitarityError()
{
    return tlm_scatter_unitarity_error(1.0f);
}

float TlmSolver::DielScatterUnitarityError(float epsr)
{
    return tlm_scatter_unitarity_error(epsr);
}

//---------------------------------------------------------------------------
// Near-to-far-field transform at f0 (the reference's Huygens NTFF,
// TlmSolver.cpp:805-916), on the device: tlm_far_field sums every face's
// equivalent currents J = n x H, M = -n x E times exp(+j k0 r.r') for each
// direction and returns U ~ |L_phi + N_theta|^2 + |L_theta - N_phi|^2, the
// directivity 4*pi*Umax / integral(U sin(theta)) (trapezoid rule on the
// (theta, phi) grid) and the peak direction.
//---------------------------------------------------------------------------
bool ComputeFarFieldFromHuygens(const std::vector<HuyFace> &huyFaces,
                                const std::vector<float> &huyDft, int huySamples, float f0,
                                ThreadPool * /*pool*/, FarFieldData &out, int nTheta, int nPhi)
{
    if (huyFaces.empty() || huySamples == 0 || nTheta < 2 || nPhi < 2)
        return false;
    if (huyDft.size() < huyFaces.size() * 8)
        return false;
    if (tlm_device_count() <= 0)
        return false;
    const size_t nh = huyFaces.size();
    std::vector<float> pos(3 * nh);
    std::vector<int32_t> ax(nh), sg(nh);
    for (size_t f = 0; f < nh; ++f)
    {
        pos[3 * f] = huyFaces[f].pos.x;
        pos[3 * f + 1] = huyFaces[f].pos.y;
        pos[3 * f + 2] = huyFaces[f].pos.z;
        ax[f] = huyFaces[f].axis;
        sg[f] = huyFaces[f].outSign;
    }
    std::vector<float> U((size_t)nTheta * nPhi);
    tlm_far_field_summary sm;
    if (tlm_far_field(0, (int)nh, pos.data(), ax.data(), sg.data(), huyDft.data(), huySamples, f0,
                      nTheta, nPhi, U.data(), &sm) != TLM_OK)
        return false;
    out.nTheta = nTheta;
    out.nPhi = nPhi;
    out.U.swap(U);
    out.uMax = sm.u_max;
    out.directivity = sm.directivity;
    out.peakThetaDeg = sm.peak_theta_deg;
    out.peakPhiDeg = sm.peak_phi_deg;
    return true;
}
