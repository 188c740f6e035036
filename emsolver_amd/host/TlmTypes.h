//---------------------------------------------------------------------------
// TlmTypes.h - data types of the TLM solver interface.
//
// Same names, fields and meaning as the reference's TlmSolver.h:59-141 and
// Geometry.h:85-113 / 141-153 / Vec3.h, so that code written against the
// reference's TlmSolver compiles against this one. Additions are marked.
//---------------------------------------------------------------------------
#ifndef EMSOLVER_AMD_TLMTYPES_H
#define EMSOLVER_AMD_TLMTYPES_H

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <vector>

struct Vec3
{
    float x = 0.0f, y = 0.0f, z = 0.0f;
    Vec3() = default;
    Vec3(float ax, float ay, float az) : x(ax), y(ay), z(az) {}
    Vec3 operator+(const Vec3 &v) const { return Vec3(x + v.x, y + v.y, z + v.z); }
    Vec3 operator-(const Vec3 &v) const { return Vec3(x - v.x, y - v.y, z - v.z); }
    Vec3 operator*(float s) const { return Vec3(x * s, y * s, z * s); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    void set(int i, float v) { (i == 0 ? x : (i == 1 ? y : z)) = v; }
    float length() const { return std::sqrt(x * x + y * y + z * z); }
};

struct Aabb
{
    Vec3 lo{1e30f, 1e30f, 1e30f};
    Vec3 hi{-1e30f, -1e30f, -1e30f};
    void grow(const Vec3 &p)
    {
        lo = Vec3(std::fmin(lo.x, p.x), std::fmin(lo.y, p.y), std::fmin(lo.z, p.z));
        hi = Vec3(std::fmax(hi.x, p.x), std::fmax(hi.y, p.y), std::fmax(hi.z, p.z));
    }
    bool valid() const { return lo.x <= hi.x; }
    Vec3 size() const { return hi - lo; }
    Vec3 center() const { return (lo + hi) * 0.5f; }
};

// Uniform voxel grid (Geometry.h:85-113): cell (i,j,k) spans
// origin + (i..i+1, j..j+1, k..k+1) * dl.
struct VoxelGridSpec
{
    int nx = 0, ny = 0, nz = 0;
    Vec3 origin;
    float dl = 1e-3f;
    size_t cellIndex(int i, int j, int k) const { return ((size_t)k * ny + j) * nx + i; }
    bool inGrid(int i, int j, int k) const
    {
        return i >= 0 && j >= 0 && k >= 0 && i < nx && j < ny && k < nz;
    }
    Vec3 cellCenter(int i, int j, int k) const
    {
        return Vec3(origin.x + (i + 0.5f) * dl, origin.y + (j + 0.5f) * dl,
                    origin.z + (k + 0.5f) * dl);
    }
    void cellOf(const Vec3 &p, int &i, int &j, int &k) const
    {
        i = (int)std::floor((p.x - origin.x) / dl);
        j = (int)std::floor((p.y - origin.y) / dl);
        k = (int)std::floor((p.z - origin.z) / dl);
    }
};

// PEC surface face (Geometry.h:141-149)
struct SurfaceFace
{
    int airCell;
    Vec3 center;
    int8_t axis;
    int8_t sign;
};

// BuildPecSurfaceFaces (Geometry.h:151-153): faces between a PEC cell and a
// non-PEC neighbour, in (k, j, i, axis, side) order.
void BuildPecSurfaceFaces(const VoxelGridSpec &g, const std::vector<uint8_t> &mat,
                          uint8_t matPec, std::vector<SurfaceFace> &out);

enum : uint8_t { MAT_AIR = 0, MAT_PEC = 1, MAT_DIEL0 = 2 };

struct MatProps
{
    float epsr = 1.0f;
    float sigma = 0.0f;
};

// TlmSolver.h:70 plus GaussianDerivative (addition): sv(n) = -sqrt(2e) u
// exp(-u^2), u = (n - delay)/width in steps (TlmConfig::gdDelaySteps/Width).
enum class WaveformType { CwRamped, GaussianSine, GaussianPulse, GaussianDerivative };

struct SourceCell
{
    size_t cell;
    int polAxis;
    float amp;
};

struct TlmPort
{
    std::vector<size_t> cells;
    int polAxis = 2;
    float amp = 1.0f;
    std::vector<float> vRec, iRec;
};

struct HuyFace
{
    int cell;
    Vec3 pos;
    int8_t axis;
    int8_t outSign;
};

struct FarFieldData
{
    int nTheta = 0, nPhi = 0;
    std::vector<float> U;
    float uMax = 0.0f;
    float directivity = 0.0f;
    float peakThetaDeg = 0, peakPhiDeg = 0;
};

struct TlmConfig
{
    float f0 = 1e9f;
    WaveformType waveform = WaveformType::CwRamped;
    int totalSteps = 2000;
    int settleSteps = 600;
    float boundaryRho = 0.0f;
    int huyOffset = 0;
    bool useGpu = true; // the only engine here is the GPU
    int recordEvery = 0;
    float gdWidthSteps = 20.0f; // addition: Gaussian-derivative width
    float gdDelaySteps = 60.0f; // addition: Gaussian-derivative delay
};

struct VizFrame
{
    int step = 0;
    std::vector<float> js;
    std::vector<float> plane;
    int planeAxis = -1, planeIdx = 0, planeN1 = 0, planeN2 = 0;
};

// E-field probe (TlmSolver.h:136-141) with running FP64 DFT accumulators
// (addition): dftRe/dftIm[j * nfreq + q] for component j at frequency q.
struct FieldProbe
{
    size_t cell = 0;
    Vec3 pos;
    std::vector<float> ex, ey, ez;
    std::vector<double> dftRe, dftIm;
};

#endif
