//---------------------------------------------------------------------------
// TlmSolver.h - SCN TLM solver facade over the gfx950 step.
//
// The public interface is the reference's IFieldSolver / TlmSolver
// (TlmSolver.h:148-264): same methods, argument meaning and results, so a
// caller written for the reference (its UI, its SelfTest) drives this one
// unchanged. The time loop runs through RunGpuTlm (GpuTlm.h), which the
// reference also calls from TlmSolver::run (TlmSolver.cpp:445-446).
//
// Differences, all deliberate:
//  * there is no CPU time loop: run() has one engine, the HIP step. If the
//    device path fails, run() records the reason in gpuStatus(), leaves
//    curStep at the last completed step and ranOnGpu() false. (In the
//    reference's own TlmSolver the CPU loop stays as the fallback.)
//  * probes run on the GPU (the reference forces the CPU path when probes
//    exist, TlmSolver.cpp:443-448) and carry running FP64 DFTs;
//  * addSoftSource() adds one soft source cell (the benchmark scenes);
//  * the final pulses V / Vs are written back after a run.
//---------------------------------------------------------------------------
#ifndef EMSOLVER_AMD_TLMSOLVER_H
#define EMSOLVER_AMD_TLMSOLVER_H

#include "TlmTypes.h"

#include <atomic>
#include <mutex>
#include <string>
#include <vector>

class IFieldSolver
{
public:
    virtual ~IFieldSolver() {}
    virtual void setup(const VoxelGridSpec &grid, std::vector<uint8_t> materials,
                       std::vector<MatProps> matTable, const TlmConfig &cfg) = 0;
    virtual int addPort(const std::vector<size_t> &cells, int polAxis, float amp) = 0;
    virtual void addPlaneWave(int propAxis, int planeIndex, int polAxis, float amp) = 0;
    virtual void run() = 0;
    virtual void requestStop() = 0;
    virtual bool isRunning() const = 0;
    virtual bool isFinished() const = 0;
    virtual int currentStep() const = 0;
    virtual int totalSteps() const = 0;
    virtual float currentEnergy() const = 0;
    virtual bool ranOnGpu() const = 0;
    virtual const std::string &gpuStatus() const = 0;
    virtual const wchar_t *solverName() const = 0;
    virtual const VoxelGridSpec &grid() const = 0;
    virtual const TlmConfig &cfg() const = 0;
    virtual const std::vector<SurfaceFace> &faces() const = 0;
    virtual const std::vector<TlmPort> &ports() const = 0;
    virtual float timestep() const = 0;
    virtual void getJsInstant(std::vector<float> &out) = 0;
    virtual bool getJsDft(std::vector<float> &out) = 0;
    virtual void setFieldPlane(int axis, int index) = 0;
    virtual bool getFieldPlane(std::vector<float> &out, int &axis, int &index, int &n1,
                               int &n2) = 0;
    virtual bool computeFarField(FarFieldData &out, int nTheta = 37, int nPhi = 73) = 0;
    virtual void setRecordEvery(int n) = 0;
    virtual int frameCount() = 0;
    virtual bool getFrame(int idx, VizFrame &out) = 0;
    virtual void getEnergyHistory(std::vector<int> &steps, std::vector<float> &vals) = 0;
    virtual int addProbe(const Vec3 &worldPos) = 0;
    virtual int probeCount() const = 0;
    virtual const FieldProbe &probe(int idx) const = 0;
};

// The reference's worker pool (TlmSolver.h:34-54). The device transform
// below takes the reference's signature; it needs no host pool.
class ThreadPool;

// Near-to-far-field transform of the Huygens phasors, with the reference's
// signature (TlmSolver.h:196-199). It runs on the HIP device; `pool` is
// accepted for source compatibility and not used. Without a device it
// returns false, as RunGpuTlm does.
bool ComputeFarFieldFromHuygens(const std::vector<HuyFace> &huyFaces,
                                const std::vector<float> &huyDft, int huySamples, float f0,
                                ThreadPool *pool, FarFieldData &out, int nTheta, int nPhi);

class TlmSolver : public IFieldSolver
{
public:
    TlmSolver();
    ~TlmSolver();

    void setup(const VoxelGridSpec &grid, std::vector<uint8_t> materials,
               std::vector<MatProps> matTable, const TlmConfig &cfg);
    int addPort(const std::vector<size_t> &cells, int polAxis, float amp);
    void addPlaneWave(int propAxis, int planeIndex, int polAxis, float amp);
    void addSoftSource(size_t cell, int polAxis, float amp);       // addition
    void setProbeDft(const std::vector<double> &freqsHz);          // addition

    void run();
    void requestStop() { stopFlag = true; }
    bool isRunning() const { return running; }
    bool isFinished() const { return finished; }
    int currentStep() const { return curStep; }
    int totalSteps() const { return config.totalSteps; }
    float currentEnergy() const { return energy; }
    bool ranOnGpu() const { return usedGpu; }
    const std::string &gpuStatus() const { return gpuMsg; }
    const wchar_t *solverName() const { return L"TLM"; }

    const VoxelGridSpec &grid() const { return g; }
    const TlmConfig &cfg() const { return config; }
    const std::vector<uint8_t> &materials() const { return mat; }
    const std::vector<SurfaceFace> &faces() const { return surfFaces; }
    const std::vector<TlmPort> &ports() const { return portList; }
    float timestep() const { return dt; }

    void getJsInstant(std::vector<float> &out);
    bool getJsDft(std::vector<float> &out);
    void setFieldPlane(int axis, int index);
    bool getFieldPlane(std::vector<float> &out, int &axis, int &index, int &n1, int &n2);
    bool computeFarField(FarFieldData &out, int nTheta = 37, int nPhi = 73);
    void setRecordEvery(int n) { config.recordEvery = n; }
    int frameCount();
    bool getFrame(int idx, VizFrame &out);
    void getEnergyHistory(std::vector<int> &steps, std::vector<float> &vals);
    int addProbe(const Vec3 &worldPos);
    int probeCount() const { return (int)probes.size(); }
    const FieldProbe &probe(int idx) const { return probes[idx]; }

    // Final link / stub pulses (12 and 3 floats per cell) after a run.
    const std::vector<float> &pulses() const { return V; }
    const std::vector<float> &stubPulses() const { return Vs; }

    // Self tests, evaluated with the device scatter.
    static float ScatterUnitarityError();
    static float DielScatterUnitarityError(float epsr);

private:
    friend class GpuTlm;

    void finalizeDft();
    float waveformValue(int n) const;
    void buildSurfaceFaces();
    void buildHuygens();

    VoxelGridSpec g;
    TlmConfig config;
    std::vector<uint8_t> mat;
    std::vector<MatProps> matTable;
    std::vector<float> yhat, ghat;
    std::vector<float> V;
    std::vector<float> Vs;
    std::vector<SourceCell> sources;
    std::vector<TlmPort> portList;
    std::vector<FieldProbe> probes;
    std::vector<double> dftFreqs;
    std::vector<SurfaceFace> surfFaces;

    std::vector<float> jsInstant;
    std::vector<float> jsDftRe, jsDftIm;
    int dftSamples = 0;
    std::vector<float> jsDftMag;

    std::vector<HuyFace> huyFaces;
    std::vector<float> huyDft;
    int huySamples = 0;

    std::atomic<int> planeAxis{-1};
    int planeIndex = 0;
    std::vector<float> planeBuf;
    int planeN1 = 0, planeN2 = 0;

    std::vector<VizFrame> frames;
    std::vector<int> energySteps;
    std::vector<float> energyVals;

    std::mutex vizMutex;
    std::atomic<bool> stopFlag{false};
    std::atomic<bool> running{false};
    std::atomic<bool> finished{false};
    std::atomic<int> curStep{0};
    std::atomic<float> energy{0.0f};
    bool usedGpu = false;
    std::string gpuMsg;
    float dt = 0.0f;
};

#endif
