//---------------------------------------------------------------------------
// tlm_stop.cpp - cancellation through the TlmSolver facade.
//
// The reference stops a run when requestStop() sets stopFlag from another
// thread: the time loop tests it (TlmSolver.cpp:459, GpuTlm.cpp:714), run()
// still calls finalizeDft() and marks the solver finished
// (TlmSolver.cpp:439-454). This program runs a scene with every monitor on a
// worker thread, requests a stop once 300 steps are done, and checks:
//   * run() returned: isFinished(), !isRunning(), 0 < currentStep() < total;
//   * every record holds exactly currentStep() steps;
//   * the stopped solver equals, bit for bit, a second solver whose run is
//     configured to end at that same step (pulses, stubs, port V/I, probe
//     series and FP64 DFTs, energy history, finalised |Js| DFT, far field),
//     so the monitors of a stopped run are those of the steps it ran.
// Prints PASS/FAIL lines and exits 1 on a failure; SKIP without a device.
//---------------------------------------------------------------------------
#include "GpuTlmHip.h"
#include "TlmSolver.h"

#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

namespace {

int g_fail = 0;

void check(bool ok, const char *what, long long a, long long b)
{
    std::printf("%s  %-58s (%lld / %lld)\n", ok ? "PASS" : "FAIL", what, a, b);
    if (!ok)
        ++g_fail;
}

template <class T> bool same(const std::vector<T> &a, const std::vector<T> &b)
{
    return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0);
}

// 96^3 cells: an eps_r = 4 lossy box, a PEC plate (Js faces), a z port of
// two cells, a soft source, two probes with a 4-frequency DFT, a Huygens box.
std::unique_ptr<TlmSolver> makeSolver(int totalSteps)
{
    VoxelGridSpec g;
    g.nx = g.ny = g.nz = 96;
    g.dl = 0.0149896f;
    g.origin = Vec3(0, 0, 0);
    std::vector<uint8_t> mat((size_t)g.nx * g.ny * g.nz, MAT_AIR);
    for (int k = 20; k < 40; ++k)
        for (int j = 30; j < 70; ++j)
            for (int i = 55; i < 80; ++i)
                mat[g.cellIndex(i, j, k)] = MAT_DIEL0;
    for (int j = 20; j < 76; ++j)
        for (int i = 20; i < 76; ++i)
            mat[g.cellIndex(i, j, 70)] = MAT_PEC;
    std::vector<MatProps> table = {{4.0f, 0.02f}};
    TlmConfig cfg;
    cfg.f0 = 1e9f;
    cfg.waveform = WaveformType::GaussianSine;
    cfg.totalSteps = totalSteps;
    cfg.settleSteps = 200;
    cfg.boundaryRho = 0.0f;
    cfg.huyOffset = 6;
    std::unique_ptr<TlmSolver> s(new TlmSolver());
    s->setup(g, mat, table, cfg);
    s->addPort({g.cellIndex(48, 48, 47), g.cellIndex(48, 48, 48)}, 2, 1.0f);
    s->addSoftSource(g.cellIndex(30, 60, 50), 1, 0.5f);
    s->setProbeDft({5e8, 8e8, 1e9, 1.3e9});
    s->addProbe(g.cellCenter(60, 40, 30));
    s->addProbe(g.cellCenter(40, 50, 60));
    return s;
}

} // namespace

int main()
{
    if (tlm_device_count() <= 0)
    {
        std::printf("SKIP  cancellation: no HIP device\n");
        return 0;
    }
    const int total = 40000;
    std::unique_ptr<TlmSolver> a = makeSolver(total);
    std::thread worker([&] { a->run(); });
    // request the stop once the run is past the DFT settling steps (200)
    const auto t0 = std::chrono::steady_clock::now();
    while (a->currentStep() < 300 && !a->isFinished() &&
           std::chrono::steady_clock::now() - t0 < std::chrono::seconds(60))
        std::this_thread::sleep_for(std::chrono::microseconds(100));
    a->requestStop();
    worker.join();
    const int S = a->currentStep();
    std::printf("stopped at step %d of %d on %s (%s)\n", S, total, tlm_device_name(0),
                a->gpuStatus().c_str());
    check(a->ranOnGpu(), "the run used the device", a->ranOnGpu(), 1);
    check(a->isFinished() && !a->isRunning(), "run() returned: finished, not running",
          a->isFinished(), a->isRunning());
    check(S >= 300 && S < total, "stopped after the requested step, before the last", S, total);
    check((int)a->ports()[0].vRec.size() == S && (int)a->ports()[0].iRec.size() == S,
          "port records hold the steps run", (long long)a->ports()[0].vRec.size(), S);
    check((int)a->probe(0).ez.size() == S, "probe records hold the steps run",
          (long long)a->probe(0).ez.size(), S);

    // the same scene configured to end at step S
    std::unique_ptr<TlmSolver> b = makeSolver(S);
    b->run();
    check(b->currentStep() == S && b->isFinished(), "reference run ends at the same step",
          b->currentStep(), S);
    check(same(a->pulses(), b->pulses()), "final link pulses bit-identical", (long long)a->pulses().size(),
          (long long)b->pulses().size());
    check(same(a->stubPulses(), b->stubPulses()), "final stub pulses bit-identical",
          (long long)a->stubPulses().size(), (long long)b->stubPulses().size());
    check(same(a->ports()[0].vRec, b->ports()[0].vRec) && same(a->ports()[0].iRec, b->ports()[0].iRec),
          "port V/I records bit-identical", S, S);
    bool pr = true;
    for (int p = 0; p < a->probeCount(); ++p)
        pr = pr && same(a->probe(p).ex, b->probe(p).ex) && same(a->probe(p).ey, b->probe(p).ey) &&
             same(a->probe(p).ez, b->probe(p).ez) && same(a->probe(p).dftRe, b->probe(p).dftRe) &&
             same(a->probe(p).dftIm, b->probe(p).dftIm);
    check(pr, "probe series and FP64 DFTs bit-identical", a->probeCount(), b->probeCount());
    std::vector<int> sa, sb;
    std::vector<float> va, vb;
    a->getEnergyHistory(sa, va);
    b->getEnergyHistory(sb, vb);
    check(same(sa, sb) && same(va, vb) && !sa.empty(), "energy history identical", (long long)sa.size(),
          (long long)sb.size());
    std::vector<float> ja, jb;
    const bool ha = a->getJsDft(ja), hb = b->getJsDft(jb);
    check(ha && hb && same(ja, jb) && !ja.empty(), "finalised |Js| DFT bit-identical", (long long)ja.size(),
          (long long)jb.size());
    FarFieldData fa, fb;
    const bool ffa = a->computeFarField(fa), ffb = b->computeFarField(fb);
    check(ffa && ffb && same(fa.U, fb.U), "far field from the Huygens DFT identical", ffa, ffb);

    // run() again on the stopped solver: the stop request stands until the
    // next setup() (TlmSolver.cpp:293), so the second run advances nothing
    // and leaves every result as the stopped run left it (the drop-in enters
    // with curStep > 0 and must neither restart nor add to the sums)
    const std::vector<float> pa = a->pulses(), sta = a->stubPulses();
    const std::vector<float> vra = a->ports()[0].vRec, eza = a->probe(0).ez;
    const std::vector<double> dra = a->probe(0).dftRe;
    a->run();
    check(a->currentStep() == S && a->isFinished() && !a->isRunning(),
          "second run() after the stop: same step, finished", a->currentStep(), S);
    check(same(a->pulses(), pa) && same(a->stubPulses(), sta), "second run(): pulses and stubs unchanged",
          (long long)a->pulses().size(), (long long)pa.size());
    check(same(a->ports()[0].vRec, vra) && same(a->probe(0).ez, eza) && same(a->probe(0).dftRe, dra),
          "second run(): records and probe DFT unchanged", (long long)a->ports()[0].vRec.size(),
          (long long)vra.size());
    std::vector<int> sa2;
    std::vector<float> va2, ja2;
    a->getEnergyHistory(sa2, va2);
    a->getJsDft(ja2);
    check(same(sa2, sa) && same(va2, va), "second run(): energy history unchanged", (long long)sa2.size(),
          (long long)sa.size());
    check(same(ja2, ja), "second run(): finalised |Js| DFT unchanged", (long long)ja2.size(),
          (long long)ja.size());
    std::printf(g_fail ? "STOPTEST FAILED (%d)\n" : "STOPTEST PASSED\n", g_fail);
    return g_fail ? 1 : 0;
}
