//---------------------------------------------------------------------------
// GpuTlm.cpp - RunGpuTlm on the HIP C ABI (include/GpuTlmHip.h).
//
// Replaces the OpenGL implementation of the reference (GpuTlm.cpp:526-884):
// marshals the solver's state and monitors into one tlm_ctx, runs the step
// loop in blocks (checking stopFlag between blocks and publishing energy,
// |Js| and the cut plane for the UI), then reads every monitor back into the
// solver and writes the final pulses back into V / Vs.
//
// It reaches TlmSolver's private state as `friend class GpuTlm`, as the
// reference's GPU path does (TlmSolver.h:267).
//---------------------------------------------------------------------------
#include "GpuTlm.h"

#include "GpuTlmHip.h"
#include "TlmSolver.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <utility>
#include <vector>

class GpuTlm
{
public:
    static bool run(TlmSolver &s, std::string &msg);

private:
    // The running FP64 probe DFT is an addition of this build's facade
    // (TlmSolver::dftFreqs, FieldProbe::dftRe / dftIm). Against the
    // reference's own TlmSolver.h, which has neither, these overloads select
    // the versions without it, so this file compiles against both headers.
    // (Members of a friend class may name TlmSolver's private members.)
    template <class S>
    static auto dft_freqs(const S &s, int) -> decltype(s.dftFreqs, std::vector<double>())
    {
        return std::vector<double>(s.dftFreqs.begin(), s.dftFreqs.end());
    }
    template <class S> static std::vector<double> dft_freqs(const S &, long)
    {
        return {};
    }
    typedef std::pair<std::vector<double> *, std::vector<double> *> DftPtrs;
    template <class P> static auto probe_dft(P &p, int) -> decltype(p.dftRe, p.dftIm, DftPtrs())
    {
        return {&p.dftRe, &p.dftIm};
    }
    template <class P> static DftPtrs probe_dft(P &, long)
    {
        return {nullptr, nullptr};
    }
};

bool RunGpuTlm(TlmSolver &s, std::string &msg)
{
    return GpuTlm::run(s, msg);
}

namespace {

struct CtxGuard
{
    tlm_ctx *c = nullptr;
    ~CtxGuard()
    {
        if (c)
            tlm_destroy(c);
    }
};

bool failed(std::string &msg, const char *what, tlm_ctx *c)
{
    msg = std::string(what) + ": " + tlm_last_error(c);
    return false;
}

// Steps are queued in blocks; between blocks the host publishes energy and
// |Js| for the UI and checks stopFlag. A block is sized to take about
// kBlockMs of device time (measured on the blocks already run), at most
// kBlockSteps steps, so a stop request is honoured within about kBlockMs on
// any grid (the reference's loops test stopFlag every step:
// TlmSolver.cpp:459, GpuTlm.cpp:714).
const int kBlockSteps = 64;
const double kBlockMs = 20.0;

} // namespace

bool GpuTlm::run(TlmSolver &s, std::string &msg)
{
    const size_t n = (size_t)s.g.nx * s.g.ny * s.g.nz;
    if (n == 0)
    {
        msg = "empty grid";
        return false;
    }
    if (tlm_device_count() <= 0)
    {
        msg = "no HIP device available";
        return false;
    }
    tlm_grid_desc d = {s.g.nx, s.g.ny, s.g.nz, s.g.dl, s.config.boundaryRho, TLM_PULSE_F32, 0, 0};
    CtxGuard guard;
    if (tlm_create(&d, &guard.c) != TLM_OK)
        return failed(msg, "tlm_create", nullptr);
    tlm_ctx *c = guard.c;

    if (tlm_upload(c, s.mat.data(), s.yhat.data(), s.ghat.data(), (int)s.yhat.size(), s.V.data(),
                   s.Vs.empty() ? nullptr : s.Vs.data()) != TLM_OK)
        return failed(msg, "tlm_upload", c);

    // injection order: plain sources, then every port cell (TlmSolver.cpp:501-505)
    std::vector<int64_t> cells;
    std::vector<int32_t> pols;
    std::vector<float> amps;
    for (const auto &sc : s.sources)
    {
        cells.push_back((int64_t)sc.cell);
        pols.push_back(sc.polAxis);
        amps.push_back(sc.amp);
    }
    for (const auto &p : s.portList)
        for (size_t cc : p.cells)
        {
            cells.push_back((int64_t)cc);
            pols.push_back(p.polAxis);
            amps.push_back(p.amp);
        }
    if (tlm_set_sources(c, cells.data(), pols.data(), amps.data(), (int)cells.size()) != TLM_OK)
        return failed(msg, "tlm_set_sources", c);

    const int total = s.config.totalSteps;
    const int start = s.curStep;
    if (total <= start)
    {
        msg = tlm_device_name(0);
        return true;
    }
    // per-step host scalars, computed the way the reference's loop does
    // (GpuTlm.cpp:716-719, TlmSolver.cpp:619-620)
    std::vector<float> sv(total), cph(total), sph(total);
    for (int k = 0; k < total; ++k)
    {
        sv[k] = s.waveformValue(k);
        const float phase = 2.0f * (float)M_PI * s.config.f0 * (k * s.dt);
        cph[k] = std::cos(phase);
        sph[k] = std::sin(phase);
    }
    if (tlm_set_schedule(c, sv.data(), cph.data(), sph.data(), total, s.config.settleSteps) != TLM_OK)
        return failed(msg, "tlm_set_schedule", c);
    if (tlm_set_energy_every(c, 16) != TLM_OK)
        return failed(msg, "tlm_set_energy_every", c);

    const int nf = (int)s.surfFaces.size();
    if (nf > 0)
    {
        std::vector<int64_t> fc(nf);
        std::vector<int32_t> fa(nf);
        for (int f = 0; f < nf; ++f)
        {
            fc[f] = s.surfFaces[f].airCell;
            fa[f] = s.surfFaces[f].axis;
        }
        if (tlm_set_faces(c, fc.data(), fa.data(), nf) != TLM_OK)
            return failed(msg, "tlm_set_faces", c);
    }
    const int nh = (int)s.huyFaces.size();
    if (nh > 0)
    {
        std::vector<int64_t> hc(nh);
        std::vector<int32_t> ha(nh);
        for (int f = 0; f < nh; ++f)
        {
            hc[f] = s.huyFaces[f].cell;
            ha[f] = s.huyFaces[f].axis;
        }
        if (tlm_set_huygens(c, hc.data(), ha.data(), nh) != TLM_OK)
            return failed(msg, "tlm_set_huygens", c);
    }
    const int np = (int)s.portList.size();
    if (np > 0)
    {
        std::vector<int64_t> pc;
        std::vector<int32_t> off(1, 0), pp;
        for (const auto &p : s.portList)
        {
            for (size_t cc : p.cells)
                pc.push_back((int64_t)cc);
            off.push_back((int32_t)pc.size());
            pp.push_back(p.polAxis);
        }
        if (tlm_set_ports(c, pc.data(), off.data(), pp.data(), np) != TLM_OK)
            return failed(msg, "tlm_set_ports", c);
    }
    const int npr = (int)s.probes.size();
    const std::vector<double> freqs = dft_freqs(s, 0);
    const int nfreq = npr > 0 ? (int)freqs.size() : 0;
    std::vector<double> ct, st; // FP64 DFT phasors of steps [0, total)
    if (npr > 0)
    {
        std::vector<int64_t> pc(npr);
        for (int p = 0; p < npr; ++p)
            pc[p] = (int64_t)s.probes[p].cell;
        if (tlm_set_probes(c, pc.data(), npr, -1.0f / s.g.dl) != TLM_OK)
            return failed(msg, "tlm_set_probes", c);
        if (nfreq > 0)
        {
            ct.resize((size_t)total * nfreq);
            st.resize((size_t)total * nfreq);
            tlm_dft_tables(freqs.data(), nfreq, (double)s.dt, total, ct.data(), st.data());
            // the device sums a run from step 0; a run resumed at a later step
            // continues the solver's sums on the host (same FP64 terms in the
            // same step order)
            if (start == 0 && tlm_set_probe_dft(c, ct.data(), st.data(), nfreq) != TLM_OK)
                return failed(msg, "tlm_set_probe_dft", c);
        }
    }
    const int pa = s.planeAxis.load();
    if (pa >= 0 && tlm_set_plane(c, pa, s.planeIndex) != TLM_OK)
        return failed(msg, "tlm_set_plane", c);
    const int every = s.config.recordEvery;
    if (every > 0 && tlm_set_frames(c, every, total / every + 1) != TLM_OK)
        return failed(msg, "tlm_set_frames", c);

    // ---- time loop, in blocks ----
    int done = start;
    std::vector<int32_t> esteps;
    std::vector<double> evals;
    int block = 1; // the first block times one step
    while (done < total && !s.stopFlag)
    {
        const int nb = std::min(block, total - done);
        const auto t0 = std::chrono::steady_clock::now();
        if (tlm_run(c, done, nb) != TLM_OK || tlm_sync(c) != TLM_OK)
            return failed(msg, "tlm_run", c);
        const double ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        block = std::max(1, std::min(kBlockSteps, (int)(kBlockMs * nb / std::max(ms, 1e-3))));
        done += nb;
        s.curStep = done;
        const int ne = tlm_read_energy(c, nullptr, nullptr, 0);
        if (ne > 0)
        {
            esteps.resize(ne);
            evals.resize(ne);
            tlm_read_energy(c, esteps.data(), evals.data(), ne);
            s.energy = (float)evals.back();
        }
        std::lock_guard<std::mutex> lk(s.vizMutex);
        if (nf > 0)
            tlm_read_js(c, s.jsInstant.data(), nullptr);
    }

    // ---- final readback into the solver ----
    {
        std::lock_guard<std::mutex> lk(s.vizMutex);
        const int ran = done - start;
        if (nf > 0)
        {
            std::vector<float> d4((size_t)nf * 4);
            if (tlm_read_js(c, s.jsInstant.data(), d4.data()) != TLM_OK)
                return failed(msg, "tlm_read_js", c);
            // the device sums start from 0 at `start`: added to the solver's
            // sums (zero after setup), so a resumed run continues them and
            // dftSamples counts the same steps (TlmSolver.cpp:467-482 divides
            // one by the other)
            for (int f = 0; f < nf; ++f)
            {
                s.jsDftRe[f * 2] += d4[f * 4];
                s.jsDftIm[f * 2] += d4[f * 4 + 1];
                s.jsDftRe[f * 2 + 1] += d4[f * 4 + 2];
                s.jsDftIm[f * 2 + 1] += d4[f * 4 + 3];
            }
        }
        const int dftSteps = std::max(0, done - std::max(start, s.config.settleSteps));
        s.dftSamples += dftSteps;
        if (nh > 0)
        {
            std::vector<float> h8(s.huyDft.size());
            if (tlm_read_huygens(c, h8.data()) != TLM_OK)
                return failed(msg, "tlm_read_huygens", c);
            for (size_t q = 0; q < h8.size(); ++q)
                s.huyDft[q] += h8[q];
            s.huySamples += dftSteps;
        }
        // records of earlier runs (steps before `start`) are kept
        while (!s.energySteps.empty() && s.energySteps.back() >= start)
        {
            s.energySteps.pop_back();
            s.energyVals.pop_back();
        }
        for (size_t q = 0; q < esteps.size(); ++q)
        {
            s.energySteps.push_back(esteps[q]);
            s.energyVals.push_back((float)evals[q]);
        }
        if (pa >= 0)
        {
            int n1 = 0, n2 = 0, pstep = -1;
            tlm_read_plane(c, nullptr, &n1, &n2, &pstep);
            if (pstep >= 0)
            {
                s.planeN1 = n1;
                s.planeN2 = n2;
                s.planeBuf.assign((size_t)n1 * n2, 0.0f);
                tlm_read_plane(c, s.planeBuf.data(), &n1, &n2, &pstep);
            }
        }
        if (every > 0)
        {
            const int nfr = tlm_read_frames(c, nullptr, nullptr, nullptr, 0);
            if (nfr > 0)
            {
                int n1 = 0, n2 = 0, pstep = -1;
                tlm_read_plane(c, nullptr, &n1, &n2, &pstep);
                std::vector<int32_t> fs(nfr);
                std::vector<float> fj((size_t)nfr * std::max(1, nf));
                std::vector<float> fp(pa >= 0 ? (size_t)nfr * n1 * n2 : 0);
                tlm_read_frames(c, fs.data(), fj.data(), fp.empty() ? nullptr : fp.data(), nfr);
                for (int f = 0; f < nfr; ++f)
                {
                    VizFrame fr;
                    fr.step = fs[f];
                    fr.js.assign(fj.begin() + (size_t)f * nf, fj.begin() + (size_t)(f + 1) * nf);
                    if (pa >= 0)
                    {
                        fr.plane.assign(fp.begin() + (size_t)f * n1 * n2,
                                        fp.begin() + (size_t)(f + 1) * n1 * n2);
                        fr.planeAxis = pa;
                        fr.planeIdx = s.planeIndex;
                        fr.planeN1 = n1;
                        fr.planeN2 = n2;
                    }
                    s.frames.push_back(std::move(fr));
                }
            }
        }
        (void)ran;
    }
    if (np > 0 && done > start)
    {
        std::vector<float> rec((size_t)(done - start) * np * 2);
        if (tlm_read_ports(c, start, done - start, rec.data()) != TLM_OK)
            return failed(msg, "tlm_read_ports", c);
        // records of steps [0, done) (GpuTlm.cpp:864-879 resizes to the
        // steps run and assigns)
        for (int p = 0; p < np; ++p)
        {
            auto &port = s.portList[p];
            port.vRec.resize(done);
            port.iRec.resize(done);
            for (int t = 0; t < done - start; ++t)
            {
                port.vRec[start + t] = rec[((size_t)t * np + p) * 2];
                port.iRec[start + t] = rec[((size_t)t * np + p) * 2 + 1];
            }
        }
    }
    if (npr > 0 && done > start)
    {
        std::vector<float> rec((size_t)(done - start) * npr * 3);
        if (tlm_read_probes(c, start, done - start, rec.data()) != TLM_OK)
            return failed(msg, "tlm_read_probes", c);
        for (int p = 0; p < npr; ++p)
        {
            auto &pr = s.probes[p];
            pr.ex.resize(done);
            pr.ey.resize(done);
            pr.ez.resize(done);
            for (int t = 0; t < done - start; ++t)
            {
                pr.ex[start + t] = rec[((size_t)t * npr + p) * 3];
                pr.ey[start + t] = rec[((size_t)t * npr + p) * 3 + 1];
                pr.ez[start + t] = rec[((size_t)t * npr + p) * 3 + 2];
            }
        }
        if (nfreq > 0)
        {
            const size_t per = (size_t)3 * nfreq;
            std::vector<double> re((size_t)npr * per), im(re.size());
            if (start == 0)
            {
                if (tlm_read_probe_dft(c, re.data(), im.data()) != TLM_OK)
                    return failed(msg, "tlm_read_probe_dft", c);
            }
            else
            {
                // continue the solver's sums over [start, done): re += x cos,
                // im -= x sin per step, as the device kernel adds them
                for (int p = 0; p < npr; ++p)
                {
                    const auto d = probe_dft(s.probes[p], 0);
                    if (d.first && d.first->size() == per && d.second->size() == per)
                    {
                        std::copy(d.first->begin(), d.first->end(), re.begin() + p * per);
                        std::copy(d.second->begin(), d.second->end(), im.begin() + p * per);
                    }
                }
                for (int p = 0; p < npr; ++p)
                    for (int comp = 0; comp < 3; ++comp)
                        for (int q = 0; q < nfreq; ++q)
                        {
                            double ar = re[p * per + comp * nfreq + q], ai = im[p * per + comp * nfreq + q];
                            for (int t = start; t < done; ++t)
                            {
                                const double x = (double)rec[((size_t)(t - start) * npr + p) * 3 + comp];
                                ar += x * ct[(size_t)t * nfreq + q];
                                ai -= x * st[(size_t)t * nfreq + q];
                            }
                            re[p * per + comp * nfreq + q] = ar;
                            im[p * per + comp * nfreq + q] = ai;
                        }
            }
            for (int p = 0; p < npr; ++p)
            {
                const auto d = probe_dft(s.probes[p], 0);
                if (!d.first)
                    continue;
                d.first->assign(re.begin() + p * per, re.begin() + (p + 1) * per);
                d.second->assign(im.begin() + p * per, im.begin() + (p + 1) * per);
            }
        }
    }
    if (tlm_read_state(c, s.V.data(), s.Vs.empty() ? nullptr : s.Vs.data()) != TLM_OK)
        return failed(msg, "tlm_read_state", c);
    msg = tlm_device_name(0);
    return true;
}
