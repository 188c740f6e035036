//---------------------------------------------------------------------------
// GpuTlm.h - GPU execution of the TLM time loop (HIP, gfx950).
//
// Same entry point as the reference's GpuTlm.h:15. Returns true when the
// configured time loop ran on the device and every monitor was filled into
// the solver; false with the reason in msg otherwise.
//---------------------------------------------------------------------------
#ifndef EMSOLVER_AMD_GPUTLM_H
#define EMSOLVER_AMD_GPUTLM_H

#include <string>

class TlmSolver;

bool RunGpuTlm(TlmSolver &s, std::string &msg);

#endif
