// A problem's user_source given as HIP source text, compiled at run time for
// the context's GPU (hipRTC) into a kernel that applies it to every valid
// cell of every box after the flux assembly, as CNS::compute_rhs calls it
// (src/tim/CNS_advance.cpp:397-407).
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <string>
#include <vector>

#include "user_source.h"

namespace amr {

namespace {

// Definitions the user's code sees.  amr_closures_t is the public header's
// struct, repeated verbatim (a CPU test compares the two texts); the views
// read the state and read/write dSdt of the cell's box as Array4 does.
const char* kPrelude = R"(
typedef int int32_t;
typedef long long int64_t;
typedef struct {
  double gamma, mw, Ru, cv, cp, Rspec, Cshock, Cdamp;
  int32_t visc_law; /* 0 constant, 1 Sutherland */
  double visc_ref, Tvisc_ref;
  int32_t cond_law; /* 0 constant, 1 Sutherland, 2 mu*cp/Pr */
  double cond_ref, Tcond_ref, Pr;
} amr_closures_t;
namespace amr {
constexpr int URHO = 0, UMX = 1, UMY = 2, UMZ = 3, UET = 4;
struct BoxLite {
  int lo[3], hi[3];
  double dxi[3];
  double dx[3];
};
struct FabLite {
  double* p;
  int64_t off, js, ks, ns;
  __device__ double& at(int i, int j, int k, int n) const { return p[off + i + j * js + k * ks + n * ns]; }
};
struct StateView {
  FabLite f;
  __device__ double operator()(int i, int j, int k, int n) const { return f.at(i, j, k, n); }
};
struct RhsView {
  FabLite f;
  __device__ double& operator()(int i, int j, int k, int n) const { return f.at(i, j, k, n); }
};
}  // namespace amr
using amr::URHO;
using amr::UMX;
using amr::UMY;
using amr::UMZ;
using amr::UET;
using amr::StateView;
using amr::RhsView;
)";

const char* kKernel = R"(
extern "C" __global__ void __launch_bounds__(256) amr_user_source_kernel(
    const amr::BoxLite* boxes, const amr::FabLite* st, const amr::FabLite* rhs, const double* pparm,
    amr_closures_t cls, const double* box_dt) {
  const int b = blockIdx.y;
  if (box_dt != nullptr && !(box_dt[b] > 0.0)) return;
  const amr::BoxLite bx = boxes[b];
  const int nx = bx.hi[0] - bx.lo[0] + 1, ny = bx.hi[1] - bx.lo[1] + 1, nz = bx.hi[2] - bx.lo[2] + 1;
  const int64_t n = int64_t(nx) * ny * nz;
  const StateView sv{st[b]};
  const RhsView rv{rhs[b]};
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
    const int i = bx.lo[0] + int(t % nx);
    const int64_t r = t / nx;
    const int j = bx.lo[1] + int(r % ny);
    const int k = bx.lo[2] + int(r / ny);
    user_source(i, j, k, sv, rv, pparm, cls, bx.dx);
  }
}
)";

}  // namespace

const char* user_source_prelude() { return kPrelude; }

int compile_user_source(const char* code, int device, hipModule_t* mod, hipFunction_t* fn, std::string* log) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    *log = "hipGetDeviceProperties failed";
    return 1;
  }
  std::string arch = prop.gcnArchName;
  const size_t colon = arch.find(':');
  if (colon != std::string::npos) arch = arch.substr(0, colon);
  const std::string src = std::string(kPrelude) + "\n#line 1 \"user_source\"\n" + code + "\n" + kKernel;
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "user_source.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
    *log = "hiprtcCreateProgram failed";
    return 1;
  }
  const std::string archopt = "--offload-arch=" + arch;
  // the library's own device code is built the same way: every operation
  // rounded (no contraction into FMAs), no packed FP32
  const char* opts[] = {archopt.c_str(), "-O3", "-std=c++17", "-ffp-contract=off",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult r = hiprtcCompileProgram(prog, 8, opts);
  size_t ls = 0;
  hiprtcGetProgramLogSize(prog, &ls);
  std::vector<char> lg(ls + 1, '\0');
  if (ls) hiprtcGetProgramLog(prog, lg.data());
  *log = lg.data();
  if (r != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return 1;
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  std::vector<char> codebuf(cs);
  hiprtcGetCode(prog, codebuf.data());
  hiprtcDestroyProgram(&prog);
  if (hipModuleLoadData(mod, codebuf.data()) != hipSuccess) {
    *log = "hipModuleLoadData failed";
    return 1;
  }
  if (hipModuleGetFunction(fn, *mod, "amr_user_source_kernel") != hipSuccess) {
    (void)hipModuleUnload(*mod);
    *mod = nullptr;
    *log = "the compiled module has no amr_user_source_kernel";
    return 1;
  }
  return 0;
}

}  // namespace amr
