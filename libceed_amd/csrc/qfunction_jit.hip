// Runtime compilation of user and gallery QFunctions with hipRTC for gfx950.
//
// A QFunction source is a C file defining `CEED_QFUNCTION(name)(void *ctx,
// CeedInt Q, const CeedScalar *const *in, CeedScalar *const *out)`.  The
// wrapper compiles it as a device function (CEED_Q_VLA = 1) and calls it once
// per quadrature point from a grid-stride kernel.  Field arrays are
// element-major: field f, component k of point q in element e sits at
// ((e * size_f + k) * Q + q), the layout the backend's basis kernels produce.
// Local `#include "..."` files are inlined; `ceed` and system headers are
// replaced by the JIT prelude.
#include <hip/hiprtc.h>
#include <string.h>

#include <fstream>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "mi355_common.h"

struct CeedMi355QFunction_private {
  hipModule_t   module = nullptr;
  hipFunction_t kernel = nullptr;
  int           num_in = 0, num_out = 0;
};

namespace {

constexpr int kMaxFields = 16;

struct QFFields {
  const double *in[kMaxFields];
  double       *out[kMaxFields];
};

std::string dirname_of(const std::string &path) {
  const size_t s = path.find_last_of('/');
  return s == std::string::npos ? std::string(".") : path.substr(0, s);
}

// Source text with local includes inlined (once each) and other includes dropped.
int load_source(const std::string &path, std::set<std::string> &seen, std::string &out, int depth) {
  MI355_CHECK(depth < 16, CEED_MI355_ERROR_BACKEND, "include depth exceeded at %s", path.c_str());
  if (!seen.insert(path).second) return CEED_MI355_SUCCESS;
  std::ifstream f(path);
  MI355_CHECK(f.good(), CEED_MI355_ERROR_BACKEND, "cannot read QFunction source %s", path.c_str());
  std::string line;
  while (std::getline(f, line)) {
    size_t p = line.find_first_not_of(" \t");
    if (p != std::string::npos && line.compare(p, 8, "#include") == 0) {
      const size_t q1 = line.find('"', p), q2 = q1 == std::string::npos ? q1 : line.find('"', q1 + 1);
      if (q1 != std::string::npos && q2 != std::string::npos) {
        const std::string inc = line.substr(q1 + 1, q2 - q1 - 1);
        if (inc.rfind("ceed", 0) != 0) {  // a local header of the user's project
          int ierr = load_source(dirname_of(path) + "/" + inc, seen, out, depth + 1);
          if (ierr) return ierr;
        }
      }
      continue;  // <...> and ceed headers: provided by the prelude
    }
    out += line;
    out += '\n';
  }
  return CEED_MI355_SUCCESS;
}

const char *kPrelude = R"(
#define CEED_RUNNING_JIT_PASS 1
typedef signed char int8_t;
typedef int int32_t;
typedef long int64_t;
typedef unsigned int uint32_t;
typedef unsigned long uint64_t;
typedef long ptrdiff_t;
typedef unsigned long size_t;
#define CEED_QFUNCTION_ATTR
#define CEED_QFUNCTION_HELPER_ATTR
#define CEED_QFUNCTION(name) static __device__ __forceinline__ int name
#define CEED_QFUNCTION_HELPER static __device__ __forceinline__
#define CEED_Q_VLA 1
#define CeedPragmaSIMD
typedef int32_t CeedInt;
typedef ptrdiff_t CeedSize;
typedef signed char CeedInt8;
typedef double CeedScalar;
#define CEED_EPSILON 1e-16
#define CEED_ERROR_SUCCESS 0
)";

}  // namespace

int mi355_jit_user_source(const char *source_path, std::string &out) {
  std::set<std::string> seen;
  std::string           user;
  int                   ierr = load_source(source_path, seen, user, 0);
  if (ierr) return ierr;
  out = std::string(kPrelude) + "\n#line 1 \"" + source_path + "\"\n" + user + "\n";
  return CEED_MI355_SUCCESS;
}

int mi355_jit_build(const std::string &code, const char *what, const char *kernel_name, hipModule_t *module, hipFunction_t *fn) {
  hiprtcProgram prog;
  MI355_CHECK(hiprtcCreateProgram(&prog, code.c_str(), "ceed-mi355-jit.hip", 0, nullptr, nullptr) == HIPRTC_SUCCESS, CEED_MI355_ERROR_BACKEND,
              "hiprtcCreateProgram failed");
  const char  *opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=fast", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  hiprtcResult rc     = hiprtcCompileProgram(prog, 8, opts);
  if (rc != HIPRTC_SUCCESS) {
    size_t log_size = 0;
    hiprtcGetProgramLogSize(prog, &log_size);
    std::string log(log_size + 1, '\0');
    hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    MI355_CHECK(false, CEED_MI355_ERROR_BACKEND, "hipRTC failed for %s: %s", what, log.c_str());
  }
  size_t code_size = 0;
  hiprtcGetCodeSize(prog, &code_size);
  std::vector<char> binary(code_size);
  hiprtcGetCode(prog, binary.data());
  hiprtcDestroyProgram(&prog);
  *module = nullptr;
  if (hipModuleLoadData(module, binary.data()) != hipSuccess || hipModuleGetFunction(fn, *module, kernel_name) != hipSuccess) {
    if (*module) (void)hipModuleUnload(*module);
    *module = nullptr;
    MI355_CHECK(false, CEED_MI355_ERROR_BACKEND, "loading the compiled kernel %s failed", what);
  }
  return CEED_MI355_SUCCESS;
}

extern "C" int CeedMi355QFunctionCompile(const char *source_path, const char *function_name, int32_t num_in, const int32_t *in_sizes,
                                         int32_t num_out, const int32_t *out_sizes, CeedMi355QFunction *qf_out) {
  *qf_out = nullptr;
  MI355_CHECK(num_in >= 0 && num_in <= kMaxFields && num_out >= 0 && num_out <= kMaxFields, CEED_MI355_ERROR_UNSUPPORTED,
              "QFunction with %d inputs / %d outputs (max %d)", num_in, num_out, kMaxFields);
  std::set<std::string> seen;
  std::string           user;
  int                   ierr = load_source(source_path, seen, user, 0);
  if (ierr) return ierr;

  std::ostringstream src;
  src << kPrelude << "\n#line 1 \"" << source_path << "\"\n" << user << "\n";
  src << "struct CeedMi355QFFields { const double *in[" << kMaxFields << "]; double *out[" << kMaxFields << "]; };\n";
  src << "extern \"C\" __global__ void __launch_bounds__(256) CeedMi355QFKernel(long num_elem, int Q, void *ctx, CeedMi355QFFields f) {\n";
  src << "  const long n = num_elem * Q;\n";
  src << "  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x) {\n";
  src << "    const long e = g / Q; const int q = (int)(g - e * Q);\n";
  for (int i = 0; i < num_in; i++) {
    src << "    CeedScalar in_" << i << "[" << in_sizes[i] << "];\n";
    src << "    for (int k = 0; k < " << in_sizes[i] << "; k++) in_" << i << "[k] = f.in[" << i << "][(e * " << in_sizes[i] << " + k) * Q + q];\n";
  }
  for (int i = 0; i < num_out; i++) src << "    CeedScalar out_" << i << "[" << out_sizes[i] << "];\n";
  src << "    const CeedScalar *in[" << (num_in > 0 ? num_in : 1) << "] = {";
  for (int i = 0; i < num_in; i++) src << (i ? ", " : "") << "in_" << i;
  src << "};\n    CeedScalar *out[" << (num_out > 0 ? num_out : 1) << "] = {";
  for (int i = 0; i < num_out; i++) src << (i ? ", " : "") << "out_" << i;
  src << "};\n    " << function_name << "(ctx, 1, in, out);\n";
  for (int i = 0; i < num_out; i++)
    src << "    for (int k = 0; k < " << out_sizes[i] << "; k++) f.out[" << i << "][(e * " << out_sizes[i] << " + k) * Q + q] = out_" << i << "[k];\n";
  src << "  }\n}\n";
  const std::string code = src.str();

  hiprtcProgram prog;
  MI355_CHECK(hiprtcCreateProgram(&prog, code.c_str(), "ceed-mi355-qfunction.hip", 0, nullptr, nullptr) == HIPRTC_SUCCESS,
              CEED_MI355_ERROR_BACKEND, "hiprtcCreateProgram failed");
  const char  *opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=fast", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  hiprtcResult rc     = hiprtcCompileProgram(prog, 8, opts);
  if (rc != HIPRTC_SUCCESS) {
    size_t log_size = 0;
    hiprtcGetProgramLogSize(prog, &log_size);
    std::string log(log_size + 1, '\0');
    hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    MI355_CHECK(false, CEED_MI355_ERROR_BACKEND, "hipRTC failed for %s:%s: %s", source_path, function_name, log.c_str());
  }
  size_t code_size = 0;
  hiprtcGetCodeSize(prog, &code_size);
  std::vector<char> binary(code_size);
  hiprtcGetCode(prog, binary.data());
  hiprtcDestroyProgram(&prog);

  CeedMi355QFunction qf = new CeedMi355QFunction_private();
  qf->num_in            = num_in;
  qf->num_out           = num_out;
  if (hipModuleLoadData(&qf->module, binary.data()) != hipSuccess ||
      hipModuleGetFunction(&qf->kernel, qf->module, "CeedMi355QFKernel") != hipSuccess) {
    delete qf;
    MI355_CHECK(false, CEED_MI355_ERROR_BACKEND, "loading the compiled QFunction %s failed", function_name);
  }
  *qf_out = qf;
  return CEED_MI355_SUCCESS;
}

extern "C" int CeedMi355QFunctionApply(CeedMi355QFunction qf, int64_t num_elem, int32_t Q, void *ctx_device, const double *const *in,
                                       double *const *out, void *stream) {
  MI355_CHECK(qf, CEED_MI355_ERROR_MAJOR, "null QFunction");
  if (num_elem <= 0 || Q <= 0) return CEED_MI355_SUCCESS;
  QFFields f;
  memset(&f, 0, sizeof(f));
  for (int i = 0; i < qf->num_in; i++) f.in[i] = in[i];
  for (int i = 0; i < qf->num_out; i++) f.out[i] = out[i];
  long      ne = (long)num_elem;
  int       q  = Q;
  void     *args[] = {&ne, &q, &ctx_device, &f};
  const int blocks = mi355_blocks(num_elem * (int64_t)Q, 256, 1 << 16);
  MI355_HIP(hipModuleLaunchKernel(qf->kernel, blocks, 1, 1, 256, 1, 1, 0, (hipStream_t)stream, args, nullptr));
  return CEED_MI355_SUCCESS;
}

extern "C" int CeedMi355QFunctionDestroy(CeedMi355QFunction *qf) {
  if (!qf || !*qf) return CEED_MI355_SUCCESS;
  if ((*qf)->module) (void)hipModuleUnload((*qf)->module);
  delete *qf;
  *qf = nullptr;
  return CEED_MI355_SUCCESS;
}
