// Fused hex operator plans whose pointwise operator is a user QFunction,
// compiled at run time with hipRTC into the line-layout kernel: E, the
// sum-factorised interpolation and collocated gradient, the QFunction at
// every quadrature point, the transposes and the plan's E^T in one kernel.
// This is what the reference's hip-gen backend does for arbitrary
// QFunctions (backends/hip-gen/ceed-hip-gen-operator-build.cpp generates a
// fused kernel around the QFunction source); here the generated part is only
// the per-point glue that hands the QFunction its field arrays in field
// order, the kernel body is fixed (the line layout of hex_lines.hpp).
//
// Supported shape: scalar (1-component) active fields on one basis and one
// restriction, active inputs with INTERP and/or GRAD, active outputs with
// INTERP and/or GRAD, at most one passive EVAL_NONE input (q_comp values per
// point, the plan's q_data slot), and the QFunction context.  FP64.
#include <math.h>
#include <string.h>

#include <sstream>
#include <string>

#include "hex_plan.hpp"
#include "mi355_common.h"

int mi355_jit_user_source(const char *source_path, std::string &out);
int mi355_jit_build(const std::string &code, const char *what, const char *kernel_name, hipModule_t *module, hipFunction_t *fn);

struct CeedMi355JitKernel {
  hipModule_t   module = nullptr;
  hipFunction_t kernel = nullptr;
  void         *ctx    = nullptr;  // device copy of the QFunction context (owned by the caller)
};

void mi355_jit_kernel_destroy(CeedMi355JitKernel *k) {
  if (!k) return;
  if (k->module) (void)hipModuleUnload(k->module);
  delete k;
}

namespace {

// The kernel body, compiled by hipRTC after the user's QFunction source and
// the generated glue (mi355_point).  Template parameters: P, Q, which
// active input/output evaluations the QFunction uses, QC passive values per
// point.  Slabs [z][y][x] with strides (Q*Q, Q, 1), three per element.
const char *kBody = R"(
template <typename T> using mi355_cptr = const __attribute__((address_space(4))) T *;
template <typename T> __device__ __forceinline__ mi355_cptr<T> mi355_phase(const T *p) { asm volatile("" : "+s"(p)); return (mi355_cptr<T>)p; }
__device__ __forceinline__ long mi355_xcd_chunk(long b, long grid) {
  const long x = b % 8, k = b / 8, q = grid / 8, r = grid % 8;
  return x * q + (x < r ? x : r) + k;
}
enum { kAdd = 0, kOverwrite = 1, kAtomic = 2, kEvec = 3, kSplit = 4 };

template <int P, int Q, bool IN_U, bool IN_G, bool OUT_U, bool OUT_G, int QC>
__device__ __forceinline__ void mi355_fused(long elem_begin, long elem_end, int scatter, const int *__restrict__ offsets,
                                            const double *__restrict__ q_data, const double *__restrict__ mats, const double *__restrict__ u,
                                            double *__restrict__ v, double *__restrict__ evec, const int *__restrict__ tpos, const int *__restrict__ elem_sys,
                                            const int *__restrict__ sys_active, void *ctx, double *lds) {
  constexpr int  NT = Q * Q, SLAB = Q * Q * Q, P3 = P * P * P, Q3 = Q * Q * Q;
  constexpr bool kRowWise = Q >= 7;
  const int epb = blockDim.x / NT;
  const int le = threadIdx.x / NT, t = threadIdx.x - le * NT;
  double   *S0 = lds + (long)le * 3 * SLAB, *S1 = S0 + SLAB, *S2 = S1 + SLAB;
  auto      at = [](int z, int y, int x) { return (z * Q + y) * Q + x; };
  const long e    = elem_begin + mi355_xcd_chunk(blockIdx.x, gridDim.x) * epb + le;
  const bool live = le < epb && e < elem_end && (!elem_sys || sys_active[elem_sys[e]]);
  if (elem_sys && !__syncthreads_or(live)) return;
  const int  cx = t % Q, cy = t / Q, ny = t % P, nz = t / P;
  const bool xline = t < P * P, node = live && cx < P && cy < P;
  int off[P];
  if (node) {
#pragma unroll
    for (int z = 0; z < P; z++) off[z] = offsets[e * P3 + (z * P + cy) * P + cx];
  }
  const double *qd = q_data + (live ? e : 0) * (long)(QC > 0 ? QC : 1) * Q3 + (cy * Q + cx);
  // up to 7 passive values per point: the column's values all loaded at the
  // element start, in flight together while E and the interpolation run
  constexpr bool kEarly = QC > 0 && QC <= 7;
  double         sq[kEarly ? Q : 1][kEarly ? QC : 1];
  if (kEarly) {
#pragma unroll
    for (int qz = 0; qz < Q; qz++)
#pragma unroll
      for (int k = 0; k < QC; k++) sq[qz][k] = __builtin_nontemporal_load(&qd[k * Q3 + qz * Q * Q]);
  }

  // ---- E, then B_x, B_y by lines and B_z on the column
  if (node) {
#pragma unroll
    for (int z = 0; z < P; z++) S0[at(z, cy, cx)] = u[off[z] >> 1];
  }
  __syncthreads();
  if (xline) {
    const auto gB = mi355_phase(mats);
    double     in[P];
#pragma unroll
    for (int x = 0; x < P; x++) in[x] = S0[at(nz, ny, x)];
#pragma unroll
    for (int qx = 0; qx < Q; qx++) {
      const auto m = kRowWise ? mi355_phase(mats + qx * P) : gB + qx * P;
      double     s = 0;
#pragma unroll
      for (int b = 0; b < P; b++) s += m[b] * in[b];
      S1[at(nz, ny, qx)] = s;
    }
  }
  __syncthreads();
  if (t < Q * P) {
    const auto gB = mi355_phase(mats);
    const int  qx = t % Q, z = t / Q;
    double     in[P];
#pragma unroll
    for (int b = 0; b < P; b++) in[b] = S1[at(z, b, qx)];
#pragma unroll
    for (int qy = 0; qy < Q; qy++) {
      const auto m = kRowWise ? mi355_phase(mats + qy * P) : gB + qy * P;
      double     s = 0;
#pragma unroll
      for (int b = 0; b < P; b++) s += m[b] * in[b];
      S0[at(z, qy, qx)] = s;
    }
  }
  __syncthreads();
  double uq[Q];
  {
    const auto gB = mi355_phase(mats);
    double     col[P];
#pragma unroll
    for (int b = 0; b < P; b++) col[b] = S0[at(b, cy, cx)];
#pragma unroll
    for (int qz = 0; qz < Q; qz++) {
      const auto m = kRowWise ? mi355_phase(mats + qz * P) : gB + qz * P;
      double     s = 0;
#pragma unroll
      for (int b = 0; b < P; b++) s += m[b] * col[b];
      uq[qz] = s;
    }
  }
  // ---- collocated gradient: z on the column, x (into S2) and y (into S0) by lines through S1
  double dz[Q];
  if (IN_G) {
    const auto gD = mi355_phase(mats + Q * P);
#pragma unroll
    for (int qz = 0; qz < Q; qz++) {
      const auto m = kRowWise ? mi355_phase(mats + Q * P + qz * Q) : gD + qz * Q;
      double     s = 0;
#pragma unroll
      for (int c = 0; c < Q; c++) s += m[c] * uq[c];
      dz[qz]             = s;
      S1[at(qz, cy, cx)] = uq[qz];
    }
    __syncthreads();
    {
      const auto gD2 = mi355_phase(mats + Q * P);
      const int  a = t % Q, qz = t / Q;
      double     in[Q];
#pragma unroll
      for (int c = 0; c < Q; c++) in[c] = S1[at(qz, a, c)];
#pragma unroll
      for (int qx = 0; qx < Q; qx++) {
        const auto m = kRowWise ? mi355_phase(mats + Q * P + qx * Q) : gD2 + qx * Q;
        double     s = 0;
#pragma unroll
        for (int c = 0; c < Q; c++) s += m[c] * in[c];
        S2[at(qz, a, qx)] = s;
      }
#pragma unroll
      for (int c = 0; c < Q; c++) in[c] = S1[at(qz, c, a)];
#pragma unroll
      for (int qy = 0; qy < Q; qy++) {
        const auto m = kRowWise ? mi355_phase(mats + Q * P + qy * Q) : gD2 + qy * Q;
        double     s = 0;
#pragma unroll
        for (int c = 0; c < Q; c++) s += m[c] * in[c];
        S0[at(qz, qy, a)] = s;
      }
    }
    __syncthreads();
  }
  // ---- the QFunction at every point of the column
  double wu[Q], vz[Q];
#pragma unroll
  for (int qz = 0; qz < Q; qz++) {
    double fu[1] = {uq[qz]}, fg[3] = {0, 0, 0}, fq[QC > 0 ? QC : 1], ou[1] = {0}, og[3] = {0, 0, 0};
    if (IN_G) {
      fg[0] = S2[at(qz, cy, cx)];
      fg[1] = S0[at(qz, cy, cx)];
      fg[2] = dz[qz];
    }
#pragma unroll
    for (int k = 0; k < QC; k++) fq[k] = kEarly ? sq[kEarly ? qz : 0][kEarly ? k : 0] : __builtin_nontemporal_load(&qd[k * Q3 + qz * Q * Q]);
    mi355_point(ctx, fu, fg, fq, ou, og);
    if (!live) ou[0] = og[0] = og[1] = og[2] = 0;
    wu[qz] = ou[0];
    if (OUT_G) {
      S2[at(qz, cy, cx)] = og[0];
      S0[at(qz, cy, cx)] = og[1];
      vz[qz]             = og[2];
    }
  }
  // ---- transposes: D_x^T and D_y^T by lines, then the column sum and B_z^T into S0
  double w[Q];
  if (OUT_G) {
    __syncthreads();
    {  // D_y^T waits in registers until every x-line has read S2, then takes S2
      const auto gD = mi355_phase(mats + Q * P);
      const int  a = t % Q, qz = t / Q;
      double     in[Q], ty[Q];
#pragma unroll
      for (int c = 0; c < Q; c++) in[c] = S2[at(qz, a, c)];
#pragma unroll
      for (int qx = 0; qx < Q; qx++) {
        const auto m = kRowWise ? mi355_phase(mats + Q * P + qx) : gD + qx;
        double     s = 0;
#pragma unroll
        for (int c = 0; c < Q; c++) s += m[c * Q] * in[c];
        S1[at(qz, a, qx)] = s;
      }
#pragma unroll
      for (int c = 0; c < Q; c++) in[c] = S0[at(qz, c, a)];
#pragma unroll
      for (int qy = 0; qy < Q; qy++) {
        const auto m = kRowWise ? mi355_phase(mats + Q * P + qy) : gD + qy;
        double     s = 0;
#pragma unroll
        for (int c = 0; c < Q; c++) s += m[c * Q] * in[c];
        ty[qy] = s;
      }
      __syncthreads();
#pragma unroll
      for (int qy = 0; qy < Q; qy++) S2[at(qz, qy, a)] = ty[qy];
    }
    __syncthreads();
    const auto gDz = mi355_phase(mats + Q * P);
#pragma unroll
    for (int qz = 0; qz < Q; qz++) {
      const auto m = kRowWise ? mi355_phase(mats + Q * P + qz) : gDz + qz;
      double     s = S1[at(qz, cy, cx)] + S2[at(qz, cy, cx)];
#pragma unroll
      for (int c = 0; c < Q; c++) s += m[c * Q] * vz[c];
      w[qz] = OUT_U ? s + wu[qz] : s;
    }
  } else {
#pragma unroll
    for (int qz = 0; qz < Q; qz++) w[qz] = wu[qz];
  }
  {
    const auto gBz = mi355_phase(mats);
#pragma unroll
    for (int b = 0; b < P; b++) {
      const auto m = kRowWise ? mi355_phase(mats + b) : gBz + b;
      double     s = 0;
#pragma unroll
      for (int qz = 0; qz < Q; qz++) s += m[qz * P] * w[qz];
      S0[at(b, cy, cx)] = s;
    }
  }
  __syncthreads();
  if (t < Q * P) {  // B_y^T into S1
    const auto gB = mi355_phase(mats);
    const int  qx = t % Q, z = t / Q;
    double     in[Q];
#pragma unroll
    for (int a = 0; a < Q; a++) in[a] = S0[at(z, a, qx)];
#pragma unroll
    for (int y = 0; y < P; y++) {
      const auto m = kRowWise ? mi355_phase(mats + y) : gB + y;
      double     s = 0;
#pragma unroll
      for (int a = 0; a < Q; a++) s += m[a * P] * in[a];
      S1[at(z, y, qx)] = s;
    }
  }
  __syncthreads();
  if (xline) {  // B_x^T into S0
    const auto gB = mi355_phase(mats);
    double     in[Q];
#pragma unroll
    for (int a = 0; a < Q; a++) in[a] = S1[at(nz, ny, a)];
#pragma unroll
    for (int x = 0; x < P; x++) {
      const auto m = kRowWise ? mi355_phase(mats + x) : gB + x;
      double     s = 0;
#pragma unroll
      for (int a = 0; a < Q; a++) s += m[a * P] * in[a];
      S0[at(nz, ny, x)] = s;
    }
  }
  __syncthreads();
  if (node) {  // E^T from the node columns
    double s[P];
#pragma unroll
    for (int z = 0; z < P; z++) s[z] = S0[at(z, cy, cx)];
    if (scatter == kEvec) {
#pragma unroll
      for (int z = 0; z < P; z++) evec[tpos[e * (long)P3 + (z * P + cy) * P + cx]] = s[z];
    } else if (scatter == kSplit) {  // shared entries to their E-vector slot, the rest onto the zeroed target
#pragma unroll
      for (int z = 0; z < P; z++) {
        if (off[z] & 1) evec[tpos[e * (long)P3 + (z * P + cy) * P + cx]] = s[z];
        else atomicAdd(&v[off[z] >> 1], s[z]);
      }
    } else if (scatter == kAtomic) {
#pragma unroll
      for (int z = 0; z < P; z++) atomicAdd(&v[off[z] >> 1], s[z]);
    } else {
      const bool overwrite = scatter == kOverwrite;
      double     old[P];
#pragma unroll
      for (int z = 0; z < P; z++) old[z] = (overwrite && (off[z] & 1)) ? 0.0 : v[off[z] >> 1];
#pragma unroll
      for (int z = 0; z < P; z++) v[off[z] >> 1] = (overwrite && (off[z] & 1)) ? s[z] : old[z] + s[z];
    }
  }
}
)";

struct JitArgs {
  long           elem_begin, elem_end;
  int            scatter;
  const int32_t *offsets;
  const double  *q_data, *mats, *u;
  double        *v;
  const int32_t *tpos, *elem_sys, *sys_active;
  void          *ctx;
};

int launch_jit(const CeedMi355Operator_private *op, const void *u, void *v, int overwrite, hipStream_t s) {
  const int    Q = op->Q, nt = Q * Q, epb = op->epb;
  const size_t lds  = sizeof(double) * (size_t)epb * 3 * Q * Q * Q;
  const bool   tmap = op->scatter == 0, split = op->scatter == CEED_MI355_PLAN_SPLIT;
  // the split mode's atomic adds land on a zeroed target: v in Apply, the scratch vector in ApplyAdd
  double *target = split && !overwrite ? (double *)op->d_scratch : (double *)v;
  if (split && !overwrite) {
    const int err = mi355::fill_zero(op, target, s);
    if (err) return err;
  }
  for (size_t c = 0; c + 1 < op->color_start.size(); c++) {
    const int64_t b = op->color_start[c], e = op->color_start[c + 1];
    if (e <= b) continue;
    const int64_t grid    = (e - b + epb - 1) / epb;
    int           scatter = tmap ? 3 : split ? 4 : (op->scatter == CEED_MI355_PLAN_ATOMIC ? 2 : (overwrite ? 1 : 0));
    long          eb = (long)b, ee = (long)e;
    const int32_t *offsets = op->d_offsets, *tpos = op->d_tpos, *elem_sys = op->d_elem_sys, *sys_active = op->d_sys_active;
    const double  *q_data = (const double *)op->d_qdata, *mats = (const double *)op->d_mats, *uu = (const double *)u;
    double        *vv = target, *evec = (double *)op->d_evec;
    void          *ctx    = op->jit->ctx;
    void          *args[] = {&eb, &ee, &scatter, &offsets, &q_data, &mats, &uu, &vv, &evec, &tpos, &elem_sys, &sys_active, &ctx};
    MI355_HIP(hipModuleLaunchKernel(op->jit->kernel, (unsigned)grid, 1, 1, epb * nt, 1, 1, (unsigned)lds, s, args, nullptr));
  }
  if (tmap) return mi355::transpose_gather(op, v, overwrite, s);
  if (split) return mi355::split_finish(op, v, overwrite, s);
  return CEED_MI355_SUCCESS;
}

}  // namespace

extern "C" int CeedMi355OperatorCreateHexJit(const char *source_path, const char *function_name, int32_t num_in, const int32_t *in_kinds,
                                             int32_t num_out, const int32_t *out_kinds, int32_t q_comp, int32_t num_elem, int32_t P_1d,
                                             int32_t Q_1d, const double *interp_1d_host, const double *grad_1d_host, const int32_t *offsets_host,
                                             int64_t l_size, int flags, CeedMi355Operator *op_out) {
  *op_out = nullptr;
  MI355_CHECK(source_path && function_name, CEED_MI355_ERROR_MAJOR, "null QFunction source or name");
  MI355_CHECK(P_1d >= 2 && P_1d <= 9 && Q_1d >= P_1d && Q_1d <= 10, CEED_MI355_ERROR_UNSUPPORTED, "fused run-time kernel for P=%d Q=%d", P_1d, Q_1d);
  MI355_CHECK(num_in >= 1 && num_in <= 3 && num_out >= 1 && num_out <= 2, CEED_MI355_ERROR_UNSUPPORTED, "%d inputs / %d outputs", num_in, num_out);
  bool in_u = false, in_g = false, out_u = false, out_g = false;
  int  passive = 0;
  for (int i = 0; i < num_in; i++) {
    const int k = in_kinds[i];
    MI355_CHECK(k == CEED_MI355_FIELD_INTERP || k == CEED_MI355_FIELD_GRAD || k == CEED_MI355_FIELD_PASSIVE, CEED_MI355_ERROR_UNSUPPORTED,
                "input field kind %d", k);
    MI355_CHECK(!(k == CEED_MI355_FIELD_INTERP && in_u) && !(k == CEED_MI355_FIELD_GRAD && in_g) && !(k == CEED_MI355_FIELD_PASSIVE && passive),
                CEED_MI355_ERROR_UNSUPPORTED, "two input fields of kind %d", k);
    in_u |= k == CEED_MI355_FIELD_INTERP;
    in_g |= k == CEED_MI355_FIELD_GRAD;
    passive += k == CEED_MI355_FIELD_PASSIVE;
  }
  for (int i = 0; i < num_out; i++) {
    const int k = out_kinds[i];
    MI355_CHECK(k == CEED_MI355_FIELD_INTERP || k == CEED_MI355_FIELD_GRAD, CEED_MI355_ERROR_UNSUPPORTED, "output field kind %d", k);
    MI355_CHECK(!(k == CEED_MI355_FIELD_INTERP && out_u) && !(k == CEED_MI355_FIELD_GRAD && out_g), CEED_MI355_ERROR_UNSUPPORTED,
                "two output fields of kind %d", k);
    out_u |= k == CEED_MI355_FIELD_INTERP;
    out_g |= k == CEED_MI355_FIELD_GRAD;
  }
  MI355_CHECK(in_u || in_g, CEED_MI355_ERROR_UNSUPPORTED, "no active input");
  MI355_CHECK(passive ? (q_comp >= 1 && q_comp <= 32) : q_comp == 0, CEED_MI355_ERROR_UNSUPPORTED, "passive field of %d values per point", q_comp);

  std::string src;
  int         ierr = mi355_jit_user_source(source_path, src);
  if (ierr) return ierr;
  std::ostringstream code;
  code << src << kBody;
  // per-point glue: the QFunction's field arrays in its field order
  code << "__device__ __forceinline__ void mi355_point(void *ctx, const double *fu, const double *fg, const double *fq, double *ou, double *og) {\n";
  code << "  (void)fu; (void)fg; (void)fq; (void)ou; (void)og;\n";
  code << "  const CeedScalar *in[" << num_in << "] = {";
  for (int i = 0; i < num_in; i++)
    code << (i ? ", " : "") << (in_kinds[i] == CEED_MI355_FIELD_INTERP ? "fu" : in_kinds[i] == CEED_MI355_FIELD_GRAD ? "fg" : "fq");
  code << "};\n  CeedScalar *out[" << num_out << "] = {";
  for (int i = 0; i < num_out; i++) code << (i ? ", " : "") << (out_kinds[i] == CEED_MI355_FIELD_INTERP ? "ou" : "og");
  code << "};\n  " << function_name << "(ctx, 1, in, out);\n}\n";
  code << "extern \"C\" __global__ void __launch_bounds__(256) CeedMi355FusedJit(long eb, long ee, int scatter, const int *offsets, "
          "const double *q_data, const double *mats, const double *u, double *v, double *evec, const int *tpos, const int *elem_sys, "
          "const int *sys_active, "
          "void *ctx) {\n  extern __shared__ double mi355_lds[];\n";
  code << "  mi355_fused<" << P_1d << ", " << Q_1d << ", " << (in_u ? "true" : "false") << ", " << (in_g ? "true" : "false") << ", "
       << (out_u ? "true" : "false") << ", " << (out_g ? "true" : "false") << ", " << q_comp
       << ">(eb, ee, scatter, offsets, q_data, mats, u, v, evec, tpos, elem_sys, sys_active, ctx, mi355_lds);\n}\n";
  // mi355_point is used before it is defined in the text order: declare it first
  std::string text = code.str();
  const std::string decl =
      "__device__ __forceinline__ void mi355_point(void *ctx, const double *fu, const double *fg, const double *fq, double *ou, double *og);\n";
  text.insert(src.size(), decl);

  CeedMi355JitKernel *k = new CeedMi355JitKernel();
  ierr                  = mi355_jit_build(text, function_name, "CeedMi355FusedJit", &k->module, &k->kernel);
  if (ierr) {
    delete k;
    return ierr;
  }
  CeedMi355Operator op = nullptr;
  ierr = mi355::create_plan(mi355::kQfJit, CEED_MI355_FP64, 1, num_elem, P_1d, Q_1d, interp_1d_host, grad_1d_host, offsets_host, 1, l_size, flags,
                            q_comp, launch_jit, &op);
  if (ierr) {
    mi355_jit_kernel_destroy(k);
    return ierr;
  }
  op->jit  = k;
  *op_out  = op;
  return CEED_MI355_SUCCESS;
}

extern "C" int CeedMi355OperatorSetContext(CeedMi355Operator op, void *ctx_device) {
  MI355_CHECK(op && op->jit, CEED_MI355_ERROR_BACKEND, "the context belongs to plans of run-time compiled QFunctions");
  op->jit->ctx = ctx_device;
  return CEED_MI355_SUCCESS;
}
