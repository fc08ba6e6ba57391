// Per-program coefficient kernels compiled with hipRTC.
//
// The interpreter (assemble.hip, run_program) keeps its register file in
// per-lane scratch memory, so every instruction waits on a memory round trip.
// Here the program of a bucket becomes one straight-line kernel: each
// instruction is a statement on a fresh value (static single assignment), the
// compiler keeps the values in VGPRs, and the program's arithmetic is the
// interpreter's, in the same order and precision (no contraction into FMAs).
// Geometry and field interpolation are the shared device helpers of
// kernels/coef_device.h, compiled with the library's flags; the compiler may
// contract them differently in the two kernels, so coefficients agree to
// rounding (measured: 1e-16 relative in FP64, 5e-8 in FP32), not bitwise.
// Kernels are cached per (device, source) for the life of the process.
#include "coefjit.h"
#include <hip/hiprtc.h>
#include <cstdio>
#include <algorithm>
#include <map>
#include <mutex>
#include <sstream>
#include <vector>

#include "jit_prelude.inc"

namespace slx {
namespace {

std::mutex g_mu;
std::map<std::string, hipFunction_t> g_cache;  // key: device id + source

}  // namespace

std::string coef_kernel_source(const slx_instr* prog, int n, bool fp64, bool runs, const std::vector<int>& field_nff) {
    std::ostringstream o;
    // fields whose values a thread gathers once per element (runs of Gauss
    // points, at most SLX_JIT_GATHER_NFF form functions)
    std::vector<int> pre;
    if (runs)
        for (int pc = 0; pc < n; pc++)
            if (prog[pc].op == SLX_OP_FIELD) {
                const int f = prog[pc].imm;
                if (f >= 0 && f < (int)field_nff.size() && field_nff[f] > 0 && field_nff[f] <= SLX_JIT_GATHER_NFF &&
                    std::find(pre.begin(), pre.end(), f) == pre.end())
                    pre.push_back(f);
            }
    auto nff_of = [&](int f) { return std::find(pre.begin(), pre.end(), f) == pre.end() ? 0 : field_nff[f]; };
    o << "namespace slx {\ntypedef " << (fp64 ? "double" : "float") << " T;\n";
    // runs: one thread per (element, run of qc Gauss points).  Gauss point
    // fastest layout (qfast): qc = 1 and consecutive threads take consecutive
    // points of one element.  Element fastest layout: consecutive threads take
    // consecutive elements at the same points, each thread loads its element's
    // nodes once and walks its qc points.  Without runs (one point per element
    // or per thread): the (element, Gauss point) of the thread, qc unused.
    o << "extern \"C\" __global__ void __launch_bounds__(256) slx_coef(DevBucket b, DevInputs in, int nq, int qfast, int qc, T* __restrict__ ctab) {\n";
    if (runs)
        o << "    const int nch = qfast ? nq : (nq + qc - 1) / qc;\n"
             "    const int64_t total = (int64_t)b.n_elem * nch;\n"
             "    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;\n"
             "    if (idx >= total) return;\n"
             "    const int e = qfast ? (int)(idx / nq) : (int)(idx % b.n_elem);\n"
             "    const int q0 = qfast ? (int)(idx % nq) : (int)(idx / b.n_elem) * qc;\n"
             "    const int q1 = qfast ? q0 + 1 : min(nq, q0 + qc);\n"
             "    GeoX gx;\n"
             "    geo_load(b, in.xyz, e, gx);\n";
    for (int f : pre)
        o << "    double fv" << f << "[" << field_nff[f] << "];\n    bool fo" << f << "[" << field_nff[f] << "];\n"
          << "    field_gather<" << field_nff[f] << ">(b, in, " << f << ", e, fv" << f << ", fo" << f << ");\n";
    if (runs)
        o << "    const int64_t stride = (int64_t)nq * b.n_elem;\n"
             "    for (int q = q0; q < q1; q++) {\n"
             "    Geo g;\n"
             "    geo_eval(b, gx, q, g);\n";
    else
        o << "    const int64_t total = (int64_t)b.n_elem * nq;\n"
             "    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;\n"
             "    if (idx >= total) return;\n"
             "    const int e = qfast ? (int)(idx / nq) : (int)(idx % b.n_elem);\n"
             "    const int q = qfast ? (int)(idx % nq) : (int)(idx / b.n_elem);\n"
             "    const int64_t stride = (int64_t)nq * b.n_elem;\n"
             "    {\n"
             "    Geo g;\n"
             "    geometry(b, in.xyz, e, q, g);\n";
    o << "    const size_t oi = qfast ? (size_t)e * nq + q : (size_t)q * b.n_elem + e;\n"
         "    const T adet = (T)fabs(g.det);\n"
         "    {\n"
         "#pragma clang fp contract(off)\n";
    std::vector<std::string> cur(SLX_MAX_PROG_REGS, "(T)0");
    for (int pc = 0; pc < n; pc++) {
        const slx_instr& x = prog[pc];
        const std::string a = (x.a >= 0 && x.a < SLX_MAX_PROG_REGS) ? cur[x.a] : "(T)0";
        const std::string bb = (x.b >= 0 && x.b < SLX_MAX_PROG_REGS) ? cur[x.b] : "(T)0";
        const std::string v = "v" + std::to_string(pc);
        std::string rhs;
        switch (x.op) {
            case SLX_OP_CONST: rhs = "(T)in.consts[" + std::to_string(x.imm) + "]"; break;
            case SLX_OP_PARAM:
                rhs = "(T)in.params[(size_t)" + std::to_string(x.imm) + " * in.param_len + b.elem_ids[e]]";
                break;
            case SLX_OP_JAC: rhs = "(T)g.J[" + std::to_string(3 * x.a + x.b) + "]"; break;
            case SLX_OP_INVJAC: rhs = "(T)g.inv[" + std::to_string(3 * x.a + x.b) + "]"; break;
            case SLX_OP_DETJ: rhs = "(T)g.det"; break;
            case SLX_OP_COORD:
                o << "        T " << v << ";\n        {\n            double s = 0.0;\n"
                  << "            for (int k = 0; k < b.n_geo; k++) {\n"
                  << "                const int node = b.geo_nodes[(size_t)e * b.n_geo + k];\n"
                  << "                s += in.xyz[3 * (size_t)node + " << x.a << "] * b.geo_tab[(size_t)k * b.n_gp + q];\n"
                  << "            }\n            " << v << " = (T)s;\n        }\n";
                cur[x.dst] = v;
                continue;
            case SLX_OP_FIELD:
                if (nff_of(x.imm) > 0)
                    rhs = "(T)field_value_pre<" + std::to_string(nff_of(x.imm)) + ">(b, " + std::to_string(x.imm) + ", fv" +
                          std::to_string(x.imm) + ", fo" + std::to_string(x.imm) + ", " + std::to_string(x.a) + ", q, g)";
                else
                    rhs = "(T)field_value(b, in, " + std::to_string(x.imm) + ", " + std::to_string(x.a) + ", e, q, g)";
                break;
            case SLX_OP_ADD: rhs = a + " + " + bb; break;
            case SLX_OP_MUL: rhs = a + " * " + bb; break;
            case SLX_OP_POW: rhs = "pow(" + a + ", " + bb + ")"; break;
            case SLX_OP_OUT:
                o << "        ctab[(size_t)" << x.imm << " * stride + oi] = " << a << " * adet;\n";
                continue;
            default: rhs = "apply_unary<T>(" + std::to_string(x.op) + ", " + a + ")"; break;
        }
        o << "        const T " << v << " = " << rhs << ";\n";
        if (x.dst >= 0 && x.dst < SLX_MAX_PROG_REGS) cur[x.dst] = v;
    }
    o << "    }\n    }\n}\n}  // namespace slx\n";
    return std::string(kJitPrelude) + o.str();
}

bool coef_compile(const slx_instr* prog, int n_instr, bool fp64, bool runs, const std::vector<int>& field_nff, hipFunction_t* fn,
                  std::string* err) {
    *fn = nullptr;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        if (err) *err = "no current device";
        return false;
    }
    const std::string src = coef_kernel_source(prog, n_instr, fp64, runs, field_nff);
    const std::string key = std::to_string(dev) + "\n" + src;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_cache.find(key);
        if (it != g_cache.end()) {
            *fn = it->second;
            return true;
        }
    }
    hiprtcProgram rp;
    if (hiprtcCreateProgram(&rp, src.c_str(), "slx_coef.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        if (err) *err = "hiprtcCreateProgram failed";
        return false;
    }
    // the library's own device flags (packed FP32 instructions disabled)
    const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=fast",
                          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult cr = hiprtcCompileProgram(rp, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    if (cr != HIPRTC_SUCCESS) {
        size_t ls = 0;
        hiprtcGetProgramLogSize(rp, &ls);
        std::string log(ls + 1, '\0');
        if (ls) hiprtcGetProgramLog(rp, &log[0]);
        if (err) *err = std::string("hiprtcCompileProgram: ") + hiprtcGetErrorString(cr) + "\n" + log;
        hiprtcDestroyProgram(&rp);
        return false;
    }
    size_t cs = 0;
    hiprtcGetCodeSize(rp, &cs);
    std::vector<char> code(cs);
    hiprtcGetCode(rp, code.data());
    hiprtcDestroyProgram(&rp);
    hipModule_t mod;
    hipError_t e = hipModuleLoadData(&mod, code.data());
    if (e != hipSuccess) {
        if (err) *err = std::string("hipModuleLoadData: ") + hipGetErrorString(e);
        return false;
    }
    hipFunction_t f;
    e = hipModuleGetFunction(&f, mod, "slx_coef");
    if (e != hipSuccess) {
        if (err) *err = std::string("hipModuleGetFunction: ") + hipGetErrorString(e);
        return false;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) {
        *fn = it->second;  // compiled meanwhile by another thread; keep the first
        return true;
    }
    g_cache[key] = f;
    *fn = f;
    return true;
}

hipError_t coef_launch(hipFunction_t fn, const DevBucket& b, const DevInputs& in, int nq, int qfast, void* ctab, hipStream_t st) {
    const int64_t pairs = (int64_t)b.n_elem * nq;
    if (pairs == 0 || b.n_coefs == 0) return hipSuccess;
    // element-fastest layout: runs of qc points per thread, as long as about
    // 256K threads remain to fill the GPU
    // (a kernel compiled without runs has nq = 1 here, so qc = 1 and the grid
    // is one thread per element)
    int qc = 1;
    if (!qfast) qc = (int)std::max<int64_t>(1, std::min<int64_t>(nq, pairs / (256 * 1024)));
    const int64_t nch = qfast ? nq : (nq + qc - 1) / qc;
    const int64_t total = (int64_t)b.n_elem * nch;
    DevBucket bb = b;
    DevInputs ii = in;
    void* args[] = {&bb, &ii, &nq, &qfast, &qc, &ctab};
    return hipModuleLaunchKernel(fn, (unsigned)((total + 255) / 256), 1, 1, 256, 1, 1, 0, st, args, nullptr);
}

}  // namespace slx
