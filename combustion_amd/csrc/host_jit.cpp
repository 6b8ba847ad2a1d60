// The per-state program compiled for the host CPU: the counterpart of the reference's Cranelift path
// (ir/lib.rs:225-263 compile + assemble: one `extern fn(*const T, *mut T)` per mechanism, a state's inputs in and its
// outputs out) and of its CPU loop (device.rs:28-48: per state, gather the inputs, call, scatter the outputs). The
// program is emitted as straight-line C in FP64 (libm exp/log, no contraction, no fast-math), compiled by the host
// compiler at -O2 into a shared object (cached beside the GPU code objects) and dlopen'ed. This is the CPU baseline of
// the benchmark; no GPU entry point ever falls back to it.
#include "host_jit.hpp"

#include <dlfcn.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <thread>

#include "device.hpp"
#include "forms.hpp"

namespace combustion {

namespace {

std::string literal(double x) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%a", x);
  return buf;
}

bool executable(const std::string& path) { return access(path.c_str(), X_OK) == 0; }

}  // namespace

std::string emit_host(const ast::Linear& l, const std::string& name) {
  using ast::LOp;
  std::ostringstream s;
  s << "#include <math.h>\n"
       "void "
    << name << "(const double* restrict in, double* restrict out) {\n";
  for (int i = 0; i < l.n_input; i++) s << "  const double r" << i << " = in[" << i << "];\n";
  for (const ast::Instruction& i : l.code) {
    const std::string d = "r" + std::to_string(i.dst), a = "r" + std::to_string(i.a), b = "r" + std::to_string(i.b),
                      c = "r" + std::to_string(i.c);
    if (i.op == LOp::LessOrEqual) {
      s << "  const int " << d << " = " << a << " <= " << b << ";\n";
      continue;
    }
    s << "  const double " << d << " = ";
    switch (i.op) {
      case LOp::Const: s << literal(i.k); break;
      case LOp::Neg: s << "-" << a; break;
      case LOp::Min: s << "fmin(" << a << ", " << b << ")"; break;
      case LOp::Max: s << "fmax(" << a << ", " << b << ")"; break;
      case LOp::Add: s << a << " + " << b; break;
      case LOp::Sub: s << a << " - " << b; break;
      case LOp::Mul: s << a << " * " << b; break;
      case LOp::Div: s << a << " / " << b; break;
      case LOp::Sqrt: s << "sqrt(" << a << ")"; break;
      case LOp::Exp: s << "exp(" << a << ")"; break;
      case LOp::Ln: s << "log(" << a << ")"; break;
      case LOp::Sq: s << a << " * " << a; break;
      case LOp::Select: s << a << " ? " << b << " : " << c; break;
      default: break;
    }
    s << ";\n";
  }
  for (size_t o = 0; o < l.output.size(); o++) s << "  out[" << o << "] = r" << l.output[o] << ";\n";
  s << "}\n";
  return s.str();
}

std::string host_compiler() {
  if (const char* cc = std::getenv("COMBUSTION_AMD_HOST_CC")) return cc;
  const char* rocm = std::getenv("ROCM_PATH");
  const std::string clang = std::string(rocm ? rocm : "/opt/rocm") + "/llvm/bin/clang";
  if (executable(clang)) return clang;
  return "cc";
}

HostProgram::HostProgram(const ast::Linear& l) : n_in_(l.n_input), n_out_((int)l.output.size()) {
  const std::string name = "combustion_program";
  const std::string source = emit_host(l, name);
  const std::string cc = host_compiler();
  const std::vector<std::string> flags = {"-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-fno-math-errno"};
  std::vector<std::string> key_parts = flags;
  key_parts.push_back(cc);
  const std::string dir = device::cache_directory();
  const std::string so = dir + "/host-" + device::cache_key(source, key_parts) + ".so";
  if (access(so.c_str(), R_OK) != 0) {
    mkdir(dir.c_str(), 0755);
    const std::string tag = std::to_string(::getpid()) + "-" + std::to_string((uintptr_t)this);
    const std::string src = so + "." + tag + ".c", tmp = so + "." + tag + ".tmp";
    {
      std::ofstream f(src);
      if (!f) throw device::Error("cannot write " + src);
      f << source;
    }
    std::string cmd = "'" + cc + "'";
    for (const auto& f : flags) cmd += " " + f;
    cmd += " -o '" + tmp + "' '" + src + "' -lm 2>&1";
    std::string log;
    FILE* p = popen(cmd.c_str(), "r");
    if (!p) throw CompileError("cannot run the host compiler " + cc);
    char buf[512];
    while (fgets(buf, sizeof buf, p)) log += buf;
    const int status = pclose(p);
    std::remove(src.c_str());
    if (status != 0) {
      std::remove(tmp.c_str());
      throw CompileError("host compiler failed (" + cc + "): " + log.substr(0, 2000));
    }
    std::rename(tmp.c_str(), so.c_str());
  }
  handle_ = dlopen(so.c_str(), RTLD_NOW | RTLD_LOCAL);
  if (!handle_) throw CompileError(std::string("dlopen: ") + dlerror());
  fn_ = (Fn)dlsym(handle_, name.c_str());
  if (!fn_) throw CompileError("dlsym " + name);
}

HostProgram::~HostProgram() {
  if (handle_) dlclose(handle_);
}

// device.rs:35-45: per state, the inputs gathered into one AoS buffer, one call, the outputs scattered back; threads
// take contiguous ranges of states.
void HostProgram::call(const double* const* inputs, size_t n, double* const* outputs, int threads) const {
  auto range = [&](size_t begin, size_t end) {
    std::vector<double> in(n_in_), out(n_out_);
    for (size_t s = begin; s < end; s++) {
      for (int i = 0; i < n_in_; i++) in[i] = inputs[i][s];
      fn_(in.data(), out.data());
      for (int o = 0; o < n_out_; o++) outputs[o][s] = out[o];
    }
  };
  threads = std::max(1, std::min<int>(threads, (int)std::max<size_t>(1, n)));
  if (threads == 1) return range(0, n);
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; t++) pool.emplace_back(range, n * t / threads, n * (t + 1) / threads);
  for (auto& t : pool) t.join();
}

}  // namespace combustion
