#include "device.hpp"

#include <dlfcn.h>
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <mutex>
#include <sstream>

namespace combustion::device {

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e));
}

static uint64_t fnv1a(const std::string& s, uint64_t h = 1469598103934665603ull) {
  for (unsigned char c : s) {
    h ^= c;
    h *= 1099511628211ull;
  }
  return h;
}

std::string cache_key(const std::string& source, const std::vector<std::string>& options) {
  int version_major = 0, version_minor = 0;
  hiprtcVersion(&version_major, &version_minor);
  uint64_t h = fnv1a(source);
  for (const auto& o : options) h = fnv1a(o + "\x1f", h);
  h = fnv1a(std::to_string(version_major) + "." + std::to_string(version_minor), h);
  char buf[32];
  std::snprintf(buf, sizeof buf, "%016llx", (unsigned long long)h);
  return buf;
}

std::string cache_directory() {
  if (const char* env = std::getenv("COMBUSTION_AMD_KERNEL_CACHE")) return env;
  Dl_info info;
  if (dladdr((void*)&cache_directory, &info) && info.dli_fname) {
    std::string path = info.dli_fname;
    size_t slash = path.rfind('/');
    if (slash != std::string::npos) return path.substr(0, slash) + "/_kernels";
  }
  return "_kernels";
}

static bool read_file(const std::string& path, std::vector<char>& out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  return !out.empty();
}

std::vector<char> compile(const std::string& source, const std::string& name, const std::vector<std::string>& options) {
  const std::string dir = cache_directory();
  const std::string path = dir + "/" + name + "-" + cache_key(source, options) + ".co";
  std::vector<char> code;
  if (read_file(path, code)) return code;

  hiprtcProgram program;
  if (hiprtcCreateProgram(&program, source.c_str(), (name + ".hip").c_str(), 0, nullptr, nullptr) != HIPRTC_SUCCESS)
    throw Error("hiprtcCreateProgram failed");
  std::vector<const char*> opts;
  for (const auto& o : options) opts.push_back(o.c_str());
  hiprtcResult r = hiprtcCompileProgram(program, (int)opts.size(), opts.data());
  size_t log_size = 0;
  hiprtcGetProgramLogSize(program, &log_size);
  std::string log(log_size, '\0');
  if (log_size) hiprtcGetProgramLog(program, log.data());
  if (r != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&program);
    throw Error("hiprtc: " + std::string(hiprtcGetErrorString(r)) + "\n" + log.substr(0, 4000));
  }
  size_t size = 0;
  hiprtcGetCodeSize(program, &size);
  code.resize(size);
  hiprtcGetCode(program, code.data());
  hiprtcDestroyProgram(&program);

  mkdir(dir.c_str(), 0755);
  const std::string tmp = path + ".tmp" + std::to_string(::getpid());
  {
    std::ofstream f(tmp, std::ios::binary);
    if (f) f.write(code.data(), (std::streamsize)code.size());
  }
  std::rename(tmp.c_str(), path.c_str());
  return code;
}

Kernel::~Kernel() {
  if (module) (void)hipModuleUnload(module);
}

void Kernel::load(std::vector<char> code, const std::string& name) {
  if (module) throw Error("Kernel::load: a module is already loaded");
  image = std::move(code);
  check(hipModuleLoadData(&module, image.data()), "hipModuleLoadData");
  check(hipModuleGetFunction(&function, module, name.c_str()), ("hipModuleGetFunction " + name).c_str());
}

}  // namespace combustion::device
