// HIP runtime side: hiprtc compilation (with an on-disk code-object cache), module loading and launches.
// Plays the role of device.rs:21-101 (assemble/call) for one GPU.
#pragma once
#include <hip/hip_runtime_api.h>

#include <stdexcept>
#include <string>
#include <vector>

namespace combustion::device {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void check(hipError_t e, const char* what);

// Compiles HIP source for gfx950 with hiprtc; code objects are cached by a hash of source and options in
// the directory named by COMBUSTION_AMD_KERNEL_CACHE, else <package>/_kernels. Needs no GPU.
std::vector<char> compile(const std::string& source, const std::string& name, const std::vector<std::string>& options);
std::string cache_directory();
std::string cache_key(const std::string& source, const std::vector<std::string>& options);

// A loaded code object and one of its kernels. The code object's bytes stay owned by the Kernel for as long as
// the module is loaded: hipModuleLoadData is not documented to copy the image it is given.
struct Kernel {
  hipModule_t module = nullptr;
  hipFunction_t function = nullptr;
  std::vector<char> image;
  Kernel() = default;
  Kernel(const Kernel&) = delete;
  Kernel& operator=(const Kernel&) = delete;
  ~Kernel();
  void load(std::vector<char> code, const std::string& name);
};

}  // namespace combustion::device
