// Allocation-free calls of a module (cb_hip_call_static), for callers that cannot allocate device memory: code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc. The memory that cb_hip_call allocates per
// module (the device table of the arrays' pointers, and the surrogate form's fallback counter and state list) comes
// instead from one file-scope device array per GPU, `scratch`, held until the process ends.
//
// Layout of `scratch` (kScratchBytes = 16 KiB + 256 B + 8 MiB):
//   [0, 16 KiB)            the pointer table: the input arrays' pointers, then the outputs' (at most 2048 pointers);
//   [16 KiB, 16 KiB + 256) the surrogate's fallback counter (cleared by a kernel at the start of every call);
//   [16 KiB + 256, end)     the surrogate's list of fallback states (at most 2^21 states per call).
// Every call on a GPU uses that GPU's one array: calls that overlap on one GPU (on different streams, or from several
// host threads) race on it. Calls on one stream are ordered and safe. Each call rewrites the table and clears the
// counter, so nothing from an earlier call is read.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "device.hpp"
#include "forms.hpp"
#include "noalloc.hpp"

namespace combustion {

namespace {

constexpr size_t kTableBytes = 16 * 1024;
constexpr size_t kMaxPointers = kTableBytes / sizeof(void*);
constexpr size_t kWorkBytes = 256 + (size_t(1) << 21) * sizeof(unsigned);
constexpr size_t kScratchBytes = kTableBytes + kWorkBytes;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Up to 32 pointers, passed by value in the kernel's arguments: the table is written by a kernel on the caller's
// stream (ordered with the call's other kernels, and capturable in a graph), not copied from host memory.
constexpr int kChunk = 32;
struct PointerChunk {
  const void* p[kChunk];
};

__global__ void __launch_bounds__(kChunk) write_pointer_table(const void** __restrict__ table, PointerChunk chunk, int count) {
  const int i = threadIdx.x;
  if (i < count) table[i] = chunk.p[i];
}

// The address of `scratch` on each GPU, from hipGetSymbolAddress with that GPU current. Host code uses these
// addresses, never the array's name (which in host code is only a stand-in for the symbol).
constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
void* scratch_address[kMaxDevices] = {};

unsigned char* scratch_of_current_device() {
  int dev = -1;
  device::check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) throw ArgumentError("device index out of range for the static scratch");
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_address[dev]) {
    void* ptr = nullptr;
    device::check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress scratch");
    if (!ptr) throw device::Error("hipGetSymbolAddress returned no address for the static scratch");
    scratch_address[dev] = ptr;
  }
  return (unsigned char*)scratch_address[dev];
}

}  // namespace

void call_static(Form& form, const double* const* inputs, size_t n, double* const* outputs, hipStream_t stream) {
  const int n_in = form.n_inputs(), n_out = form.n_outputs();
  const size_t n_ptr = (size_t)n_in + (size_t)n_out;
  if (n_ptr > kMaxPointers)
    throw Unsupported("cb_hip_call_static: " + std::to_string(n_ptr) + " arrays exceed the static pointer table (at most " +
                      std::to_string(kMaxPointers) + ")");
  const size_t work = form.work_bytes(n);
  if (work > kWorkBytes)
    throw ArgumentError("cb_hip_call_static: " + std::to_string(n) + " states need " + std::to_string(work) +
                        " bytes of work memory, above the static scratch's " + std::to_string(kWorkBytes) +
                        " (a surrogate module takes at most 2^21 states per call)");
  if (n == 0) return;
  unsigned char* base = scratch_of_current_device();
  const void** table = (const void**)base;
  for (size_t i = 0; i < n_ptr; i += kChunk) {
    PointerChunk chunk;
    const int count = (int)(n_ptr - i < (size_t)kChunk ? n_ptr - i : (size_t)kChunk);
    for (int j = 0; j < kChunk; j++) {
      const size_t k = i + (size_t)j;
      chunk.p[j] = j >= count ? nullptr : k < (size_t)n_in ? (const void*)inputs[k] : (const void*)outputs[k - (size_t)n_in];
    }
    write_pointer_table<<<dim3(1), dim3(kChunk), 0, stream>>>(table + i, chunk, count);
    device::check(hipGetLastError(), "launch write_pointer_table");
  }
  form.launch_only((const double* const*)table, (double* const*)(table + n_in), n, stream, work ? base + kTableBytes : nullptr);
}

}  // namespace combustion
