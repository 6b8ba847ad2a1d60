// Launch plumbing shared by every form: SoA pointer tables in device memory, host-buffer staging and
// kernel timing with HIP events on the caller's stream.
#pragma once
#include <vector>

#include "device.hpp"
#include "forms.hpp"

namespace combustion {

class SoAForm : public Form {
 public:
  SoAForm(int n_inputs, int n_outputs) : n_in_(n_inputs), n_out_(n_outputs) {}
  ~SoAForm() override {
    if (table_) (void)hipFree(table_);
    if (start_) (void)hipEventDestroy(start_);
    if (stop_) (void)hipEventDestroy(stop_);
  }
  int n_inputs() const override { return n_in_; }
  int n_outputs() const override { return n_out_; }

  void call(const double* const* inputs, size_t n, double* const* outputs, bool on_device, hipStream_t stream, float* kernel_ms) override {
    if (n == 0) {
      if (kernel_ms) *kernel_ms = 0;
      return;
    }
    if (!on_device) {
      std::vector<double*> din(n_in_, nullptr), dout(n_out_, nullptr);
      struct Free {
        std::vector<double*>& a;
        std::vector<double*>& b;
        ~Free() {
          for (double* p : a) (void)hipFree(p);
          for (double* p : b) (void)hipFree(p);
        }
      } release{din, dout};
      for (int i = 0; i < n_in_; i++) {
        device::check(hipMalloc(&din[i], n * sizeof(double)), "hipMalloc");
        device::check(hipMemcpy(din[i], inputs[i], n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy H2D");
      }
      for (int i = 0; i < n_out_; i++) device::check(hipMalloc(&dout[i], n * sizeof(double)), "hipMalloc");
      run(din.data(), n, dout.data(), stream, kernel_ms);
      device::check(hipStreamSynchronize(stream), "hipStreamSynchronize");
      for (int i = 0; i < n_out_; i++)
        device::check(hipMemcpy(outputs[i], dout[i], n * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy D2H");
      return;
    }
    run(inputs, n, outputs, stream, kernel_ms);
  }

 protected:
  // in/out: device tables of device pointers.
  virtual void launch(const double* const* in, double* const* out, size_t n, hipStream_t stream) = 0;

 private:
  int n_in_, n_out_;
  void* table_ = nullptr;
  std::vector<const void*> last_;
  hipEvent_t start_ = nullptr, stop_ = nullptr;

  void run(const double* const* inputs, size_t n, double* const* outputs, hipStream_t stream, float* kernel_ms) {
    std::vector<const void*> ptrs;
    for (int i = 0; i < n_in_; i++) ptrs.push_back(inputs[i]);
    for (int i = 0; i < n_out_; i++) ptrs.push_back(outputs[i]);
    if (!table_) device::check(hipMalloc(&table_, ptrs.size() * sizeof(void*)), "hipMalloc table");
    if (ptrs != last_) {
      // The table may still be read by an earlier launch on this stream.
      device::check(hipStreamSynchronize(stream), "hipStreamSynchronize");
      device::check(hipMemcpy(table_, ptrs.data(), ptrs.size() * sizeof(void*), hipMemcpyHostToDevice), "hipMemcpy table");
      last_ = ptrs;
    }
    auto in = (const double* const*)table_;
    auto out = (double* const*)((const void**)table_ + n_in_);
    if (kernel_ms) {
      if (!start_) {
        device::check(hipEventCreate(&start_), "hipEventCreate");
        device::check(hipEventCreate(&stop_), "hipEventCreate");
      }
      device::check(hipEventRecord(start_, stream), "hipEventRecord");
      launch(in, out, n, stream);
      device::check(hipEventRecord(stop_, stream), "hipEventRecord");
      device::check(hipEventSynchronize(stop_), "hipEventSynchronize");
      device::check(hipEventElapsedTime(kernel_ms, start_, stop_), "hipEventElapsedTime");
    } else {
      launch(in, out, n, stream);
    }
  }
};

}  // namespace combustion
