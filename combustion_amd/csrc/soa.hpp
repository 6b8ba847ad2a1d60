// Launch plumbing shared by every form: SoA pointer tables in device memory, host-buffer staging, the ordering
// of a module's calls across streams and kernel timing with HIP events on the caller's stream.
#pragma once
#include <map>
#include <vector>

#include "device.hpp"
#include "forms.hpp"

namespace combustion {

class SoAForm : public Form {
 public:
  SoAForm(int n_inputs, int n_outputs) : n_in_(n_inputs), n_out_(n_outputs) {}
  ~SoAForm() override {
    for (auto& t : tables_) (void)hipFree(t.second);
    if (start_) (void)hipEventDestroy(start_);
    if (stop_) (void)hipEventDestroy(stop_);
    if (done_) (void)hipEventDestroy(done_);
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
      forget_table(din.data(), dout.data());  // the staging arrays are freed on return
      for (int i = 0; i < n_out_; i++)
        device::check(hipMemcpy(outputs[i], dout[i], n * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy D2H");
      return;
    }
    run(inputs, n, outputs, stream, kernel_ms);
  }

 protected:
  void launch_only(const double* const* in, double* const* out, size_t n, hipStream_t stream, void* work) override {
    if (n) launch(in, out, n, Launch{stream, true, work});
  }

 protected:
  // in/out: device tables of device pointers.
  virtual void launch(const double* const* in, double* const* out, size_t n, const Launch& how) = 0;

  // Orders this call after the module's previous call when the two come on different streams, so that state the
  // module keeps between its kernels (the surrogate's fallback list) is never shared by two calls in flight. Not
  // while the stream is being captured into a graph: the caller orders captured work.
  void order_after_previous_call(hipStream_t stream) {
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    device::check(hipStreamIsCapturing(stream, &capture), "hipStreamIsCapturing");
    if (capture != hipStreamCaptureStatusNone) return;
    if (done_ && stream != last_stream_) device::check(hipStreamWaitEvent(stream, done_, 0), "hipStreamWaitEvent");
  }
  void mark_call_done(hipStream_t stream) {
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    device::check(hipStreamIsCapturing(stream, &capture), "hipStreamIsCapturing");
    if (capture != hipStreamCaptureStatusNone) return;
    if (!done_) device::check(hipEventCreateWithFlags(&done_, hipEventDisableTiming), "hipEventCreateWithFlags");
    device::check(hipEventRecord(done_, stream), "hipEventRecord");
    last_stream_ = stream;
  }

 private:
  // One device table of the arrays' pointers per distinct set of arrays, written once when the set is first seen and
  // never rewritten while the module lives: a launch still in flight on any stream, or a graph captured from this
  // module, keeps reading the table it was given. Past kMaxTables sets the device is synchronized and the tables
  // are freed (graphs captured before that point must then be captured again).
  static constexpr size_t kMaxTables = 1024;
  int n_in_, n_out_;
  std::map<std::vector<const void*>, void*> tables_;
  hipEvent_t start_ = nullptr, stop_ = nullptr, done_ = nullptr;
  hipStream_t last_stream_ = nullptr;

  void* table_for(const std::vector<const void*>& ptrs) {
    auto it = tables_.find(ptrs);
    if (it != tables_.end()) return it->second;
    if (tables_.size() >= kMaxTables) {
      device::check(hipDeviceSynchronize(), "hipDeviceSynchronize");
      for (auto& t : tables_) device::check(hipFree(t.second), "hipFree table");
      tables_.clear();
    }
    void* table = nullptr;
    device::check(hipMalloc(&table, ptrs.size() * sizeof(void*)), "hipMalloc table");
    const hipError_t e = hipMemcpy(table, ptrs.data(), ptrs.size() * sizeof(void*), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(table);
      device::check(e, "hipMemcpy table");
    }
    tables_.emplace(ptrs, table);
    return table;
  }

  std::vector<const void*> pointer_set(const double* const* inputs, double* const* outputs) const {
    std::vector<const void*> ptrs;
    for (int i = 0; i < n_in_; i++) ptrs.push_back(inputs[i]);
    for (int i = 0; i < n_out_; i++) ptrs.push_back(outputs[i]);
    return ptrs;
  }

  // Frees the table of a pointer set whose launches have completed.
  void forget_table(const double* const* inputs, double* const* outputs) {
    auto it = tables_.find(pointer_set(inputs, outputs));
    if (it == tables_.end()) return;
    device::check(hipFree(it->second), "hipFree table");
    tables_.erase(it);
  }

  void run(const double* const* inputs, size_t n, double* const* outputs, hipStream_t stream, float* kernel_ms) {
    void* table = table_for(pointer_set(inputs, outputs));
    auto in = (const double* const*)table;
    auto out = (double* const*)((const void**)table + n_in_);
    order_after_previous_call(stream);
    if (kernel_ms) {
      if (!start_) {
        device::check(hipEventCreate(&start_), "hipEventCreate");
        device::check(hipEventCreate(&stop_), "hipEventCreate");
      }
      device::check(hipEventRecord(start_, stream), "hipEventRecord");
      launch(in, out, n, Launch{stream, false, nullptr});
      device::check(hipEventRecord(stop_, stream), "hipEventRecord");
      mark_call_done(stream);
      device::check(hipEventSynchronize(stop_), "hipEventSynchronize");
      device::check(hipEventElapsedTime(kernel_ms, start_, stop_), "hipEventElapsedTime");
    } else {
      launch(in, out, n, Launch{stream, false, nullptr});
      mark_call_done(stream);
    }
  }
};

}  // namespace combustion
