// Multilayer perceptron training kernels' host side (see mlp.cpp).
#pragma once
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "device.hpp"

namespace combustion {

// Flat parameter layout of the network in -> width (SiLU) x depth -> out: per layer W [out][in] then b [out],
// with the input and output widths padded to multiples of 64.
struct MlpLayout {
  struct Layer {
    long in, out, w, b;
  };
  MlpLayout(int n_in, int width, int depth, int n_out);
  long in_pad, out_pad, n_params;
  int width, depth;
  std::vector<Layer> layers;
};

class Mlp {
 public:
  struct Layer {
    long in, out, w, b;  // padded input/output widths, offsets of W [out][in] and b [out] in the flat parameters
  };
  Mlp(int n_in, int width, int depth, int n_out);
  ~Mlp();
  Mlp(const Mlp&) = delete;
  Mlp& operator=(const Mlp&) = delete;

  long n_params() const { return n_params_; }
  long in_pad() const { return in_pad_; }
  long out_pad() const { return out_pad_; }
  const std::vector<Layer>& layers() const { return layers_; }
  size_t workspace_bytes(long n) const;

  // x [n][in_pad] FP32 -> y [n][out_pad] FP32; keeps the activations in ws for backward_mse.
  void forward(const float* params, const uint16_t* params16, const float* x, long n, float* y, void* ws, hipStream_t s) const;
  // Gradient of mean((y - target)²) over the n × n_out valid outputs into grad (flat, FP32), after forward.
  void backward_mse(const uint16_t* params16, const float* y, const float* target, long n, void* ws, float* grad, hipStream_t s) const;
  void adam(float* params, uint16_t* params16, const float* grad, float* m1, float* m2, float lr, float beta1, float beta2, float eps,
            int step, hipStream_t s) const;

 private:
  struct Views {
    std::vector<uint16_t*> h;
    std::vector<float*> z;
    uint16_t* g[2];
    float* partial;
  };
  Views views(void* ws, long n) const;
  long max_layer_params() const;
  void gemm(int mode, long M, long N, long Kd, const uint16_t* A, long sam, long sak, const uint16_t* B, long sbn, long sbk, const float* bias,
            const float* zprev, float* zout, uint16_t* hout, long ldc, float* partial, unsigned parts, hipStream_t s) const;

  int n_in_, n_out_, width_, depth_;
  long in_pad_, out_pad_, n_params_;
  std::vector<Layer> layers_;
  hipModule_t module_ = nullptr;
  std::vector<char> image_;
  hipFunction_t gemm_, reduce_, colsum_, to_bf16_, mse_, adam_;
};

}  // namespace combustion
