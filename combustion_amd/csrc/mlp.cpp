// Multilayer perceptron of the transport surrogate: parameter layout, workspace and the launch sequences of the
// forward pass, the backward pass of a mean-squared-error loss and Adam (kernels/mlp.hip). The network is
// in -> width (SiLU) -> ... -> width (SiLU) -> out, with `depth` hidden layers; feature dimensions are padded to
// multiples of 64 (padding columns of inputs, targets and weights are zero and stay zero).
#include "mlp.hpp"

#include <algorithm>

#include "emit.hpp"
#include "forms.hpp"

namespace combustion {

namespace {
constexpr long kChunk = 1024;  // states per weight-gradient partial sum
long pad64(long x) { return (x + 63) / 64 * 64; }
unsigned blocks(long n, unsigned per) { return (unsigned)((n + per - 1) / per); }
}  // namespace

MlpLayout::MlpLayout(int n_in, int width_, int depth_, int n_out) : width(width_), depth(depth_) {
  if (n_in < 1 || n_out < 1 || depth < 1 || width < 64 || width % 64) throw ArgumentError("mlp: width must be a positive multiple of 64");
  in_pad = pad64(n_in);
  out_pad = pad64(n_out);
  long off = 0;
  for (int l = 0; l <= depth; l++) {
    Layer L;
    L.in = l == 0 ? in_pad : width;
    L.out = l == depth ? out_pad : width;
    L.w = off;
    L.b = off + L.out * L.in;
    off = L.b + L.out;
    layers.push_back(L);
  }
  n_params = off;
}

Mlp::Mlp(int n_in, int width, int depth, int n_out) : n_in_(n_in), n_out_(n_out), width_(width), depth_(depth) {
  const MlpLayout lay(n_in, width, depth, n_out);
  in_pad_ = lay.in_pad;
  out_pad_ = lay.out_pad;
  n_params_ = lay.n_params;
  for (const auto& L : lay.layers) layers_.push_back({L.in, L.out, L.w, L.b});
  image_ = device::compile(mlp_source, "mlp", hip_options(true));  // kept while the module is loaded
  device::check(hipModuleLoadData(&module_, image_.data()), "hipModuleLoadData mlp");
  auto get = [&](hipFunction_t& f, const char* name) { device::check(hipModuleGetFunction(&f, module_, name), name); };
  get(gemm_, "mlp_gemm");
  get(reduce_, "mlp_reduce");
  get(colsum_, "mlp_column_sum");
  get(to_bf16_, "mlp_to_bf16");
  get(mse_, "mlp_mse_grad");
  get(adam_, "mlp_adam");
}

Mlp::~Mlp() {
  if (module_) (void)hipModuleUnload(module_);
}

long Mlp::max_layer_params() const {
  long m = 0;
  for (const Layer& L : layers_) m = std::max(m, L.out * L.in);
  return m;
}

// Workspace: H_0 (BF16 [n][in]), per hidden layer Z_l (FP32 [n][width]) and H_{l+1} (BF16 [n][width]), two BF16
// gradient buffers [n][max(width, out)] and the FP32 weight-gradient partials [chunks][max layer size].
size_t Mlp::workspace_bytes(long n) const {
  const long g = std::max<long>(width_, out_pad_);
  size_t b = (size_t)n * in_pad_ * 2 + (size_t)depth_ * n * width_ * (4 + 2) + 2 * (size_t)n * g * 2;
  b += (size_t)blocks(n, kChunk) * max_layer_params() * 4;
  return b + 256 * 8;
}

Mlp::Views Mlp::views(void* ws, long n) const {
  Views v;
  auto align = [](char* p) { return (char*)(((uintptr_t)p + 255) & ~(uintptr_t)255); };
  char* p = align((char*)ws);
  v.h.push_back((uint16_t*)p);
  p = align(p + (size_t)n * in_pad_ * 2);
  for (int l = 0; l < depth_; l++) {
    v.z.push_back((float*)p);
    p = align(p + (size_t)n * width_ * 4);
    v.h.push_back((uint16_t*)p);
    p = align(p + (size_t)n * width_ * 2);
  }
  const long g = std::max<long>(width_, out_pad_);
  v.g[0] = (uint16_t*)p;
  p = align(p + (size_t)n * g * 2);
  v.g[1] = (uint16_t*)p;
  p = align(p + (size_t)n * g * 2);
  v.partial = (float*)p;
  return v;
}

void Mlp::gemm(int mode, long M, long N, long Kd, const uint16_t* A, long sam, long sak, const uint16_t* B, long sbn, long sbk,
               const float* bias, const float* zprev, float* zout, uint16_t* hout, long ldc, float* partial, unsigned parts,
               hipStream_t s) const {
  long kchunk = kChunk;
  void* args[] = {&mode, &M, &N, &Kd, &A, &sam, &sak, &B, &sbn, &sbk, &bias, &zprev, &zout, &hout, &ldc, &partial, &kchunk};
  device::check(hipModuleLaunchKernel(gemm_, blocks(M, 64), blocks(N, 64), parts, 256, 1, 1, 0, s, args, nullptr), "launch mlp_gemm");
}

void Mlp::forward(const float* params, const uint16_t* params16, const float* x, long n, float* y, void* ws, hipStream_t s) const {
  if (n <= 0) return;
  Views v = views(ws, n);
  long total = n * in_pad_;
  void* a0[] = {&x, &total, &v.h[0]};
  device::check(hipModuleLaunchKernel(to_bf16_, blocks(total, 256), 1, 1, 256, 1, 1, 0, s, a0, nullptr), "launch mlp_to_bf16");
  for (int l = 0; l <= depth_; l++) {
    const Layer& L = layers_[l];
    const bool last = l == depth_;
    gemm(last ? 1 : 0, n, L.out, L.in, v.h[l], L.in, 1, params16 + L.w, L.in, 1, params + L.b, nullptr, last ? y : v.z[l],
         last ? nullptr : v.h[l + 1], L.out, nullptr, 1, s);
  }
}

void Mlp::backward_mse(const uint16_t* params16, const float* y, const float* target, long n, void* ws, float* grad, hipStream_t s) const {
  if (n <= 0) return;
  Views v = views(ws, n);
  long total = n * out_pad_, ld = out_pad_, valid = n_out_;
  void* a0[] = {&y, &target, &n, &ld, &valid, &v.g[0]};
  device::check(hipModuleLaunchKernel(mse_, blocks(total, 256), 1, 1, 256, 1, 1, 0, s, a0, nullptr), "launch mlp_mse_grad");
  int cur = 0;
  const unsigned parts = blocks(n, kChunk);
  for (int l = depth_; l >= 0; l--) {
    const Layer& L = layers_[l];
    // dW_l = dZ_lᵀ H_l (chunks of states, then a fixed-order sum), db_l = Σ_states dZ_l
    gemm(3, L.out, L.in, n, v.g[cur], 1, L.out, v.h[l], 1, L.in, nullptr, nullptr, nullptr, nullptr, 0, v.partial, parts, s);
    long size = L.out * L.in;
    int p = (int)parts;
    float* gw = grad + L.w;
    void* a1[] = {&v.partial, &size, &p, &gw};
    device::check(hipModuleLaunchKernel(reduce_, blocks(size, 256), 1, 1, 256, 1, 1, 0, s, a1, nullptr), "launch mlp_reduce");
    long out = L.out, chunk = kChunk;
    float* gb = grad + L.b;
    void* a2[] = {&v.g[cur], &n, &out, &out, &chunk, &v.partial};
    device::check(hipModuleLaunchKernel(colsum_, blocks(out, 64), 1, parts, 256, 1, 1, 0, s, a2, nullptr), "launch mlp_column_sum");
    void* a3[] = {&v.partial, &out, &p, &gb};
    device::check(hipModuleLaunchKernel(reduce_, blocks(out, 256), 1, 1, 256, 1, 1, 0, s, a3, nullptr), "launch mlp_reduce");
    if (l > 0) {
      // dZ_{l-1} = (dZ_l W_l) · SiLU'(Z_{l-1})
      gemm(2, n, L.in, L.out, v.g[cur], L.out, 1, params16 + L.w, 1, L.in, nullptr, v.z[l - 1], nullptr, v.g[1 - cur], L.in, nullptr, 1, s);
      cur = 1 - cur;
    }
  }
}

void Mlp::adam(float* params, uint16_t* params16, const float* grad, float* m1, float* m2, float lr, float beta1, float beta2, float eps,
               int step, hipStream_t s) const {
  long n = n_params_;
  float c1 = 1.f - std::pow(beta1, (float)step), c2 = 1.f - std::pow(beta2, (float)step);
  void* a[] = {&params, &params16, &grad, &m1, &m2, &n, &lr, &beta1, &beta2, &eps, &c1, &c2};
  device::check(hipModuleLaunchKernel(adam_, blocks(n, 256), 1, 1, 256, 1, 1, 0, s, a, nullptr), "launch mlp_adam");
}

}  // namespace combustion
