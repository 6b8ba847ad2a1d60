// HIP emitter for the straight-line program: one lane per state, SoA inputs/outputs, every AST value a
// register. Plays the role of cuda.rs:101-134 (compile) for gfx950; hiprtc turns the text into a code object.
// Differences from the CUDA template: a bounds guard on the state index (the reference's generated kernels
// have none, cuda.rs:128-133), LessOrEqual printed as '<=' (cuda.rs:41 prints '<'), and constants printed as
// exact hexadecimal literals.
#include <cstdio>
#include <sstream>

#include "emit.hpp"

namespace combustion {

namespace {
std::string literal(double x) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%a", x);
  return buf;
}
}  // namespace

std::string emit_straight(const ast::Linear& l, const std::string& name, const StraightOptions& o) {
  using ast::LOp;
  std::ostringstream s;
  if (o.reciprocal && !o.f32) {
    s << "__device__ __forceinline__ double rcp64(double b) {\n"
         "  double r = __builtin_amdgcn_rcp(b);\n";
    for (unsigned i = 0; i < o.newton_steps; i++)
      s << "  { const double e = __builtin_fma(-b, r, 1.0);\n"
           "    r = __builtin_fma(r, e, r); }\n";
    s << "  return r;\n"
         "}\n";
  }
  if (o.fast_transcendentals && !o.f32)
    s << "__device__ __forceinline__ double exp_wide(double x) {\n"
         "  const double y = x * 1.4426950408889634;\n"
         "  const double n = __builtin_fmin(__builtin_fmax(__builtin_rint(y), -1100.0), 1100.0);\n"
         "  return __builtin_ldexp((double)__builtin_amdgcn_exp2f((float)(y - n)), (int)n);\n"
         "}\n"
         "__device__ __forceinline__ double log_wide(double x) {\n"
         "  const long long b = __double_as_longlong(x);\n"
         "  const int e = (int)((b >> 52) & 0x7FF) - 1023;\n"
         "  const double m = __longlong_as_double((b & 0x000FFFFFFFFFFFFFLL) | 0x3FF0000000000000LL);\n"
         "  return ((double)e + (double)__builtin_amdgcn_logf((float)m)) * 0.6931471805599453;\n"
         "}\n";
  if (o.exp_poly && !o.f32 && !o.fast_transcendentals)
    s << "__device__ __forceinline__ double exp64(double x) {\n"
         "  x = __builtin_fmin(__builtin_fmax(x, -746.0), 710.0);\n"
         "  const double t = __builtin_fma(x, 0x1.71547652b82fep+0, 0x1.8p+52);\n"
         "  const double n = t - 0x1.8p+52;\n"
         "  double r = __builtin_fma(-n, 0x1.62e42fee00000p-1, x);\n"
         "  r = __builtin_fma(-n, 0x1.a39ef35793c76p-33, r);\n"
         "  double p = 0x1.28a2c0a7209fbp-22;\n"
         "  p = __builtin_fma(p, r, 0x1.72faf024b693bp-19);\n"
         "  p = __builtin_fma(p, r, 0x1.a019a66a75dd4p-16);\n"
         "  p = __builtin_fma(p, r, 0x1.a01978c6baf81p-13);\n"
         "  p = __builtin_fma(p, r, 0x1.6c16c17f43a58p-10);\n"
         "  p = __builtin_fma(p, r, 0x1.1111112dd67c5p-7);\n"
         "  p = __builtin_fma(p, r, 0x1.55555555520afp-5);\n"
         "  p = __builtin_fma(p, r, 0x1.555555554b757p-3);\n"
         "  p = __builtin_fma(p, r, 0x1.0000000000005p-1);\n"
         "  p = __builtin_fma(p, r, 0x1.000000000001ep+0);\n"
         "  p = __builtin_fma(p, r, 1.0);\n"
         "  return __builtin_ldexp(p, (int)(unsigned)__double_as_longlong(t));\n"
         "}\n";
  const char* F = o.f32 ? "float" : "double";
  if (o.f32)
    s << "__device__ __forceinline__ float rcp32(float b) { return __builtin_amdgcn_rcpf(b); }\n"
         "__device__ __forceinline__ float exp32(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }\n"
         "__device__ __forceinline__ float log32(float x) { return __builtin_amdgcn_logf(x) * 0.6931471805599453f; }\n";
  s << "extern \"C\" __global__ void __launch_bounds__(" << o.block_size;
  if (o.min_waves) s << ", " << o.min_waves;
  s << ") ";
  s << name
    << "(const double* const* __restrict__ in, double* const* __restrict__ out, unsigned n, unsigned offset) {\n"
    << "  const unsigned id = offset + blockIdx.x * blockDim.x + threadIdx.x;\n";
  const bool grouped = o.group_table >= 0;
  auto option = [&](const std::vector<int32_t>& v, int32_t r) { return r < (int32_t)v.size() ? v[r] : -1; };
  if (grouped && (o.group_size == 0 || o.block_size % o.group_size != 0 || o.group_stride == 0))
    throw ast::Error("emit: the group size must divide the block size, and groups need values");
  if (o.skip_table >= 0 && (!grouped || o.group_size != o.block_size))
    throw ast::Error("emit: skipping groups needs groups of one block each");
  for (int32_t r = 0; r < l.n_input; r++)
    if (option(o.group_offset, r) >= (int32_t)o.group_stride || option(o.increment_offset, r) >= (int32_t)o.group_stride ||
        ((option(o.group_offset, r) >= 0 || option(o.increment_offset, r) >= 0) && !grouped) ||
        (option(o.increment_offset, r) >= 0 &&
         (option(o.group_offset, r) < 0 || option(o.increment_state, r) < 0 || option(o.increment_state, r) >= (int32_t)o.group_size)))
      throw ast::Error("emit: a group offset past the group's values, without a group, or an increment for no member");
  if (!grouped) s << "  if (id >= n) return;\n";
  // LDS-resident registers: slot of each, -1 for the others.
  std::vector<int32_t> slot(l.n_registers, -1);
  int32_t n_slots = 0;
  for (int32_t r : o.lds_registers) {
    if (r < l.n_input || r >= l.n_registers) throw ast::Error("emit: an LDS register must be an instruction result");
    if (slot[r] < 0) slot[r] = n_slots++;
  }
  if (n_slots) {
    if ((size_t)n_slots * o.block_size * 8 > 160 * 1024) throw ast::Error("emit: LDS-resident registers exceed 160 KB per block");
    s << "  __shared__ double lds_v[" << n_slots * o.block_size << "];\n"
      << "  const unsigned tid = threadIdx.x;\n";
  }
  auto operand = [&](int32_t r) {
    return slot[r] >= 0 ? "lds_v[" + std::to_string(slot[r] * o.block_size) + " + tid]" : "r" + std::to_string(r);
  };
  if (grouped) {
    // the block's groups staged in LDS by the whole block at once (one round trip to memory), then read where first
    // used (the lanes of a group read one address: a broadcast); before the threads past n leave, and no further
    // than the last group (every block holds at least one state; the launch offset is a multiple of the block)
    const unsigned ng = o.group_stride;
    const unsigned per_block = o.block_size / o.group_size;
    s << "  __shared__ double group_values[" << per_block * ng << "];\n"
      << "  const unsigned first_group = (offset + blockIdx.x * blockDim.x) / " << o.group_size << "u;\n";
    if (o.skip_table >= 0)  // the whole block leaves together, before its barrier
      s << "  if (((const int*)in[" << o.skip_table << "])[first_group] != 0) return;\n";
    s << "  const unsigned staged_values = min(" << per_block << "u, (n + " << o.group_size - 1 << "u) / " << o.group_size
      << "u - first_group) * " << ng << "u;\n"
      << "  const double* __restrict__ block_groups = in[" << o.group_table << "] + (size_t)first_group * " << ng << "u;\n"
      << "  for (unsigned k = threadIdx.x; k < staged_values; k += blockDim.x) group_values[k] = block_groups[k];\n"
      << "  __syncthreads();\n"
      << "  if (id >= n) return;\n";
    if (per_block == 1)  // the block's one group: uniform addresses
      s << "  const double* grp = group_values;\n";
    else
      s << "  const double* grp = group_values + (id / " << o.group_size << "u - first_group) * " << ng << "u;\n";
    bool increments = false;
    for (int32_t r = 0; r < l.n_input; r++) increments |= option(o.increment_offset, r) >= 0;
    if (increments) s << "  const unsigned member = id % " << o.group_size << "u;\n";
  }
  const std::string out_index = o.output_stride == 1 ? "[id]" : "[(size_t)id * " + std::to_string(o.output_stride) + "u]";
  for (int i = 0; i < l.n_input; i++)
    if (option(o.group_offset, i) < 0) s << "  const " << F << " r" << i << " = (" << F << ")in[" << i << "][id];\n";
  std::vector<char> loaded(l.n_input, 0);
  auto load_grouped = [&](int32_t r) {  // a grouped input, loaded before its first use
    if (r >= l.n_input || option(o.group_offset, r) < 0 || loaded[r]) return;
    loaded[r] = 1;
    const int32_t inc = option(o.increment_offset, r);
    s << "  const " << F << " r" << r << " = (" << F << ")";
    if (inc >= 0)  // selected rather than adding zero, which would turn -0 into +0
      s << "(member == " << option(o.increment_state, r) << "u ? grp[" << option(o.group_offset, r) << "] + grp[" << inc
        << "] : grp[" << option(o.group_offset, r) << "]);\n";
    else
      s << "grp[" << option(o.group_offset, r) << "];\n";
  };
  std::vector<bool> is_bool(l.n_registers, false);
  // Outputs are stored as soon as their register is computed, ending their live range early.
  std::vector<std::vector<size_t>> stores(l.n_registers);
  for (size_t o = 0; o < l.output.size(); o++)
    if (l.output[o] >= l.n_input) stores[l.output[o]].push_back(o);
  size_t emitted = 0;
  for (const ast::Instruction& i : l.code) {
    if (n_slots && o.fence_every && ++emitted % o.fence_every == 0) s << "  asm volatile(\"\" ::: \"memory\");\n";
    // Operands by arity: a constant has none, Neg/Sqrt/Exp/Ln/Sq one, Select three, the rest two. An LDS-resident
    // operand reads its slot (the thread index is below the block size, which __launch_bounds__ enforces).
    const bool none = i.op == LOp::Const;
    const bool one = i.op == LOp::Neg || i.op == LOp::Sqrt || i.op == LOp::Exp || i.op == LOp::Ln || i.op == LOp::Sq;
    if (!none) load_grouped(i.a);
    if (!none && !one) load_grouped(i.b);
    if (i.op == LOp::Select) load_grouped(i.c);
    const std::string d = "r" + std::to_string(i.dst), a = none ? "" : operand(i.a), b = none || one ? "" : operand(i.b),
                      c = i.op == LOp::Select ? operand(i.c) : "";
    if (i.op == LOp::LessOrEqual) {
      is_bool[i.dst] = true;
      s << "  const bool " << d << " = " << a << " <= " << b << ";\n";
      continue;
    }
    s << "  const " << F << " " << d << " = ";
    switch (i.op) {
      case LOp::Const: s << literal(i.k) << (o.f32 ? "f" : ""); break;
      case LOp::Neg: s << "-" << a; break;
      case LOp::Min: s << (o.f32 ? "fminf(" : "fmin(") << a << ", " << b << ")"; break;
      case LOp::Max: s << (o.f32 ? "fmaxf(" : "fmax(") << a << ", " << b << ")"; break;
      case LOp::Add: s << a << " + " << b; break;
      case LOp::Sub: s << a << " - " << b; break;
      case LOp::Mul: s << a << " * " << b; break;
      case LOp::Div:
        if (o.f32)
          s << a << " * rcp32(" << b << ")";
        else if (o.reciprocal)
          s << a << " * rcp64(" << b << ")";
        else
          s << a << " / " << b;
        break;
      case LOp::Sqrt: s << (o.f32 ? "__builtin_sqrtf(" : "sqrt(") << a << ")"; break;
      case LOp::Exp: s << (o.f32 ? "exp32(" : o.fast_transcendentals ? "exp_wide(" : o.exp_poly ? "exp64(" : "exp(") << a << ")"; break;
      case LOp::Ln: s << (o.f32 ? "log32(" : o.fast_transcendentals ? "log_wide(" : "log(") << a << ")"; break;
      case LOp::Sq: s << a << " * " << a; break;
      case LOp::Select: s << a << " ? " << b << " : " << c; break;
      default: break;
    }
    s << ";\n";
    if (slot[i.dst] >= 0) s << "  " << operand(i.dst) << " = " << d << ";\n";
    for (size_t o : stores[i.dst]) s << "  out[" << o << "]" << out_index << " = (double)" << d << ";\n";
  }
  for (size_t o = 0; o < l.output.size(); o++)
    if (l.output[o] < l.n_input) {
      load_grouped(l.output[o]);
      s << "  out[" << o << "]" << out_index << " = r" << l.output[o] << ";\n";
    }
  s << "}\n";
  return s.str();
}

std::string emit_device_function(const ast::Linear& l, const std::vector<std::string>& scalar_names, const std::string& array,
                                 const std::string& name, DeviceOutput mode) {
  using ast::LOp;
  const int n_scalar = (int)scalar_names.size();
  if (n_scalar > l.n_input) throw ast::Error("device function: more scalar names than inputs");
  if (mode == DeviceOutput::Return && l.output.size() != 1) throw ast::Error("device function: Return needs one output");
  auto reg = [&](int32_t r) -> std::string {
    if (r < 0) return "";  // an operand the instruction does not have
    if (r < n_scalar) return scalar_names[r];
    if (r < l.n_input) return array + "[" + std::to_string(r - n_scalar) + "]";
    return "r" + std::to_string(r);
  };
  auto num = [](double x) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.17g", x);
    std::string s = buf;
    if (s.find_first_of(".eEn") == std::string::npos) s += ".";  // a floating literal (inf/nan never occur)
    return s;
  };
  std::ostringstream s;
  s << "__NEKRK_DEVICE__ " << (mode == DeviceOutput::Return ? "dfloat " : "void ") << name << "(";
  std::vector<std::string> params;
  if (mode == DeviceOutput::Strided) params.push_back("unsigned int id");
  for (const std::string& n : scalar_names) params.push_back("dfloat " + n);
  if (l.n_input > n_scalar) params.push_back("const dfloat " + array + "[]");
  if (mode != DeviceOutput::Return) params.push_back("dfloat out[]");
  if (mode == DeviceOutput::Strided) params.push_back("unsigned int n_states");
  for (size_t i = 0; i < params.size(); i++) s << (i ? ", " : "") << params[i];
  s << ") {\n";
  std::vector<std::vector<size_t>> stores(l.n_registers);
  for (size_t o = 0; o < l.output.size(); o++)
    if (l.output[o] >= l.n_input) stores[l.output[o]].push_back(o);
  auto store = [&](size_t o, const std::string& v) {
    if (mode == DeviceOutput::Return) s << "  return " << v << ";\n";
    else if (mode == DeviceOutput::Out) s << "  out[" << o << "] = " << v << ";\n";
    else s << "  out[" << o << "*n_states+id] = " << v << ";\n";
  };
  for (const ast::Instruction& i : l.code) {
    const std::string d = reg(i.dst), a = reg(i.a), b = reg(i.b), c = reg(i.c);
    if (i.op == LOp::LessOrEqual) {
      s << "  const bool " << d << " = " << a << " <= " << b << ";\n";
      continue;
    }
    s << "  const dfloat " << d << " = ";
    switch (i.op) {
      case LOp::Const: s << num(i.k); break;
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
    if (mode != DeviceOutput::Return)
      for (size_t o : stores[i.dst]) store(o, d);
  }
  for (size_t o = 0; o < l.output.size(); o++)
    if (l.output[o] < l.n_input || mode == DeviceOutput::Return) store(o, reg(l.output[o]));
  s << "}\n";
  return s.str();
}

}  // namespace combustion
