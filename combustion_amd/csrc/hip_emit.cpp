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

std::string emit_straight(const ast::Linear& l, const std::string& name, unsigned block_size) {
  using ast::LOp;
  std::ostringstream s;
  s << "extern \"C\" __global__ void __launch_bounds__(" << block_size << ") " << name
    << "(const double* const* __restrict__ in, double* const* __restrict__ out, unsigned n, unsigned offset) {\n"
    << "  const unsigned id = offset + blockIdx.x * blockDim.x + threadIdx.x;\n"
    << "  if (id >= n) return;\n";
  for (int i = 0; i < l.n_input; i++) s << "  const double r" << i << " = in[" << i << "][id];\n";
  std::vector<bool> is_bool(l.n_registers, false);
  for (const ast::Instruction& i : l.code) {
    const std::string d = "r" + std::to_string(i.dst), a = "r" + std::to_string(i.a), b = "r" + std::to_string(i.b),
                      c = "r" + std::to_string(i.c);
    if (i.op == LOp::LessOrEqual) {
      is_bool[i.dst] = true;
      s << "  const bool " << d << " = " << a << " <= " << b << ";\n";
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
  for (size_t o = 0; o < l.output.size(); o++) s << "  out[" << o << "][id] = r" << l.output[o] << ";\n";
  s << "}\n";
  return s.str();
}

}  // namespace combustion
