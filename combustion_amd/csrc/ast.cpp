#include "ast.hpp"

#include <cmath>
#include <cstring>
#include <algorithm>
#include <functional>

namespace combustion::ast {

const char* op_name(Op op) {
  static const char* names[] = {"const", "value", "neg", "min", "max", "add", "sub", "le", "mul", "div", "sqrt", "exp", "ln", "sq"};
  return names[(int)op];
}

Block::Block(std::vector<std::string> input_names) {
  f_.n_input = (int32_t)input_names.size();
  f_.values = std::move(input_names);
}

Function Block::finish(std::vector<E> outputs) {
  f_.output = std::move(outputs);
  return std::move(f_);
}

E Block::make(Op op, E a, E b, double c) {
  uint64_t bits;
  std::memcpy(&bits, &c, 8);
  uint64_t h = (uint64_t)op * 0x9E3779B97F4A7C15ull;
  h ^= (uint64_t)(uint32_t)a + 0x7F4A7C159E3779B9ull + (h << 6) + (h >> 2);
  h ^= (uint64_t)(uint32_t)b + 0x94D049BB133111EBull + (h << 6) + (h >> 2);
  h ^= bits + 0xBF58476D1CE4E5B9ull + (h << 6) + (h >> 2);
  auto& bucket = intern_[h];
  for (E e : bucket) {
    const Node& n = f_.nodes[e];
    uint64_t nb;
    std::memcpy(&nb, &n.c, 8);
    if (n.op == op && n.a == a && n.b == b && nb == bits) return e;
  }
  f_.nodes.push_back(Node{op, a, b, c});
  E e = (E)f_.nodes.size() - 1;
  bucket.push_back(e);
  return e;
}

// ast/lib.rs:155 f64(): finite, and finite once demoted to f32.
E Block::c(double x) {
  if (!std::isfinite(x)) throw Error("non-finite constant");
  if (!std::isfinite((float)x)) throw Error("constant " + std::to_string(x) + " overflows f32");
  return make(Op::Const, -1, -1, x);
}
E Block::v(int32_t id) { return make(Op::Value, id, -1); }

std::optional<double> Block::f64(E e) const {
  if (f_.nodes[e].op == Op::Const) return f_.nodes[e].c;
  return std::nullopt;
}
bool Block::is_leaf(E e) const { return f_.nodes[e].op == Op::Const || f_.nodes[e].op == Op::Value; }
// Expr::f32() comparisons of ast/lib.rs:198-229 test the constant after demotion to f32.
bool Block::is_f32(E e, float x) const { return f_.nodes[e].op == Op::Const && (float)f_.nodes[e].c == x; }

E Block::neg(E x) { return make(Op::Neg, x, -1); }

E Block::add(E a, E b) {
  if (f64(a) && f64(b)) return c(*f64(a) + *f64(b));
  if (is_f32(a, 0.f)) return b;
  if (is_f32(b, 0.f)) return a;
  return make(Op::Add, a, b);
}

E Block::sub(E a, E b) {
  if (f64(a) && f64(b)) return c(*f64(a) - *f64(b));
  if (is_f32(a, 0.f)) return neg(b);
  if (is_f32(b, 0.f)) return a;
  return make(Op::Sub, a, b);
}

E Block::mul(E a, E b) {
  if (f64(a) && f64(b)) return c(*f64(a) * *f64(b));
  if (is_f32(a, 0.f) || is_f32(b, 0.f)) return c(0.);
  if (is_f32(a, 1.f)) return b;
  if (is_f32(a, -1.f)) return neg(b);
  if (is_f32(b, 1.f)) return a;
  if (is_f32(b, -1.f)) return neg(a);
  return make(Op::Mul, a, b);
}

// ast/lib.rs:225-232: division by a constant becomes multiplication by its reciprocal.
E Block::div(E a, E b) {
  if (f64(b)) return mul(a, c(1. / *f64(b)));
  if (is_f32(a, 0.f)) throw Error("division of constant zero");
  return make(Op::Div, a, b);
}

E Block::min(E a, E b) { return make(Op::Min, a, b); }
E Block::max(E a, E b) { return make(Op::Max, a, b); }
E Block::le(E a, E b) { return make(Op::LessOrEqual, a, b); }
E Block::sqrt(E x) { return make(Op::Sqrt, x, -1); }
E Block::sq(E x) { return make(Op::Sq, x, -1); }
E Block::exp(E x) {
  if (f64(x)) throw Error("exp of a constant");  // ast/lib.rs:375
  return make(Op::Exp, x, -1);
}
E Block::ln(E x) {
  if (f64(x)) return c(std::log(*f64(x)));
  return make(Op::Ln, x, -1);
}

std::optional<E> Block::sigma(const std::vector<E>& terms) {
  std::optional<E> s;
  for (E t : terms) {
    if (is_f32(t, 0.f)) continue;
    s = s ? add(*s, t) : t;
  }
  return s;
}
E Block::sum(const std::vector<E>& terms) {
  auto s = sigma(terms);
  if (!s) throw Error("empty sum");
  return *s;
}

int32_t Block::value(const std::string& name) {
  f_.values.push_back(name);
  return (int32_t)f_.values.size() - 1;
}

int32_t Block::def(E e, const std::string& name) {
  if (is_leaf(e)) throw Error("definition of a leaf: " + name);
  auto it = defined_.find(e);
  if (it != defined_.end()) return it->second;
  int32_t id = value(name);
  Statement s;
  s.id = id;
  s.value = e;
  f_.statements.push_back(std::move(s));
  defined_[e] = id;
  return id;
}

void Block::select(E condition, std::vector<E> t, std::vector<E> f, std::vector<int32_t> results) {
  Statement s;
  s.select = true;
  s.condition = condition;
  s.t = std::move(t);
  s.f = std::move(f);
  s.results = std::move(results);
  f_.statements.push_back(std::move(s));
}

Linear linearize(const Function& f) {
  Linear l;
  l.n_input = f.n_input;
  std::vector<int32_t> value_reg(f.values.size(), -1);
  for (int32_t i = 0; i < f.n_input; i++) value_reg[i] = i;
  int32_t next = f.n_input;
  std::vector<int32_t> node_reg(f.nodes.size(), -1);
  auto emit = [&](LOp op, int32_t a, int32_t b, int32_t c, double k) {
    l.code.push_back(Instruction{op, next, a, b, c, k});
    return next++;
  };
  std::function<int32_t(E)> eval = [&](E e) -> int32_t {
    if (node_reg[e] >= 0) return node_reg[e];
    const Node& n = f.nodes[e];
    int32_t r;
    switch (n.op) {
      case Op::Const: r = emit(LOp::Const, -1, -1, -1, n.c); break;
      case Op::Value:
        if (value_reg[n.a] < 0) throw Error("use of undefined value " + f.values[n.a]);
        return value_reg[n.a];  // not memoized: Select results are assigned after their first reference is built
      case Op::Neg: r = emit(LOp::Neg, eval(n.a), -1, -1, 0); break;
      case Op::Sqrt: r = emit(LOp::Sqrt, eval(n.a), -1, -1, 0); break;
      case Op::Exp: r = emit(LOp::Exp, eval(n.a), -1, -1, 0); break;
      case Op::Ln: r = emit(LOp::Ln, eval(n.a), -1, -1, 0); break;
      case Op::Sq: r = emit(LOp::Sq, eval(n.a), -1, -1, 0); break;
      default: {
        int32_t a = eval(n.a);
        int32_t b = eval(n.b);
        LOp op = n.op == Op::Min ? LOp::Min : n.op == Op::Max ? LOp::Max : n.op == Op::Add ? LOp::Add : n.op == Op::Sub ? LOp::Sub
                 : n.op == Op::LessOrEqual ? LOp::LessOrEqual : n.op == Op::Mul ? LOp::Mul : LOp::Div;
        r = emit(op, a, b, -1, 0);
      }
    }
    node_reg[e] = r;
    return r;
  };
  // Each output is evaluated right after the last statement it depends on, so that the values only it
  // needs (e.g. the viscosity sums) do not stay live through later statements.
  std::vector<int32_t> defined_at(f.values.size(), -1);
  for (size_t i = 0; i < f.statements.size(); i++) {
    const Statement& s = f.statements[i];
    if (!s.select) defined_at[s.id] = (int32_t)i;
    for (int32_t r : s.results) defined_at[r] = (int32_t)i;
  }
  std::vector<int32_t> ready(f.nodes.size(), -2);  // latest statement index a node depends on
  std::function<int32_t(E)> depends = [&](E e) -> int32_t {
    if (ready[e] != -2) return ready[e];
    const Node& n = f.nodes[e];
    int32_t d = -1;
    if (n.op == Op::Value)
      d = defined_at[n.a];
    else if (n.op != Op::Const) {
      d = depends(n.a);
      if (n.b >= 0) d = std::max(d, depends(n.b));
    }
    return ready[e] = d;
  };
  std::vector<std::vector<size_t>> outputs_after(f.statements.size() + 1);
  for (size_t o = 0; o < f.output.size(); o++) outputs_after[depends(f.output[o]) + 1].push_back(o);
  l.output.assign(f.output.size(), -1);
  for (size_t o : outputs_after[0]) l.output[o] = eval(f.output[o]);
  for (size_t i = 0; i < f.statements.size(); i++) {
    const Statement& s = f.statements[i];
    if (!s.select) {
      value_reg[s.id] = eval(s.value);
    } else {
      int32_t cond = eval(s.condition);
      for (size_t i = 0; i < s.results.size(); i++) {
        int32_t t = eval(s.t[i]);
        int32_t fr = eval(s.f[i]);
        value_reg[s.results[i]] = emit(LOp::Select, cond, t, fr, 0);
      }
    }
    for (size_t o : outputs_after[i + 1]) l.output[o] = eval(f.output[o]);
  }
  l.n_registers = next;
  return l;
}

OpCounts count(const Linear& l) {
  OpCounts c;
  for (const Instruction& i : l.code) {
    switch (i.op) {
      case LOp::Const: c.constants++; break;
      case LOp::Neg: c.neg++; break;
      case LOp::Select: c.select++; break;
      case LOp::Div: c.div++; break;
      case LOp::Sqrt: c.sqrt++; break;
      case LOp::Exp: c.exp++; break;
      case LOp::Ln: c.ln++; break;
      default: c.alu++;
    }
  }
  return c;
}

int32_t peak_live(const Linear& l) {
  std::vector<int64_t> last(l.n_registers, -1);
  std::vector<char> constant(l.n_registers, 0);
  const int64_t end = (int64_t)l.code.size();
  for (int64_t t = 0; t < end; t++) {
    const Instruction& i = l.code[t];
    if (i.op == LOp::Const) constant[i.dst] = 1;
    const bool unary = i.op == LOp::Neg || i.op == LOp::Sqrt || i.op == LOp::Exp || i.op == LOp::Ln || i.op == LOp::Sq;
    if (i.op != LOp::Const) last[i.a] = t;
    if (i.op != LOp::Const && !unary) last[i.b] = t;
    if (i.op == LOp::Select) last[i.c] = t;
  }
  for (int32_t r : l.output) last[r] = end;
  // +1 at a definition (inputs at -1), -1 after the last use
  std::vector<int32_t> delta(end + 2, 0);
  for (int32_t r = 0; r < l.n_registers; r++) {
    if (constant[r] || last[r] < 0) continue;
    delta[0] += 1;  // placeholder, corrected below for instruction results
    delta[last[r] + 1] -= 1;
  }
  for (int64_t t = 0; t < end; t++) {
    const Instruction& i = l.code[t];
    if (constant[i.dst] || last[i.dst] < 0) continue;
    delta[0] -= 1;
    delta[t] += 1;
  }
  int32_t live = 0, peak = 0;
  for (int64_t t = 0; t <= end; t++) {
    live += delta[t];
    peak = std::max(peak, live);
  }
  return peak;
}

namespace {
// The registers an instruction reads, by arity: none for a constant, one for Neg/Sqrt/Exp/Ln/Sq, three for Select.
int operands(const Instruction& i, int32_t (&r)[3]) {
  if (i.op == LOp::Const) return 0;
  r[0] = i.a;
  if (i.op == LOp::Neg || i.op == LOp::Sqrt || i.op == LOp::Exp || i.op == LOp::Ln || i.op == LOp::Sq) return 1;
  r[1] = i.b;
  if (i.op != LOp::Select) return 2;
  r[2] = i.c;
  return 3;
}
}  // namespace

Split split_at(const Linear& l, int32_t first_varying) {
  if (first_varying < 0 || first_varying > l.n_input) throw Error("split: the first varying input is out of range");
  const int32_t R = l.n_registers;
  std::vector<char> varies(R, 0), constant(R, 0);
  std::vector<const Instruction*> def(R, nullptr);
  for (int32_t r = first_varying; r < l.n_input; r++) varies[r] = 1;
  int32_t ops[3];
  for (const Instruction& i : l.code) {
    def[i.dst] = &i;
    if (i.op == LOp::Const) constant[i.dst] = 1;
    const int n = operands(i, ops);
    for (int k = 0; k < n; k++) varies[i.dst] |= varies[ops[k]];
  }
  // Shared values the varying part reads (neither inputs nor constants), in program order.
  std::vector<int32_t> boundary_index(R, -1), boundary;
  auto read = [&](int32_t r) {
    if (varies[r] || constant[r] || r < l.n_input || boundary_index[r] >= 0) return;
    boundary_index[r] = (int32_t)boundary.size();
    boundary.push_back(r);
  };
  for (const Instruction& i : l.code) {
    if (!varies[i.dst]) continue;
    const int n = operands(i, ops);
    for (int k = 0; k < n; k++) read(ops[k]);
  }
  for (int32_t r : l.output) read(r);
  std::sort(boundary.begin(), boundary.end());
  for (size_t b = 0; b < boundary.size(); b++) boundary_index[boundary[b]] = (int32_t)b;

  Split s;
  // shared: the instructions the boundary values need, in program order, on the original registers
  std::vector<char> needed(R, 0);
  for (int32_t r : boundary) needed[r] = 1;
  for (auto it = l.code.rbegin(); it != l.code.rend(); ++it) {
    if (!needed[it->dst]) continue;
    const int n = operands(*it, ops);
    for (int k = 0; k < n; k++) needed[ops[k]] = 1;
  }
  s.shared.n_input = first_varying;
  s.shared.n_registers = R;
  for (const Instruction& i : l.code)
    if (needed[i.dst]) s.shared.code.push_back(i);
  s.shared.output = boundary;

  // varying: inputs, then the boundary values, then its own instructions renumbered in program order
  const int32_t n_in = l.n_input + (int32_t)boundary.size();
  std::vector<int32_t> map(R, -1);
  for (int32_t r = 0; r < l.n_input; r++) map[r] = r;
  for (size_t b = 0; b < boundary.size(); b++) map[boundary[b]] = l.n_input + (int32_t)b;
  int32_t next = n_in;
  auto operand = [&](int32_t r) {
    if (map[r] < 0) {  // a constant read for the first time: emitted again here
      if (!constant[r]) throw Error("split: an operand is neither an input, a shared value nor a constant");
      Instruction c = *def[r];
      c.dst = map[r] = next++;
      s.varying.code.push_back(c);
    }
    return map[r];
  };
  for (const Instruction& i : l.code) {
    if (!varies[i.dst]) continue;
    Instruction v = i;
    const int n = operands(i, ops);
    if (n > 0) v.a = operand(i.a);
    if (n > 1) v.b = operand(i.b);
    if (n > 2) v.c = operand(i.c);
    v.dst = map[i.dst] = next++;
    s.varying.code.push_back(v);
  }
  for (int32_t r : l.output) s.varying.output.push_back(operand(r));
  s.varying.n_input = n_in;
  s.varying.n_registers = next;
  return s;
}

}  // namespace combustion::ast
